"""The C ABI without a GPU: the library loads, exports every symbol the header
declares, and the ctypes structures match the C layout."""
import os
import re
import subprocess

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "feram_hip.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(fhip_[A-Za-z_0-9]+)\s*\(", text)))


def test_header_and_binding_agree():
    from feram_amd import _lib
    assert sorted(_lib.EXPORTS) == declared_functions()


def test_library_exports_every_declared_symbol():
    from feram_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libferam_hip.so not built (run python -m feram_amd.build)")
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\s[TW]\s+(fhip_\w+)", out))
    missing = [f for f in declared_functions() if f not in exported]
    assert not missing, missing
    L = _lib.load()
    for f in declared_functions():
        assert hasattr(L, f)


def test_struct_layout_matches_c(tmp_path):
    import ctypes
    from feram_amd import _lib
    src = tmp_path / "sz.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "feram_hip.h"\n'
                   'int main(void){printf("%zu %zu %zu %zu %zu %zu %zu %zu\\n", sizeof(fhip_params), '
                   'sizeof(fhip_energy), offsetof(fhip_params, dt), offsetof(fhip_params, j), '
                   'offsetof(fhip_params, external_E_field), sizeof(fhip_average), '
                   'offsetof(fhip_average, pp), offsetof(fhip_average, n_add));return 0;}\n')
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True,
                                          check=True).stdout.split()]
    P, E, A = _lib.FhipParams, _lib.FhipEnergy, _lib.FhipAverage
    assert got == [ctypes.sizeof(P), ctypes.sizeof(E), P.dt.offset, P.j.offset,
                   P.external_E_field.offset, ctypes.sizeof(A), A.pp.offset, A.n_add.offset]


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from feram_amd import _lib
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "absent.so"))
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        _lib.load()


def test_feram_parser_roundtrip():
    from feram_amd.params import Param
    text = """
    # comment
    method = 'vs'
    GPa = -5.0
    L = 16 24 32
    mass_amu = 39.0
    acoustic_mass_amu = 46.44
    P_kappa2 = 5.502 [eV/Angstrom^2]
    j = -2.648 3.894 0.898 -0.789 0.562 0.358 0.179
    Z_star = 9.0 9.5 10.0
    external_E_field = 0.0 0.0 0.001
    """
    p = Param.from_feram(text)
    assert (p.Lx, p.Ly, p.Lz) == (16, 24, 32)
    assert p.method == "vs" and p.GPa == [-5.0] * 3
    assert p.acoustic_md and p.Z_star == [9.0, 9.5, 10.0]
    assert not p.E_field_always_zero
    with pytest.raises(ValueError, match="No such parameter"):
        Param.from_feram("bogus = 1")
    with pytest.raises(ValueError, match="acoustic_mass_amu"):
        Param.from_feram("method = 'md'\nacoustic_mass_amu = 46.44")
