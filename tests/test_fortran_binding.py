"""The ISO_C_BINDING module a Fortran driver uses compiles, links against the
library and calls it (no GPU needed: the size check fails before any HIP call)."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

FLANG = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang" if os.path.exists("/opt/rocm/bin/amdflang") else None)

PROG = """
program t
  use feram_hip_binding
  use, intrinsic :: iso_c_binding
  implicit none
  type(fhip_params) :: p
  type(fhip_energy) :: e
  type(c_ptr) :: ctx
  integer(c_int) :: rc
  p%Lx = 10; p%Ly = 8; p%Lz = 8; p%padding_y = 1
  p%method = 0; p%bulk_or_film = 0; p%precision = 0; p%E_field_always_zero = 1
  p%gap_id = 0; p%reserved = 0
  p%dt = 0.002d0; p%a0 = 3.94d0; p%epsilon_inf = 5.24d0; p%mass_dipo = 0.004d0; p%mass_acou = -1d0
  rc = fhip_create(p, ctx)
  print '(i0,1x,i0,1x,i0,1x,a)', rc, c_sizeof(p), c_sizeof(e), fhip_error_message(ctx)
  call fhip_destroy(ctx)
end program t
"""


@pytest.mark.skipif(FLANG is None, reason="no Fortran compiler")
def test_fortran_binding_calls_library(tmp_path):
    from feram_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libferam_hip.so not built")
    src = os.path.join(ROOT, "feram_amd", "fortran", "feram_hip_binding.f90")
    (tmp_path / "prog.f90").write_text(PROG)
    subprocess.run([FLANG, "-c", src], cwd=tmp_path, check=True)
    libdir = os.path.dirname(_lib.LIB_PATH)
    subprocess.run([FLANG, "prog.f90", "feram_hip_binding.o", f"-L{libdir}", "-lferam_hip",
                    f"-Wl,-rpath,{libdir}", "-o", "prog"], cwd=tmp_path, check=True)
    out = subprocess.run([str(tmp_path / "prog")], capture_output=True, text=True, check=True).stdout
    rc, psize, esize, msg = out.split(maxsplit=3)
    import ctypes
    assert int(rc) == -2
    assert int(psize) == ctypes.sizeof(_lib.FhipParams)
    assert int(esize) == ctypes.sizeof(_lib.FhipEnergy)
    assert "supercell edges" in msg
