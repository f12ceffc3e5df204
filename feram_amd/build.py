"""Build the gfx950 shared library libferam_hip.so in-tree with hipcc.

    python -m feram_amd.build            # compile (always recompiles)

Flags: -O3 for gfx950 only; packed FP32 instructions are disabled
(-Xclang -target-feature -Xclang -packed-fp32-ops); host code keeps IEEE
semantics with -ffp-contract=off on the host side.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libferam_hip.so")
SOURCES = [os.path.join(CSRC, "feram_hip.hip"), os.path.join(CSRC, "fhip_batch.hip"),
           os.path.join(CSRC, "fhip_mlp.hip"), os.path.join(CSRC, "fhip_stages.hip")]
# fhip_stages.hip holds the per-edge launchers and is compiled once per group
# of edges (FHIP_STAGE_GROUP), so the edges build in parallel
STAGE_GROUPS = 13
HEADERS = [os.path.join(CSRC, "fhip_fft.hpp"), os.path.join(CSRC, "fhip_kernels.hpp"),
           os.path.join(CSRC, "fhip_host.hpp"), os.path.join(ROOT, "include", "feram_hip.h")]
ARCH = os.environ.get("FHIP_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def flags() -> list:
    return [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Xarch_host", "-ffp-contract=off",
            "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
            "-I", os.path.join(ROOT, "include")]


def _units() -> list:
    """(source, extra flags, object name) of every translation unit."""
    units = []
    for src in SOURCES:
        base = os.path.basename(src)
        if base == "fhip_stages.hip":
            units += [(src, [f"-DFHIP_STAGE_GROUP={g}"], f"{base}.{g}.o") for g in range(STAGE_GROUPS)]
        else:
            units.append((src, [], base + ".o"))
    return units


def compile_commands(objdir: str) -> list:
    """One hipcc -c per translation unit (run in parallel)."""
    return [[hipcc()] + flags() + extra + ["-c", src, "-o", os.path.join(objdir, obj)]
            for src, extra, obj in _units()]


def link_command(objdir: str, out: str = LIB) -> list:
    return [hipcc(), f"--offload-arch={ARCH}", "-fPIC", "-shared", "-o", out] + \
        [os.path.join(objdir, obj) for _, _, obj in _units()]


def _source_stamps() -> dict:
    return {p: os.stat(p).st_mtime_ns for p in SOURCES + HEADERS}


def build(verbose: bool = True) -> str:
    tmp = LIB + ".tmp"
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    t0 = time.time()
    # hipcc reads a source twice, for the device pass and minutes later for the
    # host pass: a source edited in between gives a library whose host launch
    # stubs do not match its device code, so such a build is refused
    stamps = _source_stamps()
    # the longest units first, at most `jobs` compilers at a time
    cmds = sorted(compile_commands(objdir), key=lambda c: "fhip_stages" not in " ".join(c))
    jobs = max(1, int(os.environ.get("MAX_JOBS", os.cpu_count() or 4)))
    errs, running = [], []

    def reap(pr):
        out, err = pr.communicate()
        errs.append(err)
        if pr.returncode != 0:
            for other in running:
                other.kill()
            sys.stderr.write(out + err)
            raise RuntimeError(f"hipcc failed with status {pr.returncode}")

    for cmd in cmds:
        while len(running) >= jobs:
            done = [pr for pr in running if pr.poll() is not None]
            if not done:
                time.sleep(0.2)
                continue
            for pr in done:
                running.remove(pr)
                reap(pr)
        if verbose:
            print("[feram_amd.build]", " ".join(cmd), flush=True)
        running.append(subprocess.Popen(cmd, cwd=ROOT, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True))
    while running:
        reap(running.pop(0))
    cmd = link_command(objdir, tmp)
    if verbose:
        print("[feram_amd.build]", " ".join(cmd), flush=True)
    res = subprocess.run(cmd, cwd=ROOT, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc link failed with status {res.returncode}")
    changed = [os.path.relpath(p, ROOT) for p, t in _source_stamps().items() if stamps[p] != t]
    if changed:
        os.remove(tmp)
        raise RuntimeError(f"sources changed while compiling ({', '.join(changed)}): build again")
    os.replace(tmp, LIB)
    if verbose:
        warn = [l for e in errs for l in e.splitlines() if "warning" in l and "packed-fp32" not in l]
        for l in warn[:20]:
            print(l)
        print(f"[feram_amd.build] built {LIB} in {time.time() - t0:.1f} s", flush=True)
    return LIB

FORTRAN = os.path.join(HERE, "fortran")
FORTRAN_BUILD = os.path.join(FORTRAN, "_build")
FORTRAN_DRIVER = os.path.join(ROOT, "tests", "host", "_build", "fortran_np")


def flang():
    for cand in (os.environ.get("FLANG"), "/opt/rocm/bin/amdflang", shutil.which("amdflang")):
        if cand and os.path.exists(cand):
            return cand
    return None


def build_fortran(verbose: bool = True):
    """The ISO_C_BINDING module and the context module (feram_amd/fortran/
    _build: objects and .mod files for a feram build to link), and the
    Fortran test driver of the device step (tests/host/fortran_np.f90)
    linked against libferam_hip.so.  Returns the driver's path, or None
    without a Fortran compiler."""
    fc = flang()
    if fc is None:
        if verbose:
            print("[feram_amd.build] no Fortran compiler: Fortran module not built", flush=True)
        return None
    os.makedirs(FORTRAN_BUILD, exist_ok=True)
    os.makedirs(os.path.dirname(FORTRAN_DRIVER), exist_ok=True)
    objs = []
    for name in ("feram_hip_binding", "feram_gpu"):
        obj = os.path.join(FORTRAN_BUILD, name + ".o")
        cmd = [fc, "-O2", "-module-dir", FORTRAN_BUILD, "-c", os.path.join(FORTRAN, name + ".f90"), "-o", obj]
        if verbose:
            print("[feram_amd.build]", " ".join(cmd), flush=True)
        subprocess.run(cmd, cwd=ROOT, check=True)
        objs.append(obj)
    # the driver finds the library next to the package wherever the tree is
    rel = os.path.relpath(HERE, os.path.dirname(FORTRAN_DRIVER))
    cmd = [fc, "-O2", "-module-dir", FORTRAN_BUILD, os.path.join(ROOT, "tests", "host", "fortran_np.f90")] + objs + \
          [f"-L{HERE}", "-lferam_hip", f"-Wl,-rpath,$ORIGIN/{rel}", "-o", FORTRAN_DRIVER]
    if verbose:
        print("[feram_amd.build]", " ".join(cmd), flush=True)
    subprocess.run(cmd, cwd=ROOT, check=True)
    return FORTRAN_DRIVER


if __name__ == "__main__":
    build()
    build_fortran()
