"""Host-side bounds check of the z pass's index maps (tests/host/zpass_bounds.hip).

The checker is built with hipcc from the kernels' own headers and run on the
CPU: every LDS, side-buffer and HBM index of zpass_body / zpair_body for
every compiled edge, spectral type and z group stays in bounds at the NT
threads per workgroup the kernels are built and launched with.  The same maps
enumerated for a 512-thread launch of the NT = 256 body run past the tile
(rows t >= LZ, so HBM offsets past the spectrum): a launch with more threads
than the body is compiled for is what puts a z kernel out of bounds.
"""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "host", "zpass_bounds.hip")
HEADERS = [os.path.join(ROOT, "feram_amd", "csrc", h) for h in ("fhip_kernels.hpp", "fhip_fft.hpp", "fhip_host.hpp")]


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    hipcc = "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    exe = str(tmp_path_factory.mktemp("zb") / "zpass_bounds")
    subprocess.run([hipcc, "--offload-arch=gfx950", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"),
                    SRC, "-o", exe], check=True, capture_output=True, text=True)
    return exe


def test_zpass_indices_in_bounds_at_nt(checker):
    r = subprocess.run([checker], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "checks passed (NT = 256, 256 threads per workgroup)" in r.stdout
    with open(HEADERS[2]) as f:
        sizes = re.search(r"#define FHIP_SIZES\(X\)(.*?)\n#define", f.read(), re.S).group(1)
    assert r.stdout.count("in bounds") == 2 * len(re.findall(r"X\(\d+\)", sizes))


def test_oversized_launch_leaves_the_tile(checker):
    r = subprocess.run([checker, "512"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 1
    assert "thread 256 iteration 0 maps to row" in r.stderr and "passes Mh" in r.stderr
