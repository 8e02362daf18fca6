"""The double functions of the double transforms (include/rocrand/rocrand_amd_math.h:
log_cr, exp_cr, sincos_cr) compiled for the host from the same source as the device
(tests/dmath_twin.cpp) and checked against quad precision (libquadmath): results within
0.5001 ulp of the exact value, i.e. the correctly rounded double but for arguments within
about 2^-66 (relative) of a rounding boundary (measured 0 to 3 per million), and the special
values the transforms can produce."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = "/opt/rocm/bin/hipcc"
GCC_LIB = "/usr/lib/gcc/x86_64-linux-gnu/11"


@pytest.fixture(scope="module")
def twin(tmp_path_factory):
    if not os.path.exists(HIPCC) or not os.path.exists(os.path.join(GCC_LIB, "include", "quadmath.h")):
        pytest.skip("hipcc or libquadmath not available")
    exe = tmp_path_factory.mktemp("dtwin") / "dmath_twin"
    subprocess.run([HIPCC, "-std=c++17", "-O2", "-ffp-contract=off", f"-I{os.path.join(ROOT, 'include')}",
                    "-isystem", os.path.join(GCC_LIB, "include"), "-x", "hip", "--offload-arch=gfx950",
                    os.path.join(HERE, "dmath_twin.cpp"), "-o", str(exe), f"-L{GCC_LIB}", "-lquadmath"],
                   check=True, capture_output=True)
    return exe


def test_double_functions_nearly_correctly_rounded(twin):
    out = subprocess.run([str(twin), "400000"], check=True, capture_output=True, text=True).stdout
    rows = {ln.split()[0]: [int(v) for v in ln.split()[1:]] for ln in out.strip().splitlines()}
    for name in ("log", "log_wide", "exp", "exp_wide", "sin", "cos"):
        n, bad, worst = rows[name]
        assert n == 400000, name
        assert bad <= 8, (name, bad)          # misrounded only next to a rounding boundary
        assert worst <= 500100, (name, worst)  # within 0.5001 ulp
    assert rows["special"] == [1] * 9
