"""Build librocrand.so (the C-ABI library with its gfx950 kernels) in-tree.

    python -m rocrand_amd.build            # compile every source for gfx950

Objects go to rocrand_amd/_build/, the library to rocrand_amd/lib/librocrand.so.
Every HIP translation unit is compiled with contraction off (the reference host
path rounds every product separately) and without packed-FP32 instructions.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "librocrand.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("ROCRAND_AMD_ARCH", "gfx950")

COMMON = [
    f"--offload-arch={ARCH}", "-std=c++17", "-O3", "-fPIC", "-fvisibility=hidden",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
]


DATA = os.path.join(HERE, "data")


def sources():
    return sorted(f for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp", ".S")))


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    if src.endswith(".S"):
        # read-only data tables (.incbin of rocrand_amd/data/*.bin)
        cmd = ["gcc", "-c", "-fPIC", "-x", "assembler-with-cpp", f"-Wa,-I{DATA}",
               os.path.join(CSRC, src), "-o", obj]
    else:
        cmd = [HIPCC, *COMMON, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{r.stderr[-4000:]}")
    return obj


def build(jobs: int | None = None, verbose: bool = True) -> str:
    if shutil.which(HIPCC) is None and not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(8, len(srcs))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    try:
        build()
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
