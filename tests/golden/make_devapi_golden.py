"""Writes tests/golden/devapi_golden.npz: outputs of the reference's device API
(library/include/rocrand/rocrand_kernel.h, host branch, built into oracle/_ref/libref.so
by oracle/build_ref.sh) for every engine and call of tests/devapi/program.h.

Each entry "<engine>/<FN>[/<lambda>]" holds the [threads, calls, components] result slots
(32-bit results and floats as their bit patterns in uint32, 64-bit ones in uint64) of 12
threads making 5
calls with seed 0x0123456789ABCDEF, the (subsequence, offset) pairs of
tests/devapi/cases.thread_params(12), log-normal (mean, stddev) = (0.3, 0.7), and the
discrete table of cases.discrete_table().

    python tests/golden/make_devapi_golden.py      (needs oracle/_ref/libref.so)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "devapi"))
import cases as C  # noqa: E402

THREADS, CALLS = 12, 5
OUT = os.path.join(HERE, "devapi_golden.npz")


def main():
    ref = C.reference()
    if ref is None:
        sys.exit("oracle/_ref/libref.so is not built")
    sub, off = C.thread_params(THREADS)
    data = {"sub": sub, "off": off}
    for engine in C.ENGINES:
        for fn, name in enumerate(C.FNS):
            k = C.components(engine, fn)
            lams = C.lambdas_for(engine) if name.startswith("POISSON") else [None]
            for lam in lams:
                a = lam if lam is not None else C.MEAN
                got = ref.run(engine, fn, sub, off, CALLS, a=a)
                if got is None:
                    continue
                key = f"{engine}/{name}" + (f"/{lam:g}" if lam is not None else "")
                vals = got[:, :, :k]
                data[key] = vals.astype(np.uint32) if vals.max() < 2**32 else vals.copy()
    np.savez_compressed(OUT, **data)
    print(f"wrote {OUT}: {len(data) - 2} cases, {os.path.getsize(OUT)} bytes")


if __name__ == "__main__":
    main()
