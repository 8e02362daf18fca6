"""Oracle checks of request-server outputs at full request sizes.

A fresh request's output is one plain call (rocrand_serving.h).  Where the oracle is fast
(MRG32k3a: the interleaved engines) the whole output is compared; for Philox4x32-10 and
Sobol64, whose oracles run at a few M values/s, windows are compared:

* Philox4x32-10 into 256-byte aligned memory has no head, so value i of the call is the
  one made from word (offset + i * w) on (w = 2 for doubles, else 1): a window [i0, i0+k)
  is a fresh oracle generator at offset i0 * w (i0 a multiple of 4, so the window is
  aligned as well);
* Sobol64 with D dimensions writes m = n / D points per dimension, dimension-major: a
  window of points [j0, j0+k) is a fresh oracle generator at offset (offset + j0) writing
  k points of every dimension.

tests/test_serving_host.py pins both window rules against whole oracle outputs on the CPU.
"""
from __future__ import annotations

import numpy as np

from oracle import oracle as O
from rocrand_amd import rocrand as R

ORACLE_KIND = {"u32": "u32", "f32": "uniform_f32", "f64": "uniform_f64"}
ADDR = 0x10000000  # any 256-byte aligned address: outputs of the server are


def philox_window(seed, base_offset, dtype, i0, k):
    w = 2 if dtype == "f64" else 1
    og = O.OracleGen(R.PHILOX4_32_10, seed=seed, offset=base_offset + i0 * w)
    st, want = og.generate(ORACLE_KIND[dtype], k, ADDR)
    assert st == 0
    return want


def sobol_window(offset, dims, dtype, j0, k):
    og = O.OracleGen(R.SOBOL64, dims=dims, offset=offset + j0)
    st, want = og.generate(ORACLE_KIND[dtype], dims * k, ADDR)
    assert st == 0
    return want.reshape(dims, k)


def window_starts(m, k, quantum, rs):
    """Window starts (multiples of `quantum`) over [0, m): the first, the last and two others."""
    if m <= 3 * k:
        return [0]
    last = (m - k) // quantum * quantum
    mids = sorted({int(rs.integers(0, last // quantum + 1)) * quantum for _ in range(2)})
    return sorted({0, last, *mids})


def check_request(req, got, rs, k=4096):
    """`got`: the request's output on the host; `req`: rocrand_amd.workloads.Request (fresh)."""
    assert got.size == req.n, (got.size, req.n)
    if req.rng_type == R.MRG32K3A:
        st, want = O.OracleGen(R.MRG32K3A, seed=req.seed).generate(ORACLE_KIND[req.dtype], req.n, ADDR)
        assert st == 0
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), f"MRG32k3a request {req}"
    elif req.rng_type == R.PHILOX4_32_10:
        for i0 in window_starts(req.n, k, 4, rs):
            kk = min(k, req.n - i0)
            want = philox_window(req.seed, 0, req.dtype, i0, kk)
            assert np.array_equal(got[i0:i0 + kk].view(np.uint8), want.view(np.uint8)), f"Philox {req} at {i0}"
    else:
        m = req.n // req.dims
        pts = got.reshape(req.dims, m)
        for j0 in window_starts(m, k, 1, rs):
            kk = min(k, m - j0)
            want = sobol_window(req.seed % (1 << 20), req.dims, req.dtype, j0, kk)
            assert np.array_equal(pts[:, j0:j0 + kk], want), f"Sobol64 {req} at {j0}"
