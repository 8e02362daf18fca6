"""Shared test setup: repo root on sys.path, the `gpu` marker, and helpers."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through librocrand.so on cuda:0)")


def ulp_distance(a, b):
    """Elementwise distance in units in the last place between same-dtype float arrays.

    Maps each float's bit pattern to a signed integer that is monotone in the
    value (negative floats to minus their magnitude bits), then subtracts.
    """
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    itype, bits = {np.dtype(np.float16): (np.int16, 16), np.dtype(np.float32): (np.int32, 32),
                   np.dtype(np.float64): (np.int64, 64)}[a.dtype]
    lo = -(1 << (bits - 1))
    ia = a.view(itype).astype(np.int64)
    ib = b.view(itype).astype(np.int64)
    if bits < 64:
        fa = np.where(ia >= 0, ia, lo - ia)
        fb = np.where(ib >= 0, ib, lo - ib)
        return np.abs(fa - fb).astype(np.float64)
    lo64 = np.int64(lo)
    with np.errstate(over="ignore"):
        fa = np.where(ia >= 0, ia, lo64 - ia)
        fb = np.where(ib >= 0, ib, lo64 - ib)
        d = fa - fb
    # Differences beyond 2^62 (opposite extremes) saturate instead of wrapping.
    far = (fa > 0) != (fb > 0)
    return np.where(far & (np.abs(fa.astype(np.float64) - fb.astype(np.float64)) > 2.0**62), 2.0**62,
                    np.abs(d).astype(np.float64))


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O
