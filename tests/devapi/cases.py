"""Device-API parity cases shared by the fixture generator and the tests: engine and
function ids of tests/devapi/program.h, the per-thread (subsequence, offset) pairs, the
Poisson lambdas and the discrete table, and ctypes bindings of the three drivers
(reference host build, our host branch, our GPU kernels)."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libref.so")
HOST_LIB = os.path.join(HERE, "libdevapi_host.so")
GPU_LIB = os.path.join(HERE, "libdevapi_gpu.so")

ENGINES = ["philox", "mrg32k3a", "mrg31k3p", "xorwow", "lfsr113", "threefry2x32", "threefry2x64",
           "threefry4x32", "threefry4x64", "sobol32", "sobol64", "scrambled_sobol32", "scrambled_sobol64"]
FNS = ["WORD", "WORDV", "UNIFORM", "UNIFORM2", "UNIFORM4", "UNIFORM_D", "UNIFORM_D2", "UNIFORM_D4", "NORMAL",
       "NORMAL2", "NORMAL4", "NORMAL_D", "NORMAL_D2", "NORMAL_D4", "LOGN", "LOGN2", "LOGN4", "LOGN_D", "LOGN_D2",
       "LOGN_D4", "POISSON", "POISSON4", "DISCRETE", "DISCRETE4", "MIXED", "SKIP"]
E = {n: i for i, n in enumerate(ENGINES)}
F = {n: i for i, n in enumerate(FNS)}
# engines whose single normals use the inverse error function and whose Poisson inverts
ONE_WORD = {"lfsr113", "threefry2x32", "threefry2x64", "threefry4x32", "threefry4x64", "sobol32", "sobol64",
            "scrambled_sobol32", "scrambled_sobol64"}
WORD64 = {"threefry2x64", "threefry4x64", "sobol64", "scrambled_sobol64"}
SOBOL = {"sobol32", "sobol64", "scrambled_sobol32", "scrambled_sobol64"}

SEED = 0x0123456789ABCDEF
MEAN, SD = 0.3, 0.7
LAMBDAS = (0.5, 7.0, 63.9, 64.0, 300.0, 500.0, 999.0, 1000.0, 4000.0, 4001.0, 20000.0)
# the reference's inversion does not terminate in reasonable time for 500 < lambda < 1000
REF_ITR_GAP = (500.0, 1000.0)


def lambdas_for(engine):
    if engine in ONE_WORD:
        return [x for x in LAMBDAS if not REF_ITR_GAP[0] < x < REF_ITR_GAP[1]]
    return list(LAMBDAS)


def thread_params(n, seed=7):
    """Per-thread subsequences and offsets: edge values first, then random 64-bit ones."""
    rng = np.random.default_rng(seed)
    sub = rng.integers(0, 2**64 - 1, n, dtype=np.uint64, endpoint=True)
    off = rng.integers(0, 2**64 - 1, n, dtype=np.uint64, endpoint=True)
    edges_sub = [0, 1, 2, 3, 2**32 - 1, 2**32, 2**63, 2**64 - 1]
    edges_off = [0, 1, 3, 4, 5, 2**32 + 3, 2**62 + 1, 2**64 - 1]
    k = min(n, len(edges_sub))
    sub[:k] = np.array(edges_sub[:k], dtype=np.uint64)
    off[:k] = np.array(edges_off[:k], dtype=np.uint64)
    return sub, off


def discrete_table(size=37, seed=11):
    """(probability, alias, cdf, offset): any probability in [0, 1) and alias below size
    (parity concerns the sampler only) and an increasing CDF ending at 1."""
    rng = np.random.default_rng(seed)
    w = rng.random(size) + 0.01
    cdf = np.cumsum(w) / w.sum()
    cdf[-1] = 1.0
    return rng.random(size), rng.integers(0, size, size).astype(np.uint32), cdf, 5


def sobol_vectors(dims=16, seed=13):
    """Random direction vectors: the state arithmetic is the same for any table."""
    rng = np.random.default_rng(seed)
    v32 = rng.integers(0, 2**32, dims * 32, dtype=np.uint64).astype(np.uint32)
    v64 = rng.integers(0, 2**64 - 1, dims * 64, dtype=np.uint64, endpoint=True)
    return v32, v64


def components(engine, fn):
    name = FNS[fn]
    if name == "WORDV":
        return 2 if engine.startswith("threefry2") else 4
    if name.endswith("4"):
        return 4
    if name.endswith("2"):
        return 2
    return 1


def _sobol32_safe(engine, off):
    if engine in ("sobol32", "scrambled_sobol32"):
        # keep the 32-bit point index below 2^32 - 256: at index 2^32 - 1 the
        # reference's host branch reads past the direction vectors
        # (rocrand_sobol32.h rightmost_zero_bit returns 32 for an all-ones index)
        off = off.copy()
        off[(off & np.uint64(0xFFFFFFFF)) >= np.uint64(0xFFFFFF00)] ^= np.uint64(0x100)
    return off


def _bind(lib, name):
    f = getattr(lib, name)
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_ulonglong, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                  ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_uint, ctypes.c_uint, ctypes.c_void_p,
                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint,
                  ctypes.c_void_p]
    return f


class Driver:
    """One of the three drivers; run() returns the [n, calls, 4] uint64 slots or None when
    the state type has no such call."""

    def __init__(self, path, symbol):
        self.lib = ctypes.CDLL(path)
        self.fn = _bind(self.lib, symbol)

    def run(self, engine, fn, sub, off, calls, a=MEAN, b=SD, seed=SEED, table=None, vectors=None):
        n = len(sub)
        sub = np.ascontiguousarray(sub, dtype=np.uint64)
        off = np.ascontiguousarray(off, dtype=np.uint64)
        off = _sobol32_safe(engine, off)
        prob, alias, cdf, doff = table if table is not None else discrete_table()
        prob = np.ascontiguousarray(prob, dtype=np.float64)
        alias = np.ascontiguousarray(alias, dtype=np.uint32)
        cdf = np.ascontiguousarray(cdf, dtype=np.float64)
        v32, v64 = vectors if vectors is not None else sobol_vectors()
        v32 = np.ascontiguousarray(v32, dtype=np.uint32)
        v64 = np.ascontiguousarray(v64, dtype=np.uint64)
        out = np.zeros(n * calls * 4, dtype=np.uint64)
        rc = self.fn(E[engine], fn, seed, sub.ctypes.data, off.ctypes.data, n, calls, a, b, len(prob), doff,
                     prob.ctypes.data, alias.ctypes.data, cdf.ctypes.data, v32.ctypes.data, v64.ctypes.data,
                     len(v32) // 32, out.ctypes.data)
        if rc == -1:
            return None
        assert rc == 0, f"driver returned {rc}"
        return out.reshape(n, calls, 4)


class HandleDriver:
    """The GPU kernels over a device-resident discrete distribution handle (DISCRETE and
    DISCRETE4 only)."""

    def __init__(self):
        self.lib = ctypes.CDLL(GPU_LIB)
        f = self.lib.devapi_gpu_run_handle
        f.restype = ctypes.c_int
        f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_ulonglong, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                      ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint,
                      ctypes.c_void_p]
        self.fn = f

    def run(self, engine, fn, sub, off, calls, handle, seed=SEED, vectors=None):
        n = len(sub)
        sub = np.ascontiguousarray(sub, dtype=np.uint64)
        off = _sobol32_safe(engine, np.ascontiguousarray(off, dtype=np.uint64))
        v32, v64 = vectors if vectors is not None else sobol_vectors()
        v32 = np.ascontiguousarray(v32, dtype=np.uint32)
        v64 = np.ascontiguousarray(v64, dtype=np.uint64)
        out = np.zeros(n * calls * 4, dtype=np.uint64)
        rc = self.fn(E[engine], fn, seed, sub.ctypes.data, off.ctypes.data, n, calls, handle, v32.ctypes.data,
                     v64.ctypes.data, len(v32) // 32, out.ctypes.data)
        if rc == -1:
            return None
        assert rc == 0, f"driver returned {rc}"
        return out.reshape(n, calls, 4)


def reference():
    return Driver(REF_LIB, "ref_devapi_run") if os.path.exists(REF_LIB) else None


def host():
    return Driver(HOST_LIB, "devapi_host_run")


def gpu():
    return Driver(GPU_LIB, "devapi_gpu_run")


# 32-bit words of each state that rocrand_init defines (the Box-Muller cache values and
# padding are left unset by init, as in the reference)
STATE_WORDS = {
    "philox": list(range(13)),            # counter, result, key, substate, cache flags
    "mrg32k3a": list(range(8)),           # g1, g2, cache flags
    "mrg31k3p": list(range(8)),
    "xorwow": [0, 1, 2] + list(range(6, 11)),  # d, cache flags, x[5]
    "lfsr113": list(range(8)),            # z, subsequence seed
    "threefry2x32": list(range(7)),       # counter, key, result, substate
    "threefry2x64": list(range(13)),
    "threefry4x32": list(range(13)),
    "threefry4x64": list(range(25)),
}


def _bind_state(lib, name):
    f = getattr(lib, name)
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int, ctypes.c_ulonglong, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                  ctypes.c_void_p]
    return f


def state_words(path, symbol, engine, seed, sub, off, words=32):
    """[threads, words] uint32 state bytes after rocrand_init, and sizeof(state)."""
    f = _bind_state(ctypes.CDLL(path), symbol)
    sub = np.ascontiguousarray(sub, dtype=np.uint64)
    off = np.ascontiguousarray(off, dtype=np.uint64)
    out = np.zeros(len(sub) * words, dtype=np.uint32)
    size = f(E[engine], seed, sub.ctypes.data, off.ctypes.data, len(sub), words, out.ctypes.data)
    assert size > 0, size
    return out.reshape(len(sub), words), size


# The reference's device-API state checks (test_rocrand_kernel_philox4x32_10.cpp:227-271,
# test_rocrand_kernel_xorwow.cpp:211-250).
def philox_kat(words):
    n = words.shape[0]
    assert np.all(words[:, 8] == 0xBEEFDEAD) and np.all(words[:, 9] == 0xDEADBEEF)     # key
    assert np.all(words[:, 0] == 0) and np.all(words[:, 1] == 17)                     # counter.x, .y
    assert np.array_equal(words[:, 2], np.arange(n, dtype=np.uint32))                  # counter.z = subsequence
    assert np.all(words[:, 3] == 0)
    assert np.all(words[:, 4:8].any(axis=1))                                          # result set
    assert np.all(words[:, 10] == 0)                                                  # substate


def xorwow_kat(words, seed, offset):
    x = words[:, 6:11]
    assert np.all(x[1:] != x[:-1])
    assert np.all(words[:, 0] == words[0, 0])
    # d = 6615241 + t1 + t0 + 362437 * offset (mod 2^32), rocrand_xorwow.h:100-145
    s0 = (seed & 0xFFFFFFFF) ^ 0x2C7F967F
    s1 = (seed >> 32) ^ 0xA03697CB
    d = (6615241 + (2073658381 * s1) + (1228688033 * s0) + 362437 * (offset & 0xFFFFFFFF)) & 0xFFFFFFFF
    assert int(words[0, 0]) == d


KAT_SEED = 0xDEADBEEFBEEFDEAD
PHILOX_KAT_OFFSET = 4 * ((0xFFFFFFFF * 17) + 17)
