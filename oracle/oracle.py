"""ctypes binding of the oracle (oracle/liboracle.so) and of oracle/_ref/libref.so.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product (rocrand_amd) never
imports this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref.so")
DATA = os.path.join(os.path.dirname(HERE), "rocrand_amd", "data")

# rocrand_rng_type values (library/include/rocrand/rocrand.h:77-101)
XORWOW, MRG32K3A, MTGP32, PHILOX4_32_10, MRG31K3P, LFSR113, MT19937 = 401, 402, 403, 404, 405, 406, 407
THREEFRY2_32_20, THREEFRY2_64_20, THREEFRY4_32_20, THREEFRY4_64_20 = 408, 409, 410, 411
SOBOL32, SCRAMBLED_SOBOL32, SOBOL64, SCRAMBLED_SOBOL64 = 501, 502, 504, 505

# orc_kind values (oracle.h) and their numpy dtypes
KINDS = {
    "u8": (0, np.uint8), "u16": (1, np.uint16), "u32": (2, np.uint32), "u64": (3, np.uint64),
    "uniform_f32": (4, np.float32), "uniform_f64": (5, np.float64), "uniform_f16": (6, np.float16),
    "normal_f32": (7, np.float32), "normal_f64": (8, np.float64), "normal_f16": (9, np.float16),
    "lognormal_f32": (10, np.float32), "lognormal_f64": (11, np.float64), "lognormal_f16": (12, np.float16),
    "poisson": (13, np.uint32),
}


def build():
    """Compile liboracle.so (gcc, baseline x86-64, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _load(path):
    if not os.path.exists(path):
        return None
    return ctypes.CDLL(path)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        L = _lib
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.c_int]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_set_seed.argtypes = [ctypes.c_void_p, ctypes.c_ulonglong]
        L.orc_set_seed_uint4.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint)]
        L.orc_set_offset.argtypes = [ctypes.c_void_p, ctypes.c_ulonglong]
        L.orc_set_dimensions.argtypes = [ctypes.c_void_p, ctypes.c_uint]
        L.orc_set_grid.argtypes = [ctypes.c_void_p, ctypes.c_uint, ctypes.c_uint]
        L.orc_generate.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t,
                                   ctypes.c_uint64, ctypes.c_double, ctypes.c_double]
        L.orc_engine_stream.argtypes = [ctypes.c_int, ctypes.c_ulonglong, ctypes.c_ulonglong,
                                        ctypes.c_ulonglong, ctypes.c_void_p, ctypes.c_size_t]
        L.orc_mrg_matrix.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
        L.orc_poisson_table.argtypes = [ctypes.c_double, ctypes.POINTER(ctypes.c_uint), ctypes.c_void_p,
                                        ctypes.c_void_p]
        L.orc_discrete_tables.argtypes = [ctypes.c_void_p, ctypes.c_uint, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_void_p]
        L.orc_float2half.argtypes = [ctypes.c_float]
        L.orc_float2half.restype = ctypes.c_ushort
        for name, ct in [("orc_philox4x32_10_block", ctypes.c_uint), ("orc_threefry4x32_20_block", ctypes.c_uint),
                         ("orc_threefry2x32_20_block", ctypes.c_uint),
                         ("orc_threefry4x64_20_block", ctypes.c_ulonglong),
                         ("orc_threefry2x64_20_block", ctypes.c_ulonglong)]:
            getattr(L, name).argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_sobol_load.argtypes = [ctypes.c_char_p]
        L.orc_mtgp32_load.argtypes = [ctypes.c_char_p]
        for name in ["orc_sobol32_vectors", "orc_scrambled_sobol32_vectors", "orc_scramble32_constants"]:
            getattr(L, name).restype = ctypes.POINTER(ctypes.c_uint)
        for name in ["orc_sobol64_vectors", "orc_scrambled_sobol64_vectors", "orc_scramble64_constants"]:
            getattr(L, name).restype = ctypes.POINTER(ctypes.c_ulonglong)
        # Published input tables shared with the library: Joe-Kuo direction numbers and the
        # MTGP32 parameter sets (each derived by the scripts beside them).
        if L.orc_sobol_load(os.path.join(DATA, "joe_kuo_20000.bin").encode()) != 0:
            raise RuntimeError("oracle: cannot load the Joe-Kuo direction numbers")
        if L.orc_mtgp32_load(os.path.join(DATA, "mtgp32_11213.bin").encode()) != 0:
            raise RuntimeError("oracle: cannot load the MTGP32 parameter sets")
        L.orc_mt19937_load.argtypes = [ctypes.c_char_p]
        if L.orc_mt19937_load(os.path.join(DATA, "mt19937_jumps.bin").encode()) != 0:
            raise RuntimeError("oracle: cannot load the MT19937 jump polynomials")
    return _lib


def sobol_tables():
    """(vectors32, scrambled32, vectors64, scrambled64, constants32, constants64) as numpy arrays."""
    L = lib()
    n32, n64 = 20000 * 32, 20000 * 64
    return (np.ctypeslib.as_array(L.orc_sobol32_vectors(), (n32,)).copy(),
            np.ctypeslib.as_array(L.orc_scrambled_sobol32_vectors(), (n32,)).copy(),
            np.ctypeslib.as_array(L.orc_sobol64_vectors(), (n64,)).copy(),
            np.ctypeslib.as_array(L.orc_scrambled_sobol64_vectors(), (n64,)).copy(),
            np.ctypeslib.as_array(L.orc_scramble32_constants(), (20000,)).copy(),
            np.ctypeslib.as_array(L.orc_scramble64_constants(), (20000,)).copy())


def ref():
    """The reference kernel bodies driven on the CPU, or None where not built."""
    global _ref
    if _ref is None:
        _ref = _load(REF_PATH)
        if _ref is None:
            return None
        R = _ref
        R.ref_create.restype = ctypes.c_void_p
        R.ref_create.argtypes = [ctypes.c_int]
        R.ref_destroy.argtypes = [ctypes.c_void_p]
        R.ref_set_seed.argtypes = [ctypes.c_void_p, ctypes.c_ulonglong]
        R.ref_set_seed_uint4.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint)]
        R.ref_set_offset.argtypes = [ctypes.c_void_p, ctypes.c_ulonglong]
        R.ref_generate.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t,
                                   ctypes.c_uint64, ctypes.c_double, ctypes.c_double]
        R.ref_engine_stream.argtypes = [ctypes.c_int, ctypes.c_ulonglong, ctypes.c_ulonglong,
                                        ctypes.c_ulonglong, ctypes.c_void_p, ctypes.c_size_t]
        R.ref_set_dimensions.argtypes = [ctypes.c_void_p, ctypes.c_uint]
        R.ref_set_grid.argtypes = [ctypes.c_void_p, ctypes.c_uint, ctypes.c_uint]
        R.ref_sobol64_set_vectors.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        for name in ["ref_sobol32_vectors", "ref_scrambled_sobol32_vectors", "ref_scramble32_constants"]:
            getattr(R, name).restype = ctypes.POINTER(ctypes.c_uint)
        R.ref_scramble64_constants.restype = ctypes.POINTER(ctypes.c_ulonglong)
        R.ref_mt19937_jump.restype = ctypes.POINTER(ctypes.c_uint)
        R.ref_mt19937_set_limit.argtypes = [ctypes.c_void_p, ctypes.c_uint]
        R.ref_discrete_tables.argtypes = [ctypes.c_void_p, ctypes.c_uint, ctypes.c_uint, ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_void_p]
        R.ref_poisson_tables.restype = ctypes.c_uint
        R.ref_poisson_tables.argtypes = [ctypes.c_double, ctypes.c_uint, ctypes.POINTER(ctypes.c_uint),
                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        # the reference tree has no 64-bit direction-vector tables: hand it the oracle's
        t = sobol_tables()
        _ref_keep.extend([t[2], t[3]])
        R.ref_sobol64_set_vectors(t[2].ctypes.data, t[3].ctypes.data)
    return _ref


_ref_keep = []


def ref_sobol_tables():
    """The reference's own (vectors32, scrambled32, constants32, constants64)."""
    R = ref()
    n32 = 20000 * 32
    return (np.ctypeslib.as_array(R.ref_sobol32_vectors(), (n32,)).copy(),
            np.ctypeslib.as_array(R.ref_scrambled_sobol32_vectors(), (n32,)).copy(),
            np.ctypeslib.as_array(R.ref_scramble32_constants(), (20000,)).copy(),
            np.ctypeslib.as_array(R.ref_scramble64_constants(), (20000,)).copy())


class _Gen:
    _prefix = "orc"

    def __init__(self, rng_type, seed=None, offset=None, dims=None, seed4=None, grid=None):
        self.L = self._library()
        self.h = getattr(self.L, self._prefix + "_create")(rng_type)
        if not self.h:
            raise ValueError(f"unsupported rng type {rng_type}")
        if grid is not None:
            st = self.set_grid(*grid)
            if st != 0:
                raise ValueError(f"grid {grid}: status {st}")
        if seed is not None:
            self.set_seed(seed)
        if seed4 is not None:
            self.set_seed_uint4(seed4)
        if offset is not None:
            self.set_offset(offset)
        if dims is not None:
            self.set_dimensions(dims)

    def _library(self):
        return lib()

    def __del__(self):
        try:
            getattr(self.L, self._prefix + "_destroy")(self.h)
        except Exception:
            pass

    def set_seed(self, seed):
        return getattr(self.L, self._prefix + "_set_seed")(self.h, seed)

    def set_seed_uint4(self, s4):
        arr = (ctypes.c_uint * 4)(*s4)
        return getattr(self.L, self._prefix + "_set_seed_uint4")(self.h, arr)

    def set_offset(self, off):
        return getattr(self.L, self._prefix + "_set_offset")(self.h, off)

    def set_dimensions(self, d):
        return getattr(self.L, self._prefix + "_set_dimensions")(self.h, d)

    def set_grid(self, threads, blocks):
        """The generate kernels' launch grid (threads x blocks): the engine count of the
        interleaved engines, the grid stride of the counter engines (a tuned,
        DYNAMIC-ordering configuration)."""
        return getattr(self.L, self._prefix + "_set_grid")(self.h, threads, blocks)

    def generate(self, kind, n, addr=0x10000000, a=0.0, b=1.0):
        """Return (status, array) for n values of `kind` at device address `addr`."""
        k, dt = KINDS[kind]
        out = np.zeros(max(n, 1), dtype=dt)
        st = getattr(self.L, self._prefix + "_generate")(self.h, k, out.ctypes.data, n, addr, a, b)
        return st, out[:n]


class OracleGen(_Gen):
    """Stateful oracle generator mirroring one rocrand_generator."""


class RefGen(_Gen):
    """Same interface over the reference kernel bodies (oracle/_ref/libref.so)."""

    _prefix = "ref"

    def _library(self):
        r = ref()
        if r is None:
            raise RuntimeError("oracle/_ref/libref.so not built")
        return r


def engine_stream(rng_type, seed, sub, offset, n, use_ref=False):
    out = np.zeros(n, dtype=np.uint32)
    L = ref() if use_ref else lib()
    fn = L.ref_engine_stream if use_ref else L.orc_engine_stream
    rc = fn(rng_type, seed, sub, offset, out.ctypes.data, n)
    if rc != 0:
        raise ValueError("unsupported engine")
    return out


def poisson_table(lam):
    off = ctypes.c_uint(0)
    prob = np.zeros(8192, dtype=np.float64)
    alias = np.zeros(8192, dtype=np.uint32)
    size = lib().orc_poisson_table(lam, ctypes.byref(off), prob.ctypes.data, alias.ctypes.data)
    return off.value, prob[:size].copy(), alias[:size].copy()


def discrete_tables(weights):
    """(prob, alias, cdf) of rocrand_create_discrete_distribution for `weights`."""
    w = np.ascontiguousarray(weights, dtype=np.float64)
    prob, cdf = np.zeros(w.size), np.zeros(w.size)
    alias = np.zeros(w.size, dtype=np.uint32)
    if lib().orc_discrete_tables(w.ctypes.data, w.size, prob.ctypes.data, alias.ctypes.data, cdf.ctypes.data):
        raise MemoryError("orc_discrete_tables")
    return prob, alias, cdf


def ref_discrete_tables(weights, offset=0):
    """The reference factory's (prob, alias, cdf) for `weights` (oracle/_ref/libref.so)."""
    w = np.ascontiguousarray(weights, dtype=np.float64)
    prob, cdf = np.zeros(w.size), np.zeros(w.size)
    alias = np.zeros(w.size, dtype=np.uint32)
    if ref().ref_discrete_tables(w.ctypes.data, w.size, offset, prob.ctypes.data, alias.ctypes.data,
                                 cdf.ctypes.data):
        raise RuntimeError("reference discrete factory failed")
    return prob, alias, cdf


def ref_poisson_tables(lam, cap=1 << 16):
    """The reference's (offset, prob, alias, cdf) for a Poisson distribution of mean `lam`."""
    off = ctypes.c_uint(0)
    prob, cdf = np.zeros(cap), np.zeros(cap)
    alias = np.zeros(cap, dtype=np.uint32)
    size = ref().ref_poisson_tables(lam, cap, ctypes.byref(off), prob.ctypes.data, alias.ctypes.data,
                                    cdf.ctypes.data)
    if size == 0 or size > cap:
        raise RuntimeError(f"reference Poisson table for {lam}: size {size}")
    return off.value, prob[:size].copy(), alias[:size].copy(), cdf[:size].copy()
