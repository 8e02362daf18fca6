"""Python interface over librocrand.so, mirroring the reference wrapper
python/rocrand/rocrand/rocrand.py (RNG / PRNG / QRNG, RocRandError, get_version).

Buffers are torch tensors: device tensors for device generators, CPU tensors
(or numpy arrays) for host generators.  Every call goes through the C ABI of
rocrand_amd/lib/librocrand.so; there is no fallback path, and loading fails
loudly when the library has not been built.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "librocrand.so")

# rocrand_rng_type (include/rocrand.h)
DEFAULT = 400
XORWOW = 401
MRG32K3A = 402
MTGP32 = 403
PHILOX4_32_10 = 404
MRG31K3P = 405
LFSR113 = 406
MT19937 = 407
THREEFRY2_32_20 = 408
THREEFRY2_64_20 = 409
THREEFRY4_32_20 = 410
THREEFRY4_64_20 = 411
QUASI_DEFAULT = 500
SOBOL32 = 501
SCRAMBLED_SOBOL32 = 502
SOBOL64 = 504
SCRAMBLED_SOBOL64 = 505

# rocrand_ordering
ORDERING_PSEUDO_BEST = 100
ORDERING_PSEUDO_DEFAULT = 101
ORDERING_PSEUDO_SEEDED = 102
ORDERING_PSEUDO_LEGACY = 103
ORDERING_PSEUDO_DYNAMIC = 104
ORDERING_QUASI_DEFAULT = 201

STATUS = {
    0: "SUCCESS", 100: "VERSION_MISMATCH", 101: "NOT_CREATED", 102: "ALLOCATION_FAILED",
    103: "TYPE_ERROR", 104: "OUT_OF_RANGE", 105: "LENGTH_NOT_MULTIPLE",
    106: "DOUBLE_PRECISION_REQUIRED", 107: "LAUNCH_FAILURE", 108: "INTERNAL_ERROR",
}

_lib = None
_hiplib = None


def _hip():
    """The HIP runtime librocrand.so links (for stream synchronization of host generators)."""
    global _hiplib
    if _hiplib is None:
        _hiplib = ctypes.CDLL("libamdhip64.so", mode=ctypes.RTLD_GLOBAL)
        _hiplib.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    return _hiplib


def _hip_check(err):
    if err != 0:
        raise RuntimeError(f"HIP error {err}")


class _uint4(ctypes.Structure):
    _fields_ = [("x", ctypes.c_uint), ("y", ctypes.c_uint), ("z", ctypes.c_uint), ("w", ctypes.c_uint)]


def load():
    """Load librocrand.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run `python -m rocrand_amd.build` first")
    L = ctypes.CDLL(LIB_PATH)
    vp, sz, u64 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_ulonglong
    L.rocrand_create_generator.argtypes = [ctypes.POINTER(vp), ctypes.c_int]
    L.rocrand_create_generator_host.argtypes = [ctypes.POINTER(vp), ctypes.c_int]
    L.rocrand_create_generator_host_blocking.argtypes = [ctypes.POINTER(vp), ctypes.c_int]
    L.rocrand_destroy_generator.argtypes = [vp]
    for name in ["rocrand_generate", "rocrand_generate_long_long", "rocrand_generate_char",
                 "rocrand_generate_short", "rocrand_generate_uniform", "rocrand_generate_uniform_double",
                 "rocrand_generate_uniform_half"]:
        getattr(L, name).argtypes = [vp, vp, sz]
    for name, t in [("rocrand_generate_normal", ctypes.c_float), ("rocrand_generate_log_normal", ctypes.c_float),
                    ("rocrand_generate_normal_double", ctypes.c_double),
                    ("rocrand_generate_log_normal_double", ctypes.c_double),
                    ("rocrand_generate_normal_half", ctypes.c_uint16),
                    ("rocrand_generate_log_normal_half", ctypes.c_uint16)]:
        getattr(L, name).argtypes = [vp, vp, sz, t, t]
    L.rocrand_generate_poisson.argtypes = [vp, vp, sz, ctypes.c_double]
    L.rocrand_initialize_generator.argtypes = [vp]
    L.rocrand_set_stream.argtypes = [vp, vp]
    L.rocrand_set_seed.argtypes = [vp, u64]
    L.rocrand_set_seed_uint4.argtypes = [vp, _uint4]
    L.rocrand_set_offset.argtypes = [vp, u64]
    L.rocrand_set_ordering.argtypes = [vp, ctypes.c_int]
    L.rocrand_set_quasi_random_generator_dimensions.argtypes = [vp, ctypes.c_uint]
    L.rocrand_get_version.argtypes = [ctypes.POINTER(ctypes.c_int)]
    L.rocrand_create_poisson_distribution.argtypes = [ctypes.c_double, ctypes.POINTER(vp)]
    L.rocrand_create_discrete_distribution.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.c_uint,
                                                       ctypes.c_uint, ctypes.POINTER(vp)]
    L.rocrand_destroy_discrete_distribution.argtypes = [vp]
    L.rocrand_get_direction_vectors32.argtypes = [ctypes.POINTER(ctypes.POINTER(ctypes.c_uint)), ctypes.c_int]
    L.rocrand_get_direction_vectors64.argtypes = [ctypes.POINTER(ctypes.POINTER(ctypes.c_ulonglong)), ctypes.c_int]
    L.rocrand_get_scramble_constants32.argtypes = [ctypes.POINTER(ctypes.POINTER(ctypes.c_uint))]
    L.rocrand_get_scramble_constants64.argtypes = [ctypes.POINTER(ctypes.POINTER(ctypes.c_ulonglong))]
    L.rocrand_batch_plan_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(BatchStream), sz, vp]
    L.rocrand_batch_plan_launch.argtypes = [vp]
    L.rocrand_batch_plan_destroy.argtypes = [vp]
    L.rocrand_generate_batched.argtypes = [ctypes.POINTER(BatchStream), sz, vp]
    L.rocrand_generate_batched_noalloc.argtypes = [ctypes.POINTER(BatchStream), sz, vp]
    L.rocrand_amd_set_dynamic_grid.argtypes = [ctypes.c_int, ctypes.c_uint, ctypes.c_uint]
    L.rocrand_amd_get_dynamic_grid.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_uint),
                                               ctypes.POINTER(ctypes.c_uint)]
    _lib = L
    return L


# rocrand_batched.h: one fresh generator's single generate call per stream
BATCH_UINT, BATCH_UNIFORM_FLOAT, BATCH_UNIFORM_DOUBLE, BATCH_NORMAL_FLOAT = 0, 1, 2, 3
BATCH_NORMAL_DOUBLE, BATCH_LOG_NORMAL_FLOAT, BATCH_LOG_NORMAL_DOUBLE, BATCH_POISSON = 4, 5, 6, 7


class BatchStream(ctypes.Structure):
    _fields_ = [("rng_type", ctypes.c_int), ("kind", ctypes.c_int), ("seed", ctypes.c_ulonglong),
                ("offset", ctypes.c_ulonglong), ("n", ctypes.c_size_t), ("a", ctypes.c_double),
                ("b", ctypes.c_double), ("output", ctypes.c_void_p)]


def _stream_handle(stream):
    if stream is None:
        return None
    return ctypes.c_void_p(getattr(stream, "cuda_stream", stream))


def _batch_array(streams):
    arr = (BatchStream * len(streams))()
    for i, st in enumerate(streams):
        arr[i] = st if isinstance(st, BatchStream) else BatchStream(*st)
    return arr


def generate_batched(streams, stream=None):
    """One-shot batched generation (rocrand_generate_batched); `streams` are BatchStream
    or (rng_type, kind, seed, offset, n, a, b, output_pointer) tuples."""
    arr = _batch_array(streams)
    check_rocrand(load().rocrand_generate_batched(arr, len(streams), _stream_handle(stream)))


class BatchPlan:
    """rocrand_batch_plan: a batch laid out once, replayed as one graph launch."""

    def __init__(self, streams, stream=None):
        self._arr = _batch_array(streams)
        self._plan = ctypes.c_void_p()
        check_rocrand(load().rocrand_batch_plan_create(ctypes.byref(self._plan), self._arr, len(streams),
                                                       _stream_handle(stream)))

    def launch(self):
        check_rocrand(load().rocrand_batch_plan_launch(self._plan))

    def close(self):
        if self._plan:
            check_rocrand(load().rocrand_batch_plan_destroy(self._plan))
            self._plan = ctypes.c_void_p()

    def __del__(self):
        try:
            if self._plan and _lib is not None:
                _lib.rocrand_batch_plan_destroy(self._plan)
        except Exception:
            pass


class RocRandError(Exception):
    """A non-SUCCESS rocrand_status (python/rocrand/rocrand/rocrand.py:134-146)."""

    def __init__(self, value):
        self.value = value

    def __str__(self):
        return STATUS.get(self.value, "UNKNOWN") + f" ({self.value})"


def check_rocrand(status):
    if status != 0:
        raise RocRandError(status)


def set_dynamic_grid(rngtype, threads, blocks):
    """rocrand_amd_set_dynamic_grid (include/rocrand_amd_tuning.h): the DYNAMIC-ordering grid of
    an engine in this process; (0, 0) restores the library's table."""
    check_rocrand(load().rocrand_amd_set_dynamic_grid(rngtype, threads, blocks))


def get_dynamic_grid(rngtype):
    """(threads, blocks) a DYNAMIC-ordering generator of this engine uses on the current device."""
    t, b = ctypes.c_uint(), ctypes.c_uint()
    check_rocrand(load().rocrand_amd_get_dynamic_grid(rngtype, ctypes.byref(t), ctypes.byref(b)))
    return t.value, b.value


def get_version():
    v = ctypes.c_int(0)
    check_rocrand(load().rocrand_get_version(ctypes.byref(v)))
    return v.value


def _half_bits(x):
    return int(np.array([x], dtype=np.float16).view(np.uint16)[0])


def _ptr_len(ary, size):
    """(address, element count, dtype) of a torch tensor or numpy array."""
    if hasattr(ary, "data_ptr"):
        import torch  # noqa: F401  (tensors carry their own dtype mapping)
        n = ary.numel() if size is None else size
        return ary.data_ptr(), n, str(ary.dtype).replace("torch.", "")
    arr = ary
    n = arr.size if size is None else size
    return arr.ctypes.data, n, str(arr.dtype)


class RNG:
    """Base of PRNG and QRNG: one rocrand_generator handle."""

    def __init__(self, rngtype, offset=None, stream=None, is_host=False):
        """is_host=True: rocrand_create_generator_host (values written in stream order; each
        call here then synchronizes the stream, as the reference wrapper does);
        is_host="blocking": rocrand_create_generator_host_blocking."""
        L = load()
        self._gen = ctypes.c_void_p()
        if is_host == "blocking":
            create = L.rocrand_create_generator_host_blocking
        elif is_host:
            create = L.rocrand_create_generator_host
        else:
            create = L.rocrand_create_generator
        check_rocrand(create(ctypes.byref(self._gen), rngtype))
        self.is_host = is_host
        self.rngtype = rngtype
        self._offset = 0
        self._stream = None
        if offset is not None:
            self.offset = offset
        if stream is not None:
            self.stream = stream

    def __del__(self):
        if getattr(self, "_gen", None) and _lib is not None:
            _lib.rocrand_destroy_generator(self._gen)
            self._gen = None

    @property
    def handle(self):
        return self._gen.value

    @property
    def offset(self):
        return self._offset

    @offset.setter
    def offset(self, offset):
        check_rocrand(load().rocrand_set_offset(self._gen, offset))
        self._offset = offset

    @property
    def stream(self):
        return self._stream

    @stream.setter
    def stream(self, stream):
        """stream: a raw hipStream_t (int) or a torch.cuda.Stream."""
        raw = getattr(stream, "cuda_stream", stream)
        check_rocrand(load().rocrand_set_stream(self._gen, raw or None))
        self._stream = stream

    def _done(self):
        """A stream-ordered host generator has written the values once its stream drains
        (python/rocrand/rocrand/rocrand.py:215-231 synchronizes there too)."""
        if self.is_host is True:
            raw = getattr(self._stream, "cuda_stream", self._stream)
            _hip_check(_hip().hipStreamSynchronize(ctypes.c_void_p(raw or None)))

    def set_ordering(self, order):
        check_rocrand(load().rocrand_set_ordering(self._gen, order))

    def initialize(self):
        check_rocrand(load().rocrand_initialize_generator(self._gen))

    def generate(self, ary, size=None):
        """Raw bits: uint8/uint16/uint32/uint64 by the buffer's dtype."""
        p, n, dt = _ptr_len(ary, size)
        L = load()
        fn = {"uint8": L.rocrand_generate_char, "uint16": L.rocrand_generate_short,
              "int16": L.rocrand_generate_short, "uint32": L.rocrand_generate,
              "int32": L.rocrand_generate, "uint64": L.rocrand_generate_long_long,
              "int64": L.rocrand_generate_long_long}[dt]
        check_rocrand(fn(self._gen, p, n))
        self._done()

    def uniform(self, ary, size=None):
        p, n, dt = _ptr_len(ary, size)
        L = load()
        fn = {"float32": L.rocrand_generate_uniform, "float64": L.rocrand_generate_uniform_double,
              "float16": L.rocrand_generate_uniform_half}[dt]
        check_rocrand(fn(self._gen, p, n))
        self._done()

    def normal(self, ary, mean, stddev, size=None):
        p, n, dt = _ptr_len(ary, size)
        L = load()
        if dt == "float16":
            check_rocrand(L.rocrand_generate_normal_half(self._gen, p, n, _half_bits(mean), _half_bits(stddev)))
            self._done()
            return
        fn = {"float32": L.rocrand_generate_normal, "float64": L.rocrand_generate_normal_double}[dt]
        check_rocrand(fn(self._gen, p, n, mean, stddev))
        self._done()

    def lognormal(self, ary, mean, stddev, size=None):
        p, n, dt = _ptr_len(ary, size)
        L = load()
        if dt == "float16":
            check_rocrand(L.rocrand_generate_log_normal_half(self._gen, p, n, _half_bits(mean), _half_bits(stddev)))
            self._done()
            return
        fn = {"float32": L.rocrand_generate_log_normal, "float64": L.rocrand_generate_log_normal_double}[dt]
        check_rocrand(fn(self._gen, p, n, mean, stddev))
        self._done()

    def poisson(self, ary, lmbd, size=None):
        p, n, dt = _ptr_len(ary, size)
        if dt not in ("uint32", "int32"):
            raise RocRandError(103)
        check_rocrand(load().rocrand_generate_poisson(self._gen, p, n, lmbd))
        self._done()


class PRNG(RNG):
    """Pseudo-random generator (python/rocrand/rocrand/rocrand.py:359-441)."""

    def __init__(self, rngtype=DEFAULT, seed=None, offset=None, stream=None, is_host=False):
        super().__init__(rngtype, offset=offset, stream=stream, is_host=is_host)
        self._seed = None
        if seed is not None:
            self.seed = seed

    @property
    def seed(self):
        return self._seed

    @seed.setter
    def seed(self, seed):
        check_rocrand(load().rocrand_set_seed(self._gen, seed))
        self._seed = seed

    def set_seed_uint4(self, x, y, z, w):
        check_rocrand(load().rocrand_set_seed_uint4(self._gen, _uint4(x, y, z, w)))


class QRNG(RNG):
    """Quasi-random generator (python/rocrand/rocrand/rocrand.py:444-510)."""

    def __init__(self, rngtype=QUASI_DEFAULT, ndim=None, offset=None, stream=None, is_host=False):
        super().__init__(rngtype, offset=offset, stream=stream, is_host=is_host)
        self._ndim = 1
        if ndim is not None:
            self.ndim = ndim

    @property
    def ndim(self):
        return self._ndim

    @ndim.setter
    def ndim(self, ndim):
        check_rocrand(load().rocrand_set_quasi_random_generator_dimensions(self._gen, ndim))
        self._ndim = ndim
