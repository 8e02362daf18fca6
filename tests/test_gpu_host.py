"""Host generators (rocrand_create_generator_host, _host_blocking) against the oracle.

The caller's buffer is host memory.  The reference runs its kernel bodies on the host,
from a hipLaunchHostFunc callback on the generator's stream for `_host` and directly for
`_host_blocking` (library/src/rng/system.hpp:140-214); here the device path writes staging
memory whose address has the host buffer's alignment modulo 128 bytes, and the values
reach the caller's buffer by a blocking copy or by a host-function node on the stream.
Each case generates twice in a row into misaligned numpy views (the head/tail split
follows the host pointer), and compares with the oracle given the same seed, offset and
host address.  Bar: tests/parity_checks.py.
"""
import ctypes

import numpy as np
import pytest

from parity_checks import check

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import oracle as O  # noqa: E402
from rocrand_amd import rocrand as R  # noqa: E402

PSEUDO = {"philox": R.PHILOX4_32_10, "threefry4x32": R.THREEFRY4_32_20, "threefry2x64": R.THREEFRY2_64_20,
          "mrg32k3a": R.MRG32K3A, "mrg31k3p": R.MRG31K3P, "xorwow": R.XORWOW, "lfsr113": R.LFSR113,
          "mtgp32": R.MTGP32, "mt19937": R.MT19937}
QUASI = {"sobol32": R.SOBOL32, "scrambled_sobol64": R.SCRAMBLED_SOBOL64}
KINDS = [("u32", np.uint32, "generate", ()), ("uniform_f32", np.float32, "uniform", ()),
         ("uniform_f64", np.float64, "uniform", ()), ("normal_f32", np.float32, "normal", (0.0, 1.0)),
         ("lognormal_f32", np.float32, "lognormal", (0.3, 0.7)), ("poisson", np.uint32, "poisson", (10.0,))]


def _host_views(dtype, n, shift, calls):
    """Distinct misaligned host views (each a slice of its own array, kept alive)."""
    bufs = [np.zeros(n + 64, dtype=dtype) for _ in range(calls)]
    return bufs, [b[shift:shift + n] for b in bufs]


@pytest.mark.parametrize("mode", [True, "blocking"])
@pytest.mark.parametrize("name", list(PSEUDO))
@pytest.mark.parametrize("kind,dtype,meth,args", KINDS)
def test_pseudo_host_generators(mode, name, kind, dtype, meth, args):
    rng = PSEUDO[name]
    mt = rng in (R.MTGP32, R.MT19937)
    for seed, offset, n, shift in [(12345, 0 if mt else 1048583, 100003, 1), (7, 0 if mt else 3, 13, 3)]:
        print(f"case host={mode} {name} {kind} seed={seed} offset={offset} n={n} shift={shift}", flush=True)
        g = R.PRNG(rng, seed=seed, offset=None if mt else offset, is_host=mode)
        og = O.OracleGen(rng, seed=seed, offset=None if mt else offset)
        bufs, views = _host_views(dtype, n, shift, 2)
        for v in views:
            getattr(g, meth)(v, *args)
        a0 = args[0] if args else 0.0
        a1 = args[1] if len(args) > 1 else 0.0
        for v in views:
            st, want = og.generate(kind, n, v.ctypes.data, a0, a1)
            assert st == 0
            check(kind, v, want, args if "normal" in kind else (0.0, 1.0))


@pytest.mark.parametrize("mode", [True, "blocking"])
@pytest.mark.parametrize("name", list(QUASI))
def test_quasi_host_generators(mode, name):
    rng = QUASI[name]
    dims, npd, shift = 3, 1001, 1
    g = R.QRNG(rng, ndim=dims, offset=1000, is_host=mode)
    og = O.OracleGen(rng, dims=dims, offset=1000)
    for kind, dtype, meth, args in [("u32", np.uint32, "generate", ()), ("uniform_f64", np.float64, "uniform", ())]:
        bufs, views = _host_views(dtype, dims * npd, shift, 2)
        for v in views:
            getattr(g, meth)(v, *args)
        for v in views:
            st, want = og.generate(kind, dims * npd, v.ctypes.data, 0.0, 0.0)
            assert st == 0
            check(kind, v, want)


def test_stream_ordered_host_generator_writes_in_stream_order():
    """rocrand_create_generator_host returns before the values are written when its stream
    is busy: a second stream-ordered call into a fresh buffer, then one synchronization,
    gives both calls' values; the raw C ABI call (no wrapper synchronization) leaves the
    buffer to the stream."""
    L = R.load()
    gen = ctypes.c_void_p()
    R.check_rocrand(L.rocrand_create_generator_host(ctypes.byref(gen), R.PHILOX4_32_10))
    stream = torch.cuda.Stream()
    try:
        R.check_rocrand(L.rocrand_set_seed(gen, 99))
        R.check_rocrand(L.rocrand_set_stream(gen, ctypes.c_void_p(stream.cuda_stream)))
        n = 1 << 22
        a = np.zeros(n + 1, dtype=np.float32)[1:]
        b = np.zeros(n + 3, dtype=np.float32)[3:]
        big = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
        with torch.cuda.stream(stream):
            for _ in range(4):
                big.uniform_()  # keeps the stream busy while the generator calls are queued
        R.check_rocrand(L.rocrand_generate_uniform(gen, ctypes.c_void_p(a.ctypes.data), n))
        R.check_rocrand(L.rocrand_generate_uniform(gen, ctypes.c_void_p(b.ctypes.data), n))
        stream.synchronize()
        og = O.OracleGen(O.PHILOX4_32_10, seed=99)
        for v in (a, b):
            st, want = og.generate("uniform_f32", n, v.ctypes.data)
            assert st == 0
            check("uniform_f32", v, want)
        del big
    finally:
        L.rocrand_destroy_generator(gen)
