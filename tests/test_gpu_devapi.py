"""Device API (include/rocrand/rocrand_kernel.h) in gfx950 kernels: one GPU thread per call
sequence of tests/devapi/program.h (tests/devapi/libdevapi_gpu.so).

* Against the reference's own device API: the committed fixtures
  tests/golden/devapi_golden.npz (every engine and call) and, at 4096 threads, the
  reference build oracle/_ref/libref.so where it is present (else our host branch, which
  tests/test_devapi_host.py pins to that build bit for bit).
* Bars: words, uniforms, Poisson, discrete and skipahead bit-exact; Box-Muller and
  inverse-error-function normals and log-normals within the bars of parity_checks.py
  (the device's transcendentals are correctly rounded, the host libm's are not always).
* Philox4x32-10 and ThreeFry 2x32 / 4x32 states at (seed, 0, offset) yield the bulk
  generator's output for that seed (rocrand_generate, legacy ordering).
"""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "devapi"))
import cases as C  # noqa: E402
from parity_checks import check_box_muller, check_erfinv, check_exact  # noqa: E402

GOLDEN = os.path.join(HERE, "golden", "devapi_golden.npz")

FLOAT_FNS = {"UNIFORM", "UNIFORM2", "UNIFORM4", "NORMAL", "NORMAL2", "NORMAL4", "LOGN", "LOGN2", "LOGN4"}
DOUBLE_FNS = {"UNIFORM_D", "UNIFORM_D2", "UNIFORM_D4", "NORMAL_D", "NORMAL_D2", "NORMAL_D4", "LOGN_D", "LOGN_D2",
              "LOGN_D4"}


@pytest.fixture(scope="module")
def gpu():
    if not os.path.exists(C.GPU_LIB):
        pytest.fail("tests/devapi/libdevapi_gpu.so not built")
    import torch

    assert torch.cuda.is_available()
    torch.cuda.init()
    return C.gpu()


@pytest.fixture(scope="module")
def checker():
    ref = C.reference()
    return ref if ref is not None else C.host()


def _int_view(engine, slots):
    return slots.astype(np.uint64 if engine in C.WORD64 else np.uint32)


def _compare_slots(engine, name, got, want, mean=C.MEAN, sd=C.SD, ctx="", cached=False):
    """got, want: [threads, calls, components] uint64 slots."""
    if name in ("WORD", "WORDV", "SKIP"):
        check_exact(_int_view(engine, got).ravel(), _int_view(engine, want).ravel())
        return
    if name.startswith(("POISSON", "DISCRETE")):
        check_exact(got.astype(np.uint32).ravel(), want.astype(np.uint32).ravel())
        return
    if name == "MIXED":
        calls = got.shape[1]
        for r in range(5):
            sel = [k for k in range(calls) if k % 5 == r]
            if not sel:
                continue
            sub_name = {0: "NORMAL", 1: "UNIFORM", 2: "NORMAL_D", 3: "LOGN", 4: "NORMAL"}[r]
            # call 3's log-normal takes the value call 0's rocrand_normal cached
            _compare_slots(engine, sub_name, got[:, sel, :1], want[:, sel, :1], mean, sd, ctx + f" k%5={r}",
                           cached=r == 3)
        return
    if name in FLOAT_FNS:
        g = got.astype(np.uint32).view(np.float32).ravel()
        w = want.astype(np.uint32).view(np.float32).ravel()
        suffix = "f32"
    else:
        g = got.view(np.float64).ravel()
        w = want.astype(np.uint64).view(np.float64).ravel()
        suffix = "f64"
    if name.startswith("UNIFORM"):
        check_exact(g, w)
        return
    log = name.startswith("LOGN")
    kind = ("log_normal_" if log else "normal_") + suffix
    params = (mean, sd) if log else (0.0, 1.0)
    single = name in ("NORMAL", "NORMAL_D", "LOGN", "LOGN_D")
    if single and engine in C.ONE_WORD:
        check_erfinv(kind, g, w, *params)
    else:
        check_box_muller(kind, g, w, *params, count_bar=not cached)


@pytest.fixture(scope="module")
def golden():
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("engine", C.ENGINES)
def test_gpu_equals_reference_fixtures(gpu, golden, engine):
    sub, off = golden["sub"], golden["off"]
    for key, want in golden.items():
        if not key.startswith(engine + "/"):
            continue
        parts = key.split("/")
        name = parts[1]
        lam = float(parts[2]) if len(parts) > 2 else None
        print(f"case {key}", flush=True)
        got = gpu.run(engine, C.F[name], sub, off, want.shape[1], a=lam if lam is not None else C.MEAN)
        assert got is not None, key
        _compare_slots(engine, name, got[:, :, :want.shape[2]], want.astype(np.uint64), ctx=key)


@pytest.mark.parametrize("engine", C.ENGINES)
def test_gpu_equals_checker_4096_threads(gpu, checker, engine):
    n = 4096
    sub, off = C.thread_params(n, seed=21)
    table = C.discrete_table(size=257, seed=3)
    for fn, name in enumerate(C.FNS):
        lams = C.lambdas_for(engine) if name.startswith("POISSON") else [None]
        for lam in lams:
            calls = 2 if name.startswith("POISSON") and lam >= 64 else 8
            a = lam if lam is not None else 0.25
            print(f"case {engine} {name} lambda={lam}", flush=True)
            want = checker.run(engine, fn, sub, off, calls, a=a, b=1.5, seed=0xC0FFEE, table=table)
            got = gpu.run(engine, fn, sub, off, calls, a=a, b=1.5, seed=0xC0FFEE, table=table)
            assert (got is None) == (want is None), name
            if got is not None:
                k = C.components(engine, fn)
                _compare_slots(engine, name, got[:, :, :k], want[:, :, :k], mean=0.25, sd=1.5, ctx=name)


@pytest.mark.parametrize("engine", ["lfsr113", "threefry2x64", "threefry4x32"])
def test_gpu_inversion_poisson_between_500_and_1000(gpu, engine):
    """No reference values here (its inversion does not terminate); equal to our host
    branch and with Poisson moments."""
    n, calls = 4096, 2
    sub = np.arange(n, dtype=np.uint64)
    off = np.full(n, 17, dtype=np.uint64)
    for lam in (600.0, 999.0):
        got = gpu.run(engine, C.F["POISSON"], sub, off, calls, a=lam)
        want = C.host().run(engine, C.F["POISSON"], sub, off, calls, a=lam)
        check_exact(got[:, :, 0].astype(np.uint32).ravel(), want[:, :, 0].astype(np.uint32).ravel())
        k = got[:, :, 0].astype(np.float64).ravel()
        assert abs(k.mean() - lam) < 5 * np.sqrt(lam / k.size)


@pytest.mark.parametrize("engine,rng", [("philox", "PHILOX4_32_10"), ("threefry2x32", "THREEFRY2_32_20"),
                                        ("threefry4x32", "THREEFRY4_32_20")])
def test_device_states_continue_the_bulk_stream(gpu, engine, rng):
    import torch

    from rocrand_amd import rocrand as R

    threads, calls = 64, 256
    seed = 0x1234ABCD5678
    g = R.PRNG(getattr(R, rng), seed=seed)
    bulk = torch.empty(threads * calls, dtype=torch.uint32, device="cuda")
    g.generate(bulk)
    torch.cuda.synchronize()
    off = np.arange(threads, dtype=np.uint64) * calls
    got = gpu.run(engine, C.F["WORD"], np.zeros(threads, np.uint64), off, calls, seed=seed)[:, :, 0]
    assert np.array_equal(got.astype(np.uint32).ravel(), bulk.cpu().numpy())


@pytest.mark.parametrize("engine,rng,bits", [("sobol32", "SOBOL32", 32), ("sobol64", "SOBOL64", 64)])
def test_sobol_states_equal_the_bulk_quasi_generator(gpu, engine, rng, bits):
    """States built from rocrand_get_direction_vectors32/64 (the Joe-Kuo tables the bulk
    generator uses) at offset O give point O + i of their dimension, which is what
    rocrand_generate writes at [dimension * n + i] for a generator with that offset."""
    import ctypes

    import torch

    from rocrand_amd import rocrand as R

    L = R.load()
    dims, n, offset = 6, 1000, 12345
    if bits == 32:
        p = ctypes.POINTER(ctypes.c_uint)()
        R.check_rocrand(L.rocrand_get_direction_vectors32(ctypes.byref(p), 101))
        v = np.ctypeslib.as_array(p, shape=(dims * 32,)).copy()
        vectors = (v, np.zeros(dims * 64, np.uint64))
        dtype = torch.uint32
    else:
        p = ctypes.POINTER(ctypes.c_ulonglong)()
        R.check_rocrand(L.rocrand_get_direction_vectors64(ctypes.byref(p), 103))
        v = np.ctypeslib.as_array(p, shape=(dims * 64,)).copy()
        vectors = (np.zeros(dims * 32, np.uint32), v)
        dtype = torch.uint64
    g = R.QRNG(getattr(R, rng), ndim=dims, offset=offset)
    bulk = torch.empty(dims * n, dtype=dtype, device="cuda")
    g.generate(bulk)
    torch.cuda.synchronize()
    sub = np.arange(dims, dtype=np.uint64)
    off = np.full(dims, offset, dtype=np.uint64)
    got = gpu.run(engine, C.F["WORD"], sub, off, n, vectors=vectors)[:, :, 0]
    want = bulk.cpu().numpy().reshape(dims, n).astype(np.uint64)
    assert np.array_equal(got, want)


def _mtgp32(fn, seed, calls, use_constant=0, blocks=512, threads=256):
    import ctypes

    from rocrand_amd import rocrand as R

    L = R.load()
    params = ctypes.addressof(ctypes.c_char.in_dll(L, "mtgp32dc_params_fast_11213"))
    lib = ctypes.CDLL(C.GPU_LIB)
    f = lib.devapi_gpu_mtgp32
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int, ctypes.c_ulonglong, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                  ctypes.c_double, ctypes.c_uint, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                  ctypes.c_void_p]
    prob, alias, _, _ = C.discrete_table()
    out = np.zeros(blocks * threads * calls * 4, dtype=np.uint64)
    rc = f(fn, seed, blocks, threads, calls, C.MEAN, C.SD, len(prob), prob.ctypes.data, alias.ctypes.data, params,
           use_constant, out.ctypes.data)
    assert rc == 0, rc
    return out.reshape(blocks, threads, calls, 4)


@pytest.mark.parametrize("use_constant", [0, 1])
def test_mtgp32_states_equal_the_bulk_generator(gpu, use_constant):
    """States from rocrand_make_state_mtgp32(mtgp32dc_params_fast_11213, 512, seed), one
    256-thread block each (state in LDS via rocrand_mtgp32_block_copy; optionally tables
    reloaded by rocrand_mtgp32_set_params from rocrand_make_constant): thread t of block b
    makes the words the bulk MTGP32 generator writes at vec b*256 + t + j*131072 (the
    reference's mtgp32.hpp output map; one word per u32/uniform vec, a Box-Muller pair of
    two consecutive steps per normal vec)."""
    import torch

    from rocrand_amd import rocrand as R

    seed, K = 0x0123456789ABCDEF, 8
    n = 512 * 256 * K
    vec = np.arange(512)[:, None, None] * 256 + np.arange(256)[None, :, None]
    idx = vec + np.arange(K)[None, None, :] * 131072
    idx2 = ((vec + (np.arange(K)[None, None, :] // 2) * 131072) * 2 + np.arange(K)[None, None, :] % 2)
    for kind, fn, dt in (("u32", "WORD", torch.uint32), ("uniform", "UNIFORM", torch.float32),
                         ("normal", "NORMAL2", torch.float32)):
        g = R.PRNG(R.MTGP32, seed=seed)
        bulk = torch.empty(n, dtype=dt, device="cuda")
        if kind == "u32":
            g.generate(bulk)
        elif kind == "uniform":
            g.uniform(bulk)
        else:
            g.normal(bulk, 0.0, 1.0)
        torch.cuda.synchronize()
        want = bulk.cpu().numpy().view(np.uint32)
        print(f"case mtgp32 {kind} use_constant={use_constant}", flush=True)
        if fn == "NORMAL2":
            got = _mtgp32(C.F[fn], seed, K // 2, use_constant)[:, :, :, :2].reshape(512, 256, K)
        else:
            got = _mtgp32(C.F[fn], seed, K, use_constant)[:, :, :, 0]
        assert np.array_equal(got.astype(np.uint32), want[idx2 if fn == "NORMAL2" else idx]), kind


def test_mtgp32_calls_match_their_word_transforms(gpu):
    """Poisson, discrete and the one-word normals of an MTGP32 state are the shared
    transforms of the words the same state yields (pinned for the other states)."""
    seed, K = 99, 12
    words = _mtgp32(C.F["WORD"], seed, K, blocks=4)[:, :, :, 0].astype(np.uint32)
    u = _mtgp32(C.F["UNIFORM_D"], seed, K, blocks=4)[:, :, :, 0].view(np.float64)
    assert np.array_equal(u, 2.0**-32 + words.astype(np.float64) * 2.0**-32)
    disc = _mtgp32(C.F["DISCRETE"], seed, K, blocks=4)[:, :, :, 0].astype(np.uint32)
    prob, alias, _, _ = C.discrete_table()  # the MTGP32 driver's table has offset 0
    x = words.astype(np.float64) * 2.0**-32 * len(prob)
    i = np.floor(x).astype(np.int64)
    want = np.where(x - i < prob[i], i, alias[i]).astype(np.uint32)
    assert np.array_equal(disc, want)
    pois = _mtgp32(C.F["POISSON"], seed, 3, blocks=4)[:, :, :, 0].astype(np.float64)
    lam = C.MEAN
    assert abs(pois.mean() - lam) < 5 * np.sqrt(lam / pois.size)


def test_gpu_reference_state_kats(gpu):
    """The reference's device-API state checks (test_rocrand_kernel_philox4x32_10.cpp:227-271,
    test_rocrand_kernel_xorwow.cpp:211-250), on states initialised in a kernel."""
    sub = np.arange(256, dtype=np.uint64)
    w, size = C.state_words(C.GPU_LIB, "devapi_gpu_state", "philox", C.KAT_SEED, sub,
                            np.full(256, C.PHILOX_KAT_OFFSET, dtype=np.uint64))
    assert size == 64
    C.philox_kat(w)
    w, _ = C.state_words(C.GPU_LIB, "devapi_gpu_state", "xorwow", C.KAT_SEED, sub, np.full(256, 345678, np.uint64))
    C.xorwow_kat(w, C.KAT_SEED, 345678)


@pytest.mark.parametrize("engine", list(C.STATE_WORDS))
def test_gpu_states_equal_host_states(gpu, engine):
    sub, off = C.thread_params(4096, seed=17)
    g, size = C.state_words(C.GPU_LIB, "devapi_gpu_state", engine, 0xFEEDFACE12345678, sub, off)
    h, hsize = C.state_words(C.HOST_LIB, "devapi_host_state", engine, 0xFEEDFACE12345678, sub, off)
    assert size == hsize
    cols = C.STATE_WORDS[engine]
    assert np.array_equal(g[:, cols], h[:, cols])


def test_device_api_throughput(gpu):
    """A user kernel over the device API (one state per thread, grid-stride loop of 16-byte
    stores, 2^28 values per launch): GB/s per engine and kind, printed for the record."""
    import ctypes

    lib = ctypes.CDLL(C.GPU_LIB)
    f = lib.devapi_gpu_bench
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_float)]
    ms = (ctypes.c_float * 2)()
    assert f(C.E["philox"], 3, 50, ms) == 0
    print(f"device API loop without generation: {(1 << 28) * 4 / (ms[0] * 1e-3) / 1e9:.0f} GB/s", flush=True)
    for engine in ("philox", "threefry4x32", "mrg32k3a", "mrg31k3p", "xorwow", "lfsr113"):
        row = []
        for kind, name in enumerate(("u32", "uniform_f32", "normal_f32")):
            assert f(C.E[engine], kind, 50, ms) == 0
            gbs = (1 << 28) * 4 / (ms[0] * 1e-3) / 1e9
            row.append(f"{name} {gbs:.0f} GB/s")
            assert gbs > 100
        print(f"device API {engine}: {', '.join(row)}; init of 524288 states {ms[1]:.3f} ms", flush=True)


def _weights(kind, size, rng):
    if kind == "skewed":
        return rng.random(size) ** 8
    if kind == "sparse":
        w = rng.random(size)
        w[rng.random(size) < 0.5] = 0.0
        w[0] = max(w[0], 1e-3)
        return w
    if kind == "equal":
        return np.full(size, 3.0)
    return rng.random(size)


DISCRETE_ENGINES = ["philox", "mrg32k3a", "xorwow", "lfsr113", "threefry4x64", "sobol32", "scrambled_sobol64"]


def _sample_created(handle, table, n=512, calls=8):
    """Samples of rocrand_discrete / rocrand_discrete4 in gfx950 kernels over `handle` (a
    distribution made by the C ABI) against the reference's device API on the host over
    `table` (the reference factory's tables), for alias (pseudo) and CDF (quasi) engines."""
    ref = C.reference()
    if ref is None:
        pytest.skip("oracle/_ref/libref.so not built")
    drv = C.HandleDriver()
    sub, off = C.thread_params(n, seed=5)
    for engine in DISCRETE_ENGINES:
        for name in ("DISCRETE", "DISCRETE4"):
            fn = C.F[name]
            print(f"case {engine} {name} size={len(table[0])}", flush=True)
            got = drv.run(engine, fn, sub, off, calls, handle)
            want = ref.run(engine, fn, sub, off, calls, table=table)
            assert (got is None) == (want is None), (engine, name)
            if got is not None:
                k = C.components(engine, fn)
                assert np.array_equal(got[:, :, :k], want[:, :, :k]), (engine, name)


@pytest.mark.parametrize("size,kind,offset", [(1, "uniform", 0), (7, "skewed", 3), (100, "sparse", 5),
                                              (1000, "uniform", 0), (4096, "equal", 17)])
def test_created_discrete_distribution_equals_reference(gpu, size, kind, offset):
    """rocrand_create_discrete_distribution (library/src/rocrand.cpp:508-551,
    discrete.hpp:189-286) on random weights: the device tables sample like the reference
    factory's, bit for bit."""
    import ctypes

    from oracle import oracle as O
    from rocrand_amd import rocrand as R

    L = R.load()
    w = np.ascontiguousarray(_weights(kind, size, np.random.default_rng(size + 11)), dtype=np.float64)
    h = ctypes.c_void_p()
    assert L.rocrand_create_discrete_distribution(w.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), size, offset,
                                                  ctypes.byref(h)) == 0
    try:
        prob, alias, cdf = O.ref_discrete_tables(w, offset)
        _sample_created(h.value, (prob, alias, cdf, offset))
    finally:
        assert L.rocrand_destroy_discrete_distribution(h) == 0


@pytest.mark.parametrize("lam", [0.5, 4.0, 10.0, 1000.0, 3999.5])
def test_created_poisson_distribution_equals_reference(gpu, lam):
    """rocrand_create_poisson_distribution (library/src/rocrand.cpp:462-506,
    poisson.hpp:110-161): the device tables sample like the reference's, bit for bit."""
    import ctypes

    from oracle import oracle as O
    from rocrand_amd import rocrand as R

    L = R.load()
    h = ctypes.c_void_p()
    assert L.rocrand_create_poisson_distribution(lam, ctypes.byref(h)) == 0
    try:
        off, prob, alias, cdf = O.ref_poisson_tables(lam)
        _sample_created(h.value, (prob, alias, cdf, off))
    finally:
        assert L.rocrand_destroy_discrete_distribution(h) == 0
