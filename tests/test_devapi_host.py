"""Device API (include/rocrand/rocrand_kernel.h), host branch, on the CPU.

The per-thread call sequences of tests/devapi/program.h run through our header's host
branch (tests/devapi/libdevapi_host.so) must reproduce, bit for bit:

* the committed fixtures tests/golden/devapi_golden.npz, written from the reference's
  own device API (library/include/rocrand/rocrand_kernel.h compiled host-only) by
  tests/golden/make_devapi_golden.py;
* the same reference build live (oracle/_ref/libref.so, where the reference sources are
  present) on random seeds, subsequences and offsets, for every engine and call;
* the oracle's engine streams (oracle/orc_engines.c): rocrand() on a state initialised
  with (seed, subsequence, offset) is the engine's word stream from that position;
* for Sobol, the Gray-code definition: point i of dimension j is the XOR of the direction
  vectors selected by the bits of i ^ (i >> 1) (XOR the scramble constant).

For 500 < lambda < 1000 the reference's inversion Poisson (LFSR113, ThreeFry) does not
terminate in useful time; ours completes the inversion, checked here against the Poisson
moments.
"""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "devapi"))
import cases as C  # noqa: E402

GOLDEN = os.path.join(HERE, "golden", "devapi_golden.npz")


@pytest.fixture(scope="module")
def host():
    if not os.path.exists(C.HOST_LIB):
        pytest.skip("tests/devapi/libdevapi_host.so not built")
    return C.host()


@pytest.fixture(scope="module")
def golden():
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _golden_cases(golden):
    for key, want in golden.items():
        if key in ("sub", "off"):
            continue
        parts = key.split("/")
        engine, name = parts[0], parts[1]
        lam = float(parts[2]) if len(parts) > 2 else None
        yield key, engine, C.F[name], lam, want


def test_golden_covers_every_engine_and_call(golden):
    keys = {tuple(k.split("/")[:2]) for k in golden if "/" in k}
    for engine in C.ENGINES:
        for name in ("WORD", "UNIFORM", "UNIFORM_D", "NORMAL", "NORMAL_D", "LOGN", "LOGN_D", "POISSON", "DISCRETE",
                     "MIXED"):
            assert (engine, name) in keys, (engine, name)
        pairs = [(engine, n) in keys for n in ("NORMAL2", "NORMAL_D2", "LOGN2", "LOGN_D2")]
        assert all(pairs) if engine not in C.SOBOL else not any(pairs), engine
    for engine in C.SOBOL:
        assert (engine, "SKIP") in keys and (engine, "WORDV") not in keys
    # Philox's vector calls and the skipahead family where the reference has them
    for name in ("WORDV", "UNIFORM2", "UNIFORM4", "UNIFORM_D2", "UNIFORM_D4", "NORMAL4", "NORMAL_D4", "LOGN4",
                 "LOGN_D4", "POISSON4", "DISCRETE4", "SKIP"):
        assert ("philox", name) in keys
    for engine in ("mrg32k3a", "mrg31k3p", "xorwow", "lfsr113"):
        assert (engine, "SKIP") in keys
    for engine in ("threefry2x32", "threefry2x64", "threefry4x32", "threefry4x64"):
        assert (engine, "WORDV") in keys and (engine, "SKIP") not in keys


def test_host_branch_equals_reference_fixtures(host, golden):
    sub, off = golden["sub"], golden["off"]
    n_calls = None
    for key, engine, fn, lam, want in _golden_cases(golden):
        n_calls = want.shape[1]
        got = host.run(engine, fn, sub, off, n_calls, a=lam if lam is not None else C.MEAN)
        assert got is not None, f"{key}: call missing from our header"
        got = got[:, :, :want.shape[2]]
        assert np.array_equal(got, want.astype(np.uint64)), f"{key}: first mismatch at {np.argwhere(got != want)[:3]}"


@pytest.mark.parametrize("engine", C.ENGINES)
def test_host_branch_equals_reference_live(host, engine):
    ref = C.reference()
    if ref is None:
        pytest.skip("oracle/_ref/libref.so not built (needs the reference sources)")
    table = C.discrete_table(size=101, seed=5)
    for seed in (0, 1, 0xFFFFFFFFFFFFFFFF, 0x5DEECE66D):
        sub, off = C.thread_params(48, seed=seed & 0xFFFF)
        for fn, name in enumerate(C.FNS):
            lams = C.lambdas_for(engine) if name.startswith("POISSON") else [C.MEAN]
            for a in lams:
                calls = 4 if name.startswith("POISSON") and a >= 64 else 16
                want = ref.run(engine, fn, sub, off, calls, a=a, b=1.7, seed=seed, table=table)
                got = host.run(engine, fn, sub, off, calls, a=a, b=1.7, seed=seed, table=table)
                assert (want is None) == (got is None), f"{engine} {name}: call set differs"
                if want is not None:
                    assert np.array_equal(got, want), f"{engine} {name} a={a} seed={seed:#x}"


# rocrand_mrg32k3a.h:369 / rocrand_mrg31k3p.h:313: rocrand() maps next() in [1, m1] to 32 bits
MRG_UINT_NORM = {"mrg32k3a": 1.000000048661607, "mrg31k3p": 2.000000001396983862}


@pytest.mark.parametrize("engine,rng_type", [("philox", 404), ("threefry2x32", 408), ("threefry4x32", 410),
                                             ("mrg32k3a", 402), ("mrg31k3p", 405), ("xorwow", 401)])
def test_words_are_the_oracle_engine_stream(host, oracle, engine, rng_type):
    sub, off = C.thread_params(6, seed=3)
    calls = 300
    got = host.run(engine, C.F["WORD"], sub, off, calls, seed=C.SEED)[:, :, 0]
    for t in range(len(sub)):
        want = oracle.engine_stream(rng_type, C.SEED, int(sub[t]), int(off[t]), calls)
        if engine in MRG_UINT_NORM:
            want = ((want.astype(np.float64) - 1.0) * MRG_UINT_NORM[engine]).astype(np.uint32)
        assert np.array_equal(got[t].astype(np.uint32), want), f"{engine} thread {t}"


@pytest.mark.parametrize("engine", ["mrg32k3a", "mrg31k3p", "xorwow", "lfsr113", "philox", "sobol32", "sobol64",
                                    "scrambled_sobol64"])
def test_skipahead_equals_stepping(host, engine):
    """skipahead(k) then rocrand() is word k of the unskipped stream (offset 0 + k)."""
    calls = 1
    ks = np.array([0, 1, 2, 3, 5, 63, 64, 65, 1000, 4095, 4096, 65537], dtype=np.uint64)
    sub = np.full(len(ks), 9, dtype=np.uint64)
    stream = host.run(engine, C.F["WORD"], sub[:1], np.zeros(1, np.uint64), 70000)[0, :, 0]
    got = host.run(engine, C.F["WORD"], sub, ks, calls)[:, 0, 0]
    assert np.array_equal(got, stream[ks.astype(np.int64)])


@pytest.mark.parametrize("engine", ["lfsr113", "threefry2x32", "threefry4x64"])
@pytest.mark.parametrize("lam", [501.0, 750.0, 999.0])
def test_inversion_poisson_between_500_and_1000(host, engine, lam):
    n, calls = 2048, 4
    sub = np.arange(n, dtype=np.uint64)
    off = np.zeros(n, dtype=np.uint64)
    k = host.run(engine, C.F["POISSON"], sub, off, calls, a=lam)[:, :, 0].astype(np.float64).ravel()
    m = k.size
    assert abs(k.mean() - lam) < 5 * np.sqrt(lam / m)
    assert abs(k.var() / lam - 1.0) < 5 * np.sqrt(2.0 / m)
    assert k.min() > lam - 8 * np.sqrt(lam)


@pytest.mark.parametrize("engine,bits", [("sobol32", 32), ("sobol64", 64), ("scrambled_sobol32", 32),
                                         ("scrambled_sobol64", 64)])
def test_sobol_points_are_gray_code_xors(host, engine, bits):
    v32, v64 = C.sobol_vectors(dims=4, seed=2)
    vec = (v32 if bits == 32 else v64).reshape(4, bits).astype(object)
    sub = np.array([0, 1, 2, 3, 5], dtype=np.uint64)
    off = np.array([0, 1, 77, 2**20 + 3, 2**31 - 2], dtype=np.uint64)
    calls = 40
    got = host.run(engine, C.F["WORD"], sub, off, calls, seed=0xABCDEF0123, vectors=(v32, v64))[:, :, 0]
    mask = (1 << bits) - 1
    for t in range(len(sub)):
        dim = int(sub[t]) % 4
        scr = 0
        if engine.startswith("scrambled"):
            scr = (0xABCDEF0123 ^ (int(sub[t]) * 0x9E3779B97F4A7C15)) & mask
            scr &= 0xFFFFFFFF if bits == 32 else mask
        for k in range(calls):
            i = int(off[t]) + k
            g = i ^ (i >> 1)
            d = 0
            for b in range(bits):
                if (g >> b) & 1:
                    d ^= int(vec[dim, b])
            assert int(got[t, k]) == d ^ scr, (engine, t, k)


def test_gpu_driver_exports():
    if not os.path.exists(C.GPU_LIB):
        pytest.skip("tests/devapi/libdevapi_gpu.so not built")
    import ctypes

    lib = ctypes.CDLL(C.GPU_LIB)
    assert hasattr(lib, "devapi_gpu_run")


def test_exported_mtgp32_parameter_sets():
    """mtgp32dc_params_fast_11213 (librocrand.so, declared in rocrand_mtgp32_11213.h): 512
    mtgp32_params_fast_t records of exponent 11213 holding the library's own parameter
    sets, with the float-tempering table 0x3f800000 | tmp_tbl >> 9."""
    import ctypes

    from rocrand_amd import rocrand as R

    L = R.load()
    addr = ctypes.addressof(ctypes.c_char.in_dll(L, "mtgp32dc_params_fast_11213"))
    raw = np.frombuffer(ctypes.string_at(addr, 512 * 236), dtype=np.uint8).reshape(512, 236)
    words = raw[:, :212].copy().view("<u4")
    ours = np.fromfile(os.path.join(os.path.dirname(HERE), "rocrand_amd", "data", "mtgp32_11213.bin"),
                       dtype="<u4").reshape(512, 36)
    assert np.all(words[:, 0] == 11213)
    assert np.array_equal(words[:, 1:4], ours[:, 0:3])          # pos, sh1, sh2
    assert np.array_equal(words[:, 4:20], ours[:, 3:19])        # tbl
    assert np.array_equal(words[:, 20:36], ours[:, 19:35])      # tmp_tbl
    assert np.array_equal(words[:, 36:52], 0x3F800000 | (ours[:, 19:35] >> 9))
    assert np.array_equal(words[:, 52], ours[:, 35])            # mask
    assert np.all(raw[:, 233:] == 0) and np.all(raw[:, 212:233].any(axis=1))


def test_reference_state_kats_host(host):
    """The reference's device-API state checks (test_rocrand_kernel_philox4x32_10.cpp:227-271,
    test_rocrand_kernel_xorwow.cpp:211-250) on the host branch."""
    sub = np.arange(256, dtype=np.uint64)
    w, size = C.state_words(C.HOST_LIB, "devapi_host_state", "philox", C.KAT_SEED, sub,
                            np.full(256, C.PHILOX_KAT_OFFSET, dtype=np.uint64))
    assert size == 64
    C.philox_kat(w)
    w, _ = C.state_words(C.HOST_LIB, "devapi_host_state", "xorwow", C.KAT_SEED, sub, np.full(256, 345678, np.uint64))
    C.xorwow_kat(w, C.KAT_SEED, 345678)


@pytest.mark.parametrize("engine", list(C.STATE_WORDS))
def test_state_layout_and_values_equal_reference(host, engine):
    """The state objects hold the reference's fields in the reference's order and size:
    after rocrand_init the defined words equal the reference state's bytes."""
    if C.reference() is None:
        pytest.skip("oracle/_ref/libref.so not built (needs the reference sources)")
    sub, off = C.thread_params(64, seed=9)
    ours, size = C.state_words(C.HOST_LIB, "devapi_host_state", engine, 0x0123456789ABCDEF, sub, off)
    ref, ref_size = C.state_words(C.REF_LIB, "ref_devapi_state", engine, 0x0123456789ABCDEF, sub, off)
    assert size == ref_size
    cols = C.STATE_WORDS[engine]
    assert np.array_equal(ours[:, cols], ref[:, cols])
