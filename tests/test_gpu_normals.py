"""GPU normals against an exact Box-Muller, the plain bulk configuration at seed 12345 and
offset 2^20 + 7, and single calls of 2^28 samples.

* Error report: for each engine, normal fp32 and fp64 outputs of the device (seed 12345,
  offset 0, 2^20 values) against the Box-Muller of the same words evaluated in float64
  (fp32 outputs) or long double (fp64 outputs): max and mean error in ulps, and the ulp
  distance to the oracle (the reference's host path over this libm).  The words come from
  the same generator's u32 (or, for the MRGs, uniform fp32) output, mapped to the normal
  transform's inputs by the reference's index map (counter engines: consecutive words;
  interleaved engines: vec v from engine v mod S, words in_w q .. in_w q + in_w - 1 of it,
  q = v div S).
* The plain bulk path at the second offset of the per-engine configuration (seed 12345,
  offset 1048583) for Philox4x32-10, ThreeFry4x32-20, MRG32k3a, XORWOW and LFSR113 and
  every kind, two calls in a row, against the oracle.
* 2^28 samples in one call (Philox4x32-10 and MRG32k3a, uniform and normal fp32) checked
  at 2^16 sampled positions plus both ends: words from the oracle's engine streams,
  uniforms bit-exact, normals within 0.51 ulp of the float64 Box-Muller of those words
  and, for Philox, within 2 ulp of the oracle's own normal at the same position.
"""
import numpy as np
import pytest

from conftest import ulp_distance
from parity_checks import check

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import oracle as O  # noqa: E402
from rocrand_amd import rocrand as R  # noqa: E402

S = 131072  # engines of the interleaved generators (reference grid 512 x 256)
F32_2POW32 = np.float32(2.3283064e-10)
F32_2POW32_2PI = np.float32(1.46291807e-09)
F32_2PI = np.float32(6.2831855)
MRG_NORM = {R.MRG32K3A: 2.3283065498378288e-10, R.MRG31K3P: 4.656612875245796923e-10}

ENGINES = [("philox", R.PHILOX4_32_10, "ctr"), ("threefry4x32", R.THREEFRY4_32_20, "ctr"),
           ("mrg32k3a", R.MRG32K3A, "mrg"), ("mrg31k3p", R.MRG31K3P, "mrg"),
           ("xorwow", R.XORWOW, "inter"), ("lfsr113", R.LFSR113, "inter")]


def _device(rng, kind, n, seed=12345, offset=0, args=()):
    dt = {"u32": torch.uint32, "uniform_f32": torch.float32, "normal_f32": torch.float32,
          "normal_f64": torch.float64}[kind]
    g = R.PRNG(rng, seed=seed, offset=offset)
    buf = torch.empty(n, dtype=dt, device="cuda")
    call = {"u32": "generate", "uniform_f32": "uniform", "normal_f32": "normal", "normal_f64": "normal"}[kind]
    getattr(g, call)(buf, *args)
    torch.cuda.synchronize()
    return buf.cpu().numpy()


def _input_words(family, n_vecs, in_w, words):
    """[n_vecs, in_w] engine words feeding vec v, from the u32 (or uniform) output `words`
    of a fresh generator at offset 0."""
    v = np.arange(n_vecs)
    if family == "ctr":
        return words[: n_vecs * in_w].reshape(n_vecs, in_w)
    e, q = v % S, v // S
    idx = e[:, None] + (in_w * q[:, None] + np.arange(in_w)[None, :]) * S
    return words[idx]


def _mrg_f32(raw, rng):
    return (raw.astype(np.float64) * MRG_NORM[rng]).astype(np.float32)


def _exact_f32_pairs(u, v):
    """float64 Box-Muller of float inputs u, v: (sin v, cos v) * sqrt(-2 ln u)."""
    s = np.sqrt(-2.0 * np.log(u.astype(np.float64)))
    out = np.empty((u.size, 2))
    out[:, 0] = np.sin(v.astype(np.float64)) * s
    out[:, 1] = np.cos(v.astype(np.float64)) * s
    return out.ravel()


def _ulp_error(got, exact, ftype):
    """|got - exact| in ulps of the float nearest to exact."""
    nearest = exact.astype(ftype)
    ulp = np.spacing(np.abs(nearest)).astype(exact.dtype)
    return (np.abs(got.astype(exact.dtype) - exact) / ulp).astype(np.float64)


@pytest.mark.parametrize("name,rng,family", ENGINES)
def test_normals_error_against_exact_box_muller(name, rng, family):
    n = 1 << 20
    seed = 12345
    report = []
    # fp32: two words per pair
    if family == "mrg":
        u01 = _device(rng, "uniform_f32", 2 * n, seed)
        w = _input_words(family, n // 2, 2, u01)
        u, v = w[:, 0], w[:, 1] * F32_2PI
    else:
        raw = _device(rng, "u32", 2 * n, seed)
        w = _input_words(family, n // 2, 2, raw)
        u = F32_2POW32 + w[:, 0].astype(np.float32) * F32_2POW32
        v = F32_2POW32_2PI + w[:, 1].astype(np.float32) * F32_2POW32_2PI
    got = _device(rng, "normal_f32", n, seed, 0, (0.0, 1.0))
    err = _ulp_error(got, _exact_f32_pairs(u, v), np.float32)
    _, want = O.OracleGen(rng, seed=seed).generate("normal_f32", n, 0x1000, 0.0, 1.0)
    d = ulp_distance(got, want)
    report.append(("f32", err.max(), err.mean(), d.max(), int(np.count_nonzero(d))))
    assert err.max() <= 0.51, f"{name} normal f32 {err.max():.3f} ulp from the float64 Box-Muller"
    check("normal_f32", got, want, (0.0, 1.0))

    # fp64: four words per pair (two MRG values)
    nd = 1 << 18
    ld = np.longdouble
    if family == "mrg":
        u01 = _device(rng, "uniform_f32", 4 * nd, seed)
        w = _input_words(family, nd // 2, 2, u01)
        ud, t = w[:, 0].astype(ld), (w[:, 1].astype(np.float64) * 2.0 * 3.1415926535897932).astype(ld)
    else:
        raw = _device(rng, "u32", 4 * nd, seed).astype(np.uint64)
        w = _input_words(family, nd // 2, 4, raw)
        v1 = w[:, 0] ^ (w[:, 1] << np.uint64(21))
        v2 = w[:, 2] ^ (w[:, 3] << np.uint64(21))
        ud = (2.0**-53 + v1.astype(np.float64) * 2.0**-53).astype(ld)
        wd = 2.0**-52 + v2.astype(np.float64) * 2.0**-52
        t = (wd * 3.1415926535897932).astype(ld)  # the reference rounds w * pi to double
    s = np.sqrt(-2 * np.log(ud))
    exact = np.empty(ud.size * 2, dtype=ld)
    exact[0::2], exact[1::2] = np.sin(t) * s, np.cos(t) * s
    got = _device(rng, "normal_f64", nd, seed, 0, (0.0, 1.0))
    err = _ulp_error(got, exact, np.float64)
    _, want = O.OracleGen(rng, seed=seed).generate("normal_f64", nd, 0x1000, 0.0, 1.0)
    d = ulp_distance(got, want)
    report.append(("f64", err.max(), err.mean(), d.max(), int(np.count_nonzero(d))))
    for kind, emax, emean, dmax, ndiff in report:
        print(f"normal {name} {kind}: max {emax:.3f} ulp, mean {emean:.4f} ulp from the exact Box-Muller; "
              f"oracle distance max {int(dmax)} ulp, {ndiff} values differ", flush=True)
    check("normal_f64", got, want, (0.0, 1.0))


CONFIG2 = [("philox", R.PHILOX4_32_10), ("threefry4x32", R.THREEFRY4_32_20), ("mrg32k3a", R.MRG32K3A),
           ("xorwow", R.XORWOW), ("lfsr113", R.LFSR113)]
CONFIG2_KINDS = [("u32", torch.uint32, "generate", ()), ("uniform_f32", torch.float32, "uniform", ()),
                 ("uniform_f64", torch.float64, "uniform", ()),
                 ("normal_f32", torch.float32, "normal", (0.0, 1.0)),
                 ("normal_f64", torch.float64, "normal", (0.0, 1.0)),
                 ("lognormal_f32", torch.float32, "lognormal", (0.0, 1.0)),
                 ("poisson", torch.uint32, "poisson", (0.5,)), ("poisson", torch.uint32, "poisson", (10.0,)),
                 ("poisson", torch.uint32, "poisson", (1000.0,))]


@pytest.mark.parametrize("name,rng", CONFIG2)
@pytest.mark.parametrize("kind,dt,call,args", CONFIG2_KINDS)
def test_plain_path_seed_12345_offset_2e20_plus_7(name, rng, kind, dt, call, args):
    seed, offset, n = 12345, (1 << 20) + 7, (1 << 20) + 13
    g = R.PRNG(rng, seed=seed, offset=offset)
    og = O.OracleGen(rng, seed=seed, offset=offset)
    a, b = (args[0], 0.0) if kind == "poisson" else (args if args else (0.0, 0.0))
    for _ in range(2):
        buf = torch.empty(n, dtype=dt, device="cuda")
        getattr(g, call)(buf, *args)
        torch.cuda.synchronize()
        got = buf.cpu().numpy()
        st, want = og.generate(kind, n, buf.data_ptr(), a, b)
        assert st == 0
        check(kind, got, want, args if "normal" in kind else (0.0, 1.0))


SAMPLES = 1 << 16


def _sample_positions(n, rng_seed, align):
    r = np.random.default_rng(rng_seed)
    p = np.concatenate([np.arange(0, 4096), n - 4096 + np.arange(4096), r.integers(0, n, SAMPLES)])
    return np.unique(p // align * align)


@pytest.mark.parametrize("kind", ["uniform_f32", "normal_f32"])
def test_philox_2e28_single_call(kind):
    n, seed = 1 << 28, 12345
    out = _device(R.PHILOX4_32_10, kind, n, seed, 0, (0.0, 1.0) if kind == "normal_f32" else ())
    pos = _sample_positions(n, 1, 2)  # pairs start at even positions
    got = np.stack([out[pos], out[pos + 1]], axis=1).ravel()
    words = np.stack([O.engine_stream(O.PHILOX4_32_10, seed, 0, int(p), 2) for p in pos])
    if kind == "uniform_f32":
        want = (F32_2POW32 + words.astype(np.float32) * F32_2POW32).ravel()
        check("uniform_f32", got, want)
        return
    u = F32_2POW32 + words[:, 0].astype(np.float32) * F32_2POW32
    v = F32_2POW32_2PI + words[:, 1].astype(np.float32) * F32_2POW32_2PI
    err = _ulp_error(got, _exact_f32_pairs(u, v), np.float32)
    assert err.max() <= 0.51
    want = np.concatenate([O.OracleGen(O.PHILOX4_32_10, seed=seed, offset=int(p)).generate(
        "normal_f32", 2, 0x1000, 0.0, 1.0)[1] for p in pos[:4096]])
    check("normal_f32", got[: want.size], want, (0.0, 1.0))
    print(f"philox normal f32 2^28: {pos.size} pairs, max {err.max():.3f} ulp from the float64 Box-Muller",
          flush=True)


@pytest.mark.parametrize("kind", ["uniform_f32", "normal_f32"])
def test_mrg32k3a_2e28_single_call(kind):
    n, seed = 1 << 28, 12345
    out = _device(R.MRG32K3A, kind, n, seed, 0, (0.0, 1.0) if kind == "normal_f32" else ())
    if kind == "uniform_f32":
        pos = _sample_positions(n, 2, 1)
        raw = np.array([O.engine_stream(O.MRG32K3A, seed, int(p % S), int(p // S), 1)[0] for p in pos])
        check("uniform_f32", out[pos], _mrg_f32(raw, R.MRG32K3A))
        return
    vecs = _sample_positions(n // 2, 3, 1)  # vec v: engine v mod S, words 2q, 2q + 1
    raw = np.stack([O.engine_stream(O.MRG32K3A, seed, int(v % S), int(2 * (v // S)), 2) for v in vecs])
    u = _mrg_f32(raw[:, 0], R.MRG32K3A)
    a = _mrg_f32(raw[:, 1], R.MRG32K3A) * F32_2PI
    got = np.stack([out[2 * vecs], out[2 * vecs + 1]], axis=1).ravel()
    err = _ulp_error(got, _exact_f32_pairs(u, a), np.float32)
    assert err.max() <= 0.51
    print(f"mrg32k3a normal f32 2^28: {vecs.size} pairs, max {err.max():.3f} ulp from the float64 Box-Muller",
          flush=True)
