"""GPU parity at the plain bulk path's benchmark configuration: Philox4x32-10,
ThreeFry4x32-20, MRG32k3a, XORWOW and LFSR113 with seed 12345 at offsets 0 and 2^20+7, every
output kind of that configuration (u32, uniform f32/f64, normal f32/f64 at (0, 1),
log-normal f32 at (0, 1), Poisson at lambda 0.5, 10 and 1000), against the oracle; and
single calls of the full 2^28 values for Philox4x32-10 and MRG32k3a, uniform and normal
f32, checked at sampled windows (Philox: 64 windows of 1024 values plus both ends, the
window rule pinned on the CPU in test_serving_host.py) or whole (MRG32k3a); and the
state-heavy and quasi-random configuration below.
"""
import numpy as np
import pytest

from parity_checks import check

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import oracle as O  # noqa: E402
from rocrand_amd import rocrand as R  # noqa: E402

ENGINES = {"philox": R.PHILOX4_32_10, "threefry4x32": R.THREEFRY4_32_20, "mrg32k3a": R.MRG32K3A,
           "xorwow": R.XORWOW, "lfsr113": R.LFSR113}
SEED = 12345
OFFSETS = [0, (1 << 20) + 7]
# kind, torch dtype, wrapper method, parameters
KINDS = [("u32", torch.uint32, "generate", ()), ("uniform_f32", torch.float32, "uniform", ()),
         ("uniform_f64", torch.float64, "uniform", ()), ("normal_f32", torch.float32, "normal", (0.0, 1.0)),
         ("normal_f64", torch.float64, "normal", (0.0, 1.0)),
         ("lognormal_f32", torch.float32, "lognormal", (0.0, 1.0)),
         ("poisson", torch.uint32, "poisson", (0.5,)), ("poisson", torch.uint32, "poisson", (10.0,)),
         ("poisson", torch.uint32, "poisson", (1000.0,))]


@pytest.mark.parametrize("offset", OFFSETS)
@pytest.mark.parametrize("ename", list(ENGINES))
def test_benchmark_configuration(ename, offset):
    """2^20 + 13 values per call (past one sweep of the reference grid for the counter
    engines, several steps per engine for the interleaved ones), into a pointer one element
    off alignment, two calls per generator."""
    rng, n, shift = ENGINES[ename], (1 << 20) + 13, 1
    for kind, dt, meth, args in KINDS:
        print(f"case {ename} {kind} {args} seed={SEED} offset={offset} n={n}", flush=True)
        g = R.PRNG(rng, seed=SEED, offset=offset)
        og = O.OracleGen(rng, seed=SEED, offset=offset)
        a0 = args[0] if args else 0.0
        a1 = args[1] if len(args) > 1 else 0.0
        buf = torch.zeros(n + 16, dtype=dt, device="cuda")
        view = buf[shift:shift + n]
        for _ in range(2):
            getattr(g, meth)(view, *args)
            torch.cuda.synchronize()
            st, want = og.generate(kind, n, view.data_ptr(), a0, a1)
            assert st == 0
            check(kind, view.cpu().numpy(), want, (a0, a1))


FULL = 1 << 28


@pytest.mark.parametrize("kind", ["uniform_f32", "normal_f32"])
def test_philox_full_size_windows(kind):
    offset = (1 << 20) + 7
    g = R.PRNG(R.PHILOX4_32_10, seed=SEED, offset=offset)
    buf = torch.empty(FULL, dtype=torch.float32, device="cuda")
    if kind == "uniform_f32":
        g.uniform(buf)
    else:
        g.normal(buf, 0.0, 1.0)
    torch.cuda.synchronize()
    assert buf.data_ptr() % 256 == 0
    rs = np.random.default_rng(17)
    k = 1024
    starts = sorted({0, FULL - k, *(int(x) * 4 for x in rs.integers(0, (FULL - k) // 4, 64))})
    total = 0
    for i0 in starts:
        got = buf[i0:i0 + k].cpu().numpy()
        og = O.OracleGen(R.PHILOX4_32_10, seed=SEED, offset=offset + i0)
        st, want = og.generate(kind, k, 0x10000000, 0.0, 1.0)
        assert st == 0
        check(kind, got, want, (0.0, 1.0))
        total += k
    assert total >= 1 << 16
    print(f"Philox {kind} 2^28: {len(starts)} windows, {total} values checked", flush=True)


@pytest.mark.parametrize("kind", ["uniform_f32", "normal_f32"])
def test_mrg32k3a_full_size_whole(kind):
    g = R.PRNG(R.MRG32K3A, seed=SEED)
    buf = torch.empty(FULL, dtype=torch.float32, device="cuda")
    if kind == "uniform_f32":
        g.uniform(buf)
    else:
        g.normal(buf, 0.0, 1.0)
    torch.cuda.synchronize()
    got = buf.cpu().numpy()
    del buf
    st, want = O.OracleGen(R.MRG32K3A, seed=SEED).generate(kind, FULL, 0x10000000, 0.0, 1.0)
    assert st == 0
    check(kind, got, want, (0.0, 1.0))


# The state-heavy and quasi-random engines' configuration: MT19937 and MTGP32 at 2^28
# values (whole outputs against the oracle), Sobol32/64 and their scrambled variants with
# 1, 64 and 1000 dimensions of 2^20 points at offsets 0 and 1000 (windows of every
# checked dimension; the dimension-major layout is value d * 2^20 + j for point j).
@pytest.mark.parametrize("kind", ["u32", "uniform_f32"])
@pytest.mark.parametrize("name,rng", [("mt19937", R.MT19937), ("mtgp32", R.MTGP32)])
def test_state_heavy_full_size(name, rng, kind):
    g = R.PRNG(rng, seed=SEED)
    dt = torch.uint32 if kind == "u32" else torch.float32
    buf = torch.empty(FULL, dtype=dt, device="cuda")
    (g.generate if kind == "u32" else g.uniform)(buf)
    torch.cuda.synchronize()
    got = buf.cpu().numpy()
    del buf
    print(f"case {name} {kind} 2^28", flush=True)
    st, want = O.OracleGen(rng, seed=SEED).generate(kind, FULL, 0x10000000)
    assert st == 0
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


SOBOL = {"sobol32": R.SOBOL32, "scrambled_sobol32": R.SCRAMBLED_SOBOL32, "sobol64": R.SOBOL64,
         "scrambled_sobol64": R.SCRAMBLED_SOBOL64}
NPD = 1 << 20


@pytest.mark.parametrize("offset", [0, 1000])
@pytest.mark.parametrize("dims", [1, 64, 1000])
@pytest.mark.parametrize("name", list(SOBOL))
def test_sobol_benchmark_configuration(name, dims, offset):
    rng = SOBOL[name]
    g = R.QRNG(rng, ndim=dims, offset=offset)
    buf = torch.empty(dims * NPD, dtype=torch.uint32, device="cuda")
    g.generate(buf)
    torch.cuda.synchronize()
    k = 1024
    checked_dims = sorted({0, 1 % dims, dims // 2, dims - 1})
    print(f"case {name} dims={dims} offset={offset}", flush=True)
    for j0 in (0, NPD // 2 + 1000, NPD - k):
        og = O.OracleGen(rng, dims=dims, offset=offset + j0)
        st, want = og.generate("u32", dims * k, 0x10000000)
        assert st == 0
        want = want.reshape(dims, k)
        for d in checked_dims:
            got = buf[d * NPD + j0:d * NPD + j0 + k].cpu().numpy()
            assert np.array_equal(got, want[d]), f"dimension {d} points {j0}.."
    del buf
