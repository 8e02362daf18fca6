"""GPU parity at the plain bulk path's benchmark configuration: Philox4x32-10,
ThreeFry4x32-20, MRG32k3a, XORWOW and LFSR113 with seed 12345 at offsets 0 and 2^20+7, every
output kind of that configuration (u32, uniform f32/f64, normal f32/f64 at (0, 1),
log-normal f32 at (0, 1), Poisson at lambda 0.5, 10 and 1000), against the oracle; and
single calls of the full 2^28 values for Philox4x32-10 and MRG32k3a, uniform and normal
f32, checked at sampled windows (Philox: 64 windows of 1024 values plus both ends, the
window rule pinned on the CPU in test_serving_host.py) or whole (MRG32k3a).
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
