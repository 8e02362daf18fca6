"""GPU parity of the Sobol32/64 (plain, scrambled) and MTGP32 generators against the
oracle, through the C ABI on cuda:0.

Bar: tests/parity_checks.py.  Sobol normals go through the reference's inverse error
function (rocrand_normal.h:174-296), judged with the bar for that formula; MTGP32 uses
the Box-Muller functors of the pseudo-random engines.
"""
import numpy as np
import pytest

from parity_checks import check
from test_gpu_parity import KINDS, _check

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import oracle as O  # noqa: E402
from rocrand_amd import rocrand as R  # noqa: E402

SOBOL = {"sobol32": R.SOBOL32, "scrambled_sobol32": R.SCRAMBLED_SOBOL32,
         "sobol64": R.SOBOL64, "scrambled_sobol64": R.SCRAMBLED_SOBOL64}

# (dims, offset, points per dimension, element shift of the output pointer)
SOBOL_CONFIGS = [(1, 0, 1000, 0), (3, 1000, 1001, 1), (7, 5, 37, 3), (64, 123457, 4099, 2),
                 (1, (1 << 31) - 100, 1000, 1)]


def _sobol_gpu(rng, kind, dims, offset, npd, shift, calls=2, args=()):
    dt, meth, a = KINDS[kind]
    if args:
        a = args
    g = R.QRNG(rng, ndim=dims, offset=offset)
    n = dims * npd
    out = []
    for _ in range(calls):
        buf = torch.zeros(n + 16, dtype=dt, device="cuda")
        view = buf[shift:shift + n]
        getattr(g, meth)(view, *a)
        torch.cuda.synchronize()
        out.append((view.cpu().numpy().copy(), view.data_ptr()))
    return out


def _sobol_cases():
    for name, rng in SOBOL.items():
        for kind in KINDS:
            if kind == "u64" and "64" not in name:
                continue
            yield name, kind


@pytest.mark.parametrize("name,kind", list(_sobol_cases()))
def test_sobol_parity(name, kind):
    rng = SOBOL[name]
    a = KINDS[kind][2]
    a0 = a[0] if len(a) > 0 else 0.0
    a1 = a[1] if len(a) > 1 else 0.0
    for dims, offset, npd, shift in SOBOL_CONFIGS:
        print(f"case {name} {kind} dims={dims} offset={offset} points={npd} shift={shift}", flush=True)
        got = _sobol_gpu(rng, kind, dims, offset, npd, shift)
        og = O.OracleGen(rng, dims=dims, offset=offset)
        for g, addr in got:
            st, want = og.generate(kind, dims * npd, addr, a0, a1)
            assert st == 0
            if "normal" in kind:
                check(kind, g, want, (a0, a1), erfinv=True)
            else:
                _check(kind, g, want, (a0, a1))


@pytest.mark.parametrize("name", list(SOBOL))
@pytest.mark.parametrize("lam", [0.5, 4.0, 1000.0, 5000.0])
def test_sobol_poisson_cdf(name, lam):
    rng = SOBOL[name]
    got = _sobol_gpu(rng, "poisson", 4, 77, 5003, 1, args=(lam,))
    og = O.OracleGen(rng, dims=4, offset=77)
    for g, addr in got:
        st, want = og.generate("poisson", 4 * 5003, addr, lam, 0.0)
        assert st == 0
        if lam > 4000.0:
            assert np.max(np.abs(g.astype(np.int64) - want.astype(np.int64))) <= 1
        else:
            assert np.array_equal(g, want)


def test_sobol_length_not_multiple_and_type_error():
    g = R.QRNG(R.SOBOL32, ndim=3)
    buf = torch.zeros(10, dtype=torch.uint32, device="cuda")
    with pytest.raises(R.RocRandError) as e:
        g.generate(buf)
    assert e.value.value == 105
    b64 = torch.zeros(9, dtype=torch.uint64, device="cuda")
    with pytest.raises(R.RocRandError) as e:
        g.generate(b64)
    assert e.value.value == 103


@pytest.mark.parametrize("name,dims", [("sobol32", 1000), ("scrambled_sobol64", 16)])
def test_sobol_full_size_sampled(name, dims):
    """dims x 2^20 points (the quasi-random bench shape): sampled positions against the
    closed form x = XOR of the oracle's direction vectors over Gray(offset + p)."""
    rng = SOBOL[name]
    is64 = "64" in name
    npd, offset = 1 << 20, 1000
    dt = torch.uint64 if is64 else torch.uint32
    out = torch.empty(dims * npd, dtype=dt, device="cuda")
    g = R.QRNG(rng, ndim=dims, offset=offset)
    g.generate(out)
    torch.cuda.synchronize()
    v32, sv32, v64, sv64, _, _ = O.sobol_tables()
    vec = (sv64 if "scrambled" in name else v64) if is64 else (sv32 if "scrambled" in name else v32)
    B = 64 if is64 else 32
    vec = vec.reshape(20000, B)
    rs = np.random.default_rng(3)
    d = rs.integers(0, dims, 1 << 15)
    p = rs.integers(0, npd, 1 << 15)
    p[:dims] = 0
    d[:dims] = np.arange(dims)
    i = (p + offset).astype(np.uint64)
    gray = i ^ (i >> np.uint64(1))
    want = np.zeros(p.size, dtype=np.uint64)
    for k in range(32):
        bit = ((gray >> np.uint64(k)) & np.uint64(1)).astype(bool)
        want ^= np.where(bit, vec[d, k].astype(np.uint64), np.uint64(0))
    if "scrambled" in name:
        want ^= vec.reshape(-1)[d].astype(np.uint64)
    idx = torch.from_numpy((d * npd + p).astype(np.int64)).cuda()
    signed = out.view(torch.int64 if is64 else torch.int32)
    got = signed[idx].cpu().numpy().view(np.uint64 if is64 else np.uint32).astype(np.uint64)
    assert np.array_equal(got, want)


# ------------------------------------------------------------------ MTGP32
MTGP_CONFIGS = [(12345, 1000, 0), (0, 131072 + 777, 1), (0xFEDCBA9876543210, 13, 3), (7, 262144 * 2 + 5, 2)]


def _mtgp_cases():
    for kind in KINDS:
        if kind != "u64":
            yield kind


@pytest.mark.parametrize("kind", list(_mtgp_cases()))
def test_mtgp32_parity(kind):
    dt, meth, a = KINDS[kind]
    a0 = a[0] if len(a) > 0 else 0.0
    a1 = a[1] if len(a) > 1 else 0.0
    for seed, n, shift in MTGP_CONFIGS:
        print(f"case mtgp32 {kind} seed={seed:#x} n={n} shift={shift}", flush=True)
        g = R.PRNG(R.MTGP32, seed=seed)
        og = O.OracleGen(O.MTGP32, seed=seed)
        for _ in range(2):
            buf = torch.zeros(n + 16, dtype=dt, device="cuda")
            view = buf[shift:shift + n]
            getattr(g, meth)(view, *a)
            torch.cuda.synchronize()
            st, want = og.generate(kind, n, view.data_ptr(), a0, a1)
            assert st == 0
            _check(kind, view.cpu().numpy(), want, (a0, a1))


def test_mtgp32_large():
    """2^26 u32 then 2^26 uniform floats on one MTGP32 generator, bit-exact."""
    n = 1 << 26
    g = R.PRNG(R.MTGP32, seed=12345)
    a = torch.empty(n, dtype=torch.uint32, device="cuda")
    b = torch.empty(n, dtype=torch.float32, device="cuda")
    g.generate(a)
    g.uniform(b)
    torch.cuda.synchronize()
    og = O.OracleGen(O.MTGP32, seed=12345)
    st, wa = og.generate("u32", n, a.data_ptr())
    st2, wb = og.generate("uniform_f32", n, b.data_ptr())
    assert st == st2 == 0
    assert np.array_equal(a.cpu().numpy(), wa)
    assert np.array_equal(b.cpu().numpy(), wb)


def test_mtgp32_set_offset_is_type_error():
    g = R.PRNG(R.MTGP32)
    with pytest.raises(R.RocRandError) as e:
        g.offset = 5
    assert e.value.value == 103


# ------------------------------------------------------------------ MT19937
MT_CONFIGS = [(12345, 1000, 0), (0, 65536 + 777, 1), (0xFEDCBA9876543210, 13, 3), (7, 65536 * 80 + 5, 2)]


@pytest.mark.parametrize("kind", list(_mtgp_cases()))
def test_mt19937_parity(kind):
    dt, meth, a = KINDS[kind]
    a0 = a[0] if len(a) > 0 else 0.0
    a1 = a[1] if len(a) > 1 else 0.0
    for seed, n, shift in MT_CONFIGS:
        if n > 100000 and kind not in ("u32", "uniform_f32", "normal_f32", "uniform_f64", "normal_f64", "u8"):
            continue
        print(f"case mt19937 {kind} seed={seed:#x} n={n} shift={shift}", flush=True)
        g = R.PRNG(R.MT19937, seed=seed)
        og = O.OracleGen(O.MT19937, seed=seed)
        for _ in range(2):
            buf = torch.zeros(n + 16, dtype=dt, device="cuda")
            view = buf[shift:shift + n]
            getattr(g, meth)(view, *a)
            torch.cuda.synchronize()
            st, want = og.generate(kind, n, view.data_ptr(), a0, a1)
            assert st == 0
            _check(kind, view.cpu().numpy(), want, (a0, a1))


def test_mt19937_mixed_widths_and_rounds():
    """One generator through changing input widths (1, 2, 4 words per input) and round
    boundaries, misaligned; every call bit-exact (or within the normal bar)."""
    g = R.PRNG(R.MT19937, seed=99)
    og = O.OracleGen(O.MT19937, seed=99)
    seq = [("u32", 65536 * 77 + 3, 1), ("uniform_f64", 1001, 1), ("normal_f64", 65536 * 20 + 7, 1),
           ("u8", 4099, 3), ("uniform_f32", 65536 * 78, 0), ("normal_f32", 5, 1), ("u16", 65536 * 39 * 2 + 1, 1)]
    for kind, n, shift in seq:
        dt, meth, a = KINDS[kind]
        print(f"case mt19937 sequence {kind} n={n} shift={shift}", flush=True)
        buf = torch.zeros(n + 16, dtype=dt, device="cuda")
        view = buf[shift:shift + n]
        getattr(g, meth)(view, *a)
        torch.cuda.synchronize()
        a0 = a[0] if len(a) > 0 else 0.0
        a1 = a[1] if len(a) > 1 else 0.0
        st, want = og.generate(kind, n, view.data_ptr(), a0, a1)
        assert st == 0
        _check(kind, view.cpu().numpy(), want, (a0, a1))


def test_mt19937_large():
    """2^26 u32 then 2^26 uniform floats on one MT19937 generator, bit-exact."""
    n = 1 << 26
    g = R.PRNG(R.MT19937, seed=12345)
    a = torch.empty(n, dtype=torch.uint32, device="cuda")
    b = torch.empty(n, dtype=torch.float32, device="cuda")
    g.generate(a)
    g.uniform(b)
    torch.cuda.synchronize()
    og = O.OracleGen(O.MT19937, seed=12345)
    st, wa = og.generate("u32", n, a.data_ptr())
    st2, wb = og.generate("uniform_f32", n, b.data_ptr())
    assert st == st2 == 0
    assert np.array_equal(a.cpu().numpy(), wa)
    assert np.array_equal(b.cpu().numpy(), wb)


def test_mt19937_host_decided_errors():
    g = R.PRNG(R.MT19937)
    with pytest.raises(R.RocRandError) as e:
        g.offset = 5
    assert e.value.value == 103
    with pytest.raises(R.RocRandError) as e:
        g.set_ordering(R.ORDERING_PSEUDO_DYNAMIC)
    assert e.value.value == 104
