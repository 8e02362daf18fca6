"""GPU tests of the batched path (include/rocrand_batched.h): every stream of a batch equals
the plain C-ABI call sequence it stands for (create, set_seed, set_offset, one generate),
bit for bit, for plans replayed several times and for the one-shot call; a sample of
streams also against the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import oracle as O  # noqa: E402
from rocrand_amd import rocrand as R  # noqa: E402
from rocrand_amd.workloads import batch_mix  # noqa: E402

DT = {R.BATCH_UINT: torch.uint32, R.BATCH_UNIFORM_FLOAT: torch.float32, R.BATCH_UNIFORM_DOUBLE: torch.float64,
      R.BATCH_NORMAL_FLOAT: torch.float32, R.BATCH_NORMAL_DOUBLE: torch.float64,
      R.BATCH_LOG_NORMAL_FLOAT: torch.float32, R.BATCH_LOG_NORMAL_DOUBLE: torch.float64,
      R.BATCH_POISSON: torch.uint32}
ORACLE_KIND = {R.BATCH_UINT: "u32", R.BATCH_UNIFORM_FLOAT: "uniform_f32", R.BATCH_UNIFORM_DOUBLE: "uniform_f64",
               R.BATCH_NORMAL_FLOAT: "normal_f32", R.BATCH_NORMAL_DOUBLE: "normal_f64",
               R.BATCH_LOG_NORMAL_FLOAT: "lognormal_f32", R.BATCH_LOG_NORMAL_DOUBLE: "lognormal_f64",
               R.BATCH_POISSON: "poisson"}


def _plain(spec, view):
    g = R.PRNG(spec.rng_type, seed=spec.seed, offset=spec.offset)
    k = spec.kind
    if k == R.BATCH_UINT:
        g.generate(view)
    elif k in (R.BATCH_UNIFORM_FLOAT, R.BATCH_UNIFORM_DOUBLE):
        g.uniform(view)
    elif k in (R.BATCH_NORMAL_FLOAT, R.BATCH_NORMAL_DOUBLE):
        g.normal(view, spec.a, spec.b)
    elif k in (R.BATCH_LOG_NORMAL_FLOAT, R.BATCH_LOG_NORMAL_DOUBLE):
        g.lognormal(view, spec.a, spec.b)
    else:
        g.poisson(view, spec.a)


def _mix(count, seed, hi):
    specs = batch_mix(count, seed=seed, lo=0.0, hi=hi)
    rs = np.random.default_rng(seed)
    # every kind and both engines; shifted (misaligned) outputs, short and ragged lengths
    kinds = list(DT)
    for i, s in enumerate(specs):
        s.kind = kinds[i % len(kinds)]
        if s.kind in (R.BATCH_NORMAL_FLOAT, R.BATCH_NORMAL_DOUBLE, R.BATCH_LOG_NORMAL_FLOAT,
                      R.BATCH_LOG_NORMAL_DOUBLE):
            s.a, s.b = 0.25, 1.5
        if s.kind == R.BATCH_POISSON:
            s.a = [0.5, 4.0, 1000.0, 5000.0][i % 4]
    shifts = rs.integers(0, 4, count)
    return specs, shifts


def _buffers(specs, shifts):
    bufs, views = [], []
    for s, sh in zip(specs, shifts):
        b = torch.zeros(s.n + 8, dtype=DT[s.kind], device="cuda")
        bufs.append(b)
        views.append(b[int(sh):int(sh) + s.n])
    return bufs, views


def _streams(specs, views):
    return [R.BatchStream(s.rng_type, s.kind, s.seed, s.offset, s.n, s.a, s.b, v.data_ptr())
            for s, v in zip(specs, views)]


@pytest.mark.parametrize("seed", [1, 2])
def test_batched_equals_plain_calls(seed):
    specs, shifts = _mix(96, seed, 5.3)  # lengths 1 .. 2e5
    print(f"case batch seed={seed} streams={len(specs)} total={sum(s.n for s in specs)}", flush=True)
    _, views = _buffers(specs, shifts)
    _, pviews = _buffers(specs, shifts)
    plan = R.BatchPlan(_streams(specs, views))
    for rep in range(3):  # replays write the same values
        for v in views:
            v.zero_()
        plan.launch()
        torch.cuda.synchronize()
        if rep == 0:
            for s, pv in zip(specs, pviews):
                _plain(s, pv)
            torch.cuda.synchronize()
        for i, (v, pv) in enumerate(zip(views, pviews)):
            a = v.cpu().numpy().view(np.uint8)
            b = pv.cpu().numpy().view(np.uint8)
            assert np.array_equal(a, b), f"stream {i} ({specs[i]}) differs from its plain call on replay {rep}"
    plan.close()


def test_one_shot_equals_plan():
    specs, shifts = _mix(40, 3, 4.5)
    _, views = _buffers(specs, shifts)
    _, oviews = _buffers(specs, shifts)
    plan = R.BatchPlan(_streams(specs, views))
    plan.launch()
    R.generate_batched(_streams(specs, oviews))
    torch.cuda.synchronize()
    for v, o in zip(views, oviews):
        assert np.array_equal(v.cpu().numpy().view(np.uint8), o.cpu().numpy().view(np.uint8))
    plan.close()


def test_one_shot_async_repeated():
    """Back-to-back one-shot calls without host waits (their descriptor blocks stay in use
    until each call completes, then are recycled): every call's outputs equal the plain
    calls, also for calls issued after the pool has recycled blocks."""
    stream = torch.cuda.Stream()
    cases = []
    for seed in (5, 6, 7):
        specs, shifts = _mix(24, seed, 4.0 + 0.3 * seed)
        _, pviews = _buffers(specs, shifts)
        for s, pv in zip(specs, pviews):
            _plain(s, pv)
        cases.append((specs, shifts, pviews))
    torch.cuda.synchronize()
    for rnd in range(2):
        outs = []
        with torch.cuda.stream(stream):
            for rep in range(8):
                specs, shifts, _ = cases[rep % 3]
                _, views = _buffers(specs, shifts)
                R.generate_batched(_streams(specs, views), stream)
                outs.append((rep % 3, views))
        stream.synchronize()
        for ci, views in outs:
            for i, (v, pv) in enumerate(zip(views, cases[ci][2])):
                assert np.array_equal(v.cpu().numpy().view(np.uint8), pv.cpu().numpy().view(np.uint8)), \
                    f"round {rnd} case {ci} stream {i}"


def test_batched_against_oracle():
    specs, shifts = _mix(16, 4, 4.2)
    _, views = _buffers(specs, shifts)
    R.generate_batched(_streams(specs, views))
    torch.cuda.synchronize()
    from parity_checks import check

    for s, v in zip(specs, views):
        og = O.OracleGen(s.rng_type, seed=s.seed, offset=s.offset)
        st, want = og.generate(ORACLE_KIND[s.kind], s.n, v.data_ptr(), s.a, s.b)
        assert st == 0
        got = v.cpu().numpy()
        kind = ORACLE_KIND[s.kind]
        if kind == "poisson" and s.a > 4000.0:
            assert np.max(np.abs(got.astype(np.int64) - want.astype(np.int64)), initial=0) <= 1
        else:
            check(kind, got, want, (s.a, s.b))


def test_batch_status_codes():
    buf = torch.zeros(16, dtype=torch.float32, device="cuda")
    bad_engine = [R.BatchStream(R.XORWOW, R.BATCH_UNIFORM_FLOAT, 1, 0, 16, 0.0, 0.0, buf.data_ptr())]
    with pytest.raises(R.RocRandError) as e:
        R.generate_batched(bad_engine)
    assert e.value.value == 103
    bad_lambda = [R.BatchStream(R.PHILOX4_32_10, R.BATCH_POISSON, 1, 0, 16, 0.0, 0.0, buf.data_ptr())]
    with pytest.raises(R.RocRandError) as e:
        R.generate_batched(bad_lambda)
    assert e.value.value == 104
