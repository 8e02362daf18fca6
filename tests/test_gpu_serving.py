"""GPU tests of the continuous-batching server (rocrand_amd/serving.py): a request's output
in a mixed batch equals its output generated alone, and an open-loop run completes every
request."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from rocrand_amd import rocrand as R  # noqa: E402
from rocrand_amd.serving import Server  # noqa: E402
from rocrand_amd.workloads import serving_mix  # noqa: E402


def _small_mix(count, seed):
    reqs = serving_mix(count, 1e5, seed=seed)
    for r in reqs:  # keep the test small: 256 .. ~2^17 values
        r.n = max(r.dims, (r.n >> 9) // r.dims * r.dims)
    return reqs


def test_mixed_batch_equals_solo():
    reqs = _small_mix(60, 5)
    assert {r.rng_type for r in reqs} == {R.PHILOX4_32_10, R.MRG32K3A, R.SOBOL64}
    srv = Server(torch, 1 << 30)
    mixed = [v.cpu().numpy().copy() for v in srv.run_batch(reqs)]
    torch.cuda.synchronize()
    for i, r in enumerate(reqs):
        solo = srv.run_batch([r])[0]
        torch.cuda.synchronize()
        assert np.array_equal(solo.cpu().numpy().view(np.uint8), mixed[i].view(np.uint8)), f"request {i}: {r}"


def test_open_loop_run_completes():
    reqs = _small_mix(300, 6)
    srv = Server(torch, 1 << 28)
    res = srv.serve(reqs)
    assert len(res.latencies_s) + res.rejected == len(reqs)
    assert res.rejected == 0
    assert min(res.latencies_s) >= 0
    assert res.percentile(50) <= res.percentile(99)
    assert res.bytes == sum(r.n * {"f32": 4, "f64": 8, "u32": 4}[r.dtype] for r in reqs)
    assert srv.arena.free == [(0, 1 << 28)]  # every region returned
