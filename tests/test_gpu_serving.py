"""GPU tests of the request server (include/rocrand_serving.h, rocrand_amd/csrc/rr_serving.cpp)
through its C ABI: full-size serving mixes against the oracle, sessions that continue a
generator across requests (also across streams), the arena's admission and release, the
status codes, and an open-loop run."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import oracle as O  # noqa: E402
from rocrand_amd import rocrand as R  # noqa: E402
from rocrand_amd import serving as S  # noqa: E402
from rocrand_amd.workloads import Request, serving_mix  # noqa: E402
from serving_checks import ORACLE_KIND, check_request  # noqa: E402

TYPE_ERROR, OUT_OF_RANGE, LENGTH_NOT_MULTIPLE = 103, 104, 105
S_ADDR = 0x10000000  # the oracle's output address: server outputs are 256-byte aligned


def test_full_size_mix_against_oracle():
    """56 requests of the serving mix at full size (256 .. 2^26 values: 2.1 GB of output)
    through a 1 GiB server, so that admission waits for memory; every output is checked
    against the oracle while it is valid (whole outputs for MRG32k3a, windows for
    Philox4x32-10 and Sobol64)."""
    reqs = serving_mix(56, 1e4, seed=25)
    assert {(r.rng_type, r.dtype) for r in reqs} == {(t, d) for t in (R.PHILOX4_32_10, R.MRG32K3A, R.SOBOL64)
                                                     for d in ("u32", "f32", "f64")}
    assert max(r.n for r in reqs) > 1 << 25 and min(r.n for r in reqs) < 1 << 9
    rs = np.random.default_rng(3)
    checked = []
    srv = S.Server(1 << 30, streams=2)

    def sink(i, ptr, nbytes):
        r = reqs[i]
        assert nbytes == S.request_bytes(r)
        assert ptr % 256 == 0
        print(f"request {i}: engine {r.rng_type} {r.dtype} n={r.n}", flush=True)
        check_request(r, S.copy_to_host(ptr, nbytes, r.dtype), rs)
        checked.append(i)

    tickets = srv.run(reqs, sink)
    assert all(tickets) and tickets == sorted(tickets)
    assert sorted(checked) == list(range(len(reqs)))
    srv.close()


def _session_script():
    """(session, engine, seed, dims, dtype, n) in submission order: three sessions
    interleaved with fresh requests; ragged lengths so that a session's next request starts
    mid-vector."""
    ph, mrg, sob = (1, R.PHILOX4_32_10, 11, 1), (2, R.MRG32K3A, 12, 1), (3, R.SOBOL64, 13, 16)
    rounds = [
        [(*ph, "u32", 1001), (*mrg, "f32", 5000), (*sob, "f64", 16 * 100), (0, R.PHILOX4_32_10, 99, 1, "f32", 3000)],
        [(*ph, "f32", 4099), (*mrg, "u32", (1 << 21) + 1), (*sob, "u32", 16 * 1000),
         (0, R.MRG32K3A, 98, 1, "f64", 777)],
        [(*ph, "f64", 777), (*mrg, "f64", 333), (*sob, "f32", 16 * 7)],
        [(*ph, "u32", (1 << 20) + 3), (*mrg, "f32", 1 << 22), (0, R.SOBOL64, 97, 16, "u32", 16 * 64)],
    ]
    return rounds


def _oracle_sequence(script):
    """Expected outputs: one oracle generator per session, fresh ones otherwise."""
    gens, want = {}, []
    for rnd in script:
        for sess, eng, seed, dims, dtype, n in rnd:
            if sess == 0 or sess not in gens:
                if eng == R.SOBOL64:
                    g = O.OracleGen(eng, dims=dims, offset=seed % (1 << 20))
                else:
                    g = O.OracleGen(eng, seed=seed)
                if sess:
                    gens[sess] = g
            else:
                g = gens[sess]
            st, w = g.generate(ORACLE_KIND[dtype], n, S_ADDR)
            assert st == 0
            want.append(w)
    return want


@pytest.mark.parametrize("streams", [1, 2, 3])
def test_sessions_continue_their_generators(streams):
    """Each round is submitted and issued (one poll) before the previous rounds complete,
    so consecutive requests of a session run on different streams; outputs stay valid
    (keep_outputs) until released.  Every output equals the oracle's plain call sequence."""
    script = _session_script()
    want = _oracle_sequence(script)
    srv = S.Server(1 << 28, streams=streams, keep_outputs=True)
    tickets, done = [], {}
    for rnd in script:
        reqs = [Request(0.0, eng, dtype, n, seed, dims, sess) for (sess, eng, seed, dims, dtype, n) in rnd]
        tickets += srv.submit(reqs)
        got, _ = srv.poll()
        done.update({t: (p, b) for t, p, b in got})
    while len(done) < len(tickets):
        got, _ = srv.poll()
        done.update({t: (p, b) for t, p, b in got})
    flat = [x for rnd in script for x in rnd]
    for i, ((sess, eng, seed, dims, dtype, n), t) in enumerate(zip(flat, tickets)):
        p, b = done[t]
        got = S.copy_to_host(p, b, dtype)
        assert np.array_equal(got.view(np.uint8), want[i].view(np.uint8)), f"request {i}: session {sess} {dtype} n={n}"
        srv.release(t)
    for sess in (1, 2, 3):
        srv.end_session(sess)
    srv.close()


def test_session_generators_match_plain_calls():
    """A session's requests equal successive calls of one plain generator of the library."""
    srv = S.Server(1 << 26, streams=2)
    sizes = [(1 << 16) + 5, 12345, 1 << 18]
    reqs = [Request(0.0, R.MRG32K3A, "f32", n, 4242, 1, 77) for n in sizes]
    outs = {}
    srv.run(reqs, lambda i, p, b: outs.__setitem__(i, S.copy_to_host(p, b, "f32")))
    g = R.PRNG(R.MRG32K3A, seed=4242)
    for i, n in enumerate(sizes):
        buf = torch.empty(n, dtype=torch.float32, device="cuda")
        g.uniform(buf)
        torch.cuda.synchronize()
        assert np.array_equal(buf.cpu().numpy(), outs[i]), i
    srv.close()


def test_admission_waits_for_released_memory():
    """keep_outputs: four 256 KiB outputs fill a 1 MiB server; a 1 MiB request waits until
    all four are released in scattered order (the arena coalesces), then lands at the base."""
    srv = S.Server(1 << 20, streams=2, keep_outputs=True)
    small = [Request(0.0, R.PHILOX4_32_10, "u32", 1 << 16, 5 + i) for i in range(4)]
    tickets = srv.submit(small)
    done = {}
    while len(done) < 4:
        got, _ = srv.poll()
        done.update({t: p for t, p, _ in got})
    base = min(done.values())
    assert sorted(done.values()) == [base + i * (1 << 18) for i in range(4)]
    big = srv.submit([Request(0.0, R.PHILOX4_32_10, "u32", 1 << 18, 9)])[0]
    assert big != 0
    for k in (2, 0, 3):
        srv.release(tickets[k])
        for _ in range(3):
            got, pending = srv.poll()
            assert got == [] and pending == 1  # still waiting for memory
    srv.release(tickets[1])
    got = []
    while not got:
        got, _ = srv.poll()
    assert got[0][0] == big and got[0][1] == base
    out = S.copy_to_host(got[0][1], got[0][2], "u32")
    st, want = O.OracleGen(R.PHILOX4_32_10, seed=9).generate("u32", 1 << 18, S_ADDR)
    assert st == 0 and np.array_equal(out, want)
    srv.release(big)
    srv.close()


def test_status_codes_and_edge_requests():
    L = S._lib()
    srv = S.Server(1 << 20, streams=1)
    h = srv._h
    one = lambda r: (S.ServeRequest * 1)(r)  # noqa: E731
    t = (S.ctypes.c_ulonglong * 1)()
    req = lambda eng, dt, seed, sess, n, dims=1: S.ServeRequest(eng, dt, seed, sess, n, dims)  # noqa: E731
    assert L.rocrand_server_submit(h, one(req(R.XORWOW, 0, 1, 0, 10)), 1, t) == TYPE_ERROR
    assert L.rocrand_server_submit(h, one(req(R.PHILOX4_32_10, 3, 1, 0, 10)), 1, t) == TYPE_ERROR
    assert L.rocrand_server_submit(h, one(req(R.SOBOL64, 0, 1, 0, 17, 16)), 1, t) == LENGTH_NOT_MULTIPLE
    assert L.rocrand_server_submit(h, one(req(R.SOBOL64, 0, 1, 0, 16, 0)), 1, t) == OUT_OF_RANGE
    # larger than the whole budget: rejected with ticket 0, nothing queued
    assert L.rocrand_server_submit(h, one(req(R.PHILOX4_32_10, 2, 1, 0, 1 << 18)), 1, t) == 0
    assert t[0] == 0
    # a session keeps its engine and seed
    assert L.rocrand_server_submit(h, one(req(R.MRG32K3A, 1, 5, 8, 100)), 1, t) == 0 and t[0] != 0
    assert L.rocrand_server_submit(h, one(req(R.MRG32K3A, 1, 6, 8, 100)), 1, t) == TYPE_ERROR
    assert L.rocrand_server_submit(h, one(req(R.PHILOX4_32_10, 1, 5, 8, 100)), 1, t) == TYPE_ERROR
    assert L.rocrand_server_end_session(h, 8) == OUT_OF_RANGE  # its request is still queued
    # a call whose requests open one session with two engines queues nothing
    two = (S.ServeRequest * 2)(req(R.MRG32K3A, 1, 5, 9, 100), req(R.SOBOL64, 1, 5, 9, 16, 16))
    t2 = (S.ctypes.c_ulonglong * 2)()
    assert L.rocrand_server_submit(h, two, 2, t2) == TYPE_ERROR
    # an empty request completes with no bytes
    tickets = srv.submit([Request(0.0, R.PHILOX4_32_10, "f32", 0, 3)])
    seen = {}
    while len(seen) < 2:
        got, _ = srv.poll()
        seen.update({tk: b for tk, _, b in got})
    assert seen[tickets[0]] == 0
    srv.end_session(8)
    assert L.rocrand_server_end_session(h, 8) == OUT_OF_RANGE
    assert L.rocrand_server_release(h, tickets[0]) == OUT_OF_RANGE  # not a keep_outputs server
    srv.close()


def test_open_loop_run_completes():
    reqs = serving_mix(300, 1e5, seed=6)
    for r in reqs:  # 256 .. ~2^17 values
        r.n = max(r.dims, (r.n >> 9) // r.dims * r.dims)
    srv = S.Server(1 << 28)
    res = srv.serve(reqs)
    assert len(res.latencies_s) + res.rejected == len(reqs)
    assert res.rejected == 0
    assert min(res.latencies_s) >= 0
    assert res.percentile(50) <= res.percentile(99)
    assert res.bytes == sum(S.request_bytes(r) for r in reqs)
    srv.close()
