"""Continuous-batching request server over the batched C ABI.

Requests (rocrand_amd.workloads.Request) arrive open-loop at their timestamps.  The
server loop never waits for the device while work is pending:

* every turn it retires the batches whose completion event has fired (their outputs
  are done; their memory returns to the arena) and timestamps their requests;
* it then admits the queued requests, first come first served, whose outputs fit the
  free arena memory (one pre-allocated HBM budget, first-fit with coalescing; a request
  larger than the whole budget is rejected), and issues them as one batch: one
  asynchronous batched call (rocrand_generate_batched, include/rocrand_batched.h) for the
  Philox4x32-10 / MRG32k3a requests, one quasi-random generator call per Sobol64
  request, then a completion event;
* consecutive batches alternate between two device streams, so a batch admitted while
  earlier ones run overlaps with them on the device.

Latency = time the completion was observed - arrival time.  A request's output does
not depend on its batch: it is the output of the plain call sequence of
include/rocrand_batched.h (Sobol64: a 16-dimensional generator at offset seed mod 2^20).
"""
from __future__ import annotations

import bisect
import collections
import time
from dataclasses import dataclass, field

from . import rocrand as R

KIND = {"f32": R.BATCH_UNIFORM_FLOAT, "f64": R.BATCH_UNIFORM_DOUBLE, "u32": R.BATCH_UINT}
ITEMSIZE = {"f32": 4, "f64": 8, "u32": 4}
ALIGN = 256


@dataclass
class Result:
    latencies_s: list = field(default_factory=list)
    bytes: int = 0
    wall_s: float = 0.0
    iterations: int = 0
    rejected: int = 0
    max_batch: int = 0
    max_inflight: int = 0

    def percentile(self, q):
        if not self.latencies_s:
            return None
        xs = sorted(self.latencies_s)
        k = min(len(xs) - 1, max(0, int(round(q / 100.0 * (len(xs) - 1)))))
        return xs[k]


def request_bytes(req) -> int:
    return req.n * ITEMSIZE[req.dtype]


def padded(nbytes: int) -> int:
    return (nbytes + ALIGN - 1) // ALIGN * ALIGN


def sobol_offset(req) -> int:
    return req.seed % (1 << 20)


def batch_streams(reqs, ptrs):
    """BatchStream entries for the batchable requests (Philox / MRG32k3a)."""
    out = []
    for r, p in zip(reqs, ptrs):
        if r.rng_type in (R.PHILOX4_32_10, R.MRG32K3A):
            out.append(R.BatchStream(r.rng_type, KIND[r.dtype], r.seed, 0, r.n, 0.0, 0.0, p))
    return out


def write_sobol(req, view, stream):
    g = R.QRNG(R.SOBOL64, ndim=req.dims, offset=sobol_offset(req))
    g.stream = stream
    if req.dtype == "u32":
        g.generate(view)
    else:
        g.uniform(view)


class Arena:
    """First-fit allocator over [0, size) with coalescing of freed ranges."""

    def __init__(self, size: int):
        self.size = size
        self.free = [(0, size)]  # sorted (start, length), no two adjacent

    def alloc(self, n: int):
        for i, (s, ln) in enumerate(self.free):
            if ln >= n:
                if ln == n:
                    self.free.pop(i)
                else:
                    self.free[i] = (s + n, ln - n)
                return s
        return None

    def release(self, start: int, n: int):
        """Returns [start, start + n) to the free list, merged with its free neighbours."""
        free = self.free
        i = bisect.bisect_left(free, (start, 0))
        if i < len(free) and start + n == free[i][0]:  # merge with the next range
            n += free[i][1]
            free.pop(i)
        if i > 0 and free[i - 1][0] + free[i - 1][1] == start:  # and with the previous one
            free[i - 1] = (free[i - 1][0], free[i - 1][1] + n)
        else:
            free.insert(i, (start, n))


@dataclass
class _Inflight:
    reqs: list
    event: object
    regions: list


class Server:
    def __init__(self, torch, budget_bytes: int, streams: int = 2):
        self.torch = torch
        self.budget = budget_bytes
        self.memory = torch.empty(budget_bytes, dtype=torch.uint8, device="cuda")
        self.base = self.memory.data_ptr()
        self.arena = Arena(budget_bytes)
        self.streams = [torch.cuda.Stream() for _ in range(streams)]
        self.turn = 0

    def _view(self, start, req):
        dt = {"f32": self.torch.float32, "f64": self.torch.float64, "u32": self.torch.uint32}[req.dtype]
        return self.memory[start:start + request_bytes(req)].view(dt)

    def issue(self, reqs, regions):
        """Queue one batch on the next stream; returns its completion event."""
        torch = self.torch
        stream = self.streams[self.turn % len(self.streams)]
        self.turn += 1
        base = self.base
        streams = batch_streams(reqs, [base + s for s in regions])
        with torch.cuda.stream(stream):
            if streams:
                R.generate_batched(streams, stream)
            for r, s in zip(reqs, regions):
                if r.rng_type == R.SOBOL64:
                    write_sobol(r, self._view(s, r), stream)
            ev = torch.cuda.Event()
            ev.record(stream)
        return ev

    def run_batch(self, reqs):
        """Writes `reqs` as one batch and waits for it (outputs stay valid until the next call)."""
        self.arena = Arena(self.budget)
        regions = [self.arena.alloc(padded(request_bytes(r))) for r in reqs]
        if any(s is None for s in regions):
            raise ValueError("batch exceeds the budget")
        ev = self.issue(reqs, regions)
        views = [self._view(s, r) for s, r in zip(regions, reqs)]
        ev.synchronize()
        self.arena = Arena(self.budget)
        return views

    def serve(self, requests) -> Result:
        """Open-loop run: request i becomes visible at its arrival time."""
        res = Result()
        i = 0
        queue, inflight = collections.deque(), []
        t0 = time.perf_counter()
        while i < len(requests) or queue or inflight:
            now = time.perf_counter() - t0
            # retire finished batches
            still = []
            for b in inflight:
                if b.event.query():
                    done = time.perf_counter() - t0
                    for r in b.reqs:
                        res.latencies_s.append(done - r.arrival_s)
                        res.bytes += request_bytes(r)
                    for s, r in zip(b.regions, b.reqs):
                        self.arena.release(s, padded(request_bytes(r)))
                else:
                    still.append(b)
            inflight = still
            # arrivals
            while i < len(requests) and requests[i].arrival_s <= now:
                if padded(request_bytes(requests[i])) > self.budget:
                    res.rejected += 1
                else:
                    queue.append(requests[i])
                i += 1
            # admission: FIFO while memory is free
            batch, regions = [], []
            while queue:
                s = self.arena.alloc(padded(request_bytes(queue[0])))
                if s is None:
                    break
                batch.append(queue.popleft())
                regions.append(s)
            if batch:
                ev = self.issue(batch, regions)
                inflight.append(_Inflight(batch, ev, regions))
                res.iterations += 1
                res.max_batch = max(res.max_batch, len(batch))
                res.max_inflight = max(res.max_inflight, len(inflight))
            elif not inflight and i < len(requests):
                time.sleep(max(0.0, min(2e-4, requests[i].arrival_s - now)))
        res.wall_s = time.perf_counter() - t0
        return res
