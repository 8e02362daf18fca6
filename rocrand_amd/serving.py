"""Continuous-batching request server over the batched C ABI.

Requests (rocrand_amd.workloads.Request) arrive open-loop at their timestamps.  The
server loop repeatedly admits every arrived request, first come first served, while the
outputs fit the HBM budget (one pre-allocated arena; a request larger than the whole
budget is rejected), writes the admitted Philox4x32-10 / MRG32k3a requests with one
rocrand_generate_batched call and each Sobol64 request with its own quasi-random
generator call on the same stream, and marks them complete when that stream's work has
finished.  Latency = completion time - arrival time.

A request's output does not depend on what it was batched with: it is the output of
the plain call sequence of include/rocrand_batched.h (Sobol64: a 16-dimensional
generator at offset seed mod 2^20).
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

from . import rocrand as R

KIND = {"f32": R.BATCH_UNIFORM_FLOAT, "f64": R.BATCH_UNIFORM_DOUBLE, "u32": R.BATCH_UINT}
ITEMSIZE = {"f32": 4, "f64": 8, "u32": 4}


@dataclass
class Result:
    latencies_s: list = field(default_factory=list)
    bytes: int = 0
    wall_s: float = 0.0
    iterations: int = 0
    rejected: int = 0
    max_batch: int = 0

    def percentile(self, q):
        if not self.latencies_s:
            return None
        xs = sorted(self.latencies_s)
        k = min(len(xs) - 1, max(0, int(round(q / 100.0 * (len(xs) - 1)))))
        return xs[k]


def request_bytes(req) -> int:
    return req.n * ITEMSIZE[req.dtype]


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


class Server:
    """One stream, one arena of `budget_bytes` device memory."""

    def __init__(self, torch, budget_bytes: int, stream=None):
        self.torch = torch
        self.budget = budget_bytes
        self.arena = torch.empty(budget_bytes, dtype=torch.uint8, device="cuda")
        self.stream = stream or torch.cuda.Stream()

    def views(self, reqs):
        """Arena slices (256-byte aligned) for the admitted requests."""
        torch = self.torch
        dt = {"f32": torch.float32, "f64": torch.float64, "u32": torch.uint32}
        pos, out = 0, []
        for r in reqs:
            nb = request_bytes(r)
            v = self.arena[pos:pos + nb].view(dt[r.dtype])
            out.append(v)
            pos += (nb + 255) // 256 * 256
        return out

    def admit(self, queue, start):
        """FIFO admission while the 256-byte-rounded outputs fit the budget."""
        used, k = 0, start
        while k < len(queue):
            nb = (request_bytes(queue[k]) + 255) // 256 * 256
            if used + nb > self.budget:
                break
            used += nb
            k += 1
        return k

    def run_batch(self, reqs):
        views = self.views(reqs)
        ptrs = [v.data_ptr() for v in views]
        streams = batch_streams(reqs, ptrs)
        if streams:
            R.generate_batched(streams, self.stream)
        for r, v in zip(reqs, views):
            if r.rng_type == R.SOBOL64:
                write_sobol(r, v, self.stream)
        return views

    def serve(self, requests) -> Result:
        """Open-loop run: request i becomes visible at its arrival time."""
        torch = self.torch
        res = Result()
        i = 0
        t0 = time.perf_counter()
        queue = []
        while i < len(requests) or queue:
            now = time.perf_counter() - t0
            while i < len(requests) and requests[i].arrival_s <= now:
                if request_bytes(requests[i]) > self.budget:
                    res.rejected += 1
                else:
                    queue.append(requests[i])
                i += 1
            if not queue:
                if i < len(requests):
                    time.sleep(max(0.0, min(1e-3, requests[i].arrival_s - now)))
                continue
            k = self.admit(queue, 0)
            batch, queue = queue[:k], queue[k:]
            with torch.cuda.stream(self.stream):
                self.run_batch(batch)
            self.stream.synchronize()
            done = time.perf_counter() - t0
            for r in batch:
                res.latencies_s.append(done - r.arrival_s)
                res.bytes += request_bytes(r)
            res.iterations += 1
            res.max_batch = max(res.max_batch, len(batch))
        res.wall_s = time.perf_counter() - t0
        return res
