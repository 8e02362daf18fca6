"""Continuous-batching request server: Python client of include/rocrand_serving.h.

The server itself is C++ (rocrand_amd/csrc/rr_serving.cpp): its HBM arena (first fit with
coalescing over one allocation), FIFO admission, batch issue on alternating streams,
completion events and session generators.  This module binds it and adds the open-loop
driver used by the benchmark: requests (rocrand_amd.workloads.Request) become visible at
their arrival timestamps, every turn submits the new arrivals and polls once, and a
request's latency is the time its completion was reported minus its arrival time.  The
host never waits for the device while requests are pending.

A request's output is that of a plain call sequence (include/rocrand_serving.h): session 0
is a fresh generator's single call (Sobol64: 'dims' dimensions at offset seed mod 2^20);
a session s != 0 continues the generator of session s.
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass, field

import numpy as np

from . import rocrand as R

DTYPE = {"u32": 0, "f32": 1, "f64": 2}
ITEMSIZE = {"f32": 4, "f64": 8, "u32": 4}
NUMPY = {"f32": np.float32, "f64": np.float64, "u32": np.uint32}
ALIGN = 256


class ServeRequest(ctypes.Structure):
    _fields_ = [("rng_type", ctypes.c_int), ("dtype", ctypes.c_int), ("seed", ctypes.c_ulonglong),
                ("session", ctypes.c_ulonglong), ("n", ctypes.c_size_t), ("dims", ctypes.c_uint)]


class ServeCompletion(ctypes.Structure):
    _fields_ = [("ticket", ctypes.c_ulonglong), ("output", ctypes.c_void_p), ("bytes", ctypes.c_size_t)]


def _lib():
    L = R.load()
    if not getattr(L, "_serving_bound", False):
        vp, sz, u64 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_ulonglong
        L.rocrand_server_create.argtypes = [ctypes.POINTER(vp), sz, ctypes.c_uint, ctypes.c_int]
        L.rocrand_server_destroy.argtypes = [vp]
        L.rocrand_server_submit.argtypes = [vp, ctypes.POINTER(ServeRequest), sz, ctypes.POINTER(u64)]
        L.rocrand_server_poll.argtypes = [vp, ctypes.POINTER(ServeCompletion), sz, ctypes.POINTER(sz),
                                          ctypes.POINTER(sz)]
        L.rocrand_server_release.argtypes = [vp, u64]
        L.rocrand_server_end_session.argtypes = [vp, u64]
        L._serving_bound = True
    return L


def request_bytes(req) -> int:
    return req.n * ITEMSIZE[req.dtype]


def slot(nbytes: int) -> int:
    """Arena bytes the server reserves for an output of `nbytes` (whole 256-byte units, at least one)."""
    return max(ALIGN, (nbytes + ALIGN - 1) // ALIGN * ALIGN)


def _check(st, what):
    if st != 0:
        raise R.RocRandError(st) if hasattr(R, "RocRandError") else RuntimeError(f"{what}: status {st}")


@dataclass
class Result:
    latencies_s: list = field(default_factory=list)
    bytes: int = 0
    wall_s: float = 0.0
    polls: int = 0
    rejected: int = 0

    def percentile(self, q):
        if not self.latencies_s:
            return None
        xs = sorted(self.latencies_s)
        k = min(len(xs) - 1, max(0, int(round(q / 100.0 * (len(xs) - 1)))))
        return xs[k]


class Server:
    """One rocrand_server: `budget_bytes` of HBM, `streams` HIP streams."""

    def __init__(self, budget_bytes: int, streams: int = 2, keep_outputs: bool = False, capacity: int = 4096):
        self._L = _lib()
        self._h = ctypes.c_void_p()
        _check(self._L.rocrand_server_create(ctypes.byref(self._h), budget_bytes, streams, int(keep_outputs)),
               "rocrand_server_create")
        self.budget = budget_bytes // ALIGN * ALIGN
        self._done = (ServeCompletion * capacity)()
        self._capacity = capacity

    def close(self):
        if self._h:
            _check(self._L.rocrand_server_destroy(self._h), "rocrand_server_destroy")
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, reqs) -> list[int]:
        """Queues `reqs` (objects with rng_type, dtype, seed, n and optional session, dims);
        returns their tickets (0 = rejected: larger than the budget)."""
        if not reqs:
            return []
        arr = (ServeRequest * len(reqs))()
        for i, r in enumerate(reqs):
            arr[i] = ServeRequest(r.rng_type, DTYPE[r.dtype], r.seed & (2**64 - 1), getattr(r, "session", 0),
                                  r.n, getattr(r, "dims", 1))
        tickets = (ctypes.c_ulonglong * len(reqs))()
        _check(self._L.rocrand_server_submit(self._h, arr, len(reqs), tickets), "rocrand_server_submit")
        return list(tickets)

    def poll(self):
        """One turn: returns ([(ticket, device pointer, bytes)], pending)."""
        k, pending = ctypes.c_size_t(), ctypes.c_size_t()
        _check(self._L.rocrand_server_poll(self._h, self._done, self._capacity, ctypes.byref(k),
                                           ctypes.byref(pending)), "rocrand_server_poll")
        return [(self._done[i].ticket, self._done[i].output, self._done[i].bytes) for i in range(k.value)], \
            pending.value

    def release(self, ticket: int):
        _check(self._L.rocrand_server_release(self._h, ticket), "rocrand_server_release")

    def end_session(self, session: int):
        _check(self._L.rocrand_server_end_session(self._h, session), "rocrand_server_end_session")

    def run(self, reqs, sink=None):
        """Submits `reqs` and polls until all are reported; `sink(index, pointer, bytes)` is
        called for each completion while its output is valid.  Returns the tickets."""
        tickets = self.submit(reqs)
        index = {t: i for i, t in enumerate(tickets) if t}
        left = len(index)
        while left:
            done, _ = self.poll()
            for t, ptr, nbytes in done:
                if sink is not None:
                    sink(index[t], ptr, nbytes)
                left -= 1
            if not done:
                time.sleep(2e-5)
        return tickets

    def serve(self, requests) -> Result:
        """Open-loop run: request i is submitted once its arrival time has passed."""
        res = Result()
        i, owner, t0 = 0, {}, time.perf_counter()
        while True:
            now = time.perf_counter() - t0
            j = i
            while j < len(requests) and requests[j].arrival_s <= now:
                j += 1
            if j > i:
                for k, t in zip(range(i, j), self.submit(requests[i:j])):
                    if t:
                        owner[t] = requests[k]
                    else:
                        res.rejected += 1
                i = j
            done, pending = self.poll()
            res.polls += 1
            stamp = time.perf_counter() - t0
            for t, _, nbytes in done:
                res.latencies_s.append(stamp - owner.pop(t).arrival_s)
                res.bytes += nbytes
            if i == len(requests) and pending == 0:
                break
            if not done and pending == 0 and i < len(requests):
                time.sleep(max(0.0, min(2e-4, requests[i].arrival_s - stamp)))
        res.wall_s = time.perf_counter() - t0
        return res


def copy_to_host(ptr: int, nbytes: int, dtype: str) -> np.ndarray:
    """A synchronous device-to-host copy of a completed output."""
    out = np.empty(nbytes // ITEMSIZE[dtype], dtype=NUMPY[dtype])
    hip = R._hip()
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    R._hip_check(hip.hipMemcpy(out.ctypes.data, ptr, nbytes, 2))  # hipMemcpyDeviceToHost
    return out
