#!/usr/bin/env python3
"""Benchmark of the bulk-generation hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N] [--no-table] [--no-cpu]
    python bench.py --only ENGINE:KIND --steps K --warmup W      (one pair, for profiler runs)

A step is one rocrand_generate_uniform call (Philox4x32-10, seed 12345,
offset 0) writing N = 2^28 fp32 samples into a preallocated HBM buffer: the
plain bulk path per engine at 2^28 samples.  Before the W warm-up steps, the step
runs for 0.5 s of wall time (`settle_calls`) so that the GPU's clocks have left their
idle level.  The JSON line's `value` is that call's output bandwidth in decimal GB/s;
`per_generator` times the other engine x distribution pairs of the same workload;
`roofline` prices the kernel against MI355X's 8 TB/s HBM peak (and a measured
store-kernel ceiling); `normals_accuracy` gives the normal fp32 outputs' error against
a float64 Box-Muller of the same words; `cpu_baseline` times the reference's
host-generator path (the installed rocRAND's, on every available core, and the
reference's own kernel bodies on one core) over the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E bandwidth


def _gbs(nbytes, ms):
    return nbytes / (ms * 1e-3) / 1e9


def time_calls(fn, stream, steps, warmup):
    """Average per-call device time (ms) over `steps` calls on `stream`, via HIP events."""
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(steps):
        fn()
    end.record(stream)
    torch.cuda.synchronize()
    return start.elapsed_time(end) / steps


def settle(fn, seconds=0.5, chunk=50):
    """Run `fn` in chunks until `seconds` of wall time have passed; returns the call count.
    A GPU that was idle runs its first calls at low clocks while the power management
    raises them (the first calls of a fresh process took up to 1.6x the steady-state time);
    this phase brings it to the steady state before the warm-up and timed steps."""
    import torch

    calls, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        for _ in range(chunk):
            fn()
        torch.cuda.synchronize()
        calls += chunk
    return calls


def _pairs(R, torch):
    engines = [("philox4x32_10", R.PHILOX4_32_10), ("threefry4x32_20", R.THREEFRY4_32_20),
               ("mrg32k3a", R.MRG32K3A), ("xorwow", R.XORWOW), ("lfsr113", R.LFSR113),
               ("mrg31k3p", R.MRG31K3P), ("threefry2x32_20", R.THREEFRY2_32_20),
               ("threefry2x64_20", R.THREEFRY2_64_20), ("threefry4x64_20", R.THREEFRY4_64_20)]
    dists = [("u32", torch.uint32, lambda g, b: g.generate(b)),
             ("uniform_f32", torch.float32, lambda g, b: g.uniform(b)),
             ("uniform_f64", torch.float64, lambda g, b: g.uniform(b)),
             ("normal_f32", torch.float32, lambda g, b: g.normal(b, 0.0, 1.0)),
             ("normal_f64", torch.float64, lambda g, b: g.normal(b, 0.0, 1.0)),
             ("lognormal_f32", torch.float32, lambda g, b: g.lognormal(b, 0.0, 1.0)),
             ("poisson_10", torch.uint32, lambda g, b: g.poisson(b, 10.0))]
    engines += [("mtgp32", R.MTGP32), ("mt19937", R.MT19937), ("sobol32", R.SOBOL32), ("scrambled_sobol32", R.SCRAMBLED_SOBOL32),
                ("sobol64", R.SOBOL64), ("scrambled_sobol64", R.SCRAMBLED_SOBOL64)]
    return engines, dists


def only_pair(R, torch, stream, n, pair, steps, warmup):
    """GB/s of one ENGINE:KIND pair of the per-generator table (for profiler runs)."""
    engines, dists = _pairs(R, torch)
    ename, dname = pair.split(":")
    rng = dict(engines)[ename]
    _, dt, call = [d for d in dists if d[0] == dname][0]
    g = R.QRNG(rng) if "sobol" in ename else R.PRNG(rng, seed=12345)
    g.stream = stream
    b = torch.empty(n, dtype=dt, device="cuda")
    ms = time_calls(lambda: call(g, b), stream, steps, warmup)
    return {"pair": pair, "n": n, "steps": steps, "warmup": warmup, "ms_per_call": round(ms, 4),
            "GB/s": round(_gbs(n * b.element_size(), ms), 1)}


def per_generator_table(R, torch, stream, n, steps):
    """GB/s for each engine x distribution at n samples (same workload as the step)."""
    engines, dists = _pairs(R, torch)
    table = {}
    bufs = {}
    for ename, rng in engines:
        g = R.QRNG(rng) if "sobol" in ename else R.PRNG(rng, seed=12345)
        g.stream = stream
        row = {}
        for dname, dt, call in dists:
            if dt not in bufs:
                bufs[dt] = torch.empty(n, dtype=dt, device="cuda")
            b = bufs[dt]
            # clocks settle per kernel: ALU-heavy transforms run slower for their first calls
            ms = time_calls(lambda: call(g, b), stream, steps, 40)
            row[dname] = round(_gbs(n * b.element_size(), ms), 1)
        table[ename] = row
        print(f"# {ename}: {row}", file=sys.stderr, flush=True)
    del bufs
    return table


def store_ceiling(torch, stream, n, steps):
    """Achievable HBM write bandwidth: a plain fill kernel over the same buffer size."""
    buf = torch.empty(n, dtype=torch.float32, device="cuda")
    with torch.cuda.stream(stream):
        ms = time_calls(lambda: buf.fill_(1.0), stream, steps, 3)
    return _gbs(n * 4, ms)


def copy_calibration(torch, n):
    """A plain device copy of n floats (n*4 bytes read, n*4 written): the known-size kernel
    whose FETCH_SIZE / WRITE_SIZE counters calibrate the generator kernel's HBM traffic."""
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()


def batched_bench(R, torch, count=4096, seed=7, launches=10):
    """4096 independent Philox4x32-10 / MRG32k3a streams (rocrand_amd.workloads.batch_mix):
    sequential plain C-ABI call sequences (create, set_seed, set_offset, generate, destroy:
    per-call set-up included) against one replay of the batch plan's HIP graph."""
    import ctypes

    from rocrand_amd.workloads import batch_mix

    L = R.load()
    specs = batch_mix(count, seed)
    total_bytes = sum(sp.n * sp.itemsize for sp in specs)
    offs, pos = [], 0
    for sp in specs:
        offs.append(pos)
        pos += (sp.n * sp.itemsize + 255) // 256 * 256
    buf = torch.empty(pos, dtype=torch.uint8, device="cuda")
    base = buf.data_ptr()
    streams = [R.BatchStream(sp.rng_type, sp.kind, sp.seed, sp.offset, sp.n, sp.a, sp.b, base + o)
               for sp, o in zip(specs, offs)]
    fn = {R.BATCH_UNIFORM_FLOAT: L.rocrand_generate_uniform, R.BATCH_NORMAL_FLOAT: L.rocrand_generate_normal,
          R.BATCH_POISSON: L.rocrand_generate_poisson}

    def plain_all():
        for sp, o in zip(specs, offs):
            g = ctypes.c_void_p()
            R.check_rocrand(L.rocrand_create_generator(ctypes.byref(g), sp.rng_type))
            L.rocrand_set_seed(g, sp.seed)
            L.rocrand_set_offset(g, sp.offset)
            if sp.kind == R.BATCH_UNIFORM_FLOAT:
                st = fn[sp.kind](g, ctypes.c_void_p(base + o), sp.n)
            elif sp.kind == R.BATCH_NORMAL_FLOAT:
                st = fn[sp.kind](g, ctypes.c_void_p(base + o), sp.n, sp.a, sp.b)
            else:
                st = fn[sp.kind](g, ctypes.c_void_p(base + o), sp.n, sp.a)
            R.check_rocrand(st)
            L.rocrand_destroy_generator(g)
        torch.cuda.synchronize()

    plain_all()  # warm caches (Poisson table code, jump tables, kernels)
    t = time.perf_counter()
    plain_all()
    plain_ms = (time.perf_counter() - t) * 1e3
    ref = buf.clone()

    stream = torch.cuda.Stream()
    torch.cuda.synchronize()
    t = time.perf_counter()
    plan = R.BatchPlan(streams, stream)
    torch.cuda.synchronize()
    plan_ms = (time.perf_counter() - t) * 1e3
    buf.zero_()
    torch.cuda.synchronize()
    plan.launch()
    torch.cuda.synchronize()
    same = bool(torch.equal(buf, ref))
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(launches):
        plan.launch()
    end.record(stream)
    torch.cuda.synchronize()
    graph_ms = start.elapsed_time(end) / launches
    plan.close()
    del buf, ref
    return {"streams": count, "workload_seed": seed, "values": sum(sp.n for sp in specs),
            "bytes": total_bytes, "plain_ms": round(plain_ms, 2), "graph_launch_ms": round(graph_ms, 3),
            "plan_create_ms": round(plan_ms, 1), "speedup": round(plain_ms / graph_ms, 1),
            "graph_GBps": round(_gbs(total_bytes, graph_ms), 1), "outputs_equal_plain": same}


def serving_bench(torch, rates=(1000, 10000, 50000), seconds=1.0, budget_gb=32, streams=4):
    """Open-loop Poisson arrivals at each rate for `seconds` (rocrand_amd.workloads.serving_mix:
    sizes log-uniform in [256, 2^26], 50% Philox / 30% MRG32k3a / 20% Sobol64 x16 dims,
    f32/f64/u32) through the request server (include/rocrand_serving.h) with a 32 GB HBM
    budget and `streams` streams."""
    from rocrand_amd.serving import Server
    from rocrand_amd.workloads import serving_mix

    srv = Server(int(budget_gb * 1e9), streams=streams)
    # warm-up: kernels, tables
    srv.serve(serving_mix(50, 1e6, seed=1))
    out = []
    for rate in rates:
        reqs = serving_mix(int(rate * seconds), rate, seed=11)
        res = srv.serve(reqs)
        out.append({"rate_rps": rate, "requests": len(reqs), "p50_ms": round(res.percentile(50) * 1e3, 3),
                    "p99_ms": round(res.percentile(99) * 1e3, 3), "GBps": round(res.bytes / res.wall_s / 1e9, 1),
                    "polls": res.polls, "rejected": res.rejected,
                    "wall_s": round(res.wall_s, 3)})
        print(f"# serving: {out[-1]}", file=sys.stderr, flush=True)
    srv.close()
    return {"budget_GB": budget_gb, "streams": streams, "workload_seed": 11, "runs": out}


def _exact_box_muller_f32(u, v):
    import numpy as np

    s = np.sqrt(-2.0 * np.log(u.astype(np.float64)))
    out = np.empty((u.size, 2))
    out[:, 0] = np.sin(v.astype(np.float64)) * s
    out[:, 1] = np.cos(v.astype(np.float64)) * s
    return out.ravel()


def normals_accuracy(R, torch, n=1 << 22, seed=12345):
    """Max and mean error, in ulps, of the device's normal fp32 outputs (mean 0, stddev 1)
    against the Box-Muller of the same engine words evaluated in float64: Philox4x32-10
    (words from the u32 output, consecutive pairs) and MRG32k3a (uniforms from the uniform
    fp32 output; normal vec v comes from engine v mod 131072, words 2q, 2q + 1, q = v div
    131072)."""
    import numpy as np

    res = {}
    for name, rng in (("philox4x32_10", R.PHILOX4_32_10), ("mrg32k3a", R.MRG32K3A)):
        g = R.PRNG(rng, seed=seed)
        z = torch.empty(n, dtype=torch.float32, device="cuda")
        g.normal(z, 0.0, 1.0)
        if rng == R.PHILOX4_32_10:
            w = torch.empty(n, dtype=torch.uint32, device="cuda")
            R.PRNG(rng, seed=seed).generate(w)
            w = w.cpu().numpy().reshape(-1, 2)
            c, c2pi = np.float32(2.3283064e-10), np.float32(1.46291807e-09)
            u = c + w[:, 0].astype(np.float32) * c
            v = c2pi + w[:, 1].astype(np.float32) * c2pi
        else:
            f = torch.empty(n, dtype=torch.float32, device="cuda")
            R.PRNG(rng, seed=seed).uniform(f)
            f = f.cpu().numpy()
            S = 131072
            vec = np.arange(n // 2)
            e, q = vec % S, vec // S
            u = f[e + 2 * q * S]
            v = f[e + (2 * q + 1) * S] * np.float32(6.2831855)
        torch.cuda.synchronize()
        exact = _exact_box_muller_f32(u, v)
        got = z.cpu().numpy().astype(np.float64)
        ulp = np.spacing(np.abs(exact.astype(np.float32))).astype(np.float64)
        err = np.abs(got - exact) / ulp
        res[name] = {"max_ulp": round(float(err.max()), 4), "mean_ulp": round(float(err.mean()), 4)}
    return {"against": "float64 Box-Muller of the same words", "n": n, "seed": seed, "normal_f32": res}


def load_traffic():
    """HBM bytes per launch of the headline kernel, from the committed counter summary."""
    path = os.path.join(HERE, "profiles", "traffic_philox_uniform_f32_2e28.json")
    try:
        with open(path) as f:
            return json.load(f).get("traffic_bytes_per_launch")
    except (OSError, ValueError):
        return None


def _installed_rocrand_host(n):
    """The image's installed rocRAND (its host-generator path, rocrand_create_generator_host_blocking):
    Philox4x32-10, seed 12345, n uniform fp32 into host memory; median of 3 timed calls after
    one warm-up.  Returns (seconds, version) or None when the library is absent."""
    import ctypes

    import numpy as np

    path = "/opt/rocm/lib/librocrand.so"
    if not os.path.exists(path):
        return None
    try:
        L = ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)
    except OSError:
        return None
    gen = ctypes.c_void_p()
    if L.rocrand_create_generator_host_blocking(ctypes.byref(gen), 404) != 0:
        return None
    ver = ctypes.c_int(0)
    L.rocrand_get_version(ctypes.byref(ver))
    L.rocrand_set_seed(gen, ctypes.c_ulonglong(12345))
    out = np.empty(n, dtype=np.float32)
    ptr = ctypes.c_void_p(out.ctypes.data)
    L.rocrand_generate_uniform(gen, ptr, ctypes.c_size_t(1 << 20))
    times = []
    for _ in range(3):
        t = time.perf_counter()
        st = L.rocrand_generate_uniform(gen, ptr, ctypes.c_size_t(n))
        times.append(time.perf_counter() - t)
        if st != 0:
            L.rocrand_destroy_generator(gen)
            return None
    L.rocrand_destroy_generator(gen)
    return sorted(times)[1], ver.value


def cpu_baseline(n):
    """The reference's host-generator CPU path on this box's host cores.

    Primary: the image's installed rocRAND host generator (same path, a later release), on
    every core this process may use.  Also: the reference's own Philox kernel bodies
    (compiled here from its sources, oracle/_ref) on one core."""
    from oracle import oracle as O

    cores = len(os.sched_getaffinity(0))
    installed = _installed_rocrand_host(n)
    g, kind_ref = None, "reference"
    try:
        g = O.RefGen(O.PHILOX4_32_10, seed=12345)
    except Exception:
        g, kind_ref = O.OracleGen(O.PHILOX4_32_10, seed=12345), "port"
    g.generate("uniform_f32", 1024, 0x1000)
    t = time.perf_counter()
    st, _ = g.generate("uniform_f32", n, 0x1000)
    dt1 = time.perf_counter() - t
    assert st == 0
    one_core = {"value": round(n * 4 / dt1 / 1e9, 4), "unit": "GB/s", "cores": 1, "kind": kind_ref,
                "sample": f"reference kernel bodies compiled from its sources (oracle/_ref), Philox4x32-10 "
                          f"uniform fp32, seed 12345, n=2^{n.bit_length() - 1}, {dt1:.2f} s"}
    if installed is None:
        one_core["secondary"] = None
        return one_core
    dt, ver = installed
    return {"value": round(n * 4 / dt / 1e9, 4), "unit": "GB/s", "cores": cores, "kind": "reference",
            "sample": f"installed rocRAND {ver // 100000}.{ver // 100 % 1000}.{ver % 100} host generator "
                      f"(rocrand_create_generator_host_blocking), Philox4x32-10 uniform fp32, seed 12345, "
                      f"n=2^{n.bit_length() - 1} into host memory, median of 3 calls ({dt:.2f} s), "
                      f"{cores} cores available ({os.cpu_count()} online)",
            "one_core_reference_bodies": one_core}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=1 << 28)
    ap.add_argument("--no-table", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=1 << 28)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serving-streams", type=int, default=4)
    ap.add_argument("--calib", action="store_true",
                    help="also run a plain 2^28-float device copy (counter calibration)")
    ap.add_argument("--only", metavar="ENGINE:KIND",
                    help="time one pair of the per-generator table (e.g. philox4x32_10:normal_f32) and "
                         "print its figures only; for profiler runs")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: every generator here runs on a single MI355X and the "
              "multi-GPU layout is design-only; run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch
    from rocrand_amd import rocrand as R

    n = args.n
    if args.calib:
        copy_calibration(torch, 1 << 28)
    stream = torch.cuda.Stream()
    if args.only:
        print(json.dumps(only_pair(R, torch, stream, n, args.only, args.steps, args.warmup)), flush=True)
        return
    out = torch.empty(n, dtype=torch.float32, device="cuda")
    g = R.PRNG(R.PHILOX4_32_10, seed=12345)
    g.stream = stream
    g.initialize()

    def step():
        g.uniform(out)

    settle_calls = settle(step)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record(stream)
    for _ in range(args.steps):
        step()
    end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = start.elapsed_time(end) / args.steps
    nbytes = n * 4
    value = _gbs(nbytes, ms)

    ceiling = store_ceiling(torch, stream, n, 10)
    table = None if args.no_table else per_generator_table(R, torch, stream, n, 10)
    accuracy = normals_accuracy(R, torch)
    cpu = None if args.no_cpu else cpu_baseline(args.cpu_n)
    batched = None if args.no_batched else batched_bench(R, torch)
    if batched:
        print(f"# batched: {batched}", file=sys.stderr, flush=True)
    serving = None if args.no_serving else serving_bench(torch, streams=args.serving_streams)

    line = {
        "metric": "output GB/s & % of HBM write peak per generator; batched-vs-plain speedup",
        "value": round(value, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "settle_calls": settle_calls,
        "ms_per_step": round(ms, 4),
        "wall_ms_per_step": round(wall * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (generated output; no input data)",
        "config": {"workload": "plain bulk path, Philox4x32-10 uniform fp32, 2^28 samples, seed 12345, offset 0",
                   "n": n, "samples_per_s": round(n / (ms * 1e-3), 1)},
        "roofline": {"bound": "hbm", "achieved": round(value, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(value / HBM_PEAK_GBS, 4), "traffic": load_traffic(),
                     "measured_store_ceiling_gbs": round(ceiling, 1)},
        "cpu_baseline": cpu,
        "normals_accuracy": accuracy,
        "per_generator": table,
        "batched": batched,
        "serving": serving,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
