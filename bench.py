#!/usr/bin/env python3
"""Benchmark of the bulk-generation hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N] [--no-table] [--no-cpu]

A step is one rocrand_generate_uniform call (Philox4x32-10, seed 12345,
offset 0) writing N = 2^28 fp32 samples into a preallocated HBM buffer: the
plain bulk path per engine at 2^28 samples.  The JSON line's `value` is that
call's output bandwidth in decimal GB/s; `per_generator` times the other
engine x distribution pairs of the same workload; `roofline` prices the
kernel against MI355X's 8 TB/s HBM peak (and a measured store-kernel ceiling);
`cpu_baseline` times the reference's own kernel bodies on the host CPU over
the same workload.
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


def per_generator_table(R, torch, stream, n, steps):
    """GB/s for each engine x distribution at n samples (same workload as the step)."""
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
            ms = time_calls(lambda: call(g, b), stream, steps, 2)
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


def load_traffic():
    """HBM bytes per launch of the headline kernel, from the committed counter summary."""
    path = os.path.join(HERE, "profiles", "traffic_philox_uniform_f32_2e28.json")
    try:
        with open(path) as f:
            return json.load(f).get("traffic_bytes_per_launch")
    except (OSError, ValueError):
        return None


def cpu_baseline(n):
    """The reference's Philox kernel bodies on one host core (oracle/_ref), else the oracle port."""
    from oracle import oracle as O

    kind = "reference"
    try:
        g = O.RefGen(O.PHILOX4_32_10, seed=12345)
    except Exception:
        g = O.OracleGen(O.PHILOX4_32_10, seed=12345)
        kind = "port"
    g.generate("uniform_f32", 1024, 0x1000)
    t = time.perf_counter()
    st, _ = g.generate("uniform_f32", n, 0x1000)
    dt = time.perf_counter() - t
    assert st == 0
    return {"value": round(n * 4 / dt / 1e9, 4), "unit": "GB/s", "cores": 1, "kind": kind,
            "sample": f"Philox4x32-10 uniform fp32, seed 12345, n=2^{n.bit_length() - 1} on one host core "
                      f"({n / dt / 1e6:.1f} Msamples/s, {dt:.1f} s)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    # ~0.5 s of untimed calls: the first calls after process start run at lower clocks
    ap.add_argument("--warmup", type=int, default=3000)
    ap.add_argument("--n", type=int, default=1 << 28)
    ap.add_argument("--no-table", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=1 << 28)
    ap.add_argument("--calib", action="store_true",
                    help="also run a plain 2^28-float device copy (counter calibration)")
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
    out = torch.empty(n, dtype=torch.float32, device="cuda")
    g = R.PRNG(R.PHILOX4_32_10, seed=12345)
    g.stream = stream
    g.initialize()

    def step():
        g.uniform(out)

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
    table = None if args.no_table else per_generator_table(R, torch, stream, n, 5)
    cpu = None if args.no_cpu else cpu_baseline(args.cpu_n)

    line = {
        "metric": "output GB/s & % of HBM write peak per generator; batched-vs-plain speedup",
        "value": round(value, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
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
        "per_generator": table,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
