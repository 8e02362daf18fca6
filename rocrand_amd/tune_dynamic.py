"""Autotuner of the DYNAMIC ordering's launch grids on the current GPU.

    python -m rocrand_amd.tune_dynamic [--n N] [--reps R] [--calls C] [--inc PATH] [--report PATH]

For every engine with a DYNAMIC grid it times candidate grids through the library itself:
each candidate is installed with rocrand_amd_set_dynamic_grid and a fresh generator set to
ROCRAND_ORDERING_PSEUDO_DYNAMIC is initialised (taking that grid for its lifetime).  After
0.5 s of calls that bring the GPU's clocks to their working level, R rounds each time C
calls of every output kind of the bulk path (N values per call, HIP events on the
generators' stream) for every candidate in turn, so that clock drift falls on all
candidates alike; per (candidate, kind) the median over the rounds gives GB/s.  A
candidate's score is the geometric mean over the kinds; the best score wins, except that
the DEFAULT grid is kept unless a candidate beats it by more than 2% (so that DYNAMIC
output stays DEFAULT output where the grid makes no measurable difference).

Candidates: the interleaved engines (MRG32k3a, MRG31k3p, XORWOW, LFSR113) run one engine
per lane, so their engine count is their parallelism: 256 x {256, 512, 1024, 2048, 4096}.
The counter engines' kernels cover any call with one vec per lane, so their grid only moves
the reference's head shift; they are timed at the DEFAULT grid and twice it, for the record.

Writes the C table the library compiles in (--inc, default rocrand_amd/csrc/rr_dynamic_gfx950.inc)
and a markdown report of every measurement (--report).
"""
from __future__ import annotations

import argparse
import math
import os
import statistics
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)

DEFAULT_INC = os.path.join(HERE, "csrc", "rr_dynamic_gfx950.inc")

INTERLEAVED = ["MRG32K3A", "MRG31K3P", "XORWOW", "LFSR113"]
COUNTER = ["PHILOX4_32_10", "THREEFRY2_32_20", "THREEFRY2_64_20", "THREEFRY4_32_20", "THREEFRY4_64_20"]
ENUM = {"PHILOX4_32_10": "ROCRAND_RNG_PSEUDO_PHILOX4_32_10", "THREEFRY2_32_20": "ROCRAND_RNG_PSEUDO_THREEFRY2_32_20",
        "THREEFRY2_64_20": "ROCRAND_RNG_PSEUDO_THREEFRY2_64_20",
        "THREEFRY4_32_20": "ROCRAND_RNG_PSEUDO_THREEFRY4_32_20",
        "THREEFRY4_64_20": "ROCRAND_RNG_PSEUDO_THREEFRY4_64_20", "MRG32K3A": "ROCRAND_RNG_PSEUDO_MRG32K3A",
        "MRG31K3P": "ROCRAND_RNG_PSEUDO_MRG31K3P", "XORWOW": "ROCRAND_RNG_PSEUDO_XORWOW",
        "LFSR113": "ROCRAND_RNG_PSEUDO_LFSR113"}


def candidates(name):
    if name in INTERLEAVED:
        return [(256, b) for b in (256, 512, 1024, 2048, 4096)], (256, 512)
    return [(256, 1024), (256, 2048)], (256, 1024)


def kinds(torch):
    return [("u32", torch.uint32, lambda g, b: g.generate(b)),
            ("uniform_f32", torch.float32, lambda g, b: g.uniform(b)),
            ("uniform_f64", torch.float64, lambda g, b: g.uniform(b)),
            ("normal_f32", torch.float32, lambda g, b: g.normal(b, 0.0, 1.0)),
            ("normal_f64", torch.float64, lambda g, b: g.normal(b, 0.0, 1.0)),
            ("lognormal_f32", torch.float32, lambda g, b: g.lognormal(b, 0.0, 1.0)),
            ("poisson_10", torch.uint32, lambda g, b: g.poisson(b, 10.0))]


def timed(torch, stream, fn, calls):
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(calls):
        fn()
    end.record(stream)
    torch.cuda.synchronize()
    return start.elapsed_time(end) / calls


def measure(R, torch, stream, rng, grids, n, reps, calls, bufs):
    """{grid: {kind: GB/s}}: medians over `reps` rounds, candidates interleaved per round."""
    gens = {}
    for grid in grids:
        R.set_dynamic_grid(rng, *grid)
        g = R.PRNG(rng, seed=12345)
        g.stream = stream
        g.set_ordering(R.ORDERING_PSEUDO_DYNAMIC)
        g.initialize()
        gens[grid] = g
    R.set_dynamic_grid(rng, 0, 0)
    ks = kinds(torch)
    for _, dt, _ in ks:
        bufs.setdefault(dt, torch.empty(n, dtype=dt, device="cuda"))
    times = {grid: {k: [] for k, _, _ in ks} for grid in grids}
    for _ in range(reps):
        for grid, g in gens.items():
            for kname, dt, call in ks:
                b = bufs[dt]
                call(g, b)  # the kind's first call of the round is not timed
                times[grid][kname].append(timed(torch, stream, lambda: call(g, b), calls))
    out = {}
    for grid in grids:
        out[grid] = {k: n * bufs[dt].element_size() / (statistics.median(times[grid][k]) * 1e-3) / 1e9
                     for k, dt, _ in ks}
    return out


def settle(R, torch, stream, n, bufs, seconds=0.5):
    g = R.PRNG(R.PHILOX4_32_10, seed=1)
    g.stream = stream
    b = bufs.setdefault(torch.float32, torch.empty(n, dtype=torch.float32, device="cuda"))
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        for _ in range(20):
            g.uniform(b)
        torch.cuda.synchronize()


def geomean(row):
    return math.exp(sum(math.log(v) for v in row.values()) / len(row))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=1 << 26)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--calls", type=int, default=20)
    ap.add_argument("--inc", default=DEFAULT_INC)
    ap.add_argument("--report", default=None)
    args = ap.parse_args(argv)

    import torch

    from rocrand_amd import rocrand as R

    arch = torch.cuda.get_device_properties(0).gcnArchName.split(":")[0]
    stream = torch.cuda.Stream()
    bufs = {}
    results, chosen = {}, {}
    settle(R, torch, stream, args.n, bufs)
    for name in COUNTER + INTERLEAVED:
        rng = getattr(R, name)
        cands, default = candidates(name)
        rows = measure(R, torch, stream, rng, cands, args.n, args.reps, args.calls, bufs)
        scores = {}
        for grid in cands:
            row = rows[grid]
            scores[grid] = (geomean(row), row)
            print(f"# {name} {grid[0]}x{grid[1]}: geomean {scores[grid][0]:.1f} GB/s "
                  + " ".join(f"{k}={v:.0f}" for k, v in row.items()), flush=True)
        best = max(scores, key=lambda gr: scores[gr][0])
        if scores[best][0] <= 1.02 * scores[default][0]:
            best = default
        chosen[name] = best
        results[name] = scores

    stamp = time.strftime("%Y-%m-%d")
    lines = [f"// DYNAMIC-ordering launch grids for {arch} (threads, blocks per engine), written by",
             f"// rocrand_amd/tune_dynamic.py on {stamp} (--n {args.n}, median of {args.reps} interleaved rounds of",
             f"// {args.calls} calls per output kind; geometric mean of GB/s over the kinds; DEFAULT kept",
             "// unless beaten by 2%)."]
    for name in COUNTER + INTERLEAVED:
        t, b = chosen[name]
        lines.append(f"{{{ENUM[name]}, {t}, {b}}}, // {results[name][(t, b)][0]:.0f} GB/s")
    with open(args.inc, "w") as f:
        f.write("\n".join(lines) + "\n")
    print(f"wrote {args.inc}")

    if args.report:
        rep = [f"# DYNAMIC-ordering grid sweep on {arch}", "",
               f"Command: `python -m rocrand_amd.tune_dynamic --n {args.n} --reps {args.reps} --calls {args.calls} "
               f"--inc {args.inc} --report {args.report}` ({stamp}).",
               "GB/s per output kind (median of the interleaved rounds), geometric mean, and the grid chosen "
               "(DEFAULT kept unless beaten by more than 2%).", ""]
        for name in COUNTER + INTERLEAVED:
            knames = list(next(iter(results[name].values()))[1])
            rep.append(f"## {name} (chosen {chosen[name][0]} x {chosen[name][1]})")
            rep.append("")
            rep.append("| grid | engines / stride | geomean | " + " | ".join(knames) + " |")
            rep.append("|---|---|---|" + "---|" * len(knames))
            for grid, (gm, row) in results[name].items():
                rep.append(f"| {grid[0]} x {grid[1]} | {grid[0] * grid[1]} | {gm:.0f} | "
                           + " | ".join(f"{row[k]:.0f}" for k in knames) + " |")
            rep.append("")
        with open(args.report, "w") as f:
            f.write("\n".join(rep) + "\n")
        print(f"wrote {args.report}")


if __name__ == "__main__":
    main()
