"""Synthetic request mixes for the batched and serving paths.

Every draw comes from splitmix64 (Steele, Lea & Flood 2014) so a mix is reproducible
from its seed on any host:

* batch_mix(count, seed=7): `count` independent streams; engine uniform over
  {Philox4x32-10, MRG32k3a}; per-stream seed = successive splitmix64 outputs; offset
  uniform in [0, 2^32); length floor(10^(lo + (hi - lo) u)) (log-uniform, default
  10^3 .. 10^7); kind uniform over {uniform float, normal float (0, 1), Poisson
  lambda = 4}.
* serving_mix(count, rate, seed): open-loop Poisson arrivals at `rate` requests/s;
  sizes log-uniform in [256, 2^26]; engine 50% Philox4x32-10 / 30% MRG32k3a / 20%
  Sobol64 with 16 dimensions; output type uniform over {float, double, unsigned int}.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

from . import rocrand as R

MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self) -> float:
        """[0, 1) with 53 random bits."""
        return (self.next() >> 11) * (1.0 / (1 << 53))

    def below(self, n: int) -> int:
        return self.next() % n


@dataclass
class StreamSpec:
    rng_type: int
    kind: int
    seed: int
    offset: int
    n: int
    a: float
    b: float

    @property
    def itemsize(self) -> int:
        return 8 if self.kind in (R.BATCH_UNIFORM_DOUBLE, R.BATCH_NORMAL_DOUBLE, R.BATCH_LOG_NORMAL_DOUBLE) else 4


BATCH_KINDS = [(R.BATCH_UNIFORM_FLOAT, 0.0, 0.0), (R.BATCH_NORMAL_FLOAT, 0.0, 1.0), (R.BATCH_POISSON, 4.0, 0.0)]


def batch_mix(count: int = 4096, seed: int = 7, lo: float = 3.0, hi: float = 7.0) -> list[StreamSpec]:
    g = SplitMix64(seed)
    out = []
    for _ in range(count):
        rng = R.PHILOX4_32_10 if g.below(2) == 0 else R.MRG32K3A
        s = g.next()
        off = g.next() >> 32
        n = int(math.floor(10.0 ** (lo + (hi - lo) * g.uniform())))
        kind, a, b = BATCH_KINDS[g.below(3)]
        out.append(StreamSpec(rng, kind, s, off, n, a, b))
    return out


@dataclass
class Request:
    arrival_s: float
    rng_type: int
    dtype: str  # "f32", "f64" or "u32"
    n: int
    seed: int
    dims: int = 1
    session: int = 0  # 0: a fresh stream; else the server session it continues


def serving_mix(count: int, rate: float, seed: int = 11) -> list[Request]:
    g = SplitMix64(seed)
    t = 0.0
    out = []
    for _ in range(count):
        t += -math.log(1.0 - g.uniform()) / rate
        u = g.uniform()
        rng = R.PHILOX4_32_10 if u < 0.5 else (R.MRG32K3A if u < 0.8 else R.SOBOL64)
        dtype = ("f32", "f64", "u32")[g.below(3)]
        n = int(math.floor(2.0 ** (8 + 18 * g.uniform())))
        dims = 16 if rng == R.SOBOL64 else 1
        if dims > 1:
            n = max(dims, n - n % dims)
        out.append(Request(t, rng, dtype, n, g.next(), dims))
    return out
