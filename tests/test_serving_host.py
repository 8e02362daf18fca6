"""CPU tests of the serving loop's host logic: the arena allocator and the request mixes."""
import math
import random

from rocrand_amd import rocrand as R
from rocrand_amd.serving import Arena, padded
from rocrand_amd.workloads import SplitMix64, batch_mix, serving_mix


def test_splitmix64_reference_values():
    # splitmix64 from state 0: published first outputs
    g = SplitMix64(0)
    assert [g.next() for _ in range(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_arena_first_fit_and_coalescing():
    a = Arena(1024)
    x = a.alloc(256)
    y = a.alloc(512)
    assert (x, y) == (0, 256)
    assert a.alloc(512) is None
    a.release(x, 256)
    assert a.alloc(300) is None  # free: [0,256) and [768,1024)
    z = a.alloc(256)
    assert z == 0
    a.release(y, 512)
    a.release(z, 256)
    assert a.free == [(0, 1024)]
    assert padded(1) == 256 and padded(256) == 256 and padded(257) == 512


def test_arena_matches_bitmap_model():
    """Random alloc/release sequences: first-fit positions and the free list equal those of
    a per-byte bitmap (free list = maximal free runs)."""
    rs = random.Random(1)
    for _ in range(100):
        size = rs.randrange(1, 300)
        a, used, bitmap = Arena(size), [], [0] * size
        for _ in range(150):
            if used and rs.random() < 0.5:
                s, n = used.pop(rs.randrange(len(used)))
                a.release(s, n)
                bitmap[s:s + n] = [0] * n
            else:
                n = rs.randrange(1, 40)
                want, run = None, 0
                for j in range(size):
                    run = run + 1 if bitmap[j] == 0 else 0
                    if run == n:
                        want = j - n + 1
                        break
                got = a.alloc(n)
                assert got == want
                if got is not None:
                    used.append((got, n))
                    bitmap[got:got + n] = [1] * n
            runs, j = [], 0
            while j < size:
                if bitmap[j] == 0:
                    k = j
                    while k < size and bitmap[k] == 0:
                        k += 1
                    runs.append((j, k - j))
                    j = k
                else:
                    j += 1
            assert a.free == runs


def test_batch_mix_shape():
    specs = batch_mix(4096, seed=7)
    assert len(specs) == 4096
    assert {s.rng_type for s in specs} == {R.PHILOX4_32_10, R.MRG32K3A}
    assert all(1000 <= s.n < 10 ** 7 and 0 <= s.offset < 2 ** 32 for s in specs)
    # log-uniform lengths: the median is near 10^5
    ns = sorted(s.n for s in specs)
    assert 10 ** 4.8 < ns[len(ns) // 2] < 10 ** 5.2
    assert batch_mix(10, seed=7)[3] == specs[3]


def test_serving_mix_shape():
    reqs = serving_mix(20000, 10000.0)
    rate = len(reqs) / reqs[-1].arrival_s
    assert abs(rate - 10000.0) / 10000.0 < 0.05
    frac = {t: sum(r.rng_type == t for r in reqs) / len(reqs) for t in (R.PHILOX4_32_10, R.MRG32K3A, R.SOBOL64)}
    assert abs(frac[R.PHILOX4_32_10] - 0.5) < 0.02 and abs(frac[R.MRG32K3A] - 0.3) < 0.02
    assert all(256 <= r.n <= 2 ** 26 for r in reqs)
    assert all(r.n % r.dims == 0 for r in reqs)
    assert all(r.dims == 16 for r in reqs if r.rng_type == R.SOBOL64)
    assert abs(math.log2(sorted(r.n for r in reqs)[len(reqs) // 2]) - 17) < 0.5
