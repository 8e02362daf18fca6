"""CPU checks that pin the oracle (oracle/liboracle.so) before it is trusted as the checker.

Each test fixes a part of the restatement by something outside it:
  * published known-answer vectors (Random123 Philox4x32-10, Threefry2x32/2x64),
  * published jump matrices (L'Ecuyer's RngStreams A^(2^76), A^(2^127) for MRG32k3a),
  * brute force (jump-ahead by k == k single steps),
  * closed forms (uniform scalings, Poisson pmf recovered from the alias table),
  * FP64 Box-Muller on the same inputs,
  * invariants of the reference's index maps (offsets, continuity),
  * and, where oracle/_ref/libref.so is built, the reference's own kernel bodies.
"""
import math
import os

import numpy as np
import pytest

from conftest import ulp_distance
from oracle import oracle as O

COUNTER = [O.PHILOX4_32_10, O.THREEFRY2_32_20, O.THREEFRY4_32_20, O.THREEFRY2_64_20, O.THREEFRY4_64_20]
INTERLEAVED = [O.MRG32K3A, O.MRG31K3P, O.XORWOW, O.LFSR113]


def _block(fn, ctr, key, w64=False, n=4):
    t = np.uint64 if w64 else np.uint32
    c, k, o = np.array(ctr, dtype=t), np.array(key, dtype=t), np.zeros(n, dtype=t)
    getattr(O.lib(), fn)(c.ctypes.data, k.ctypes.data, o.ctypes.data)
    return [int(x) for x in o]


PI32 = [0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344, 0xA4093822, 0x299F31D0]


def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32 10 rounds.
    f = "orc_philox4x32_10_block"
    assert _block(f, [0] * 4, [0] * 2) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert _block(f, [0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert _block(f, PI32[:4], PI32[4:6]) == [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_threefry2_known_answers():
    # Random123 kat_vectors, threefry2x32 / threefry2x64 with 20 rounds.
    f = "orc_threefry2x32_20_block"
    assert _block(f, [0, 0], [0, 0], n=2) == [0x6B200159, 0x99BA4EFE]
    assert _block(f, [0xFFFFFFFF] * 2, [0xFFFFFFFF] * 2, n=2) == [0x1CB996FC, 0xBB002BE7]
    assert _block(f, PI32[:2], PI32[2:4], n=2) == [0xC4923A9C, 0x483DF7A0]
    assert _block("orc_threefry2x64_20_block", [0, 0], [0, 0], w64=True, n=2) == [0xC2B6E3A8C2C69865,
                                                                                0x6F81ED42F350084D]


_TF4_ROT = {32: [(10, 26), (11, 21), (13, 27), (23, 5), (6, 20), (17, 11), (25, 10), (18, 20)],
            64: [(14, 16), (52, 57), (23, 40), (5, 37), (25, 33), (46, 12), (58, 22), (32, 32)]}
_TF_PARITY = {32: 0x1BD11BDA, 64: 0x1BD11BDAA9FC1A22}


def _threefry4_block(ctr, key, w, pair_switch_bit):
    """Threefry-4xW with 20 rounds, Skein key schedule (injection every 4 rounds).

    Word pairs (0,1),(2,3) mix in a round whose `pair_switch_bit` bit is clear and (0,3),(2,1)
    otherwise.  Random123 alternates every round (bit 1); the reference switches every two
    rounds (`(round_idx & 2u) == 0`, rocrand_threefry4_impl.h:210-229).
    """
    m = (1 << w) - 1
    ks = list(key) + [_TF_PARITY[w] ^ key[0] ^ key[1] ^ key[2] ^ key[3]]
    x = [(c + k) & m for c, k in zip(ctr, key)]

    def rotl(v, r):
        return ((v << r) | (v >> (w - r))) & m

    for r in range(20):
        a, b = _TF4_ROT[w][r & 7]
        if not r & pair_switch_bit:
            x[0] = (x[0] + x[1]) & m
            x[1] = rotl(x[1], a) ^ x[0]
            x[2] = (x[2] + x[3]) & m
            x[3] = rotl(x[3], b) ^ x[2]
        else:
            x[0] = (x[0] + x[3]) & m
            x[3] = rotl(x[3], a) ^ x[0]
            x[2] = (x[2] + x[1]) & m
            x[1] = rotl(x[1], b) ^ x[2]
        if r & 3 == 3:
            i = r // 4
            x = [(x[j] + ks[(1 + i + j) % 5]) & m for j in range(4)]
            x[3] = (x[3] + 1 + i) & m
    return x


PI64 = [0x243F6A8885A308D3, 0x13198A2E03707344, 0xA4093822299F31D0, 0x082EFA98EC4E6C89,
        0x452821E638D01377, 0xBE5466CF34E90C6C, 0xC0AC29B7C97C50DD, 0x3F84D5B5B5470917]
# Random123 kat_vectors, threefry4x32 / threefry4x64 with 20 rounds.
_TF4_KATS = [
    (32, [0] * 4, [0] * 4, [0x9C6CA96A, 0xE17EAE66, 0xFC10ECD4, 0x5256A7D8]),
    (32, [0xFFFFFFFF] * 4, [0xFFFFFFFF] * 4, [0x2A881696, 0x57012287, 0xF6C7446E, 0xA16A6732]),
    (32, PI32[:4], PI32[4:6] + [0x082EFA98, 0xEC4E6C89], [0x59CD1DBB, 0xB8879579, 0x86B5D00C, 0xAC8B6D84]),
    (64, [0] * 4, [0] * 4, [0x09218EBDE6C85537, 0x55941F5266D86105, 0x4BD25E16282434DC, 0xEE29EC846BD2E40B]),
    (64, [2**64 - 1] * 4, [2**64 - 1] * 4,
     [0x29C24097942BBA1B, 0x0371BBFB0F6F4E11, 0x3C231FFA33F83A1C, 0xCD29113FDE32D168]),
]


@pytest.mark.parametrize("w,ctr,key,kat", _TF4_KATS)
def test_threefry4_rounds_pinned(w, ctr, key, kat):
    # The restatement above reproduces Random123's published vectors with Random123's
    # pairing, which fixes the rotation constants, the parity word and the key injection;
    # with the reference's pairing it must then equal the oracle's block function.  A
    # wrong rotation, injection index or pairing fails one of the two.
    assert _threefry4_block(ctr, key, w, 1) == kat
    fn = "orc_threefry4x32_20_block" if w == 32 else "orc_threefry4x64_20_block"
    assert _block(fn, ctr, key, w64=(w == 64)) == _threefry4_block(ctr, key, w, 2)
    rng = np.random.default_rng(w)
    for _ in range(8):
        c = [int(v) for v in rng.integers(0, 2**w, 4, dtype=np.uint64 if w == 64 else np.uint32)]
        k = [int(v) for v in rng.integers(0, 2**w, 4, dtype=np.uint64 if w == 64 else np.uint32)]
        assert _block(fn, c, k, w64=(w == 64)) == _threefry4_block(c, k, w, 2)


def test_philox_engine_stream_is_counter_blocks():
    # word k of engine(seed) = block(counter = k/4, key = (seed lo, seed hi))[k % 4]
    seed = 0x0123456789ABCDEF
    s = O.engine_stream(O.PHILOX4_32_10, seed, 0, 0, 64)
    for b in range(16):
        assert list(s[4 * b:4 * b + 4]) == _block("orc_philox4x32_10_block", [b, 0, 0, 0],
                                                  [seed & 0xFFFFFFFF, seed >> 32])


def _mat(t, comp, which, idx):
    o = np.zeros(9, dtype=np.uint64)
    O.lib().orc_mrg_matrix(t, comp, which, idx, o.ctypes.data)
    return o.reshape(3, 3).tolist()


def test_mrg32k3a_published_jump_matrices():
    # L'Ecuyer, Simard, Chen, Kelton (2002), RngStreams: A1^(2^76), A2^(2^76), A1^(2^127), A2^(2^127).
    assert _mat(O.MRG32K3A, 1, 1, 0) == [[82758667, 1871391091, 4127413238], [3672831523, 69195019, 1871391091],
                                         [3672091415, 3528743235, 69195019]]
    assert _mat(O.MRG32K3A, 2, 1, 0) == [[1511326704, 3759209742, 1610795712], [4292754251, 1511326704, 3889917532],
                                         [3859662829, 4292754251, 3708466080]]
    assert _mat(O.MRG32K3A, 1, 2, 0) == [[2427906178, 3580155704, 949770784], [226153695, 1230515664, 3580155704],
                                         [1988835001, 986791581, 1230515664]]
    assert _mat(O.MRG32K3A, 2, 2, 0) == [[1464411153, 277697599, 1610723613], [32183930, 1464411153, 1022607788],
                                         [2824425944, 32183930, 2093834863]]


@pytest.mark.parametrize("rng", INTERLEAVED + [O.PHILOX4_32_10, O.THREEFRY4_32_20, O.THREEFRY2_32_20])
@pytest.mark.parametrize("k", [1, 3, 64, 1000, 65537])
def test_jump_equals_stepping(rng, k):
    seed = 0x1234567 if rng != O.LFSR113 else 0x0123_4567_89AB_CDEF
    base = O.engine_stream(rng, seed, 0, 0, k + 50)
    jumped = O.engine_stream(rng, seed, 0, k, 50)
    assert np.array_equal(base[k:], jumped)


@pytest.mark.parametrize("rng", COUNTER + INTERLEAVED)
@pytest.mark.parametrize("k", [1, 4, 11, 65536, 112233])
def test_offset_invariant(rng, k):
    # gen(offset=k)[i] == gen(offset=0)[i+k] (test_rocrand_prng.hpp:703-745).
    kind = "u64" if rng in (O.THREEFRY2_64_20, O.THREEFRY4_64_20) else "u32"
    n = 5000
    _, a = O.OracleGen(rng, seed=77).generate(kind, n + k, 0x1000)
    _, b = O.OracleGen(rng, seed=77, offset=k).generate(kind, n, 0x1000)
    assert np.array_equal(a[k:], b)


@pytest.mark.parametrize("rng", COUNTER + INTERLEAVED)
def test_continuity(rng):
    # Same seed, different chunkings -> identical prefix (test_rocrand_prng.hpp:502-562).
    # The 64-bit ThreeFry engines emit two u32 per word and drop the spare half of an
    # odd-sized call, so their continuity holds on their native 64-bit output.
    kind = "u64" if rng in (O.THREEFRY2_64_20, O.THREEFRY4_64_20) else "u32"
    chunks_a = [100, 1, 24783, 3, 2, 77654]
    chunks_b = [1024, 55, 65536, 623, 30, 11111]
    total = min(sum(chunks_a), sum(chunks_b))
    outs = []
    for chunks in (chunks_a, chunks_b):
        g = O.OracleGen(rng, seed=5)
        parts = [g.generate(kind, c, 0x1000)[1] for c in chunks]
        outs.append(np.concatenate(parts)[:total])
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("rng", [O.PHILOX4_32_10, O.XORWOW, O.THREEFRY4_32_20])
def test_uniform_closed_form(rng):
    g1, g2 = O.OracleGen(rng, seed=9), O.OracleGen(rng, seed=9)
    _, u = g1.generate("u32", 4096, 0x1000)
    _, f = g2.generate("uniform_f32", 4096, 0x1000)
    inv = np.float32(2.0 ** -32)
    assert np.array_equal(f, inv + u.astype(np.float32) * inv)
    if rng == O.XORWOW:
        return  # an interleaved engine draws both words of a double from one engine
    g3, g4 = O.OracleGen(rng, seed=9), O.OracleGen(rng, seed=9)
    _, u = g3.generate("u32", 4096, 0x1000)
    _, d = g4.generate("uniform_f64", 2048, 0x1000)
    v = u[0::2].astype(np.uint64) | ((u[1::2].astype(np.uint64) >> np.uint64(11)) << np.uint64(32))
    assert np.array_equal(d, 2.0 ** -53 + v.astype(np.float64) * 2.0 ** -53)


def test_mrg_uniform_closed_form():
    g1, g2 = O.OracleGen(O.MRG32K3A, seed=3), O.OracleGen(O.MRG32K3A, seed=3)
    raw = O.engine_stream(O.MRG32K3A, 3, 0, 0, 1)  # engine 0, first value
    _, f = g2.generate("uniform_f64", 1, 0x1000)
    assert f[0] == raw[0] * (1.0 / 4294967087.0) or abs(f[0] - raw[0] / 4294967087.0) < 1e-18
    _, u = g1.generate("u32", 1, 0x1000)
    assert u[0] == int((int(raw[0]) - 1) * 1.000000048661607)


def test_normal_against_fp64_box_muller():
    g = O.OracleGen(O.PHILOX4_32_10, seed=21)
    _, w = O.OracleGen(O.PHILOX4_32_10, seed=21).generate("u32", 200000, 0x1000)
    _, z = g.generate("normal_f32", 200000, 0x1000, 0.0, 1.0)
    f32 = np.float32
    u = (f32(2.3283064e-10) + w[0::2].astype(f32) * f32(2.3283064e-10)).astype(np.float64)
    v = (f32(1.46291807e-09) + w[1::2].astype(f32) * f32(1.46291807e-09)).astype(np.float64)
    s = np.sqrt(-2.0 * np.log(u))
    exact = np.empty(z.size)
    exact[0::2] = s * np.sin(v)
    exact[1::2] = s * np.cos(v)
    d = ulp_distance(z, exact.astype(np.float32))
    assert d.max() <= 3 and d.mean() < 0.5
    # moments of the standard normal
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1.0) < 0.01


@pytest.mark.parametrize("lam", [0.5, 4.0, 10.0, 1000.0, 3999.0])
def test_poisson_alias_table_recovers_pmf(lam):
    off, prob, alias = O.poisson_table(lam)
    size = prob.size
    mass = prob.copy()
    for j in range(size):
        mass[alias[j]] += 1.0 - prob[j]
    mass /= size
    k = np.arange(off, off + size, dtype=np.float64)
    pmf = np.exp(k * math.log(lam) - np.array([math.lgamma(x + 1.0) for x in k]) - lam)
    pmf /= pmf.sum()
    assert np.max(np.abs(mass - pmf)) < 1e-12
    # truncation at 1e-12 keeps essentially all the mass
    full = sum(math.exp(x * math.log(lam) - math.lgamma(x + 1.0) - lam) for x in range(0, int(lam * 3 + 60)))
    kept = sum(math.exp(x * math.log(lam) - math.lgamma(x + 1.0) - lam) for x in range(off, off + size))
    assert abs(full - kept) < 1e-9


@pytest.mark.parametrize("rng", [O.PHILOX4_32_10, O.MRG32K3A])
def test_poisson_moments(rng):
    _, x = O.OracleGen(rng, seed=4).generate("poisson", 400000, 0x1000, 10.0)
    assert abs(x.mean() - 10.0) < 0.03 and abs(x.var() - 10.0) < 0.1


def test_float2half_round_to_nearest_even():
    L = O.lib()
    vals = np.array([0.0, 1.0, -2.5, 65504.0, 65520.0, 1e-8, 6.1e-5, 3.0e-5, 0.333333, 1.0009765625, 1.00048828125],
                    dtype=np.float32)
    got = np.array([L.orc_float2half(float(v)) for v in vals], dtype=np.uint16)
    with np.errstate(over="ignore"):
        want = vals.astype(np.float16).view(np.uint16)
    assert np.array_equal(got, want)


# --------------------------------------------------------------- vs reference
REF = O.ref() is not None
ref_only = pytest.mark.skipif(not REF, reason="oracle/_ref/libref.so not built (needs the reference sources)")

REF_KINDS = ["u8", "u32", "uniform_f32", "uniform_f64", "uniform_f16", "normal_f32", "normal_f64",
             "lognormal_f32", "normal_f16", "poisson"]


@ref_only
@pytest.mark.parametrize("rng", COUNTER + INTERLEAVED)
@pytest.mark.parametrize("kind", REF_KINDS)
def test_oracle_equals_reference_kernel_bodies(rng, kind):
    esz = np.dtype(O.KINDS[kind][1]).itemsize
    for seed, off, n, shift in [(12345, 1048583, 1001, 1), (0xAAAAAAAAAAA, 0, 33, 3)]:
        addr = 0x1000 + shift * esz
        a, b = (10.0, 0.0) if kind == "poisson" else (0.25, 1.5)
        g1, g2 = O.OracleGen(rng, seed=seed, offset=off), O.RefGen(rng, seed=seed, offset=off)
        for _ in range(2):
            s1, x1 = g1.generate(kind, n, addr, a, b)
            s2, x2 = g2.generate(kind, n, addr, a, b)
            assert s1 == s2 == 0
            assert x1.tobytes() == x2.tobytes()


@ref_only
@pytest.mark.parametrize("lam", [0.01, 0.5, 1.0, 4.0, 10.0, 63.5, 64.0, 100.0, 999.9, 1000.0, 2500.0, 3999.5, 4000.0])
def test_poisson_alias_layout_equals_reference(lam):
    # The alias table's layout (which column holds which alias, the LIFO order of the
    # small/large stacks) is not fixed by the pmf; the reference's own construction
    # (poisson.hpp:110-161, discrete.hpp:189-286) fixes it.
    off, prob, alias = O.poisson_table(lam)
    roff, rprob, ralias, rcdf = O.ref_poisson_tables(lam)
    assert off == roff
    assert prob.tobytes() == rprob.tobytes()
    assert np.array_equal(alias, ralias)
    assert rcdf[-1] == pytest.approx(1.0, abs=1e-12)


def _weights(kind, size, rng):
    if kind == "uniform":
        return rng.random(size)
    if kind == "skewed":
        return rng.random(size) ** 8
    if kind == "sparse":
        w = rng.random(size)
        w[rng.random(size) < 0.5] = 0.0
        w[0] = max(w[0], 1e-3)
        return w
    return np.full(size, 3.0)  # all equal


@ref_only
@pytest.mark.parametrize("size", [1, 2, 7, 100, 1000, 4096])
@pytest.mark.parametrize("kind", ["uniform", "skewed", "sparse", "equal"])
def test_discrete_tables_equal_reference(size, kind):
    w = _weights(kind, size, np.random.default_rng(size))
    prob, alias, cdf = O.discrete_tables(w)
    rprob, ralias, rcdf = O.ref_discrete_tables(w, 5)
    assert prob.tobytes() == rprob.tobytes()
    assert np.array_equal(alias, ralias)
    assert cdf.tobytes() == rcdf.tobytes()


@pytest.mark.parametrize("size", [1, 3, 100, 4096])
def test_discrete_tables_recover_weights(size):
    # closed form: the alias pair puts mass prob[j] on column j and 1 - prob[j] on alias[j]
    w = _weights("skewed", size, np.random.default_rng(7 + size))
    prob, alias, cdf = O.discrete_tables(w)
    mass = prob.copy()
    np.add.at(mass, alias, 1.0 - prob)
    assert np.max(np.abs(mass / size - w / w.sum())) < 1e-12
    assert np.max(np.abs(cdf - np.cumsum(w / w.sum()))) < 1e-12
    assert np.all((prob >= 0) & (prob <= 1.0)) and np.all(alias < size)


@ref_only
@pytest.mark.parametrize("rng", [O.PHILOX4_32_10, O.THREEFRY4_32_20, O.MRG32K3A, O.MRG31K3P, O.XORWOW, O.LFSR113])
def test_engine_streams_equal_reference(rng):
    for seed, sub, off in [(1, 0, 0), (0xDEADBEEF, 3, 1000), (12345, 131071, 7)]:
        a = O.engine_stream(rng, seed, sub, off, 256)
        b = O.engine_stream(rng, seed, sub, off, 256, use_ref=True)
        assert np.array_equal(a, b)


# Tuned launch grids (DYNAMIC ordering): the engine count of the interleaved engines and
# the counter engines' grid stride follow threads x blocks (config_types.hpp:314-336,
# mrg.hpp:284-337, philox4x32_10.hpp:106-214).  Sizes where the grid matters: enough
# outputs to wrap the engine count or the stride, misaligned and odd lengths.
GRIDS = [(256, 256), (512, 512)]


@ref_only
@pytest.mark.parametrize("rng", COUNTER + INTERLEAVED)
@pytest.mark.parametrize("grid", GRIDS)
def test_oracle_grid_equals_reference_kernel_bodies(rng, grid):
    size = grid[0] * grid[1]
    for kind, seed, off, n, shift in [("u32", 12345, 1048583, size + 77, 1),
                                      ("normal_f32", 7, 3 * size + 5, size + 3, 2),
                                      ("uniform_f64", 99, 17, 2 * size + 1, 1),
                                      ("u32", 3, 0, 4 * size + 13, 3)]:  # past the stride's multiples
        addr = 0x1000 + shift * np.dtype(O.KINDS[kind][1]).itemsize
        g1 = O.OracleGen(rng, seed=seed, offset=off, grid=grid)
        g2 = O.RefGen(rng, seed=seed, offset=off, grid=grid)
        for _ in range(2):
            s1, x1 = g1.generate(kind, n, addr, 0.25, 1.5)
            s2, x2 = g2.generate(kind, n, addr, 0.25, 1.5)
            assert s1 == s2 == 0
            assert x1.tobytes() == x2.tobytes(), (rng, grid, kind, seed)


@pytest.mark.parametrize("rng", INTERLEAVED)
def test_oracle_grid_changes_the_interleaving(rng):
    """A different engine count is a different stream (the grid is not value-neutral for
    the interleaved engines), and the default grid's output is unchanged by set_grid with
    the default (512 x 256)."""
    n = 300000
    _, a = O.OracleGen(rng, seed=5).generate("u32", n)
    _, b = O.OracleGen(rng, seed=5, grid=(512, 256)).generate("u32", n)
    _, c = O.OracleGen(rng, seed=5, grid=(256, 1024)).generate("u32", n)
    assert np.array_equal(a, b)
    assert np.array_equal(a[:131072], c[:131072])  # engines below 131072 agree at step 0
    assert not np.array_equal(a[131072:262144], c[131072:262144])
