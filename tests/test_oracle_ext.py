"""CPU checks pinning the oracle's Sobol and MTGP32 restatements (and the library's
host-side Sobol tables).

Pins, each chosen so that a plausible mistake fails one of them:
  * direction vectors: the reference's own 32-bit tables (digests in
    tests/golden/sobol_tables.json, made by tests/golden/make_sobol_golden.py) and
    scipy's independent Joe-Kuo implementation for the 64-bit ones;
  * point order: scipy.stats.qmc.Sobol's unscrambled points, which walk the same
    Gray-code sequence (a wrong Gray map or bit order fails);
  * scramble constants: digests of the reference's constant tables, first value
    0xd091bb5c = first output of a default-seeded std::mt19937;
  * generation (index map, head/tail, every distribution): the reference kernel
    bodies (library/src/rng/sobol.hpp:72-261, mtgp32.hpp:232-305) driven on the CPU by
    oracle/_ref, where that library is built.
"""
import ctypes
import hashlib
import json
import os

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "sobol_tables.json")))

SOBOL_TYPES = [O.SOBOL32, O.SCRAMBLED_SOBOL32, O.SOBOL64, O.SCRAMBLED_SOBOL64]
ALL_KINDS = list(O.KINDS)


def _digest(a):
    return hashlib.sha256(a.astype(a.dtype.newbyteorder("<")).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def tables():
    return O.sobol_tables()


def test_sobol32_vectors_match_reference_tables(tables):
    v32, sv32 = tables[0], tables[1]
    assert [int(x) for x in v32[32:40]] == GOLD["sobol32_vectors_dim1_head"]
    assert _digest(v32) == GOLD["sobol32_vectors_sha256"]
    assert _digest(sv32) == GOLD["scrambled_sobol32_vectors_sha256"]


def test_scramble_constants_match_reference_tables(tables):
    c32, c64 = tables[4], tables[5]
    assert c32[0] == 0xD091BB5C
    assert _digest(c32) == GOLD["scramble32_constants_sha256"]
    assert _digest(c64) == GOLD["scramble64_constants_sha256"]


def test_scrambled_vectors_are_offset_tables(tables):
    v32, sv32, v64, sv64 = tables[:4]
    assert np.array_equal(sv32, v32 + np.uint32(640000))
    assert np.array_equal(sv64, v64 + np.uint64(1280000))


def test_dimension_zero_is_van_der_corput(tables):
    v32, v64 = tables[0], tables[2]
    assert np.array_equal(v32[:32], np.uint32(1) << np.arange(31, -1, -1, dtype=np.uint32))
    assert np.array_equal(v64[:64], np.uint64(1) << np.arange(63, -1, -1, dtype=np.uint64))


@pytest.mark.parametrize("bits", [32, 64])
def test_direction_vectors_match_scipy(tables, bits):
    qmc = pytest.importorskip("scipy.stats.qmc")
    d = 20000
    sv = np.asarray(qmc.Sobol(d=d, scramble=False, bits=bits)._sv, dtype=np.uint64)
    mine = (tables[0] if bits == 32 else tables[2]).reshape(20000, bits)[:d].astype(np.uint64)
    assert np.array_equal(sv, mine)


@pytest.mark.parametrize("dims", [1, 5, 37])
def test_sobol32_points_match_scipy(dims):
    qmc = pytest.importorskip("scipy.stats.qmc")
    n = 1 << 10
    pts = qmc.Sobol(d=dims, scramble=False, bits=32).random(n)
    want = np.round(pts * 2.0 ** 32).astype(np.uint64).astype(np.uint32).T.reshape(-1)
    g = O.OracleGen(O.SOBOL32, dims=dims)
    st, got = g.generate("u32", n * dims, 0x1000)
    assert st == 0
    assert np.array_equal(got, want)


def test_sobol64_points_match_scipy():
    qmc = pytest.importorskip("scipy.stats.qmc")
    n, dims = 1 << 9, 6
    # scipy returns its 64-bit integer points scaled by 2^-64 in float64
    pts = qmc.Sobol(d=dims, scramble=False, bits=64).random(n)
    g = O.OracleGen(O.SOBOL64, dims=dims)
    st, got = g.generate("u64", n * dims, 0x1000)
    assert st == 0
    assert np.array_equal(got.astype(np.float64) * 2.0 ** -64, pts.T.reshape(-1))


def test_sobol_offset_and_continuity():
    """Offset k then n points equals points k.. of one long call; two calls continue."""
    dims, n = 3, 300
    a = O.OracleGen(O.SOBOL32, dims=dims)
    st, full = a.generate("u32", dims * (n + 100), 0x1000)
    full = full.reshape(dims, n + 100)
    b = O.OracleGen(O.SOBOL32, dims=dims, offset=100)
    st1, x1 = b.generate("u32", dims * 200, 0x1000)
    st2, x2 = b.generate("u32", dims * 100, 0x1000)
    assert st1 == st2 == 0
    assert np.array_equal(x1.reshape(dims, 200), full[:, 100:300])
    assert np.array_equal(x2.reshape(dims, 100), full[:, 300:400])


def test_sobol_status_codes():
    g = O.OracleGen(O.SOBOL32, dims=3)
    assert g.generate("u32", 10, 0x1000)[0] == 105  # LENGTH_NOT_MULTIPLE
    assert g.generate("u64", 9, 0x1000)[0] == 103   # TYPE_ERROR on 32-bit Sobol
    assert g.set_dimensions(0) == 104 and g.set_dimensions(20001) == 104
    assert g.set_seed(5) == 0


def test_mtgp32_offset_is_type_error():
    g = O.OracleGen(O.MTGP32)
    assert g.set_offset(5) == 103


def _same(kind, x, y):
    if x.dtype.kind == "f":
        return np.array_equal(x.view(np.uint8), y.view(np.uint8))
    return np.array_equal(x, y)


@pytest.mark.parametrize("rng", SOBOL_TYPES)
def test_sobol_oracle_equals_reference_kernel(rng):
    if O.ref() is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    for dims, off, n, addr in [(1, 0, 1000, 0x1000), (3, 1000, 3 * 1001, 0x1002), (7, 5, 7 * 37, 0x1001)]:
        og = O.OracleGen(rng, dims=dims, offset=off)
        rg = O.RefGen(rng, dims=dims, offset=off)
        for kind in ALL_KINDS:
            for lam in ((4.0, 5000.0) if kind == "poisson" else (None,)):
                a, b = (lam, 0.0) if kind == "poisson" else ((0.5, 1.5) if "normal" in kind else (0.0, 1.0))
                al = np.dtype(O.KINDS[kind][1]).itemsize
                ad = addr - addr % al
                s1, x = og.generate(kind, n, ad, a, b)
                s2, y = rg.generate(kind, n, ad, a, b)
                assert s1 == s2, (rng, kind)
                if s1 == 0:
                    assert _same(kind, x, y), (rng, dims, kind)


@pytest.mark.parametrize("seed", [0, 12345, 0xFEDCBA9876543210])
def test_mtgp32_oracle_equals_reference_kernel(seed):
    if O.ref() is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    og = O.OracleGen(O.MTGP32, seed=seed)
    rg = O.RefGen(O.MTGP32, seed=seed)
    # sizes around the 256-value block step and the 512 x 256 stride; odd addresses
    for n, addr in [(1000, 0x1000), (131072 + 777, 0x1002), (13, 0x1001), (262144 * 2 + 5, 0x1003)]:
        for kind in ALL_KINDS:
            if kind == "u64":
                continue
            a, b = (10.0, 0.0) if kind == "poisson" else ((0.5, 1.5) if "normal" in kind else (0.0, 1.0))
            al = np.dtype(O.KINDS[kind][1]).itemsize
            ad = addr - addr % al
            s1, x = og.generate(kind, n, ad, a, b)
            s2, y = rg.generate(kind, n, ad, a, b)
            assert s1 == s2 == 0, kind
            assert _same(kind, x, y), (seed, n, kind)


def test_library_sobol_tables_equal_oracle(tables):
    """librocrand's rocrand_get_direction_vectors*/scramble_constants* (host tables, no GPU)."""
    from rocrand_amd import rocrand as R

    L = R.load()
    v32, sv32, v64, sv64, c32, c64 = tables

    def get(fn, ct, n, *a):
        p = ctypes.POINTER(ct)()
        st = fn(ctypes.byref(p), *a)
        return st, (np.ctypeslib.as_array(p, (n,)).copy() if st == 0 else None)

    for want, fn, ct, n, s in [(v32, L.rocrand_get_direction_vectors32, ctypes.c_uint, 640000, 101),
                               (sv32, L.rocrand_get_direction_vectors32, ctypes.c_uint, 640000, 102),
                               (v64, L.rocrand_get_direction_vectors64, ctypes.c_ulonglong, 1280000, 103),
                               (sv64, L.rocrand_get_direction_vectors64, ctypes.c_ulonglong, 1280000, 104)]:
        st, got = get(fn, ct, n, s)
        assert st == 0 and np.array_equal(got, want)
    assert get(L.rocrand_get_direction_vectors32, ctypes.c_uint, 1, 103)[0] == 104
    assert get(L.rocrand_get_direction_vectors64, ctypes.c_ulonglong, 1, 101)[0] == 104
    st, got = get(L.rocrand_get_scramble_constants32, ctypes.c_uint, 20000)
    assert st == 0 and np.array_equal(got, c32)
    st, got = get(L.rocrand_get_scramble_constants64, ctypes.c_ulonglong, 20000)
    assert st == 0 and np.array_equal(got, c64)


# ------------------------------------------------------------------ MT19937
MT_DATA = os.path.join(os.path.dirname(HERE), "rocrand_amd", "data")


@pytest.fixture(scope="module")
def jump_tool(tmp_path_factory):
    """The library's jump-polynomial tool (rocrand_amd/data/make_mt19937_jumps.c), built here."""
    import subprocess

    exe = str(tmp_path_factory.mktemp("mtjump") / "make_mt19937_jumps")
    subprocess.run(["gcc", "-O2", "-o", exe, os.path.join(MT_DATA, "make_mt19937_jumps.c")], check=True)
    return exe


def test_mt19937_jump_polynomial_matches_reference_vector(jump_tool):
    """x^552552 mod phi equals the reference test's known polynomial (Berlekamp-Massey on
    the recurrence, then square-and-multiply: a wrong phi or reduction fails)."""
    import subprocess

    gold = json.load(open(os.path.join(HERE, "golden", "mt19937_jump_552552.json")))
    out = subprocess.run([jump_tool, "--power", "552552"], capture_output=True, text=True, check=True).stdout
    assert [int(x) for x in out.split()] == gold["words"]


def test_mt19937_jump_polynomial_brute_force(jump_tool):
    """Horner with x^k mod phi equals k single steps of the word recurrence (k = 1000, 100003)."""
    import subprocess

    def comp(a, b, m):
        y = (a & 0x80000000) | (b & 0x7FFFFFFF)
        return m ^ (y >> 1) ^ (0x9908B0DF if y & 1 else 0)

    s0 = [5489]
    for i in range(1, 624):
        s0.append((1812433253 * (s0[-1] ^ (s0[-1] >> 30)) + i) & 0xFFFFFFFF)
    for k in (1000, 100003):
        p = [int(x) for x in subprocess.run([jump_tool, "--power", str(k)], capture_output=True, text=True,
                                            check=True).stdout.split()]
        # Horner over the window (the reference jump kernel's evaluation), in numpy
        seq = np.array(s0 + [0] * (19937 + k), dtype=np.uint64)
        for i in range(624, 624 + max(k, 19937) + 1):
            seq[i] = comp(int(seq[i - 624]), int(seq[i - 623]), int(seq[i - 227]))
        bits = np.array([(p[j // 32] >> (j % 32)) & 1 for j in range(19937)], dtype=bool)
        idx = np.nonzero(bits)[0]
        jumped = np.array([np.bitwise_xor.reduce(seq[i + idx]) for i in range(624)], dtype=np.uint64)
        want = seq[k:k + 624]
        # the oldest word's low 31 bits never influence the sequence
        assert jumped[0] >> 31 == want[0] >> 31
        assert np.array_equal(jumped[1:], want[1:])


def test_mt19937_jump_table_matches_reference():
    if O.ref() is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    ref_jump = np.ctypeslib.as_array(O.ref().ref_mt19937_jump(), (510 * 624,)).copy()
    mine = np.fromfile(os.path.join(MT_DATA, "mt19937_jumps.bin"), dtype="<u4")
    assert np.array_equal(ref_jump, mine)


@pytest.mark.parametrize("seed", [12345, 0xFEDCBA9876543210])
def test_mt19937_generator0_is_std_mt19937(seed):
    """Generator 0 (lanes 0-7 of thread block 0) is MT19937 seeded by init_genrand((seed >> 32) ^ seed):
    numpy's RandomState implements the same recurrence, seeding and tempering."""
    og = O.OracleGen(O.MT19937, seed=seed)
    st, x = og.generate("u32", 65536 * 78, 0x1000)
    assert st == 0
    x = x.reshape(78, 65536)
    special = [0, 113, 170, 283, 340, 397, 510, 567]
    offs = [1, 57, 114, 171, 227, 284, 341, 398, 454, 511, 568]
    g0 = np.zeros(624, dtype=np.uint32)
    for lane in range(8):
        for j in range(78):
            e = special[lane] if j == 0 else offs[(j - 1) // 7] + 7 * lane + (j - 1) % 7
            g0[e] = x[j, lane]
    seedu = (seed ^ (seed >> 32)) & 0xFFFFFFFF
    want = np.random.RandomState(seedu).randint(0, 2 ** 32, size=624, dtype=np.uint64).astype(np.uint32)
    assert np.array_equal(g0, want)


def test_mt19937_offset_is_type_error():
    assert O.OracleGen(O.MT19937).set_offset(3) == 103


@pytest.mark.parametrize("seed,calls", [
    (12345, [("u32", 100), ("uniform_f32", 200), ("normal_f32", 50), ("uniform_f64", 20), ("u8", 40),
             ("normal_f64", 10), ("u32", 64), ("poisson", 33)]),
    (0xFEDCBA9876543210, [("u16", 256), ("lognormal_f32", 100), ("uniform_f16", 70), ("lognormal_f64", 6)]),
])
def test_mt19937_oracle_equals_reference_kernel_small(seed, calls):
    """Aligned calls of whole vectors inside the first 64 generators (the host form of the
    reference kernels shares one head/tail flag among eight emulated lanes)."""
    R = O.ref()
    if R is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    og = O.OracleGen(O.MT19937, seed=seed)
    rg = O.RefGen(O.MT19937, seed=seed)
    R.ref_mt19937_set_limit(rg.h, 64)
    for kind, n in calls:
        a, b = (10.0, 0.0) if kind == "poisson" else ((0.5, 1.5) if "normal" in kind else (0.0, 1.0))
        s1, x = og.generate(kind, n, 0x10000, a, b)
        s2, y = rg.generate(kind, n, 0x10000, a, b)
        assert s1 == s2 == 0
        assert _same(kind, x, y), (kind, n)


def test_mt19937_oracle_equals_reference_kernel_rounds():
    """Calls spanning several rounds of all 8192 generators, with input-width changes."""
    R = O.ref()
    if R is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    og = O.OracleGen(O.MT19937, seed=777)
    rg = O.RefGen(O.MT19937, seed=777)
    R.ref_mt19937_set_limit(rg.h, 8192)
    for kind, n in [("u32", 65536 * 78 + 65536 * 3 + 100), ("uniform_f64", 65536 * 39 * 2 + 4),
                    ("normal_f32", 1000000), ("u32", 65536 * 80)]:
        a, b = (0.5, 1.5) if "normal" in kind else (0.0, 1.0)
        s1, x = og.generate(kind, n, 0x10000, a, b)
        s2, y = rg.generate(kind, n, 0x10000, a, b)
        assert s1 == s2 == 0
        assert _same(kind, x, y), (kind, n)


def test_mtgp32_parameter_tables_are_linear():
    """The MTGP32 kernel evaluates each set's recurrence table from its entries at 1, 2,
    4 and 8 (an XOR selected by the index bits) and reads the ring at off + t + pos with
    t < 256 without wrap masks: both rest on properties of the 512 parameter sets
    (rocrand_mtgp32_11213.h), checked here on the library's data file."""
    here = os.path.dirname(os.path.abspath(__file__))
    p = np.fromfile(os.path.join(os.path.dirname(here), "rocrand_amd", "data", "mtgp32_11213.bin"),
                    dtype=np.uint32).reshape(512, 36)
    pos, tbl, tmp = p[:, 0], p[:, 3:19], p[:, 19:35]
    for t in (tbl, tmp):
        assert not t[:, 0].any()
        for k in range(16):
            want = np.zeros(512, dtype=np.uint32)
            for b in range(4):
                if k >> b & 1:
                    want ^= t[:, 1 << b]
            assert np.array_equal(t[:, k], want), k
    assert (pos > 0).all() and (pos + 256 <= 351).all()
