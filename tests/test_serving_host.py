"""CPU tests of the serving path's host logic: the request server's arena allocator
(rocrand_amd/csrc/rr_arena.hpp, compiled here with g++ behind tests/arena_shim.cpp), the
server's host-decided status codes, and the request mixes."""
import ctypes
import math
import os
import random
import shutil
import subprocess

import pytest

from rocrand_amd import rocrand as R
from rocrand_amd import serving as S
from rocrand_amd.workloads import SplitMix64, batch_mix, serving_mix

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "rocrand_amd", "csrc")


def test_splitmix64_reference_values():
    # splitmix64 from state 0: published first outputs
    g = SplitMix64(0)
    assert [g.next() for _ in range(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


@pytest.fixture(scope="module")
def shim(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    so = str(tmp_path_factory.mktemp("arena") / "libarena.so")
    subprocess.run(["g++", "-O1", "-std=c++17", "-shared", "-fPIC", "-I", CSRC,
                    os.path.join(HERE, "arena_shim.cpp"), "-o", so], check=True)
    L = ctypes.CDLL(so)
    L.arena_new.restype = ctypes.c_void_p
    L.arena_new.argtypes = [ctypes.c_size_t]
    L.arena_delete.argtypes = [ctypes.c_void_p]
    L.arena_alloc.restype = ctypes.c_longlong
    L.arena_alloc.argtypes = [ctypes.c_void_p, ctypes.c_size_t]
    L.arena_release.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t]
    L.arena_free_runs.restype = ctypes.c_size_t
    L.arena_free_runs.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t), ctypes.c_size_t]
    return L


class Arena:
    def __init__(self, L, size):
        self.L, self.h = L, L.arena_new(size)

    def alloc(self, n):
        s = self.L.arena_alloc(self.h, n)
        return None if s < 0 else s

    def release(self, s, n):
        self.L.arena_release(self.h, s, n)

    @property
    def free(self):
        buf = (ctypes.c_size_t * 2048)()
        k = self.L.arena_free_runs(self.h, buf, 1024)
        assert k <= 1024
        return [(buf[2 * i], buf[2 * i + 1]) for i in range(k)]

    def __del__(self):
        self.L.arena_delete(self.h)


def test_arena_first_fit_and_coalescing(shim):
    a = Arena(shim, 1024)
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
    assert Arena(shim, 0).free == [] and Arena(shim, 0).alloc(1) is None


def test_slot_sizes():
    assert S.slot(0) == 256 and S.slot(1) == 256 and S.slot(256) == 256 and S.slot(257) == 512


def test_arena_matches_bitmap_model(shim):
    """Random alloc/release sequences: first-fit positions and the free list equal those of
    a per-byte bitmap (free list = maximal free runs)."""
    rs = random.Random(1)
    for _ in range(100):
        size = rs.randrange(1, 300)
        a, used, bitmap = Arena(shim, size), [], [0] * size
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


def test_server_host_decided_status_codes():
    """Null handles and bad creation arguments are rejected before any HIP call."""
    L = S._lib()
    NOT_CREATED, OUT_OF_RANGE = 101, 104
    k, p = ctypes.c_size_t(), ctypes.c_size_t()
    t = (ctypes.c_ulonglong * 1)()
    req = (S.ServeRequest * 1)()
    assert L.rocrand_server_submit(None, req, 1, t) == NOT_CREATED
    assert L.rocrand_server_poll(None, None, 0, ctypes.byref(k), ctypes.byref(p)) == NOT_CREATED
    assert L.rocrand_server_release(None, 1) == NOT_CREATED
    assert L.rocrand_server_end_session(None, 1) == NOT_CREATED
    assert L.rocrand_server_destroy(None) == NOT_CREATED
    h = ctypes.c_void_p()
    assert L.rocrand_server_create(None, 1 << 20, 2, 0) == OUT_OF_RANGE
    assert L.rocrand_server_create(ctypes.byref(h), 255, 2, 0) == OUT_OF_RANGE
    assert L.rocrand_server_create(ctypes.byref(h), 1 << 20, 0, 0) == OUT_OF_RANGE
    assert L.rocrand_server_create(ctypes.byref(h), 1 << 20, 9, 0) == OUT_OF_RANGE


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


@pytest.mark.parametrize("dtype,n", [("u32", (1 << 20) + 4099), ("f64", 600_001), ("f32", 70_003)])
def test_philox_window_rule_matches_whole_output(dtype, n):
    """Philox4x32-10 windows (tests/serving_checks.py) equal slices of one whole oracle
    call, past the reference grid's stride of 2^20 words for u32 and f64."""
    import numpy as np

    from oracle import oracle as O
    from serving_checks import ADDR, ORACLE_KIND, philox_window

    st, whole = O.OracleGen(R.PHILOX4_32_10, seed=0xABCDEF12345).generate(ORACLE_KIND[dtype], n, ADDR)
    assert st == 0
    for i0 in (0, 4, 1000, n // 2 // 4 * 4, (n - 4096) // 4 * 4):
        k = min(4096, n - i0)
        assert np.array_equal(philox_window(0xABCDEF12345, 0, dtype, i0, k).view(np.uint8),
                              whole[i0:i0 + k].view(np.uint8)), i0


@pytest.mark.parametrize("dtype", ["u32", "f32", "f64"])
def test_sobol_window_rule_matches_whole_output(dtype):
    import numpy as np

    from oracle import oracle as O
    from serving_checks import ADDR, ORACLE_KIND, sobol_window

    dims, m, off = 16, 5000, 777
    st, whole = O.OracleGen(R.SOBOL64, dims=dims, offset=off).generate(ORACLE_KIND[dtype], dims * m, ADDR)
    assert st == 0
    pts = whole.reshape(dims, m)
    for j0 in (0, 1, 1234, m - 4096):
        k = min(4096, m - j0)
        assert np.array_equal(sobol_window(off, dims, dtype, j0, k), pts[:, j0:j0 + k]), j0


@pytest.mark.parametrize("kind", ["uniform_f32", "normal_f32"])
def test_philox_offset_window_rule_for_full_size_calls(kind):
    """The windows of tests/test_gpu_configs.py: value i of an aligned call at offset o is
    value 0 of an aligned call at offset o + i (i a multiple of 4), Box-Muller pairs
    included, for an odd base offset."""
    import numpy as np

    from oracle import oracle as O
    from parity_checks import check_exact

    base, n = (1 << 20) + 7, (1 << 20) + 4096
    st, whole = O.OracleGen(R.PHILOX4_32_10, seed=12345, offset=base).generate(kind, n, 0x10000000, 0.0, 1.0)
    assert st == 0
    for i0 in (0, 4, 262144 * 4 - 4, n - 1024):
        st, w = O.OracleGen(R.PHILOX4_32_10, seed=12345, offset=base + i0).generate(kind, 1024, 0x10000000, 0.0, 1.0)
        assert st == 0
        check_exact(w, whole[i0:i0 + 1024])


@pytest.mark.parametrize("rng", [R.SOBOL32, R.SCRAMBLED_SOBOL32, R.SOBOL64, R.SCRAMBLED_SOBOL64])
def test_quasi_window_rule_every_variant(rng):
    """The point-window rule the Sobol benchmark-configuration test uses: points
    [j0, j0 + k) of every dimension are a fresh generator at offset + j0."""
    import numpy as np

    from oracle import oracle as O

    dims, npd, off = 64, 4096, 1000
    st, whole = O.OracleGen(rng, dims=dims, offset=off).generate("u32", dims * npd, 0x10000000)
    assert st == 0
    whole = whole.reshape(dims, npd)
    for j0 in (0, 1234, npd - 1024):
        st, w = O.OracleGen(rng, dims=dims, offset=off + j0).generate("u32", dims * 1024, 0x10000000)
        assert st == 0
        assert np.array_equal(w.reshape(dims, 1024), whole[:, j0:j0 + 1024]), j0
