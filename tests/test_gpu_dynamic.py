"""DYNAMIC ordering (ROCRAND_ORDERING_PSEUDO_DYNAMIC) on the GPU against the oracle run on the
same launch grid (oracle.set_grid, itself pinned to the reference kernel bodies on that
grid in test_oracle.py).  The grid comes from the library's gfx950 table
(rocrand_amd_get_dynamic_grid) or from an override (rocrand_amd_set_dynamic_grid, the
autotuner's hook): engine counts below, at and above the DEFAULT 131072 for the
interleaved engines, strides for the counter engines' head shift; two calls in a row,
misaligned outputs, and long calls that take the chunked kernels."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import oracle as O  # noqa: E402
from parity_checks import check  # noqa: E402
from rocrand_amd import rocrand as R  # noqa: E402

ENGINES = [("philox", R.PHILOX4_32_10), ("threefry2x32", R.THREEFRY2_32_20), ("threefry4x64", R.THREEFRY4_64_20),
           ("mrg32k3a", R.MRG32K3A), ("mrg31k3p", R.MRG31K3P), ("xorwow", R.XORWOW), ("lfsr113", R.LFSR113)]
INTERLEAVED = {R.MRG32K3A, R.MRG31K3P, R.XORWOW, R.LFSR113}
KINDS = [("u32", torch.uint32, "generate", ()), ("uniform_f64", torch.float64, "uniform", ()),
         ("normal_f32", torch.float32, "normal", (0.25, 1.5)), ("poisson", torch.uint32, "poisson", (4.0,))]


def _length(rng, size):
    """Enough outputs to wrap the engine count twice, or to pass the stride's first
    multiple in vec indices (vecs of up to 4 outputs) for the counter engines."""
    return 2 * size + 77 if rng in INTERLEAVED else 4 * size + 77


def _run(rng, grid, seed, offset, n, shift, kinds=KINDS):
    g = R.PRNG(rng, seed=seed, offset=offset)
    g.set_ordering(R.ORDERING_PSEUDO_DYNAMIC)
    og = O.OracleGen(rng, seed=seed, offset=offset, grid=grid)
    for kind, dt, call, args in kinds:
        print(f"case dynamic rng={rng} grid={grid} kind={kind} n={n} shift={shift}", flush=True)
        for _ in range(2):  # continuity across calls
            buf = torch.zeros(n + 8, dtype=dt, device="cuda")
            view = buf[shift:shift + n]
            getattr(g, call)(view, *args)
            torch.cuda.synchronize()
            a, b = (args[0], 0.0) if kind == "poisson" else (args if args else (0.0, 0.0))
            st, want = og.generate(kind, n, view.data_ptr(), a, b)
            assert st == 0
            check(kind, view.cpu().numpy(), want, args if "normal" in kind else (0.0, 1.0))


@pytest.mark.parametrize("name,rng", ENGINES)
def test_dynamic_table_grid_equals_oracle(name, rng):
    grid = R.get_dynamic_grid(rng)
    size = grid[0] * grid[1]
    _run(rng, grid, 0x1234ABCD, 1048583, _length(rng, size), 1)


@pytest.mark.parametrize("grid", [(256, 256), (512, 1024)])
@pytest.mark.parametrize("name,rng", ENGINES)
def test_dynamic_override_grid_equals_oracle(name, rng, grid):
    R.set_dynamic_grid(rng, *grid)
    try:
        assert R.get_dynamic_grid(rng) == grid
        size = grid[0] * grid[1]
        _run(rng, grid, 12345, 3 * size + 5, _length(rng, size) - 74, 3)
    finally:
        R.set_dynamic_grid(rng, 0, 0)


@pytest.mark.parametrize("name,rng", [e for e in ENGINES if e[1] in INTERLEAVED])
def test_dynamic_long_calls_take_the_chunked_kernels(name, rng):
    """65536 engines and 65 steps per engine: MRG32k3a / MRG31k3p split every engine's
    steps into chunks, XORWOW / LFSR113 do so for Poisson."""
    grid = (256, 256)
    R.set_dynamic_grid(rng, *grid)
    try:
        kinds = [k for k in KINDS if k[0] in ("u32", "poisson")]
        _run(rng, grid, 7, 99, 65 * 65536 + 11, 2, kinds)
    finally:
        R.set_dynamic_grid(rng, 0, 0)


def test_dynamic_grid_errors():
    with pytest.raises(R.RocRandError):
        R.set_dynamic_grid(R.MRG32K3A, 100, 3)  # not a multiple of 256
    with pytest.raises(R.RocRandError):
        R.set_dynamic_grid(R.MT19937, 256, 256)  # no grid for MT19937
    with pytest.raises(R.RocRandError):
        R.set_dynamic_grid(R.MTGP32, 256, 256)  # MTGP32 keeps 256 x 512
    with pytest.raises(R.RocRandError):
        R.set_dynamic_grid(R.PHILOX4_32_10, 256, 768)  # counter strides are powers of two
    assert R.get_dynamic_grid(R.MTGP32) == (256, 512)
