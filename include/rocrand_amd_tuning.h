/*
 * rocrand_amd_tuning.h -- launch-grid control of the DYNAMIC ordering (MI355X extension).
 *
 * No counterpart in the reference's public API: the reference compiles its tuned grids in
 * (library/src/rng/config/*_config.hpp, selected by config_types.hpp:314-336).  These two
 * entry points serve the project's autotuner (rocrand_amd/tune_dynamic.py), which times
 * candidate grids on the device and writes the winners into the library's gfx950 table,
 * and tests that compare DYNAMIC-ordering output with the oracle run on the same grid.
 * Only threads x blocks matters to the values: the engine count of MRG32k3a, MRG31k3p,
 * XORWOW and LFSR113, the grid stride of Philox and ThreeFry (where the reference's head
 * shift lands).  MTGP32 keeps its DEFAULT grid (256 x 512: one block per parameter set).
 */
#ifndef ROCRAND_AMD_TUNING_H_
#define ROCRAND_AMD_TUNING_H_

#include "rocrand.h"

#if defined(__cplusplus)
extern "C" {
#endif

/* Overrides the DYNAMIC grid of rng_type in this process (threads * blocks a multiple of
 * 256, at most 2^22, and a power of two for Philox and ThreeFry; threads = blocks = 0
 * removes the override).  Takes effect when a
 * generator next initialises (after rocrand_set_ordering, set_seed or set_offset).
 * TYPE_ERROR for engines without a grid (MT19937, the quasi engines), OUT_OF_RANGE for
 * sizes the kernels do not take. */
ROCRANDAPI rocrand_status rocrand_amd_set_dynamic_grid(rocrand_rng_type rng_type,
                                                       unsigned int     threads,
                                                       unsigned int     blocks);
/* The grid a DYNAMIC-ordering generator of rng_type uses on the current device. */
ROCRANDAPI rocrand_status rocrand_amd_get_dynamic_grid(rocrand_rng_type rng_type,
                                                       unsigned int*    threads,
                                                       unsigned int*    blocks);

#if defined(__cplusplus)
}
#endif

#endif /* ROCRAND_AMD_TUNING_H_ */
