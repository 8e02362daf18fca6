// rr_dynamic.hpp -- launch grids of the DYNAMIC ordering (rr_dynamic.cpp).
#pragma once

#include <rocrand.h>

namespace rr
{
// The (threads, blocks) a DYNAMIC-ordering generator of type t uses on the current device:
// an override set by rocrand_amd_set_dynamic_grid, else gfx950's tuned entry, else the
// DEFAULT grid.  TYPE_ERROR for engines without a grid (MT19937, the quasi engines).
rocrand_status dynamic_grid(rocrand_rng_type t, unsigned& threads, unsigned& blocks);
} // namespace rr
