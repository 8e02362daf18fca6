// rr_dynamic.cpp -- launch grids of the DYNAMIC ordering (ROCRAND_ORDERING_PSEUDO_DYNAMIC).
//
// The reference selects a tuned (threads, blocks) per engine for the device's architecture
// and the DEFAULT grid for an architecture it has no entry for (config_types.hpp:314-336,
// config/*_config.hpp).  Only threads x blocks affects the output: the engine count of the
// interleaved engines (mrg.hpp:284-337, get_least_common_grid_size: one grid for every
// output type) and the grid stride that places the counter engines' head-shifted vec
// indices (philox4x32_10.hpp:106-214).  gfx950's entries come from the project's autotuner
// (rocrand_amd/tune_dynamic.py) and are compiled in from rr_dynamic_gfx950.inc; an override
// set through rocrand_amd_set_dynamic_grid takes precedence (the autotuner's candidates).
#include "rr_dynamic.hpp"

#include <rocrand_amd_tuning.h>

#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

namespace rr
{
namespace
{

struct grid_entry
{
    rocrand_rng_type type;
    unsigned         threads, blocks;
};

const grid_entry gfx950_grids[] = {
#include "rr_dynamic_gfx950.inc"
};

// DEFAULT grids (config_defaults.hpp:33-108)
bool default_grid(rocrand_rng_type t, unsigned& threads, unsigned& blocks)
{
    switch(t)
    {
        case ROCRAND_RNG_PSEUDO_PHILOX4_32_10:
        case ROCRAND_RNG_PSEUDO_THREEFRY2_32_20:
        case ROCRAND_RNG_PSEUDO_THREEFRY2_64_20:
        case ROCRAND_RNG_PSEUDO_THREEFRY4_32_20:
        case ROCRAND_RNG_PSEUDO_THREEFRY4_64_20: threads = 256; blocks = 1024; return true;
        case ROCRAND_RNG_PSEUDO_MRG32K3A:
        case ROCRAND_RNG_PSEUDO_MRG31K3P:
        case ROCRAND_RNG_PSEUDO_XORWOW:
        case ROCRAND_RNG_PSEUDO_LFSR113:
        case ROCRAND_RNG_PSEUDO_MTGP32: threads = 256; blocks = 512; return true;
        default: return false;
    }
}

struct overrides
{
    std::mutex       mu;
    rocrand_rng_type type[16];
    unsigned         threads[16], blocks[16];
    int              n = 0;
};
overrides& table()
{
    static overrides o;
    return o;
}

bool is_gfx950(int device)
{
    static std::mutex mu;
    static int        known[64] = {}; // 0 unknown, 1 gfx950, 2 other
    if(device < 0 || device >= 64)
        return false;
    std::lock_guard<std::mutex> lock(mu);
    if(known[device] == 0)
    {
        hipDeviceProp_t prop;
        if(hipGetDeviceProperties(&prop, device) != hipSuccess)
            return false;
        known[device] = std::strncmp(prop.gcnArchName, "gfx950", 6) == 0 ? 1 : 2;
    }
    return known[device] == 1;
}

// Engine count / stride limits the kernels accept: a multiple of 256 and at most 2^22
// (interleaved state memory 2 x 24 B per engine, 192 MiB at the limit).
bool valid_size(unsigned long long size) { return size != 0 && size % 256 == 0 && size <= (1ull << 22); }

} // namespace

rocrand_status dynamic_grid(rocrand_rng_type t, unsigned& threads, unsigned& blocks)
{
    if(!default_grid(t, threads, blocks))
        return ROCRAND_STATUS_TYPE_ERROR;
    {
        overrides&                  o = table();
        std::lock_guard<std::mutex> lock(o.mu);
        for(int i = 0; i < o.n; i++)
            if(o.type[i] == t)
            {
                threads = o.threads[i];
                blocks  = o.blocks[i];
                return ROCRAND_STATUS_SUCCESS;
            }
    }
    int device = 0;
    if(hipGetDevice(&device) != hipSuccess)
        return ROCRAND_STATUS_INTERNAL_ERROR;
    if(is_gfx950(device))
        for(const grid_entry& e : gfx950_grids)
            if(e.type == t)
            {
                threads = e.threads;
                blocks  = e.blocks;
            }
    return ROCRAND_STATUS_SUCCESS;
}

} // namespace rr

extern "C" {

rocrand_status ROCRANDAPI rocrand_amd_set_dynamic_grid(rocrand_rng_type rng_type, unsigned int threads,
                                                       unsigned int blocks)
{
    unsigned dt, db;
    if(!rr::default_grid(rng_type, dt, db))
        return ROCRAND_STATUS_TYPE_ERROR;
    const unsigned long long size = (unsigned long long)threads * blocks;
    if(size != 0 && !rr::valid_size(size))
        return ROCRAND_STATUS_OUT_OF_RANGE;
    // the counter engines' kernels test vec indices against the stride with a mask
    const bool counter = rng_type == ROCRAND_RNG_PSEUDO_PHILOX4_32_10 || rng_type == ROCRAND_RNG_PSEUDO_THREEFRY2_32_20
                         || rng_type == ROCRAND_RNG_PSEUDO_THREEFRY2_64_20
                         || rng_type == ROCRAND_RNG_PSEUDO_THREEFRY4_32_20
                         || rng_type == ROCRAND_RNG_PSEUDO_THREEFRY4_64_20;
    if(counter && (size & (size - 1)) != 0)
        return ROCRAND_STATUS_OUT_OF_RANGE;
    // MTGP32 runs one 256-thread block per generator and its parameter sets: its DYNAMIC
    // grid is the DEFAULT one (fewer generators would only lower its parallelism)
    if(rng_type == ROCRAND_RNG_PSEUDO_MTGP32 && size != 0 && (threads != 256 || blocks != 512))
        return ROCRAND_STATUS_OUT_OF_RANGE;
    rr::overrides&              o = rr::table();
    std::lock_guard<std::mutex> lock(o.mu);
    int                         i = 0;
    while(i < o.n && o.type[i] != rng_type)
        i++;
    if(size == 0) // clear
    {
        if(i < o.n)
        {
            o.n--;
            o.type[i]    = o.type[o.n];
            o.threads[i] = o.threads[o.n];
            o.blocks[i]  = o.blocks[o.n];
        }
        return ROCRAND_STATUS_SUCCESS;
    }
    if(i == o.n)
    {
        if(o.n == 16)
            return ROCRAND_STATUS_INTERNAL_ERROR;
        o.n++;
    }
    o.type[i]    = rng_type;
    o.threads[i] = threads;
    o.blocks[i]  = blocks;
    return ROCRAND_STATUS_SUCCESS;
}

rocrand_status ROCRANDAPI rocrand_amd_get_dynamic_grid(rocrand_rng_type rng_type, unsigned int* threads,
                                                       unsigned int* blocks)
{
    if(threads == nullptr || blocks == nullptr)
        return ROCRAND_STATUS_OUT_OF_RANGE;
    return rr::dynamic_grid(rng_type, *threads, *blocks);
}

} // extern "C"
