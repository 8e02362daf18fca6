// rr_launch.hpp -- host-side launch interface between the generator objects and
// the gfx950 kernels (one translation unit per engine family).
#pragma once

#include "rr_dist.hpp"

#include <hip/hip_runtime.h>

#include <stddef.h>
#include <stdint.h>

namespace rr
{

// Grid of the reference's DEFAULT ordering for the counter-based engines
// (config_defaults.hpp:54-108): 1024 blocks x 256 threads.  It fixes which vec
// indices take the head-shifted engine position (see counter_call).
constexpr uint32_t CTR_REF_STRIDE = 1024u * 256u;
// Engines of the interleaved generators (config_defaults.hpp:33-73): 512 x 256.
constexpr uint32_t INTERLEAVED_ENGINES = 512u * 256u;

// One counter-based generate call, fully resolved on the host.
struct counter_call
{
    ctr_key  key;
    ctr_pos  pos;        // word position of the call's first word (engine state)
    ctr_pos  vec_base;   // position of vec 0's first word: pos + H
    ctr_pos  vec_base0;  // position used by vec indices that are multiples of ref_stride: pos + C0
    int      quirk;      // vec_base0 != vec_base
    uint32_t ref_stride; // grid stride of the reference's generate kernel for the ordering
                         // (CTR_REF_STRIDE, or the DYNAMIC grid)
    uint32_t lane_shift; // vecs between the last 128-byte line boundary and vec 0
    void*    data;       // output base pointer
    size_t   n;          // number of outputs
    size_t   mis;        // misalignment (outputs before the first full vec)
    uint32_t head, tail; // head/tail output counts
    size_t   vec_n;      // full vecs
    dist_params params;
};

hipError_t launch_counter(ctr_kind k, out_kind o, bool normal_one_word, const counter_call& c, hipStream_t s);

// One interleaved generate call.
struct interleaved_call
{
    uint32_t*   states;     // structure-of-arrays engine states, word j of engine e at [j*S + e]
    uint32_t*   states_out; // where the call leaves the advanced states (same layout; a second
                            // buffer, so no block of a launch reads a state another block writes)
    uint32_t    start;    // start engine
    uint32_t    S;        // engines (INTERLEAVED_ENGINES, or the DYNAMIC grid), a multiple of 256
    uint32_t    lane_shift; // vecs between the last 128-byte line boundary and vec 0
    void*       data;
    size_t      n, mis;
    uint32_t    head, tail;
    size_t      vec_n;
    dist_params params;
};

enum class inter_kind : int
{
    mrg32k3a = 0,
    mrg31k3p = 1,
    xorwow   = 2,
    lfsr113  = 3
};

// Number of 32-bit state words per engine in the SoA layout.
constexpr int inter_state_words(inter_kind k)
{
    return k == inter_kind::lfsr113 ? 4 : 6;
}

hipError_t launch_interleaved_init(inter_kind k,
                                   uint32_t*  states,
                                   const uint32_t* seed_state, // one engine's seeded state
                                   uint32_t   start,
                                   uint64_t   base_offset,
                                   uint32_t   engines, // S
                                   hipStream_t s);
hipError_t launch_interleaved(inter_kind k, out_kind o, const interleaved_call& c, hipStream_t s);

// One Sobol generate call (rr_sobol.hip).
struct sobol_call
{
    const void* vectors;  // device direction vectors, [dims][32 or 64]
    const void* scramble; // scramble constant table indexed by dimension, or nullptr
    void*       data;
    size_t      size;     // points per dimension
    uint32_t    offset;   // index of the first point
    uint32_t    dims;
    // filled by launch_sobol
    int64_t     first_group; // may be -1 (see launch_sobol)
    uint32_t    groups_per_wave;
    uint32_t    waves;
    dist_params params;
};
hipError_t launch_sobol(bool is64, out_kind o, sobol_call c, hipStream_t s);

// Sobol tables (rr_quasi.cpp): host arrays of the reference's layout, and the
// device copies of the current device.
struct sobol_device_tables
{
    const uint32_t* v32;
    const uint32_t* sv32; // scrambled
    const uint64_t* v64;
    const uint64_t* sv64;
};
hipError_t get_sobol_device_tables(sobol_device_tables& t);

// Subsequence radix tables: digit 0 = e mod 512, digit 1 = e / 512 (e < 2^17).
constexpr int MRG_RADIX_LO_BITS = 9;
constexpr int MRG_RADIX_LO      = 1 << MRG_RADIX_LO_BITS;
constexpr int MRG_RADIX_HI      = 256;
constexpr int MRG_RADIX_WORDS   = 20;
static_assert((uint64_t)MRG_RADIX_LO * MRG_RADIX_HI >= INTERLEAVED_ENGINES, "radix tables cover every engine");

// Host-computed jump tables, uploaded once per device (rr_tables.cpp).
struct device_tables
{
    const uint32_t* mrg32_step; // [64][2][9]  A^(2^i)
    const uint32_t* mrg32_sub;  // [64][2][9]  A^(2^(76+i))
    const uint32_t* mrg31_step;
    const uint32_t* mrg31_sub;  // A^(2^(72+i))
    const uint32_t* mrg32_subr; // [512 + 256][20]  A^(d 2^76), A^(512 d 2^76)
    const uint32_t* mrg31_subr; // [512 + 256][20]  A^(d 2^72), A^(512 d 2^72)
    const uint32_t* xorwow_jump; // [32][800]  M^(4^k)
    const uint32_t* xorwow_seq;  // [32][800]  M^(2^67 * 4^k)
    const uint32_t* lfsr_jump;   // [32][512]
    const uint32_t* lfsr_seq;    // [32][512]  M^(2^55 * 4^k)
};
// Returns the tables of the current device (built and uploaded on first use).
hipError_t get_device_tables(device_tables& t);

// Host copies of the same tables (used by host-side seeding helpers and tests).
const uint32_t* host_mrg_table(int which_engine /*32 or 31*/, int sub /*0 step,1 subsequence*/);
// Host copy of the subsequence radix table (MRG_RADIX_LO + MRG_RADIX_HI entries of
// MRG_RADIX_WORDS words: X^e for e < 512, then X^(512 h) for h < 256, X = A^(2^76) or A^(2^72)).
const uint32_t* host_mrg_radix(int which_engine /*32 or 31*/);
const uint32_t* host_gf2_table(int xorwow /*1 xorwow, 0 lfsr*/, int seq);

} // namespace rr
