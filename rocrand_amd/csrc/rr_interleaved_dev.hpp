// rr_interleaved_dev.hpp -- engine wrappers of the interleaved generators (MRG32k3a,
// MRG31k3p, XORWOW, LFSR113), shared by the per-call kernels (rr_interleaved.hip) and
// the batched kernels (rr_batched.hip).  State layout: structure-of-arrays, word j of
// engine e at states[j*S + e].
#pragma once

#include "rr_launch.hpp"

namespace rr
{
namespace inter_dev
{

// Engines of the DEFAULT grid; the per-call kernels take the call's engine count (the
// DYNAMIC grid may differ), the batched kernels use this one.
constexpr uint32_t S = INTERLEAVED_ENGINES;

template<inter_kind K>
struct ieng;

// One radix-table entry (both components) applied to the state.
template<uint32_t K, uint32_t C1, uint32_t C2>
__device__ __forceinline__ void mrg_entry_apply(const uint32_t* entry, mrg_state& s)
{
    const uint4* p = reinterpret_cast<const uint4*>(entry);
    uint32_t     w[MRG_RADIX_WORDS];
#pragma unroll
    for(int i = 0; i < MRG_RADIX_WORDS / 4; i++)
    {
        const uint4 q = p[i];
        w[4 * i + 0]  = q.x;
        w[4 * i + 1]  = q.y;
        w[4 * i + 2]  = q.z;
        w[4 * i + 3]  = q.w;
    }
    mat3_vec_fold<K, C1>(w, s.g1);
    mat3_vec_fold<K, C2>(w + 9, s.g2);
}

// Jump to subsequence e mod 2^17 with the radix tables: X^(e mod 512), then
// X^(512 ((e / 512) mod 256)); the powers commute.  Same state as the binary expansion
// over the A^(2^(sub + i)) table of the reference's skipahead_subsequence.  (Engines at
// or beyond 2^17, which only a larger DYNAMIC grid has, add their high part by that
// binary expansion: see ieng::sub.)
template<uint32_t K, uint32_t C1, uint32_t C2>
__device__ __forceinline__ void mrg_sub_jump(const uint32_t* tab, uint32_t e, mrg_state& s)
{
    const uint32_t lo = e & (uint32_t)(MRG_RADIX_LO - 1);
    const uint32_t hi = (e >> MRG_RADIX_LO_BITS) & (uint32_t)(MRG_RADIX_HI - 1);
    if(lo != 0u)
        mrg_entry_apply<K, C1, C2>(tab + lo * MRG_RADIX_WORDS, s);
    if(hi != 0u)
        mrg_entry_apply<K, C1, C2>(tab + (MRG_RADIX_LO + hi) * MRG_RADIX_WORDS, s);
}

template<>
struct ieng<inter_kind::mrg32k3a>
{
    static constexpr word_family F = word_family::mrg32;
    mrg_state                    s;
    __device__ void load(const uint32_t* st, uint32_t e, uint32_t ns = S)
    {
#pragma unroll
        for(int j = 0; j < 3; j++)
        {
            s.g1[j] = st[j * ns + e];
            s.g2[j] = st[(3 + j) * ns + e];
        }
    }
    __device__ void store(uint32_t* st, uint32_t e, uint32_t ns = S) const
    {
#pragma unroll
        for(int j = 0; j < 3; j++)
        {
            st[j * ns + e]       = s.g1[j];
            st[(3 + j) * ns + e] = s.g2[j];
        }
    }
    __host__ __device__ void set(const uint32_t* w)
    {
        for(int j = 0; j < 3; j++)
        {
            s.g1[j] = w[j];
            s.g2[j] = w[3 + j];
        }
    }
    __device__ uint32_t next() { return mrg32k3a_next(s); }
    // Apply X^(sum of set bits) from a table of X^(2^i) (rocrand_mrg32k3a.h:204-263).
    __device__ void jump(const uint32_t* tab, uint64_t v)
    {
        for(int i = 0; v > 0; v >>= 1, i++)
            if(v & 1)
            {
                mat3_vec_fold<32, 209>(tab + (i * 2 + 0) * 9, s.g1);
                mat3_vec_fold<32, 22853>(tab + (i * 2 + 1) * 9, s.g2);
            }
    }
    // X^j for j < 512: one radix-table entry
    __device__ void sub_entry(const device_tables& t, uint32_t j)
    {
        mrg_entry_apply<32, 209, 22853>(t.mrg32_subr + j * MRG_RADIX_WORDS, s);
    }
    __device__ void sub(const device_tables& t, uint32_t e)
    {
        mrg_sub_jump<32, 209, 22853>(t.mrg32_subr, e, s);
        if(e >> 17) // A^(2^(76+17+i)) rows of the subsequence table
            jump(t.mrg32_sub + 17 * 2 * 9, e >> 17);
    }
    // k steps ahead (rocrand_mrg32k3a.h:246-263)
    __device__ void advance(const device_tables& t, uint64_t k) { jump(t.mrg32_step, k); }
    __device__ void init(const device_tables& t, uint32_t e, uint64_t off)
    {
        sub(t, e);
        jump(t.mrg32_step, off);
    }
};

template<>
struct ieng<inter_kind::mrg31k3p> : ieng<inter_kind::mrg32k3a>
{
    static constexpr word_family F = word_family::mrg31;
    __device__ uint32_t next() { return mrg31k3p_next(s); }
    __device__ void jump(const uint32_t* tab, uint64_t v)
    {
        for(int i = 0; v > 0; v >>= 1, i++)
            if(v & 1)
            {
                mat3_vec_fold<31, 1>(tab + (i * 2 + 0) * 9, s.g1);
                mat3_vec_fold<31, 21069>(tab + (i * 2 + 1) * 9, s.g2);
            }
    }
    __device__ void sub_entry(const device_tables& t, uint32_t j)
    {
        mrg_entry_apply<31, 1, 21069>(t.mrg31_subr + j * MRG_RADIX_WORDS, s);
    }
    __device__ void sub(const device_tables& t, uint32_t e)
    {
        mrg_sub_jump<31, 1, 21069>(t.mrg31_subr, e, s);
        if(e >> 17) // A^(2^(72+17+i)) rows of the subsequence table
            jump(t.mrg31_sub + 17 * 2 * 9, e >> 17);
    }
    __device__ void advance(const device_tables& t, uint64_t k) { jump(t.mrg31_step, k); }
    __device__ void init(const device_tables& t, uint32_t e, uint64_t off)
    {
        sub(t, e);
        jump(t.mrg31_step, off);
    }
};

// Base-4 digit expansion with matrix k = M^(c * 4^k) applied digit times
// (rocrand_xorwow.h:186-200, rocrand_lfsr113.h jump()).
template<int N>
__device__ void gf2_jump(uint32_t* v, const uint32_t* mats, uint64_t n)
{
    for(int k = 0; n > 0; n >>= 2, k++)
    {
        const uint32_t d = (uint32_t)(n & 3u);
        for(uint32_t i = 0; i < d; i++)
            gf2_mat_vec<N>(mats + (size_t)k * 32 * N * N, v);
    }
}

// gf2_jump for a whole block whose lanes share `n` (block-uniform): each 4^k power is
// staged once in LDS (`lds`: 32 N^2 words) and read from there by every lane, instead
// of every wavefront fetching the same rows through the scalar or vector caches.
// Every thread of the block must call it.
template<int N>
__device__ void gf2_jump_block(uint32_t* v, const uint32_t* mats, uint64_t n, uint32_t* lds)
{
    constexpr int SZ = 32 * N * N;
    for(int k = 0; n > 0; n >>= 2, k++)
    {
        const uint32_t d = (uint32_t)(n & 3u);
        if(d == 0)
            continue;
        __syncthreads(); // the previous matrix is no longer read
        for(int w = threadIdx.x; w < SZ; w += blockDim.x)
            lds[w] = mats[(size_t)k * SZ + w];
        __syncthreads();
        for(uint32_t i = 0; i < d; i++)
            gf2_mat_vec<N>(lds, v);
    }
}

template<>
struct ieng<inter_kind::xorwow>
{
    static constexpr word_family F = word_family::w32;
    xorwow_state                 s;
    __device__ void load(const uint32_t* st, uint32_t e, uint32_t ns = S)
    {
        s.d = st[e];
#pragma unroll
        for(int j = 0; j < 5; j++)
            s.x[j] = st[(1 + j) * ns + e];
    }
    __device__ void store(uint32_t* st, uint32_t e, uint32_t ns = S) const
    {
        st[e] = s.d;
#pragma unroll
        for(int j = 0; j < 5; j++)
            st[(1 + j) * ns + e] = s.x[j];
    }
    __host__ __device__ void set(const uint32_t* w)
    {
        s.d = w[0];
        for(int j = 0; j < 5; j++)
            s.x[j] = w[1 + j];
    }
    __device__ uint32_t next() { return xorwow_next(s); }
    __device__ void init(const device_tables& t, uint32_t e, uint64_t off)
    {
        gf2_jump<5>(s.x, t.xorwow_seq, e);
        gf2_jump<5>(s.x, t.xorwow_jump, off);
        s.d += (uint32_t)off * 362437u; // rocrand_xorwow.h:134-145
    }
    // k steps ahead: the 4^j powers for the xorshift part, the Weyl counter linearly
    __device__ void advance(const device_tables& t, uint64_t k)
    {
        gf2_jump<5>(s.x, t.xorwow_jump, k);
        s.d += (uint32_t)k * 362437u;
    }
    static constexpr int LDS_WORDS = 32 * 5 * 5;
    // the same for a block sharing k (every thread calls it)
    __device__ void advance_block(const device_tables& t, uint64_t k, uint32_t* lds)
    {
        gf2_jump_block<5>(s.x, t.xorwow_jump, k, lds);
        s.d += (uint32_t)k * 362437u;
    }
};

template<>
struct ieng<inter_kind::lfsr113>
{
    static constexpr word_family F = word_family::w32;
    lfsr113_state                s;
    __device__ void load(const uint32_t* st, uint32_t e, uint32_t ns = S)
    {
#pragma unroll
        for(int j = 0; j < 4; j++)
            s.z[j] = st[j * ns + e];
    }
    __device__ void store(uint32_t* st, uint32_t e, uint32_t ns = S) const
    {
#pragma unroll
        for(int j = 0; j < 4; j++)
            st[j * ns + e] = s.z[j];
    }
    __host__ __device__ void set(const uint32_t* w)
    {
        for(int j = 0; j < 4; j++)
            s.z[j] = w[j];
    }
    __device__ uint32_t next() { return lfsr113_next(s); }
    __device__ void init(const device_tables& t, uint32_t e, uint64_t off)
    {
        // lfsr113.hpp:45-62: the step count travels as an unsigned int.
        gf2_jump<4>(s.z, t.lfsr_seq, (uint32_t)e);
        gf2_jump<4>(s.z, t.lfsr_jump, (uint32_t)off);
    }
    __device__ void advance(const device_tables& t, uint64_t k) { gf2_jump<4>(s.z, t.lfsr_jump, k); }
    static constexpr int LDS_WORDS = 32 * 4 * 4;
    __device__ void advance_block(const device_tables& t, uint64_t k, uint32_t* lds)
    {
        gf2_jump_block<4>(s.z, t.lfsr_jump, k, lds);
    }
};


// Non-temporal store of BYTES at base + off through a raw buffer resource built from a
// wavefront-uniform base: the address arithmetic stays in scalar registers (a global
// store with a per-lane 64-bit pointer costs two VALU additions a step).  `limit` bounds
// the byte offsets (the hardware drops a store at or beyond it).
template<int BYTES>
__device__ __forceinline__ void store_uniform_base(char* base, uint32_t off, uint32_t limit, const void* src)
{
    // dword 3 of a gfx9 buffer resource: 32-bit data format, raw (stride 0) addressing
    const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(base, 0, (int)limit, 0x00020000);
    constexpr int NT = 2; // cache-policy bit 1: nt on gfx950
    using v2 = unsigned int __attribute__((ext_vector_type(2)));
    using v4 = unsigned int __attribute__((ext_vector_type(4)));
    if constexpr(BYTES == 4)
        __builtin_amdgcn_raw_buffer_store_b32(*reinterpret_cast<const uint32_t*>(src), r, (int)off, 0, NT);
    else if constexpr(BYTES == 8)
        __builtin_amdgcn_raw_buffer_store_b64(*reinterpret_cast<const v2*>(src), r, (int)off, 0, NT);
    else
        __builtin_amdgcn_raw_buffer_store_b128(*reinterpret_cast<const v4*>(src), r, (int)off, 0, NT);
}

} // namespace inter_dev
} // namespace rr
