// rr_counter_dev.hpp -- device pieces of the counter-based engines' generation, shared by
// the per-call kernels (rr_counter.hip) and the batched kernels (rr_batched.hip).
//
// Output order reproduces generate_philox (library/src/rng/philox4x32_10.hpp:106-214)
// and generate_threefry (library/src/rng/threefry.hpp:86-198) under the DEFAULT grid:
// vec index v (one engine block of DIM words, FULL outputs) starts at word position
// base + DIM*v, where base = engine position + H, except that vec indices that are
// multiples of the reference grid stride start at engine position + C0 (the words the
// head actually consumed).
#pragma once

#include <type_traits>

#include "rr_launch.hpp"

namespace rr
{
namespace ctr_dev
{



template<class W, int DIM>
__device__ __forceinline__ void interleave(const W* a, const W* b, uint32_t sub, W* o)
{
    if constexpr(DIM == 4)
    {
        switch(sub)
        {
            case 0: o[0] = a[0]; o[1] = a[1]; o[2] = a[2]; o[3] = a[3]; break;
            case 1: o[0] = a[1]; o[1] = a[2]; o[2] = a[3]; o[3] = b[0]; break;
            case 2: o[0] = a[2]; o[1] = a[3]; o[2] = b[0]; o[3] = b[1]; break;
            default: o[0] = a[3]; o[1] = b[0]; o[2] = b[1]; o[3] = b[2]; break;
        }
    }
    else
    {
        if(sub == 0)
        {
            o[0] = a[0];
            o[1] = a[1];
        }
        else
        {
            o[0] = a[1];
            o[1] = b[0];
        }
    }
}

// Sequential reader over the word stream (the engine's next()), for head/tail.
template<ctr_kind K>
struct seq_reader
{
    using T    = ctr_traits<K>;
    using W    = typename T::word;
    static constexpr int DIM = T::DIM;
    W        ctr[DIM];
    W        cur[DIM];
    uint32_t sub;
    __device__ seq_reader(const ctr_pos& p, uint64_t add_ctr, const ctr_key& key)
    {
#pragma unroll
        for(int i = 0; i < DIM; i++)
            ctr[i] = (W)p.c[i];
        ctr_add<W, DIM>(ctr, add_ctr);
        sub = p.sub;
        T::block(ctr, key, cur);
    }
    __device__ W next(const ctr_key& key)
    {
        W r = cur[0];
#pragma unroll
        for(int i = 1; i < DIM; i++)
            if((int)sub == i)
                r = cur[i];
        sub++;
        if((int)sub == DIM)
        {
            sub = 0;
            ctr_add<W, DIM>(ctr, 1);
            T::block(ctr, key, cur);
        }
        return r;
    }
};

// Non-temporal vector stores (global_store_dwordx2/x4 ... nt): the output is written
// once and not read back by the kernel.
template<int BYTES>
__device__ __forceinline__ void store_vec(void* dst, const uint32_t* u)
{
    using v2 = unsigned int __attribute__((ext_vector_type(2)));
    using v4 = unsigned int __attribute__((ext_vector_type(4)));
    if constexpr(BYTES == 8)
    {
        const v2 t = {u[0], u[1]};
        __builtin_nontemporal_store(t, reinterpret_cast<v2*>(dst));
    }
    else
    {
#pragma unroll
        for(int j = 0; j < BYTES / 16; j++)
        {
            const v4 t = {u[4 * j], u[4 * j + 1], u[4 * j + 2], u[4 * j + 3]};
            __builtin_nontemporal_store(t, reinterpret_cast<v4*>(dst) + j);
        }
    }
}

// Head and tail of a call (at most FULL-1 outputs each), written by one thread.
// stride_mask = c.ref_stride - 1 (the stride is a power of two), passed by the caller: a
// read of the field here keeps the batched kernels' per-tile copy of the call in scratch.
template<ctr_kind K, class D>
__device__ void head_tail(const counter_call& c, uint64_t stride_mask)
{
    using T                    = ctr_traits<K>;
    using W                    = typename T::word;
    using out_t                = typename D::out_t;
    constexpr int DIM          = T::DIM;
    constexpr int IN           = D::in_w;
    constexpr int OUT          = D::out_w;
    constexpr int OPT          = DIM / IN;
    out_t* const  data         = static_cast<out_t*>(c.data);
    {
        out_t tmp[OUT];
        if(c.head > 0)
        {
            // Head: consecutive words from the engine position (thread 0 of the reference grid).
            seq_reader<K> r(c.pos, 0, c.key);
            for(int s = 0; s < OPT; s++)
            {
                if((uint32_t)(s * OUT) >= c.head)
                    break;
                W in[IN];
#pragma unroll
                for(int i = 0; i < IN; i++)
                    in[i] = r.next(c.key);
                D::apply(in, tmp, c.params);
                for(int o = 0; o < OUT; o++)
                    if((uint32_t)(s * OUT + o) < c.head)
                        data[s * OUT + o] = tmp[o];
            }
        }
        if(c.tail > 0)
        {
            // Tail: the thread whose next vec index is vec_n continues from that vec's position.
            const ctr_pos  b = (c.quirk && (c.vec_n & stride_mask) == 0) ? c.vec_base0 : c.vec_base;
            seq_reader<K>  r(b, c.vec_n, c.key);
            for(int s = 0; s < OPT; s++)
            {
                if((uint32_t)(s * OUT) >= c.tail)
                    break;
                W in[IN];
#pragma unroll
                for(int i = 0; i < IN; i++)
                    in[i] = r.next(c.key);
                D::apply(in, tmp, c.params);
                for(int o = 0; o < OUT; o++)
                    if((uint32_t)(s * OUT + o) < c.tail)
                        data[c.n - c.tail + s * OUT + o] = tmp[o];
            }
        }
    }

}


// One vec of a call, any position (two engine blocks when the stream starts mid-block).
template<ctr_kind K, class D>
__device__ __forceinline__ void generic_vec(const counter_call& c, size_t v, uint64_t stride_mask)
{
    using T                    = ctr_traits<K>;
    using W                    = typename T::word;
    using out_t                = typename D::out_t;
    constexpr int DIM          = T::DIM;
    constexpr int IN           = D::in_w;
    constexpr int OUT          = D::out_w;
    constexpr int OPT          = DIM / IN;
    constexpr int FULL         = OPT * OUT;
    constexpr int BYTES        = FULL * (int)sizeof(out_t);
    static_assert(DIM % IN == 0, "input width");
    static_assert(BYTES == 8 || BYTES == 16 || BYTES == 32, "vec size");

    out_t* const   vec_data = static_cast<out_t*>(c.data) + c.mis;
    const ctr_pos  b        = (c.quirk && (v & stride_mask) == 0) ? c.vec_base0 : c.vec_base;
    W              ctr[DIM];
#pragma unroll
    for(int i = 0; i < DIM; i++)
        ctr[i] = (W)b.c[i];
    ctr_add<W, DIM>(ctr, v);
    W w[DIM];
    T::block(ctr, c.key, w);
    W words[DIM];
    if(b.sub != 0)
    {
        W w2[DIM];
        ctr_add<W, DIM>(ctr, 1);
        T::block(ctr, c.key, w2);
        interleave<W, DIM>(w, w2, b.sub, words);
    }
    else
    {
#pragma unroll
        for(int i = 0; i < DIM; i++)
            words[i] = w[i];
    }
    out_t o[FULL];
#pragma unroll
    for(int s = 0; s < OPT; s++)
        D::apply(words + s * IN, o + s * OUT, c.params);
    uint32_t u[BYTES / 4];
    __builtin_memcpy(u, o, BYTES);
    store_vec<BYTES>(vec_data + v * FULL, u);
}

// Fast path for the 32-bit-word engines when every vec shares one engine position
// modulo the block (no head-shift quirk) and the counter's words above bit 63 do not
// change within the call: each lane runs one engine block per vec.
//
// Vecs go in chunks of 64, one per lane, whose starts are congruent to -lane_shift
// modulo 64, so a full chunk's store covers whole 128-byte lines of the output (a
// wavefront store that straddles line boundaries measured 31% slower).  This wavefront
// runs `chunks` consecutive chunks from vec `vb`; lanes outside [lo, end) compute but do
// not store.  When the stream starts mid-block (SUB) vec v takes words from blocks v and
// v + 1: a lane holds block v, the next lane block v + 1 (one DPP wavefront shift per
// word), and lane 63 takes lane 0 of the next chunk's blocks, which are computed one
// chunk ahead (chunks + 1 block evaluations in all).  Every lane of the wavefront calls
// this.
template<ctr_kind K, class D, bool SUB>
__device__ __forceinline__ void fast_run(const counter_call& c, int64_t vb, uint32_t chunks, int64_t lo, int64_t end)
{
    using T                = ctr_traits<K>;
    using W                = typename T::word;
    using out_t            = typename D::out_t;
    constexpr int DIM      = T::DIM;
    constexpr int IN       = D::in_w;
    constexpr int OUT      = D::out_w;
    constexpr int OPT      = DIM / IN;
    constexpr int FULL     = OPT * OUT;
    constexpr int BYTES    = FULL * (int)sizeof(out_t);
    static_assert(sizeof(W) == 4, "32-bit-word engines only");

    const uint64_t clo  = c.vec_base.c[0] | (c.vec_base.c[1] << 32);
    const W        c2   = (W)c.vec_base.c[2];
    const W        c3   = (W)c.vec_base.c[3];
    const uint32_t sub  = c.vec_base.sub;
    const uint32_t lane = threadIdx.x & 63u;
    out_t* const   vec_data = static_cast<out_t*>(c.data) + c.mis;
    // block of vec index b (b outside the call gives words no lane stores)
    auto block_at = [&](int64_t b, W* w) {
        const uint64_t cc = clo + (uint64_t)b;
        W              ctr[DIM];
        ctr[0] = (W)cc;
        ctr[1] = (W)(cc >> 32);
        if constexpr(DIM == 4)
        {
            ctr[2] = c2;
            ctr[3] = c3;
        }
        T::block(ctr, c.key, w);
    };

    W cur[DIM];
    block_at(vb + lane, cur);
    for(uint32_t k = 0; k < chunks; k++)
    {
        const int64_t base = vb + 64 * (int64_t)k;
        if(base >= end)
            break;
        const int64_t v     = base + lane;
        const bool    ahead = SUB || k + 1 < chunks;
        W             nxt[DIM];
        if(ahead)
            block_at(v + 64, nxt);
        W words[DIM];
        if constexpr(SUB)
        {
            W nb[DIM];
#pragma unroll
            for(int i = 0; i < DIM; i++)
            {
                // wave_shl:1 (lane l reads lane l + 1); lane 63 keeps `old`, the next
                // chunk's first block word
                const int first_next = __builtin_amdgcn_readlane((int)nxt[i], 0);
                nb[i] = (W)__builtin_amdgcn_update_dpp(first_next, (int)cur[i], 0x130, 0xf, 0xf, false);
            }
            interleave<W, DIM>(cur, nb, sub, words);
        }
        else
        {
#pragma unroll
            for(int i = 0; i < DIM; i++)
                words[i] = cur[i];
        }
        if(v >= lo && v < end)
        {
            out_t o[FULL];
#pragma unroll
            for(int s = 0; s < OPT; s++)
                D::apply(words + s * IN, o + s * OUT, c.params);
            uint32_t u[BYTES / 4];
            __builtin_memcpy(u, o, BYTES);
            store_vec<BYTES>(vec_data + v * FULL, u);
        }
        if(ahead)
        {
#pragma unroll
            for(int i = 0; i < DIM; i++)
                cur[i] = nxt[i];
        }
    }
}

// The ALU-bound ThreeFry engines' mid-block start (SUB) without the look-ahead block: a
// wavefront covers 63 vecs with 64 consecutive blocks (the one-block-per-chunk extra of
// fast_run costs them more than the aligned stores gain: ThreeFry4x32 uniform f32 at
// offset 2^20 + 7 measured 3.10 TB/s with it, 3.98 this way).  `first` is the wavefront's
// first vec; every lane of the wavefront calls this.
template<ctr_kind K, class D>
__device__ __forceinline__ void fast_tile63(const counter_call& c, size_t first, size_t end)
{
    using T                = ctr_traits<K>;
    using W                = typename T::word;
    using out_t            = typename D::out_t;
    constexpr int DIM      = T::DIM;
    constexpr int IN       = D::in_w;
    constexpr int OUT      = D::out_w;
    constexpr int OPT      = DIM / IN;
    constexpr int FULL     = OPT * OUT;
    constexpr int BYTES    = FULL * (int)sizeof(out_t);
    static_assert(sizeof(W) == 4, "32-bit-word engines only");

    const uint64_t clo  = c.vec_base.c[0] | (c.vec_base.c[1] << 32);
    const uint32_t lane = threadIdx.x & 63u;
    out_t* const   vec_data = static_cast<out_t*>(c.data) + c.mis;
    const size_t   v        = first + lane;
    const uint64_t cc       = clo + v;
    W              ctr[DIM];
    ctr[0] = (W)cc;
    ctr[1] = (W)(cc >> 32);
    if constexpr(DIM == 4)
    {
        ctr[2] = (W)c.vec_base.c[2];
        ctr[3] = (W)c.vec_base.c[3];
    }
    W w[DIM];
    T::block(ctr, c.key, w);
    W nx[DIM];
#pragma unroll
    for(int i = 0; i < DIM; i++)
        nx[i] = (W)__builtin_amdgcn_update_dpp(0, (int)w[i], 0x130, 0xf, 0xf, false); // lane l + 1
    if(v >= end || lane == 63u)
        return;
    // the word offset as a constant (a run-time one can leave the eight words in scratch)
    auto finish = [&](auto sub_c) {
        constexpr uint32_t S = decltype(sub_c)::value;
        W                  words[DIM];
#pragma unroll
        for(int i = 0; i < DIM; i++)
            words[i] = (uint32_t)i + S < (uint32_t)DIM ? w[i + S] : nx[i + S - DIM];
        out_t o[FULL];
#pragma unroll
        for(int s = 0; s < OPT; s++)
            D::apply(words + s * IN, o + s * OUT, c.params);
        uint32_t u[BYTES / 4];
        __builtin_memcpy(u, o, BYTES);
        store_vec<BYTES>(vec_data + v * FULL, u);
    };
    const uint32_t sub = c.vec_base.sub; // 1 .. DIM - 1
    if(sub == 1)
        finish(std::integral_constant<uint32_t, 1>{});
    else if constexpr(DIM == 4)
    {
        if(sub == 2)
            finish(std::integral_constant<uint32_t, 2>{});
        else
            finish(std::integral_constant<uint32_t, 3>{});
    }
}

// Whether fast_run applies to a call.
template<ctr_kind K>
__host__ __device__ inline bool fast_ok(const counter_call& c)
{
    if constexpr(sizeof(typename ctr_traits<K>::word) != 4)
        return false;
    else
    {
        const uint64_t clo     = c.vec_base.c[0] | (c.vec_base.c[1] << 32);
        const bool     carries = ctr_traits<K>::DIM == 4 && clo + c.vec_n + 1 < clo;
        return !c.quirk && !carries;
    }
}

} // namespace ctr_dev
} // namespace rr
