// rr_interleaved.hip -- MRG32k3a, MRG31k3p, XORWOW and LFSR113 bulk generation on gfx950.
//
// These engines are sequential, so the reference runs S = 131072 independent
// engines (one per thread of its 512 x 256 grid) and interleaves them: output
// vec v comes from engine (v + start) mod S at that engine's (v / S)-th step
// (library/src/rng/mrg.hpp:65-157, xorwow.hpp:63-161, lfsr113.hpp:66-160).
// The engine states stay resident in HBM between calls, structure-of-arrays
// (word j of engine e at states[j*S + e]) so every state load/store and every
// output store of a wavefront is one contiguous segment.
//
// Engine e is initialised by a jump-ahead kernel: the seeded state advanced by
// e subsequences and then by offset/S + [e < offset mod S] steps, with matrix
// powers (mod-m 3x3 for the MRGs, GF(2) for XORWOW / LFSR113) computed on the
// host once per device (rr_tables.cpp).
#include "rr_interleaved_dev.hpp"

#include <algorithm>

namespace rr
{
namespace
{

using namespace inter_dev;

template<inter_kind K>
__global__ __launch_bounds__(256) void init_kernel(uint32_t* states, const device_tables t, ieng<K> seeded,
                                                   uint32_t start, uint64_t base_off, uint32_t ns)
{
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if(e >= ns)
        return;
    ieng<K> g = seeded;
    g.init(t, e, base_off + (e < start ? 1u : 0u));
    g.store(states, e, ns);
}

// Poisson: each block copies the packed alias table into dynamic LDS (size * 8 bytes, no
// table above lambda 4000) and samples from there.  Every thread of the block calls this
// before any thread leaves.
template<class D>
__device__ __forceinline__ void stage_poisson_table(interleaved_call& c)
{
    if constexpr(dist_kind<D>::value == out_kind::poisson)
    {
        extern __shared__ uint64_t lds_packed[];
        const uint32_t             size = c.params.huge ? 0u : c.params.pt.size;
        for(uint32_t i = threadIdx.x; i < size; i += blockDim.x)
            lds_packed[i] = c.params.pt.packed[i];
        __syncthreads();
        c.params.pt.packed = lds_packed;
    }
}

template<class D>
size_t poisson_lds_bytes(const interleaved_call& c)
{
    return dist_kind<D>::value == out_kind::poisson && !c.params.huge ? c.params.pt.size * sizeof(uint64_t) : 0;
}

// Lane p of an S-lane grid (0 <= p < S) runs engine (p - shift) mod S, so that the
// wavefront of lanes 64w..64w+63 stores engines 64w - shift onwards, whose first vec sits
// on a 128-byte line of the output.
__device__ __forceinline__ uint32_t lane_engine(uint32_t p, uint32_t shift, uint32_t S)
{
    return p >= shift ? p - shift : p + S - shift;
}

// The highest engine of 256-lane block b: the one with the fewest vec-steps.
__device__ __forceinline__ uint32_t last_engine(uint32_t b, uint32_t shift, uint32_t S)
{
    return b == 0 && shift != 0 ? S - 1 : b * 256u + 255u - shift;
}

template<inter_kind K, class D>
__global__ __launch_bounds__(256) void generate_kernel(interleaved_call c)
{
    using out_t          = typename D::out_t;
    constexpr int IN     = D::in_w;
    constexpr int OUT    = D::out_w;
    constexpr int BYTES  = OUT * (int)sizeof(out_t);
    static_assert(BYTES == 4 || BYTES == 8 || BYTES == 16, "vec size");

    stage_poisson_table<D>(c);
    const uint32_t S = c.S;
    // engines over lanes shifted by lane_shift, so that each wavefront's stores start on
    // a 128-byte line; the last lane_shift engines wrap round to block 0's first lanes
    // (an extra block instead left one CU with three blocks of a two-per-CU grid)
    const uint32_t id = lane_engine(blockIdx.x * blockDim.x + threadIdx.x, c.lane_shift, S);
    const uint32_t eid = (id + c.start) % S;
    ieng<K>        g;
    g.load(c.states, eid, S);

    out_t* const data     = static_cast<out_t*>(c.data);
    out_t* const vec_data = data + c.mis;
    alignas(16) out_t o[OUT];
    uint32_t     in[IN];
    size_t       index = id;
    // The transforms (and table loads) of consecutive steps overlap when unrolled, which
    // the 2 wavefronts per SIMD of an S-lane grid cannot provide; integer and uniform
    // outputs are store-bound already, except LFSR113's doubles (two of its longer steps
    // per output: unrolled x2, 4.43 -> 5.31 TB/s; XORWOW's measured 2% slower so).
    constexpr out_kind ok     = dist_kind<D>::value;
    constexpr int      UNROLL = (ok == out_kind::u8 || ok == out_kind::u16 || ok == out_kind::u32
                                 || ok == out_kind::uniform_f32 || ok == out_kind::uniform_f16)
                                    ? 1
                                    : ok == out_kind::uniform_f64 ? (K == inter_kind::lfsr113 ? 2 : 1)
                                                                  : 4;
    // The block's last engine takes the fewest vec-steps: that many run for every lane
    // under a scalar counter, storing from a block-uniform base (buffer stores), and the
    // rest (at most one) per lane.
    const uint32_t id_last = last_engine(blockIdx.x, c.lane_shift, S);
    const uint32_t ucnt = id_last < c.vec_n ? (uint32_t)((c.vec_n - id_last + S - 1) / S) : 0u;
    const uint32_t     off     = id * (uint32_t)BYTES;
    const uint32_t     lim     = S * (uint32_t)BYTES;
    char*              base    = reinterpret_cast<char*>(vec_data);
#pragma unroll UNROLL
    for(uint32_t k = 0; k < ucnt; k++, index += S, base += (size_t)S * BYTES)
    {
#pragma unroll
        for(int i = 0; i < IN; i++)
            in[i] = g.next();
        D::apply(in, o, c.params);
        store_uniform_base<BYTES>(base, off, lim, o);
    }
    for(; index < c.vec_n; index += S, base += (size_t)S * BYTES)
    {
#pragma unroll
        for(int i = 0; i < IN; i++)
            in[i] = g.next();
        D::apply(in, o, c.params);
        store_uniform_base<BYTES>(base, off, lim, o);
    }
    // The engine that would produce vec index vec_n also writes the head and
    // then the tail (mrg.hpp:114-153).
    if(OUT > 1 && index == c.vec_n)
    {
        if(c.head > 0)
        {
#pragma unroll
            for(int i = 0; i < IN; i++)
                in[i] = g.next();
            D::apply(in, o, c.params);
            for(uint32_t j = 0; j < (uint32_t)OUT; j++)
                if(j < c.head)
                    data[j] = o[j];
        }
        if(c.tail > 0)
        {
#pragma unroll
            for(int i = 0; i < IN; i++)
                in[i] = g.next();
            D::apply(in, o, c.params);
            for(uint32_t j = 0; j < (uint32_t)OUT; j++)
                if(j < c.tail)
                    data[c.n - c.tail + j] = o[j];
        }
    }
    g.store(c.states_out, eid, S);
}

// Long calls: each engine's vec-steps are split into J chunks of C, one lane per
// (engine, chunk).  A lane jumps its copy of the engine state ahead by j*C*IN words
// (MRG: the A^(2^i) tables, a handful of 3x3 modular products; XORWOW / LFSR113: the
// GF(2) 4^k powers, with C a power of 4 so that j*C*IN has at most three nonzero base-4
// digits), so S*J lanes run the serial recurrence instead of S: the per-engine
// dependency chain and the S-lane grid's two wavefronts per SIMD no longer limit the
// store rate.  Lanes are ordered chunk-major, so a wavefront's 64 lanes are 64
// consecutive engines of one chunk and each of its stores is one contiguous segment.
// The lane whose range ends at the engine's last vec-step also writes the head/tail
// (for the engine that produces vec index vec_n) and stores the final state.
template<inter_kind K, class D>
__global__ __launch_bounds__(256) void generate_chunked_kernel(interleaved_call c, const device_tables t,
                                                               uint32_t J, uint32_t C)
{
    using out_t         = typename D::out_t;
    constexpr int IN    = D::in_w;
    constexpr int OUT   = D::out_w;
    constexpr int BYTES = OUT * (int)sizeof(out_t);
    static_assert(BYTES == 4 || BYTES == 8 || BYTES == 16, "vec size");

    // S is a multiple of the block size: the chunk index is uniform per block, so the
    // jump's digit loop and its table rows stay in scalar registers
    const uint32_t     S   = c.S;
    const uint32_t     BPC = S / 256; // blocks per chunk
    const uint32_t     j   = blockIdx.x / BPC;
    // engines over lanes shifted by lane_shift, wrapping as in generate_kernel
    const uint32_t     id  = lane_engine((blockIdx.x % BPC) * 256 + threadIdx.x, c.lane_shift, S);
    if(j >= J)
        return; // the whole block
    stage_poisson_table<D>(c);
    // vec-steps of this engine: vec indices id, id + S, ... below vec_n
    const uint64_t steps  = id < c.vec_n ? (c.vec_n - id + S - 1) / S : 0;
    const uint64_t k0     = (uint64_t)j * C;
    const uint64_t end_j  = steps / C < J ? steps / C : J - 1; // chunk that ends at `steps`
    const bool     ending = j == end_j;
    const bool     idle   = k0 > steps && !ending; // nothing to write in this chunk
    const uint32_t eid    = (id + c.start) % S;
    ieng<K>        eng;
    if constexpr(K == inter_kind::xorwow || K == inter_kind::lfsr113)
    {
        // GF(2) jumps: k0 is the same for the whole block, which stages each matrix in
        // LDS once; every thread takes part (barriers) before idle ones leave
        __shared__ uint32_t mat[ieng<K>::LDS_WORDS];
        eng.load(c.states, eid, S);
        eng.advance_block(t, k0 * IN, mat);
        if(idle)
            return;
    }
    else
    {
        if(idle)
            return;
        eng.load(c.states, eid, S);
        eng.advance(t, k0 * IN);
    }

    out_t* const data     = static_cast<out_t*>(c.data);
    out_t* const vec_data = data + c.mis;
    alignas(16) out_t o[OUT];
    uint32_t     in[IN];
    const uint64_t k1  = ending ? steps : (k0 + C < steps ? k0 + C : steps);
    const uint32_t cnt = (uint32_t)(k1 - k0);
    // Steps that every working lane of the block takes: the block's last engine has the
    // fewest vec-steps (steps_lo), and every lane's count is min(C, steps_lo - k0) or one
    // more, so that many run under a scalar loop counter and the rest (at most one) per
    // lane.  Stores go to a block-uniform base (scalar registers, advanced by S vecs a
    // step) plus the lane's 32-bit byte offset.
    const uint32_t id_last = last_engine(blockIdx.x % BPC, c.lane_shift, S);
    const uint64_t steps_lo = id_last < c.vec_n ? (c.vec_n - id_last + S - 1) / S : 0;
    const uint32_t ucnt
        = steps_lo > k0 ? (uint32_t)(steps_lo - k0 < C ? steps_lo - k0 : (uint64_t)C) : 0u;
    char*          base = reinterpret_cast<char*>(vec_data + k0 * S * OUT);
    const uint32_t off  = id * (uint32_t)BYTES;
    const uint32_t lim  = S * (uint32_t)BYTES;
    uint32_t       k    = 0;
    // three steps per iteration turn the state rotation into renaming
#pragma unroll 3
    for(; k < ucnt; k++, base += (size_t)S * BYTES)
    {
#pragma unroll
        for(int i = 0; i < IN; i++)
            in[i] = eng.next();
        D::apply(in, o, c.params);
        store_uniform_base<BYTES>(base, off, lim, o);
    }
    for(; k < cnt; k++, base += (size_t)S * BYTES)
    {
#pragma unroll
        for(int i = 0; i < IN; i++)
            in[i] = eng.next();
        D::apply(in, o, c.params);
        store_uniform_base<BYTES>(base, off, lim, o);
    }
    if(!ending)
        return;
    if(OUT > 1 && id + steps * S == c.vec_n)
    {
        if(c.head > 0)
        {
#pragma unroll
            for(int i = 0; i < IN; i++)
                in[i] = eng.next();
            D::apply(in, o, c.params);
            for(uint32_t q = 0; q < (uint32_t)OUT; q++)
                if(q < c.head)
                    data[q] = o[q];
        }
        if(c.tail > 0)
        {
#pragma unroll
            for(int i = 0; i < IN; i++)
                in[i] = eng.next();
            D::apply(in, o, c.params);
            for(uint32_t q = 0; q < (uint32_t)OUT; q++)
                if(q < c.tail)
                    data[c.n - c.tail + q] = o[q];
        }
    }
    eng.store(c.states_out, eid, S);
}

template<inter_kind K, class D>
hipError_t launch_gen(const interleaved_call& c, hipStream_t s)
{
    const uint32_t S = c.S;
    if constexpr(K == inter_kind::xorwow || K == inter_kind::lfsr113)
    {
        // XORWOW / LFSR113 write cheap outputs at the store rate with S lanes (2 wavefronts
        // per SIMD).  Poisson's dependent table loads and the longest transforms need more
        // wavefronts: chunks of C a power of 4, at most JMAX per engine.  Measured at 2^28
        // outputs (A/B of builds on one box): J <= 4 takes XORWOW log-normal f32 1.34 ->
        // 2.03 TB/s, normal f64 1.90 -> 1.98, LFSR113 log-normal f32 1.33 -> 1.82, normal
        // f32 2.70 -> 2.90, normal f64 1.67 -> 1.77, while XORWOW normal f32 lost 2% (kept
        // on S lanes) and J <= 16 was slower than S lanes for every transform.
        constexpr out_kind ok   = dist_kind<D>::value;
        constexpr uint64_t JMAX = ok == out_kind::poisson ? 16
                                  : (ok == out_kind::lognormal_f32 || ok == out_kind::lognormal_f64
                                     || ok == out_kind::normal_f64
                                     || (ok == out_kind::normal_f32 && K == inter_kind::lfsr113))
                                      ? 4
                                      : 0;
        if constexpr(JMAX > 0)
        {
            const uint64_t kmax = (c.vec_n + S - 1) / S;
            if(kmax >= 64)
            {
                uint64_t C = 16;
                while((kmax + C - 1) / C > JMAX)
                    C *= 4;
                const uint32_t J = (uint32_t)((kmax + C - 1) / C);
                device_tables  t;
                hipError_t     err = get_device_tables(t);
                if(err != hipSuccess)
                    return err;
                hipLaunchKernelGGL((generate_chunked_kernel<K, D>), dim3((S / 256) * J), dim3(256),
                                   poisson_lds_bytes<D>(c), s, c, t, J, (uint32_t)C);
                return hipGetLastError();
            }
        }
    }
    if constexpr(K == inter_kind::mrg32k3a || K == inter_kind::mrg31k3p)
    {
        // chunks per engine: enough lanes to hide the recurrence's latency
        const uint64_t kmax = (c.vec_n + S - 1) / S;
        uint32_t       J    = 1;
        while(J < 8 && (uint64_t)J * 64 < kmax)
            J *= 2;
        if(J > 1)
        {
            device_tables t;
            hipError_t    err = get_device_tables(t);
            if(err != hipSuccess)
                return err;
            const uint32_t C = (uint32_t)((kmax + J - 1) / J);
            hipLaunchKernelGGL((generate_chunked_kernel<K, D>), dim3((S / 256) * J), dim3(256),
                               poisson_lds_bytes<D>(c), s, c, t, J, C);
            return hipGetLastError();
        }
    }
    hipLaunchKernelGGL((generate_kernel<K, D>), dim3(S / 256), dim3(256), poisson_lds_bytes<D>(c), s, c);
    return hipGetLastError();
}

template<inter_kind K>
hipError_t dispatch(out_kind o, const interleaved_call& c, hipStream_t s)
{
    constexpr word_family F = ieng<K>::F;
    switch(o)
    {
        case out_kind::u8: return launch_gen<K, dist<F, out_kind::u8>>(c, s);
        case out_kind::u16: return launch_gen<K, dist<F, out_kind::u16>>(c, s);
        case out_kind::u32: return launch_gen<K, dist<F, out_kind::u32>>(c, s);
        case out_kind::uniform_f32: return launch_gen<K, dist<F, out_kind::uniform_f32>>(c, s);
        case out_kind::uniform_f64: return launch_gen<K, dist<F, out_kind::uniform_f64>>(c, s);
        case out_kind::uniform_f16: return launch_gen<K, dist<F, out_kind::uniform_f16>>(c, s);
        case out_kind::normal_f32:
            return c.params.unit ? launch_gen<K, dist<F, out_kind::normal_f32, true>>(c, s)
                                 : launch_gen<K, dist<F, out_kind::normal_f32>>(c, s);
        case out_kind::normal_f64: return launch_gen<K, dist<F, out_kind::normal_f64>>(c, s);
        case out_kind::normal_f16: return launch_gen<K, dist<F, out_kind::normal_f16>>(c, s);
        case out_kind::lognormal_f32: return launch_gen<K, dist<F, out_kind::lognormal_f32>>(c, s);
        case out_kind::lognormal_f64: return launch_gen<K, dist<F, out_kind::lognormal_f64>>(c, s);
        case out_kind::lognormal_f16: return launch_gen<K, dist<F, out_kind::lognormal_f16>>(c, s);
        case out_kind::poisson: return launch_gen<K, dist<F, out_kind::poisson>>(c, s);
        case out_kind::u64: return hipErrorInvalidValue;
    }
    return hipErrorInvalidValue;
}

template<inter_kind K>
hipError_t launch_init_k(uint32_t* states, const uint32_t* seed_state, uint32_t start, uint64_t base_off,
                         uint32_t ns, hipStream_t s)
{
    device_tables t;
    hipError_t    err = get_device_tables(t);
    if(err != hipSuccess)
        return err;
    ieng<K> seeded;
    seeded.set(seed_state);
    hipLaunchKernelGGL((init_kernel<K>), dim3((ns + 255) / 256), dim3(256), 0, s, states, t, seeded, start, base_off,
                       ns);
    return hipGetLastError();
}

} // namespace

hipError_t launch_interleaved_init(inter_kind k, uint32_t* states, const uint32_t* seed_state, uint32_t start,
                                   uint64_t base_offset, uint32_t engines, hipStream_t s)
{
    if(engines == 0 || engines % 256 != 0)
        return hipErrorInvalidValue;
    switch(k)
    {
        case inter_kind::mrg32k3a:
            return launch_init_k<inter_kind::mrg32k3a>(states, seed_state, start, base_offset, engines, s);
        case inter_kind::mrg31k3p:
            return launch_init_k<inter_kind::mrg31k3p>(states, seed_state, start, base_offset, engines, s);
        case inter_kind::xorwow:
            return launch_init_k<inter_kind::xorwow>(states, seed_state, start, base_offset, engines, s);
        case inter_kind::lfsr113:
            return launch_init_k<inter_kind::lfsr113>(states, seed_state, start, base_offset, engines, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_interleaved(inter_kind k, out_kind o, const interleaved_call& c, hipStream_t s)
{
    if(c.S == 0 || c.S % 256 != 0)
        return hipErrorInvalidValue;
    switch(k)
    {
        case inter_kind::mrg32k3a: return dispatch<inter_kind::mrg32k3a>(o, c, s);
        case inter_kind::mrg31k3p: return dispatch<inter_kind::mrg31k3p>(o, c, s);
        case inter_kind::xorwow: return dispatch<inter_kind::xorwow>(o, c, s);
        case inter_kind::lfsr113: return dispatch<inter_kind::lfsr113>(o, c, s);
    }
    return hipErrorInvalidValue;
}

} // namespace rr
