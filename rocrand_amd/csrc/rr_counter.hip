// rr_counter.hip -- Philox4x32-10 and ThreeFry{2,4}x{32,64}-20 bulk generation on gfx950.
//
// The output map is described in rr_counter_dev.hpp.  The reference's grid is not this
// kernel's grid: any schedule that computes that map is equivalent.
//
// Schedule: one vec per lane over a grid that covers the whole call, so the
// resident wavefronts always write one contiguous front of the output (16-byte
// stores, 1 KiB per wavefront).  A grid-stride loop over a small grid spreads the
// concurrent stores over the whole buffer and was measured 1.5x slower on MI355X.
#include "rr_counter_dev.hpp"

namespace rr
{
namespace
{

using namespace ctr_dev;

template<ctr_kind K, class D>
__global__ __launch_bounds__(256) void counter_kernel(counter_call c)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for(size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < c.vec_n; v += stride)
        generic_vec<K, D>(c, v, (uint64_t)c.ref_stride - 1);
    if(blockIdx.x == 0 && threadIdx.x == 0)
        head_tail<K, D>(c, (uint64_t)c.ref_stride - 1);
}

// Work of one wavefront: chunks of 64 vecs on the output's line-aligned grid -- one
// without a mid-block start (every wavefront owns one chunk, so the resident wavefronts
// write one contiguous front), 8 with one (a run computes one block beyond its chunks for
// its last lane) -- except that ThreeFry with a mid-block start takes 63-vec tiles.
template<ctr_kind K, bool SUB>
constexpr bool tile63()
{
    return SUB && K != ctr_kind::philox4x32;
}
template<ctr_kind K, bool SUB>
constexpr uint32_t fast_chunks()
{
    return SUB ? 8u : 1u;
}

template<ctr_kind K, class D, bool SUB>
__global__ __launch_bounds__(256) void counter_fast_kernel(counter_call c)
{
    constexpr uint32_t CH   = fast_chunks<K, SUB>();
    const size_t       wave = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t      vb   = (int64_t)(wave * CH * 64) - (int64_t)c.lane_shift;

    if constexpr(dist_kind<D>::value == out_kind::poisson)
    {
        // the packed alias table in LDS (dynamic, size * 8 bytes): one LDS read per sample
        // instead of a global load with a per-lane address
        extern __shared__ uint64_t lds_packed[];
        const uint32_t             size = c.params.huge ? 0u : c.params.pt.size; // no table above 4000
        for(uint32_t i = threadIdx.x; i < size; i += 256)
            lds_packed[i] = c.params.pt.packed[i];
        __syncthreads();
        counter_call cl   = c;
        cl.params.pt.packed = lds_packed;
        if constexpr(tile63<K, SUB>())
        {
            if(wave * 63 < c.vec_n)
                fast_tile63<K, D>(cl, wave * 63, c.vec_n);
        }
        else if(vb < (int64_t)c.vec_n)
            fast_run<K, D, SUB>(cl, vb, CH, 0, (int64_t)c.vec_n);
    }
    else if constexpr(tile63<K, SUB>())
    {
        if(wave * 63 < c.vec_n)
            fast_tile63<K, D>(c, wave * 63, c.vec_n);
    }
    else if(vb < (int64_t)c.vec_n)
        fast_run<K, D, SUB>(c, vb, CH, 0, (int64_t)c.vec_n);
    if(blockIdx.x == 0 && threadIdx.x == 0)
        head_tail<K, D>(c, (uint64_t)c.ref_stride - 1);
}

template<ctr_kind K, class D>
hipError_t launch_one(const counter_call& c, hipStream_t s)
{
    using W = typename ctr_traits<K>::word;
    if constexpr(sizeof(W) == 4)
    {
        // Fast path unless the head shifted some vecs (quirk) or the counter's low
        // 64 bits would carry within the call (4-word engines only).
        if(fast_ok<K>(c))
        {
            const bool   t63    = c.vec_base.sub && K != ctr_kind::philox4x32;
            const size_t per    = t63 ? 63u : 64u * (c.vec_base.sub ? fast_chunks<K, true>() : fast_chunks<K, false>());
            const size_t waves  = (c.vec_n + (t63 ? 0u : c.lane_shift) + per - 1) / per; // one per wavefront
            size_t       blocks = (waves + 3) / 4;
            if(blocks == 0)
                blocks = 1;
            if(blocks > 0x7fffffffu)
                return hipErrorInvalidConfiguration;
            const size_t lds = dist_kind<D>::value == out_kind::poisson && !c.params.huge
                                   ? c.params.pt.size * sizeof(uint64_t)
                                   : 0;
            if(c.vec_base.sub)
                hipLaunchKernelGGL((counter_fast_kernel<K, D, true>), dim3((unsigned)blocks), dim3(256), lds, s, c);
            else
                hipLaunchKernelGGL((counter_fast_kernel<K, D, false>), dim3((unsigned)blocks), dim3(256), lds, s, c);
            return hipGetLastError();
        }
    }
    size_t blocks = (c.vec_n + 255) / 256;
    if(blocks > 0x7fffffffu)
        blocks = 0x7fffffffu; // the kernel strides over what one grid cannot cover
    if(blocks == 0)
        blocks = 1;
    hipLaunchKernelGGL((counter_kernel<K, D>), dim3((unsigned)blocks), dim3(256), 0, s, c);
    return hipGetLastError();
}

template<ctr_kind K>
hipError_t dispatch_kind(out_kind o, bool one_word, const counter_call& c, hipStream_t s)
{
    (void)one_word;
    constexpr word_family F = (K == ctr_kind::threefry2x64 || K == ctr_kind::threefry4x64) ? word_family::w64
                                                                                            : word_family::w32;
    switch(o)
    {
        case out_kind::u8: return launch_one<K, dist<F, out_kind::u8>>(c, s);
        case out_kind::u16: return launch_one<K, dist<F, out_kind::u16>>(c, s);
        case out_kind::u32: return launch_one<K, dist<F, out_kind::u32>>(c, s);
        case out_kind::u64:
            if constexpr(F == word_family::w64)
                return launch_one<K, dist<F, out_kind::u64>>(c, s);
            else
                return hipErrorInvalidValue;
        case out_kind::uniform_f32: return launch_one<K, dist<F, out_kind::uniform_f32>>(c, s);
        case out_kind::uniform_f64: return launch_one<K, dist<F, out_kind::uniform_f64>>(c, s);
        case out_kind::uniform_f16: return launch_one<K, dist<F, out_kind::uniform_f16>>(c, s);
        case out_kind::normal_f32:
            return c.params.unit ? launch_one<K, dist<F, out_kind::normal_f32, true>>(c, s)
                                 : launch_one<K, dist<F, out_kind::normal_f32>>(c, s);
        case out_kind::lognormal_f32: return launch_one<K, dist<F, out_kind::lognormal_f32>>(c, s);
        case out_kind::normal_f16: return launch_one<K, dist<F, out_kind::normal_f16>>(c, s);
        case out_kind::lognormal_f16: return launch_one<K, dist<F, out_kind::lognormal_f16>>(c, s);
        case out_kind::poisson: return launch_one<K, dist<F, out_kind::poisson>>(c, s);
        case out_kind::normal_f64:
            // ThreeFry2x32 doubles use the one-word erfinv transform (threefry.hpp:441-449).
            if constexpr(K == ctr_kind::threefry2x32)
                return launch_one<K, dist<F, out_kind::normal_f64, true>>(c, s);
            else
                return launch_one<K, dist<F, out_kind::normal_f64>>(c, s);
        case out_kind::lognormal_f64:
            if constexpr(K == ctr_kind::threefry2x32)
                return launch_one<K, dist<F, out_kind::lognormal_f64, true>>(c, s);
            else
                return launch_one<K, dist<F, out_kind::lognormal_f64>>(c, s);
    }
    return hipErrorInvalidValue;
}

} // namespace

hipError_t launch_counter(ctr_kind k, out_kind o, bool one_word, const counter_call& c, hipStream_t s)
{
    switch(k)
    {
        case ctr_kind::philox4x32: return dispatch_kind<ctr_kind::philox4x32>(o, one_word, c, s);
        case ctr_kind::threefry2x32: return dispatch_kind<ctr_kind::threefry2x32>(o, one_word, c, s);
        case ctr_kind::threefry2x64: return dispatch_kind<ctr_kind::threefry2x64>(o, one_word, c, s);
        case ctr_kind::threefry4x32: return dispatch_kind<ctr_kind::threefry4x32>(o, one_word, c, s);
        case ctr_kind::threefry4x64: return dispatch_kind<ctr_kind::threefry4x64>(o, one_word, c, s);
    }
    return hipErrorInvalidValue;
}

} // namespace rr
