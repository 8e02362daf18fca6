// rr_sobol.hip -- Sobol32 / Sobol64 (plain and scrambled) quasi-random generation on gfx950.
//
// Output map (library/src/rng/sobol.hpp:72-261, 618-718): element p of dimension d
// (stored at data[d * size + p]) is point i = offset + p of that dimension,
//     x(i) = XOR_k [bit k of gray(i)] * V_d[k],  gray(i) = i ^ (i >> 1),
// XOR the dimension's scramble constant for the scrambled variants; the index is
// 32-bit (the generator's offset is an unsigned int).  The reference walks i with
// Gray-code leap-frog steps; any schedule computing x(i) is equivalent.
//
// Schedule: a wavefront owns 64*M consecutive points starting at a multiple of 64.
// Writing i = 64*A + l (lane l), gray(i) = (gray(A) << 6) ^ ((A & 1) << 5) ^ gray(l), so
//     x(i) = W(A) ^ L(l),
// where L(l) (six direction vectors, per lane) is computed once and W(A) is
// wavefront-uniform and advances as W(A+1) = W(A) ^ V[6 + ctz(A+1)] ^ V[5].  Each
// output costs one XOR plus its transform; the store of a wavefront is 64
// consecutive elements.
#include "rr_launch.hpp"

namespace rr
{
namespace
{

// rocrand_discrete.h:109-165: binary search of the CDF; the draw is x * 2^-32
// (32-bit values) or x * 2^-64 (64-bit values).  The table is the block's copy in
// dynamic LDS (sobol_kernel stages it), so every probe is an LDS read.
__device__ __forceinline__ unsigned int discrete_cdf_lds(double x, const alias_table& t)
{
    extern __shared__ double lds_cdf[];
    unsigned int lo = 0, hi = t.size - 1;
    do
    {
        const unsigned int c = (lo + hi) / 2;
        if(x > lds_cdf[c])
            lo = c + 1;
        else
            hi = c;
    } while(lo != hi);
    return t.offset + lo;
}
__device__ __forceinline__ double cdf_draw(uint32_t x) { return __dmul_rn((double)x, R_2POW32_INV_DBL); }
__device__ __forceinline__ double cdf_draw(uint64_t x) { return __dmul_rn((double)x, R_2POW64_INV_DBL); }

template<out_kind K>
struct sob_out;
#define RR_SOB(K, T) \
    template<>       \
    struct sob_out<K> \
    {                \
        using type = T; \
    };
RR_SOB(out_kind::u8, uint8_t)
RR_SOB(out_kind::u16, uint16_t)
RR_SOB(out_kind::u32, uint32_t)
RR_SOB(out_kind::u64, uint64_t)
RR_SOB(out_kind::uniform_f32, float)
RR_SOB(out_kind::uniform_f64, double)
RR_SOB(out_kind::uniform_f16, __half)
RR_SOB(out_kind::normal_f32, float)
RR_SOB(out_kind::normal_f64, double)
RR_SOB(out_kind::normal_f16, __half)
RR_SOB(out_kind::lognormal_f32, float)
RR_SOB(out_kind::lognormal_f64, double)
RR_SOB(out_kind::lognormal_f16, __half)
RR_SOB(out_kind::poisson, uint32_t)
#undef RR_SOB

// sobol_*_distribution functors (uniform.hpp:351-442, normal.hpp:303-364,
// log_normal.hpp:310-370, poisson.hpp:49-108 with the CDF method).
template<out_kind K, class V>
__device__ __forceinline__ typename sob_out<K>::type sob_apply(V x, const dist_params& p)
{
    constexpr int B = sizeof(V) * 8;
    if constexpr(K == out_kind::u8)
        return (uint8_t)(x >> (B - 8));
    else if constexpr(K == out_kind::u16)
        return (uint16_t)(x >> (B - 16));
    else if constexpr(K == out_kind::u32)
        return (uint32_t)(x >> (B - 32));
    else if constexpr(K == out_kind::u64)
        return (uint64_t)x;
    else if constexpr(K == out_kind::uniform_f32)
        return sob_uf32(x);
    else if constexpr(K == out_kind::uniform_f64)
        return sob_uf64(x);
    else if constexpr(K == out_kind::uniform_f16)
        return uniform_f16((uint32_t)(x >> (B - 16)));
    else if constexpr(K == out_kind::normal_f32)
        return __fadd_rn(p.mf, __fmul_rn(sob_normal_f32(x), p.sf));
    else if constexpr(K == out_kind::normal_f64)
        return __dadd_rn(p.md, __dmul_rn(sob_normal_f64(x), p.sd));
    else if constexpr(K == out_kind::normal_f16)
        return __float2half(__fadd_rn(__half2float(p.mh), __fmul_rn(sob_normal_f32(x), __half2float(p.sh))));
    else if constexpr(K == out_kind::lognormal_f32)
        return expf_acc(__fadd_rn(p.mf, __fmul_rn(p.sf, sob_normal_f32(x))));
    else if constexpr(K == out_kind::lognormal_f64)
        return exp_cr(__dadd_rn(p.md, __dmul_rn(p.sd, sob_normal_f64(x))));
    else if constexpr(K == out_kind::lognormal_f16)
        return __float2half(
            expf_acc(__fadd_rn(__half2float(p.mh), __fmul_rn(__half2float(p.sh), sob_normal_f32(x)))));
    else
    {
        if(p.huge)
            return (uint32_t)round(__dadd_rn(__dmul_rn(p.sqrt_lambda, sob_normal_f64(x)), p.lambda));
        return discrete_cdf_lds(cdf_draw(x), p.pt);
    }
}

template<class T>
__device__ __forceinline__ void store_one(T* p, T v)
{
    __builtin_nontemporal_store(v, p);
}
__device__ __forceinline__ void store_one(__half* p, __half v)
{
    __builtin_nontemporal_store(__half_as_ushort(v), reinterpret_cast<unsigned short*>(p));
}

template<class V, out_kind K>
__global__ __launch_bounds__(256) void sobol_kernel(const sobol_call c)
{
    using T            = typename sob_out<K>::type;
    constexpr int B    = sizeof(V) * 8;
    const uint32_t dim = blockIdx.y;
    const V*       vec = static_cast<const V*>(c.vectors) + (size_t)dim * B;
    const V        sc  = c.scramble ? static_cast<const V*>(c.scramble)[dim] : V(0);
    T*             out = static_cast<T*>(c.data) + (size_t)dim * c.size;

    // Poisson: the block searches its copy of the CDF in dynamic LDS (at most 4096
    // doubles), so the binary search's dependent loads are LDS round trips; every thread
    // copies before any leaves
    if constexpr(K == out_kind::poisson)
    {
        extern __shared__ double lds_cdf[];
        if(!c.params.huge)
        {
            for(uint32_t i = threadIdx.x; i < c.params.pt.size; i += blockDim.x)
                lds_cdf[i] = c.params.pt.cdf[i];
            __syncthreads();
        }
    }

    // the wavefront index as a scalar, so that the group index below, its Gray-code
    // steps and their direction-vector loads are scalar too (a per-lane load here would
    // make every step wait for the wavefront's outstanding stores)
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(blockIdx.x * 4u + (threadIdx.x >> 6));
    if(wave >= c.waves)
        return;

    // A wavefront's 64 lanes write 64 consecutive outputs.  To make each such store start
    // on a 128-byte line (64-byte for 1-byte outputs) the lanes cover points
    // 64 G + d + lane, d = (offset - output element address) mod the line's element count:
    // lanes with lane + d < 64 are points of group G, the others of group G + 1.  A point's
    // value is W(group) ^ L(low six bits): L per lane, W for both groups per wavefront.
    constexpr uint32_t U   = (sizeof(T) >= 2 ? 128u : 64u) / (uint32_t)sizeof(T);
    const uint32_t     d   = (uint32_t)((c.offset - reinterpret_cast<uintptr_t>(out) / sizeof(T)) & (U - 1));
    const uint32_t     low = (lane + d) & 63u;
    const bool         hi  = lane + d >= 64u;

    // L: the low six Gray-code bits of the lane's point
    const uint32_t gl = low ^ (low >> 1);
    V              L  = 0;
#pragma unroll
    for(int k = 0; k < 6; k++)
        if(gl & (1u << k))
            L ^= vec[k];
    L ^= sc;

    // W(A) for the wavefront's first group.  G counts 64-point groups of the output range
    // (from one group before the first, for the shift d); the direction values use the
    // 32-bit point index, i.e. A = G mod 2^26.
    const int64_t  G  = c.first_group + (int64_t)wave * c.groups_per_wave;
    uint32_t       A  = (uint32_t)G & 0x03ffffffu;
    const uint32_t gA = A ^ (A >> 1);
    V              Wc = (A & 1u) ? vec[5] : V(0);
    for(int k = 0; k < 26; k++)
        if(gA & (1u << k))
            Wc ^= vec[6 + k];
    // W(A + 1): one Gray-code step from W(A).  The direction vectors are read-only for
    // the kernel: through the constant address space a wavefront-uniform index becomes a
    // scalar load, which the stores do not delay.
    using cvec_t = const __attribute__((address_space(4))) V*;
    const cvec_t cvec = (cvec_t)vec;
    auto step = [&](V w, uint32_t a1) { return a1 == 0 ? V(0) : V(w ^ cvec[5] ^ cvec[6 + __builtin_ctz(a1)]); };
    uint32_t An = (A + 1u) & 0x03ffffffu;
    V        Wn = step(Wc, An);

    int64_t e = G * 64 + (int64_t)(d + lane) - (int64_t)c.offset;
    for(uint32_t m = 0; m < c.groups_per_wave; m++)
    {
        if(e >= 0 && (uint64_t)e < c.size)
            store_one(out + e, sob_apply<K, V>((hi ? Wn : Wc) ^ L, c.params));
        e += 64;
        Wc = Wn;
        An = (An + 1u) & 0x03ffffffu;
        Wn = step(Wn, An);
    }
}

template<class V>
hipError_t dispatch(out_kind o, const sobol_call& c, dim3 grid, hipStream_t s)
{
    const size_t lds = o == out_kind::poisson && !c.params.huge ? c.params.pt.size * sizeof(double) : 0;
#define RR_CASE(K)                                                                            \
    case K: hipLaunchKernelGGL((sobol_kernel<V, K>), grid, dim3(256), lds, s, c); break;
    switch(o)
    {
        RR_CASE(out_kind::u8)
        RR_CASE(out_kind::u16)
        RR_CASE(out_kind::u32)
        RR_CASE(out_kind::uniform_f32)
        RR_CASE(out_kind::uniform_f64)
        RR_CASE(out_kind::uniform_f16)
        RR_CASE(out_kind::normal_f32)
        RR_CASE(out_kind::normal_f64)
        RR_CASE(out_kind::normal_f16)
        RR_CASE(out_kind::lognormal_f32)
        RR_CASE(out_kind::lognormal_f64)
        RR_CASE(out_kind::lognormal_f16)
        RR_CASE(out_kind::poisson)
        case out_kind::u64:
            if constexpr(sizeof(V) == 8)
            {
                hipLaunchKernelGGL((sobol_kernel<V, out_kind::u64>), grid, dim3(256), 0, s, c);
                break;
            }
            else
                return hipErrorInvalidValue;
    }
#undef RR_CASE
    return hipGetLastError();
}

} // namespace

hipError_t launch_sobol(bool is64, out_kind o, sobol_call c, hipStream_t s)
{
    if(c.size == 0 || c.dims == 0)
        return hipSuccess;
    // Points [offset, offset + size) of every dimension, in 64-point groups, from one
    // group earlier: the kernel shifts a dimension's groups by up to 63 points to align
    // its stores.
    const uint64_t first = c.offset / 64u;
    const uint64_t last  = ((uint64_t)c.offset + c.size + 63u) / 64u; // exclusive
    const uint64_t groups = last - first + 1;
    // 16 groups (1024 points) per wavefront keeps the uniform set-up negligible
    // while a 2^20-point dimension still spans 256 wavefronts.
    uint32_t gpw = 16;
    if(groups < 4096)
        gpw = 1;
    const uint64_t waves = (groups + gpw - 1) / gpw;
    if(waves > 0xffffffffull / 4 || c.dims > 65535)
        return hipErrorInvalidValue;
    c.first_group     = (int64_t)first - 1;
    c.groups_per_wave = gpw;
    c.waves           = (uint32_t)waves;
    const dim3 grid((uint32_t)((waves + 3) / 4), c.dims);
    return is64 ? dispatch<uint64_t>(o, c, grid, s) : dispatch<uint32_t>(o, c, grid, s);
}

} // namespace rr
