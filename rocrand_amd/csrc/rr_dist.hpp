// rr_dist.hpp -- output transforms (uniform / normal / log-normal / Poisson) on gfx950.
//
// Each functor maps IN_W engine words to OUT_W outputs, reproducing the
// reference functors of library/src/rng/distribution/{uniform,normal,log_normal,
// poisson}.hpp over the scalar math of rocrand_{uniform,normal,discrete}.h.
//
// Exactness: integer and uniform outputs are bit-exact (the scalings are exact
// powers of two or the reference's own double arithmetic).  The Box-Muller
// angle and the mean + v * stddev affine are evaluated with contraction off so
// they round exactly as the reference host path does; the transcendental
// functions are the accurate (non-native) device versions.
#pragma once

#include "rr_engines.hpp"

#include <rocrand/rocrand_amd_math.h>

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <stdint.h>

#pragma clang fp contract(off)

namespace rr
{

// ---------------------------------------------------------------- Poisson
// Alias table (rocrand_discrete_types.h layout) sampled as rocrand_discrete.h:61-90.
struct alias_table
{
    unsigned int        size;
    unsigned int        offset;
    const unsigned int* alias;
    const double*       prob;
    const double*       cdf;    // cumulative table (quasi-random generators)
    const uint64_t*     packed; // ceil(prob * 2^32) | alias << 33, see discrete_alias_u32
};

__device__ __forceinline__ unsigned int discrete_alias(double x, const alias_table& t)
{
    const double       nx  = __dmul_rn((double)t.size, x);
    const double       fnx = floor(nx);
    const double       y   = __dsub_rn(nx, fnx);
    const unsigned int i   = (unsigned int)fnx;
    return t.offset + (y < t.prob[i] ? i : t.alias[i]);
}

// discrete_alias for the draw x = r 2^-32 of a 32-bit word, in integers: size * x is
// exact in double (size <= 2^12), so floor(nx) = (size r) >> 32 and
// nx - floor(nx) = ((size r) mod 2^32) 2^-32 exactly, and for an integer lo,
// lo 2^-32 < prob  <=>  lo < ceil(prob 2^32).  One table load per sample.
__device__ __forceinline__ unsigned int discrete_alias_u32(uint32_t r, const alias_table& t)
{
    const uint64_t nx = (uint64_t)t.size * r;
    const uint32_t i  = (uint32_t)(nx >> 32);
    const uint64_t e  = t.packed[i];
    const uint64_t lo = (uint32_t)nx;
    return t.offset + (lo < (e & ((1ull << 33) - 1)) ? i : (uint32_t)(e >> 33));
}

// ---------------------------------------------------------------- functors
// Output kinds (one per rocrand_generate_* entry point).
enum class out_kind : int
{
    u8,
    u16,
    u32,
    u64,
    uniform_f32,
    uniform_f64,
    uniform_f16,
    normal_f32,
    normal_f64,
    normal_f16,
    lognormal_f32,
    lognormal_f64,
    lognormal_f16,
    poisson
};

// Runtime parameters of a transform (mean/stddev or Poisson table).
struct dist_params
{
    float       mf, sf;
    double      md, sd;
    __half      mh, sh;
    alias_table pt;
    int         huge;
    double      lambda, sqrt_lambda;
    int         unit; // mf is +0 and sf is 1 (the standard normal)
};

// Family of the engine feeding the transform: generic 32-bit words, generic
// 64-bit words (ThreeFry x64), or MRG (raw values in [1, m1], remapped).
enum class word_family : int
{
    w32,
    w64,
    mrg32,
    mrg31
};

// B selects a variant: for normal_f64 / lognormal_f64 of ThreeFry2x32 the one-word
// inverse-error-function transform; for normal_f32 the standard normal (mean +0, stddev 1),
// whose affine step is the identity but for the sign of a zero (mean + 1 * x is +0 there).
template<word_family F, out_kind K, bool B = false>
struct dist;

// mean + x * stddev in float, rounded twice as the reference does (normal.hpp), or the
// pair as it is for the standard normal (the Box-Muller pair then rounded x + 0).
template<bool UNIT>
__device__ __forceinline__ void rr_affine(float r0, float r1, float* out, const dist_params& p)
{
    if constexpr(UNIT)
    {
        (void)p;
        out[0] = r0;
        out[1] = r1;
    }
    else
    {
        out[0] = __fadd_rn(p.mf, __fmul_rn(r0, p.sf));
        out[1] = __fadd_rn(p.mf, __fmul_rn(r1, p.sf));
    }
}

// The output kind of a transform type.
template<class D>
struct dist_kind;
template<word_family F, out_kind K, bool B>
struct dist_kind<dist<F, K, B>>
{
    static constexpr out_kind value = K;
};

// Generic 32-bit engines (uniform.hpp:37-120, normal.hpp:45-108, log_normal.hpp:45-108).
#define RR_DIST(F, K, IW, OW, T) \
    template<bool B>              \
    struct dist<F, K, B>          \
    {                             \
        static constexpr int in_w = IW, out_w = OW; \
        using out_t               = T;              \
        template<class W>                           \
        __device__ __forceinline__ static void apply(const W* in, T* out, const dist_params& p)

RR_DIST(word_family::w32, out_kind::u32, 1, 1, uint32_t) { (void)p; out[0] = (uint32_t)in[0]; } };
RR_DIST(word_family::w32, out_kind::u8, 1, 4, uint8_t)
{
    (void)p;
    const uint32_t v = (uint32_t)in[0];
    out[0] = (uint8_t)v; out[1] = (uint8_t)(v >> 8); out[2] = (uint8_t)(v >> 16); out[3] = (uint8_t)(v >> 24);
} };
RR_DIST(word_family::w32, out_kind::u16, 1, 2, uint16_t)
{
    (void)p;
    const uint32_t v = (uint32_t)in[0];
    out[0] = (uint16_t)v; out[1] = (uint16_t)(v >> 16);
} };
RR_DIST(word_family::w32, out_kind::uniform_f32, 1, 1, float) { (void)p; out[0] = uniform_f32((uint32_t)in[0]); } };
RR_DIST(word_family::w32, out_kind::uniform_f64, 2, 1, double)
{
    (void)p;
    out[0] = uniform_f64((uint32_t)in[0], (uint32_t)in[1]);
} };
RR_DIST(word_family::w32, out_kind::uniform_f16, 1, 2, __half)
{
    (void)p;
    const uint32_t v = (uint32_t)in[0];
    out[0] = uniform_f16(v); out[1] = uniform_f16(v >> 16);
} };
RR_DIST(word_family::w32, out_kind::normal_f32, 2, 2, float)
{
    float r0, r1;
    box_muller<B>((uint32_t)in[0], (uint32_t)in[1], r0, r1);
    rr_affine<B>(r0, r1, out, p);
} };
RR_DIST(word_family::w32, out_kind::lognormal_f32, 2, 2, float)
{
    float r0, r1;
    box_muller_stepwise((uint32_t)in[0], (uint32_t)in[1], r0, r1);
    out[0] = expf_acc(__fadd_rn(p.mf, __fmul_rn(r0, p.sf)));
    out[1] = expf_acc(__fadd_rn(p.mf, __fmul_rn(r1, p.sf)));
} };
RR_DIST(word_family::w32, out_kind::normal_f16, 1, 2, __half)
{
    float h0, h1;
    const uint32_t v = (uint32_t)in[0];
    box_muller_half(v, v >> 16, h0, h1);
    const float m = __half2float(p.mh), s = __half2float(p.sh);
    out[0] = __float2half(__fadd_rn(m, __fmul_rn(s, h0)));
    out[1] = __float2half(__fadd_rn(m, __fmul_rn(s, h1)));
} };
RR_DIST(word_family::w32, out_kind::lognormal_f16, 1, 2, __half)
{
    float h0, h1;
    const uint32_t v = (uint32_t)in[0];
    box_muller_half(v, v >> 16, h0, h1);
    const float m = __half2float(p.mh), s = __half2float(p.sh);
    out[0] = __float2half(expf_acc(__fadd_rn(m, __fmul_rn(s, h0))));
    out[1] = __float2half(expf_acc(__fadd_rn(m, __fmul_rn(s, h1))));
} };
RR_DIST(word_family::w32, out_kind::poisson, 1, 1, uint32_t)
{
    const uint32_t r = (uint32_t)in[0];
    if(p.huge)
    {
        // poisson.hpp:80-108
        const double nd = normal_double_1(__dadd_rn(R_2POW32_INV_DBL, __dmul_rn((double)r, R_2POW32_INV_DBL)));
        out[0] = (uint32_t)round(__dadd_rn(__dmul_rn(p.sqrt_lambda, nd), p.lambda)) ;
        return;
    }
    out[0] = discrete_alias_u32(r, p.pt);
} };

// normal_f64 / lognormal_f64: four words per pair (default), or one word through
// erfinv for ThreeFry2x32 (threefry.hpp:441-449, normal.hpp:65-79).
template<>
struct dist<word_family::w32, out_kind::normal_f64, false>
{
    static constexpr int in_w = 4, out_w = 2;
    using out_t               = double;
    template<class W>
    __device__ __forceinline__ static void apply(const W* in, double* out, const dist_params& p)
    {
        double r0, r1;
        box_muller_double((uint32_t)in[0], (uint32_t)in[1], (uint32_t)in[2], (uint32_t)in[3], r0, r1);
        out[0] = __dadd_rn(p.md, __dmul_rn(r0, p.sd));
        out[1] = __dadd_rn(p.md, __dmul_rn(r1, p.sd));
    }
};
template<>
struct dist<word_family::w32, out_kind::lognormal_f64, false>
{
    static constexpr int in_w = 4, out_w = 2;
    using out_t               = double;
    template<class W>
    __device__ __forceinline__ static void apply(const W* in, double* out, const dist_params& p)
    {
        double r0, r1;
        box_muller_double((uint32_t)in[0], (uint32_t)in[1], (uint32_t)in[2], (uint32_t)in[3], r0, r1);
        out[0] = exp_cr(__dadd_rn(p.md, __dmul_rn(r0, p.sd)));
        out[1] = exp_cr(__dadd_rn(p.md, __dmul_rn(r1, p.sd)));
    }
};
template<>
struct dist<word_family::w32, out_kind::normal_f64, true>
{
    static constexpr int in_w = 1, out_w = 1;
    using out_t               = double;
    template<class W>
    __device__ __forceinline__ static void apply(const W* in, double* out, const dist_params& p)
    {
        const double v = normal_double_1(__dadd_rn(R_2POW32_INV_DBL, __dmul_rn((double)(uint32_t)in[0], R_2POW32_INV_DBL)));
        out[0] = __dadd_rn(p.md, __dmul_rn(v, p.sd));
    }
};
template<>
struct dist<word_family::w32, out_kind::lognormal_f64, true>
{
    static constexpr int in_w = 1, out_w = 1;
    using out_t               = double;
    template<class W>
    __device__ __forceinline__ static void apply(const W* in, double* out, const dist_params& p)
    {
        const double v = normal_double_1(__dadd_rn(R_2POW32_INV_DBL, __dmul_rn((double)(uint32_t)in[0], R_2POW32_INV_DBL)));
        out[0] = exp_cr(__dadd_rn(p.md, __dmul_rn(v, p.sd)));
    }
};

// 64-bit-word engines (uniform.hpp:122-218, normal.hpp:113-198).
RR_DIST(word_family::w64, out_kind::u64, 1, 1, uint64_t) { (void)p; out[0] = (uint64_t)in[0]; } };
RR_DIST(word_family::w64, out_kind::u32, 1, 2, uint32_t)
{
    (void)p;
    const uint64_t v = (uint64_t)in[0];
    out[0] = (uint32_t)v; out[1] = (uint32_t)(v >> 32);
} };
RR_DIST(word_family::w64, out_kind::u16, 1, 4, uint16_t)
{
    (void)p;
    const uint64_t v = (uint64_t)in[0];
#pragma unroll
    for(int i = 0; i < 4; i++) out[i] = (uint16_t)(v >> (16 * i));
} };
RR_DIST(word_family::w64, out_kind::u8, 1, 8, uint8_t)
{
    (void)p;
    const uint64_t v = (uint64_t)in[0];
#pragma unroll
    for(int i = 0; i < 8; i++) out[i] = (uint8_t)(v >> (8 * i));
} };
RR_DIST(word_family::w64, out_kind::uniform_f32, 1, 1, float) { (void)p; out[0] = uniform_f32((uint64_t)in[0]); } };
RR_DIST(word_family::w64, out_kind::uniform_f64, 1, 1, double) { (void)p; out[0] = uniform_f64((uint64_t)in[0]); } };
RR_DIST(word_family::w64, out_kind::uniform_f16, 1, 4, __half)
{
    (void)p;
    const uint64_t v = (uint64_t)in[0];
#pragma unroll
    for(int i = 0; i < 4; i++) out[i] = uniform_f16((uint32_t)(v >> (16 * i)));
} };
RR_DIST(word_family::w64, out_kind::normal_f32, 1, 2, float)
{
    float r0, r1;
    const uint64_t v = (uint64_t)in[0];
    box_muller<B>((uint32_t)v, (uint32_t)(v >> 32), r0, r1);
    rr_affine<B>(r0, r1, out, p);
} };
RR_DIST(word_family::w64, out_kind::lognormal_f32, 1, 2, float)
{
    float r0, r1;
    const uint64_t v = (uint64_t)in[0];
    box_muller_stepwise((uint32_t)v, (uint32_t)(v >> 32), r0, r1);
    out[0] = expf_acc(__fadd_rn(p.mf, __fmul_rn(r0, p.sf)));
    out[1] = expf_acc(__fadd_rn(p.mf, __fmul_rn(r1, p.sf)));
} };
RR_DIST(word_family::w64, out_kind::normal_f64, 2, 2, double)
{
    double r0, r1;
    const uint64_t a = (uint64_t)in[0], b = (uint64_t)in[1];
    box_muller_double((uint32_t)a, (uint32_t)(a >> 32), (uint32_t)b, (uint32_t)(b >> 32), r0, r1);
    out[0] = __dadd_rn(p.md, __dmul_rn(r0, p.sd));
    out[1] = __dadd_rn(p.md, __dmul_rn(r1, p.sd));
} };
RR_DIST(word_family::w64, out_kind::lognormal_f64, 2, 2, double)
{
    double r0, r1;
    const uint64_t a = (uint64_t)in[0], b = (uint64_t)in[1];
    box_muller_double((uint32_t)a, (uint32_t)(a >> 32), (uint32_t)b, (uint32_t)(b >> 32), r0, r1);
    out[0] = exp_cr(__dadd_rn(p.md, __dmul_rn(r0, p.sd)));
    out[1] = exp_cr(__dadd_rn(p.md, __dmul_rn(r1, p.sd)));
} };
RR_DIST(word_family::w64, out_kind::normal_f16, 1, 2, __half)
{
    float h0, h1;
    const uint64_t v = (uint64_t)in[0];
    box_muller_half((uint32_t)v, (uint32_t)(v >> 32), h0, h1);
    const float m = __half2float(p.mh), s = __half2float(p.sh);
    out[0] = __float2half(__fadd_rn(m, __fmul_rn(s, h0)));
    out[1] = __float2half(__fadd_rn(m, __fmul_rn(s, h1)));
} };
RR_DIST(word_family::w64, out_kind::lognormal_f16, 1, 2, __half)
{
    float h0, h1;
    const uint64_t v = (uint64_t)in[0];
    box_muller_half((uint32_t)v, (uint32_t)(v >> 32), h0, h1);
    const float m = __half2float(p.mh), s = __half2float(p.sh);
    out[0] = __float2half(expf_acc(__fadd_rn(m, __fmul_rn(s, h0))));
    out[1] = __float2half(expf_acc(__fadd_rn(m, __fmul_rn(s, h1))));
} };
RR_DIST(word_family::w64, out_kind::poisson, 1, 1, uint32_t)
{
    const uint64_t r = (uint64_t)in[0];
    if(p.huge)
    {
        const double nd = normal_double_1(uniform_f64(r));
        out[0] = (uint32_t)round(__dadd_rn(__dmul_rn(p.sqrt_lambda, nd), p.lambda));
        return;
    }
    out[0] = discrete_alias(__dmul_rn((double)r, R_2POW64_INV_DBL), p.pt);
} };

// MRG engines (uniform.hpp:226-311, normal.hpp:200-270, poisson.hpp:416-441).
#define RR_MRG_DISTS(F, M)                                                                       \
    RR_DIST(F, out_kind::u32, 1, 1, uint32_t) { (void)p; out[0] = mrg_uint<M>((uint32_t)in[0]); } }; \
    RR_DIST(F, out_kind::u8, 1, 4, uint8_t)                                                      \
    {                                                                                            \
        (void)p;                                                                                 \
        const uint32_t v = mrg_uint<M>((uint32_t)in[0]);                                         \
        out[0] = (uint8_t)v; out[1] = (uint8_t)(v >> 8); out[2] = (uint8_t)(v >> 16); out[3] = (uint8_t)(v >> 24); \
    } };                                                                                         \
    RR_DIST(F, out_kind::u16, 1, 2, uint16_t)                                                    \
    {                                                                                            \
        (void)p;                                                                                 \
        const uint32_t v = mrg_uint<M>((uint32_t)in[0]);                                         \
        out[0] = (uint16_t)v; out[1] = (uint16_t)(v >> 16);                                      \
    } };                                                                                         \
    RR_DIST(F, out_kind::uniform_f32, 1, 1, float) { (void)p; out[0] = mrg_f32<M>((uint32_t)in[0]); } }; \
    RR_DIST(F, out_kind::uniform_f64, 1, 1, double) { (void)p; out[0] = mrg_f64<M>((uint32_t)in[0]); } }; \
    RR_DIST(F, out_kind::uniform_f16, 1, 2, __half)                                              \
    {                                                                                            \
        (void)p;                                                                                 \
        const uint32_t v = mrg_uint<M>((uint32_t)in[0]);                                         \
        out[0] = uniform_f16(v); out[1] = uniform_f16(v >> 16);                                  \
    } };                                                                                         \
    RR_DIST(F, out_kind::normal_f32, 2, 2, float)                                                \
    {                                                                                            \
        float r0, r1;                                                                            \
        mrg_box_muller<M, B>((uint32_t)in[0], (uint32_t)in[1], r0, r1);                          \
        rr_affine<B>(r0, r1, out, p);                                                            \
    } };                                                                                         \
    RR_DIST(F, out_kind::lognormal_f32, 2, 2, float)                                             \
    {                                                                                            \
        float r0, r1;                                                                            \
        mrg_box_muller_stepwise<M>((uint32_t)in[0], (uint32_t)in[1], r0, r1);                             \
        out[0] = expf_acc(__fadd_rn(p.mf, __fmul_rn(r0, p.sf)));                                     \
        out[1] = expf_acc(__fadd_rn(p.mf, __fmul_rn(r1, p.sf)));                                     \
    } };                                                                                         \
    RR_DIST(F, out_kind::normal_f64, 2, 2, double)                                               \
    {                                                                                            \
        double r0, r1;                                                                           \
        mrg_box_muller_double<M>((uint32_t)in[0], (uint32_t)in[1], r0, r1);                      \
        out[0] = __dadd_rn(p.md, __dmul_rn(r0, p.sd));                                           \
        out[1] = __dadd_rn(p.md, __dmul_rn(r1, p.sd));                                           \
    } };                                                                                         \
    RR_DIST(F, out_kind::lognormal_f64, 2, 2, double)                                            \
    {                                                                                            \
        double r0, r1;                                                                           \
        mrg_box_muller_double<M>((uint32_t)in[0], (uint32_t)in[1], r0, r1);                      \
        out[0] = exp_cr(__dadd_rn(p.md, __dmul_rn(r0, p.sd)));                                      \
        out[1] = exp_cr(__dadd_rn(p.md, __dmul_rn(r1, p.sd)));                                      \
    } };                                                                                         \
    RR_DIST(F, out_kind::normal_f16, 1, 2, __half)                                               \
    {                                                                                            \
        float h0, h1;                                                                            \
        const uint32_t v = mrg_uint<M>((uint32_t)in[0]);                                         \
        box_muller_half(v, v >> 16, h0, h1);                                                     \
        const float m = __half2float(p.mh), s = __half2float(p.sh);                              \
        out[0] = __float2half(__fadd_rn(m, __fmul_rn(s, h0)));                                   \
        out[1] = __float2half(__fadd_rn(m, __fmul_rn(s, h1)));                                   \
    } };                                                                                         \
    RR_DIST(F, out_kind::lognormal_f16, 1, 2, __half)                                            \
    {                                                                                            \
        float h0, h1;                                                                            \
        const uint32_t v = mrg_uint<M>((uint32_t)in[0]);                                         \
        box_muller_half(v, v >> 16, h0, h1);                                                     \
        const float m = __half2float(p.mh), s = __half2float(p.sh);                              \
        out[0] = __float2half(expf_acc(__fadd_rn(m, __fmul_rn(s, h0))));                             \
        out[1] = __float2half(expf_acc(__fadd_rn(m, __fmul_rn(s, h1))));                             \
    } };                                                                                         \
    RR_DIST(F, out_kind::poisson, 1, 1, uint32_t)                                                \
    {                                                                                            \
        const uint32_t r = mrg_uint<M>((uint32_t)in[0]);                                         \
        if(p.huge)                                                                               \
        {                                                                                        \
            const double nd = normal_double_1(__dadd_rn(R_2POW32_INV_DBL, __dmul_rn((double)r, R_2POW32_INV_DBL))); \
            out[0] = (uint32_t)round(__dadd_rn(__dmul_rn(p.sqrt_lambda, nd), p.lambda));          \
            return;                                                                              \
        }                                                                                        \
        out[0] = discrete_alias_u32(r, p.pt);                                                    \
    } };

RR_MRG_DISTS(word_family::mrg32, 32)
RR_MRG_DISTS(word_family::mrg31, 31)

#undef RR_MRG_DISTS
#undef RR_DIST

} // namespace rr
