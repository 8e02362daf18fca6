// rocrand_amd_math.h -- scalar output transforms of the reference's distributions
// (rocrand_uniform.h, rocrand_normal.h, rocrand_log_normal.h) on gfx950, shared by the
// library's kernels and the device API (rocrand_kernel.h).
//
// Uniforms are bit-exact (exact power-of-two scalings or the reference's own double
// arithmetic).  The Box-Muller angle and the affine steps are evaluated operation by
// operation (no contraction) as the reference host path does; the transcendental
// functions are evaluated in double with short fixed polynomials and rounded once.
#pragma once

#include "rocrand_amd_engines.h"
#include "rocrand_amd_mathtab.h"

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <stdint.h>


namespace rocrand_amd_engine
{

// rocrand_common.h:24-37
constexpr float  R_2POW16_INV     = 1.5258789e-05f;
constexpr float  R_2POW16_INV_2PI = 9.58738e-05f;
constexpr float  R_2POW32_INV     = 2.3283064e-10f;
constexpr double R_2POW32_INV_DBL = 2.3283064365386963e-10;
constexpr double R_2POW64_INV_DBL = 5.4210108624275221700372640043497e-20;
constexpr float  R_2POW32_INV_2PI = 1.46291807e-09f;
constexpr double R_2POW53_INV_DBL = 1.1102230246251565e-16;
constexpr double R_PI_DBL         = 3.1415926535897932;
constexpr float  R_2PI            = 6.2831855f;
constexpr float  R_SQRT2          = 1.4142135f;

// ---------------------------------------------------------------- scalar math
// Float transcendentals evaluated in double and rounded once, written for the
// arguments the transforms produce, with short fixed polynomials instead of the
// general-purpose library routines:
//   logf_tab (Box-Muller radius, u a positive normal float): u = 2^k z with
//        z in [0.699, 1.398), ln u = k ln 2 - ln(1/c_i) + log1p(z/c_i - 1) over the 64-entry
//        table of rocrand_amd_mathtab.h, |z/c_i - 1| < 2^-7, series to r^6 (relative error < 2^-44);
//   log (erfinv argument, any float): x = 2^e m, m in [sqrt(1/2), sqrt(2)),
//        ln m = 2 atanh(f), f = (m-1)/(m+1), |f| < 0.172, series to f^21;
//   sin/cos for x in [0, 8): reduction by pi/32 (two parts, fused), sin/cos(k pi/32) from
//        rocrand_amd_mathtab.h and Taylor series to r^7 / r^6 on |r| <= pi/64;
//   expf_acc: reduction by ln2/32 (two parts), 2^(j/32) from rocrand_amd_mathtab.h, Taylor series
//        of exp(r) - 1 to r^5 on |r| <= ln2/64, scaling by 2^e in double.
// A host C twin of these routines (exact division in place of the refined reciprocal)
// rounds to the correctly rounded float result for every float in [2^-32, 1] but one
// (logf_tab; that one lies within 2^-48 of a rounding boundary), every float in (0, 7]
// (sin, cos) and every float in (-105, 90) but two (exp).
// So each differs from the host libm's float result by at most the host's own error
// (< 1 ulp).
constexpr double RR_LN2_HI  = 6.93147180369123816490e-01; // ln 2, top 32 bits
constexpr double RR_LN2_LO  = 1.90821492927058770002e-10; // ln 2 - RR_LN2_HI
constexpr double RR_LN2     = 6.93147180559945286227e-01; // ln 2 rounded to double

// ln u for a positive normal float u (the Box-Muller uniforms lie in [2^-32, 1]).
__device__ __forceinline__ float logf_tab(float u)
{
    ROCRAND_AMD_FP_EXACT
    const uint32_t ix  = __float_as_uint(u);
    const uint32_t tmp = ix - RR_LOGTAB_OFF;
    const uint32_t i   = (tmp >> (23 - RR_LOGTAB_BITS)) & ((1u << RR_LOGTAB_BITS) - 1u);
    const int      k   = (int)tmp >> 23;
    const float    z   = __uint_as_float(ix - (tmp & 0xff800000u));
    const double2  t   = rr_logtab[i];
    const double   r   = __builtin_fma((double)z, t.x, -1.0);
    const double   r2  = r * r;
    double         p   = __builtin_fma(r, -1.0 / 6, 1.0 / 5);
    p                  = __builtin_fma(p, r, -0.25);
    p                  = __builtin_fma(p, r, 1.0 / 3);
    p                  = __builtin_fma(p, r, -0.5);
    const double hi    = __builtin_fma((double)k, RR_LN2, t.y);
    return (float)(hi + __builtin_fma(p, r2, r));
}

__device__ __forceinline__ double log_dbl(double x)
{
    ROCRAND_AMD_FP_EXACT
    int    e = __builtin_amdgcn_frexp_exp(x);
    double m = __builtin_amdgcn_frexp_mant(x); // [0.5, 1)
    if(m < 0.70710678118654752440)
    {
        m = m * 2.0;
        e = e - 1;
    }
    const double d  = m + 1.0;
    double       r  = __builtin_amdgcn_rcp(d);
    r               = __builtin_fma(r, __builtin_fma(-d, r, 1.0), r);
    const double f  = (m - 1.0) * r;
    const double f2 = f * f;
    double       p  = 1.0 / 21;
    p               = __builtin_fma(p, f2, 1.0 / 19);
    p               = __builtin_fma(p, f2, 1.0 / 17);
    p               = __builtin_fma(p, f2, 1.0 / 15);
    p               = __builtin_fma(p, f2, 1.0 / 13);
    p               = __builtin_fma(p, f2, 1.0 / 11);
    p               = __builtin_fma(p, f2, 1.0 / 9);
    p               = __builtin_fma(p, f2, 1.0 / 7);
    p               = __builtin_fma(p, f2, 1.0 / 5);
    p               = __builtin_fma(p, f2, 1.0 / 3);
    const double tf = 2.0 * f;
    const double lm = __builtin_fma(tf * f2, p, tf);
    const double y  = __builtin_fma((double)e, RR_LN2_HI, __builtin_fma((double)e, RR_LN2_LO, lm));
    if(x == 0.0)
        return -__builtin_inf();
    if(!(x > 0.0) || x == __builtin_inf())
        return x < 0.0 ? __builtin_nan("") : x; // NaN, negative -> NaN; +inf -> +inf
    return y;
}

// sin and cos of x in [0, 8), rounded to float: x = q pi/32 + r (two-part reduction,
// |r| <= pi/64), then the angle-addition formulas with sin/cos(k pi/32), k = q mod 64,
// from rocrand_amd_mathtab.h and Taylor series of sin r (to r^7) and cos r (to r^6).  Where
// sin x or cos x is near 0 the table entry is an exact 0 / +-1, so the sum does not cancel.
__device__ __forceinline__ void sincosf_acc(float xf, float* s, float* c)
{
    ROCRAND_AMD_FP_EXACT
    const double  x  = (double)xf;
    const double  q  = __builtin_rint(x * RR_SC_INV_STEP);
    const int     k  = (int)q & ((1 << RR_SCTAB_BITS) - 1);
    double        r  = __builtin_fma(-q, RR_SC_STEP_HI, x);
    r                = __builtin_fma(-q, RR_SC_STEP_LO, r);
    const double2 t  = rr_sctab[k];
    const double  r2 = r * r;
    double        ps = __builtin_fma(r2, -1.0 / 5040, 1.0 / 120);
    ps               = __builtin_fma(ps, r2, -1.0 / 6);
    const double  sr = __builtin_fma(r * r2, ps, r);
    double        pc = __builtin_fma(r2, -1.0 / 720, 1.0 / 24);
    pc               = __builtin_fma(pc, r2, -0.5);
    const double  cr = __builtin_fma(r2, pc, 1.0);
    *s               = (float)__builtin_fma(t.x, cr, t.y * sr);
    *c               = (float)__builtin_fma(t.y, cr, -(t.x * sr));
}

// sqrtf for 0 and positive normal floats: the hardware square root (within 1 ulp) and
// the two fused-multiply-add checks of its neighbours that make it correctly rounded --
// the same instructions as the compiler's sqrtf without its rescaling of tiny inputs
// (below 2^-96) and of infinities, which the Box-Muller radius (0 or >= 2^-24) never is.
__device__ __forceinline__ float sqrtf_radius(float x)
{
    ROCRAND_AMD_FP_EXACT
    const float s  = __builtin_amdgcn_sqrtf(x);
    const float sm = __uint_as_float(__float_as_uint(s) - 1u);
    const float sp = __uint_as_float(__float_as_uint(s) + 1u);
    float       r  = __builtin_fmaf(-sm, s, x) <= 0.0f ? sm : s;
    r              = __builtin_fmaf(-sp, s, x) > 0.0f ? sp : r;
    return r;
}

__device__ __forceinline__ float logf_acc(float x) { return (float)log_dbl((double)x); }

// exp of a float, rounded to float: x = (32 e + j) ln2/32 + r, exp x = 2^e 2^(j/32) exp(r)
// with the 2^(j/32) table of rocrand_amd_mathtab.h.  Arguments are clamped to [-110, 100]
// first (the float result is then 0 or +inf already), which keeps |32 e + j| < 2^13.
__device__ __forceinline__ float expf_acc(float xf)
{
    ROCRAND_AMD_FP_EXACT
    const double x  = (double)xf;
    const double xc = x < -110.0 ? -110.0 : (x > 100.0 ? 100.0 : x);
    const double kd = __builtin_rint(xc * RR_EXP_INV_STEP);
    const int    k  = (int)kd;
    double       r  = __builtin_fma(-kd, RR_EXP_STEP_HI, xc);
    r               = __builtin_fma(-kd, RR_EXP_STEP_LO, r);
    double p        = __builtin_fma(r, 1.0 / 120, 1.0 / 24);
    p               = __builtin_fma(p, r, 1.0 / 6);
    p               = __builtin_fma(p, r, 0.5);
    p               = __builtin_fma(p, r, 1.0);
    p               = p * r;
    const double t  = rr_exptab[k & ((1 << RR_EXPTAB_BITS) - 1)];
    const float  y  = (float)__builtin_amdgcn_ldexp(__builtin_fma(t, p, t), k >> RR_EXPTAB_BITS);
    return xf != xf ? xf : y;
}

// rocrand_uniform.h:63-66: 2^-32 + v * 2^-32.  The product by a power of two is
// exact, so one fused multiply-add rounds exactly like the separate operations.
__device__ __forceinline__ float uniform_f32(uint32_t v)
{
    return __builtin_fmaf((float)v, R_2POW32_INV, R_2POW32_INV);
}
// rocrand_uniform.h:70-73
__device__ __forceinline__ float uniform_f32(uint64_t v)
{
    return __builtin_fmaf((float)(v >> 32), R_2POW32_INV, R_2POW32_INV);
}
// rocrand_uniform.h:100-107 (53-bit integer: exact in double, exact product)
__device__ __forceinline__ double uniform_f64(uint32_t v1, uint32_t v2)
{
    const uint64_t v = (uint64_t)v1 | ((uint64_t)(v2 >> 11) << 32);
    return __builtin_fma((double)v, R_2POW53_INV_DBL, R_2POW53_INV_DBL);
}
// rocrand_uniform.h:109-116
__device__ __forceinline__ double uniform_f64(uint64_t v)
{
    return __builtin_fma((double)(v >> 11), R_2POW53_INV_DBL, R_2POW53_INV_DBL);
}
// rocrand_uniform.h:154-157
__device__ __forceinline__ __half uniform_f16(uint32_t v16)
{
    return __float2half(__builtin_fmaf((float)(v16 & 0xffffu), R_2POW16_INV, R_2POW16_INV));
}

template<int MRG>
struct mrg_consts;
template<>
struct mrg_consts<32>
{
    static constexpr double norm      = MRG32_NORM_DOUBLE;
    static constexpr double uint_norm = MRG32_UINT_NORM;
};
template<>
struct mrg_consts<31>
{
    static constexpr double norm      = MRG31_NORM_DOUBLE;
    static constexpr double uint_norm = MRG31_UINT_NORM;
};

// (double)v * k, correctly rounded, in one fused multiply-add and no integer-to-double
// conversion: the double with high word 0x43300000 and low word v is 2^52 + v exactly,
// and fma(2^52 + v, k, -2^52 k) rounds (2^52 + v) k - 2^52 k = v k once (2^52 k is exact,
// a power-of-two scaling of k).
__device__ __forceinline__ double scale_u32(uint32_t v, double k)
{
    const double x = __hiloint2double(0x43300000, (int)v);
    return __builtin_fma(x, k, -0x1p52 * k);
}

// rocrand_uniform.h:161-219
template<int MRG>
__device__ __forceinline__ uint32_t mrg_uint(uint32_t v)
{
    return (uint32_t)scale_u32(v - 1, mrg_consts<MRG>::uint_norm);
}
template<int MRG>
__device__ __forceinline__ float mrg_f32(uint32_t v)
{
    return (float)scale_u32(v, mrg_consts<MRG>::norm);
}
template<int MRG>
__device__ __forceinline__ double mrg_f64(uint32_t v)
{
    return scale_u32(v, mrg_consts<MRG>::norm);
}

// The Box-Muller pair sin(v) sqrt(-2 ln u), cos(v) sqrt(-2 ln u) for a float u in
// [2^-32, 1] and a float angle v in (0, 8), evaluated as one expression in double and
// rounded to float once.  Relative error before that rounding is below 2^-36 (log1p series
// to r^4 on |r| < 2^-7: 2^-30 relative on ln u, halved by the square root; sin series to
// r^5 and cos to r^4 on |r| <= pi/64; a float reciprocal square root refined by one Newton
// step), so each output is within 0.51 ulp of the exact value.  The reference's float
// pipeline (logf, sqrtf, sinf/cosf, product: four roundings, rocrand_normal.h:51-66) lies
// within 2.4 ulp of the exact value, so the two stay within 2 ulp of each other.
// u = 1 gives ln u = +0 (table entry {1, 0}), x = -0 and the radius -0, as the reference's
// sqrtf(-2.0f * logf(1.0f)); the clamp keeps the reciprocal square root finite there.
__device__ __forceinline__ void box_muller_uv(float u, float v, float& r0, float& r1)
{
    ROCRAND_AMD_FP_EXACT
    const uint32_t ix  = __float_as_uint(u);
    const uint32_t tmp = ix - RR_LOGTAB_OFF;
    const uint32_t i   = (tmp >> (23 - RR_LOGTAB_BITS)) & ((1u << RR_LOGTAB_BITS) - 1u);
    const int      k   = (int)tmp >> 23;
    const float    z   = __uint_as_float(ix - (tmp & 0xff800000u));
    const double2  lt  = rr_logtab[i];
    const double   r   = __builtin_fma((double)z, lt.x, -1.0);
    const double   r2  = r * r;
    double         q   = __builtin_fma(r, -0.25, 1.0 / 3);
    q                  = __builtin_fma(q, r, -0.5);
    const double   lnu = __builtin_fma((double)k, RR_LN2, lt.y) + __builtin_fma(q, r2, r);
    const double   x   = -2.0 * lnu;
    const double   y   = (double)__builtin_amdgcn_rsqf(__builtin_fmaxf((float)x, 0x1p-126f));
    const double   t   = x * y;
    const double   h   = __builtin_fma(-t, y, 1.0);
    const double   s   = __builtin_fma(t * h, 0.5, t);

    const double  vd = (double)v;
    const double  qv = __builtin_rint(vd * RR_SC_INV_STEP);
    const int     kv = (int)qv & ((1 << RR_SCTAB_BITS) - 1);
    double        a  = __builtin_fma(-qv, RR_SC_STEP_HI, vd);
    a                = __builtin_fma(-qv, RR_SC_STEP_LO, a);
    const double2 st = rr_sctab[kv];
    const double  a2 = a * a;
    const double  sa = __builtin_fma(a * a2, __builtin_fma(a2, 1.0 / 120, -1.0 / 6), a);
    const double  ca = __builtin_fma(a2, __builtin_fma(a2, 1.0 / 24, -0.5), 1.0);
    const double  sn = __builtin_fma(st.x, ca, st.y * sa);
    const double  cs = __builtin_fma(st.y, ca, -(st.x * sa));
    r0               = (float)(sn * s);
    r1               = (float)(cs * s);
}

// rocrand_normal.h:51-66: u = 2^-32 + x 2^-32, angle 2^-32 2pi + y 2^-32 2pi (rounded as
// the reference rounds them), then the pair above.
__device__ __forceinline__ void box_muller(uint32_t x, uint32_t y, float& r0, float& r1)
{
    ROCRAND_AMD_FP_EXACT
    const float u = uniform_f32(x);
    const float v = R_2POW32_INV_2PI + (float)y * R_2POW32_INV_2PI;
    box_muller_uv(u, v, r0, r1);
}

// The same pair through the reference's own steps -- logf, sqrtf, sinf/cosf, each rounded
// to float (correctly) -- for the log-normal transforms: exp magnifies a one-ulp change
// of the variate to several ulps of the output, and these steps reproduce the reference
// host path's variate bit for bit wherever its libm rounds correctly.
__device__ __forceinline__ void box_muller_uv_stepwise(float u, float v, float& r0, float& r1)
{
    ROCRAND_AMD_FP_EXACT
    const float s = sqrtf_radius(-2.0f * logf_tab(u));
    float       sn, cs;
    sincosf_acc(v, &sn, &cs);
    r0 = sn * s;
    r1 = cs * s;
}
__device__ __forceinline__ void box_muller_stepwise(uint32_t x, uint32_t y, float& r0, float& r1)
{
    ROCRAND_AMD_FP_EXACT
    const float u = uniform_f32(x);
    const float v = R_2POW32_INV_2PI + (float)y * R_2POW32_INV_2PI;
    box_muller_uv_stepwise(u, v, r0, r1);
}
// rocrand_normal.h:76-96 (host branch: sin(w * pi), not sincospi).
__device__ __forceinline__ void box_muller_double(uint32_t a, uint32_t b, uint32_t c, uint32_t d, double& r0,
                                                  double& r1)
{
    ROCRAND_AMD_FP_EXACT
    const uint64_t v1 = (uint64_t)a ^ ((uint64_t)b << 21);
    const double   u  = R_2POW53_INV_DBL + (double)v1 * R_2POW53_INV_DBL;
    const uint64_t v2 = (uint64_t)c ^ ((uint64_t)d << 21);
    const double   w  = R_2POW53_INV_DBL * 2.0 + (double)v2 * (R_2POW53_INV_DBL * 2.0);
    const double   s  = sqrt(-2.0 * log(u));
    const double   t  = w * R_PI_DBL;
    double         sn, cs;
    sincos(t, &sn, &cs);
    r0 = sn * s;
    r1 = cs * s;
}
// rocrand_normal.h:108-135, host (float math) branch.
__device__ __forceinline__ void box_muller_half(uint32_t x16, uint32_t y16, float& h0, float& h1)
{
    ROCRAND_AMD_FP_EXACT
    const float u = R_2POW16_INV + (float)(x16 & 0xffffu) * R_2POW16_INV;
    const float v = R_2POW16_INV_2PI + (float)(y16 & 0xffffu) * R_2POW16_INV_2PI;
    const float s = sqrtf_radius(-2.0f * logf_tab(u));
    float       sn, cs;
    sincosf_acc(v, &sn, &cs);
    h0 = __half2float(__float2half(sn * s));
    h1 = __half2float(__float2half(cs * s));
}
// rocrand_normal.h:138-154
template<int MRG>
__device__ __forceinline__ void mrg_box_muller(uint32_t x, uint32_t y, float& r0, float& r1)
{
    ROCRAND_AMD_FP_EXACT
    box_muller_uv(mrg_f32<MRG>(x), mrg_f32<MRG>(y) * R_2PI, r0, r1);
}
template<int MRG>
__device__ __forceinline__ void mrg_box_muller_stepwise(uint32_t x, uint32_t y, float& r0, float& r1)
{
    ROCRAND_AMD_FP_EXACT
    box_muller_uv_stepwise(mrg_f32<MRG>(x), mrg_f32<MRG>(y) * R_2PI, r0, r1);
}
// rocrand_normal.h:156-172: float-precision uniforms widened to double.
template<int MRG>
__device__ __forceinline__ void mrg_box_muller_double(uint32_t x, uint32_t y, double& r0, double& r1)
{
    ROCRAND_AMD_FP_EXACT
    const double u = (double)mrg_f32<MRG>(x);
    const double v = (double)mrg_f32<MRG>(y) * 2.0;
    const double s = sqrt(-2.0 * log(u));
    double       sn, cs;
    sincos(v * R_PI_DBL, &sn, &cs);
    r0 = sn * s;
    r1 = cs * s;
}
// rocrand_normal.h:199-221 roc_d_erfinv and :263-268 normal_distribution_double.
__device__ __forceinline__ double roc_d_erfinv(double x)
{
    ROCRAND_AMD_FP_EXACT
    const double sgn = (x < 0.0) ? -1.0 : 1.0;
    x                = (1.0 - x) * (1.0 + x);
    const double lnx = log(x);
    if(isnan(lnx))
        return 1.0;
    if(isinf(lnx))
        return 0.0;
    const double tt1 = 2.0 / (R_PI_DBL * 0.147) + 0.5 * lnx;
    const double tt2 = 1.0 / 0.147 * lnx;
    return sgn * sqrt(-tt1 + sqrt(tt1 * tt1 - tt2));
}
__device__ __forceinline__ double normal_double_1(double p)
{
    ROCRAND_AMD_FP_EXACT
    return (double)R_SQRT2 * roc_d_erfinv(2.0 * p - 1.0);
}

// rocrand_normal.h:174-197 roc_f_erfinv, float arithmetic rounded step by step.
__device__ __forceinline__ float roc_f_erfinv(float x)
{
    ROCRAND_AMD_FP_EXACT
    const float sgn = (x < 0.0f) ? -1.0f : 1.0f;
    x               = (1.0f - x) * (1.0f + x);
    // (1 - x)(1 + x) is a positive normal float or 0 here (|x| <= 1)
    const float lnx = x > 0.0f ? logf_tab(x) : logf_acc(x);
    if(isnan(lnx))
        return 1.0f;
    if(isinf(lnx))
        return 0.0f;
    constexpr float c1  = 2.0f / (3.141592653f * 0.147f);
    constexpr float c2  = 1.0f / 0.147f;
    const float     tt1 = c1 + 0.5f * lnx;
    const float     tt2 = c2 * lnx;
    return sgn * sqrtf(-tt1 + sqrtf(tt1 * tt1 - tt2));
}

// Uniforms of a direction value (rocrand_uniform.h:63-116 on the 32- or 64-bit value).
__device__ __forceinline__ float sob_uf32(uint32_t x) { return uniform_f32(x); }
__device__ __forceinline__ float sob_uf32(uint64_t x) { return uniform_f32(x); }
__device__ __forceinline__ double sob_uf64(uint32_t x)
{
    return __builtin_fma((double)x, R_2POW32_INV_DBL, R_2POW32_INV_DBL);
}
__device__ __forceinline__ double sob_uf64(uint64_t x) { return uniform_f64(x); }

// rocrand_normal.h:228-296 normal_distribution(_double) on a direction value.
template<class V>
__device__ __forceinline__ float sob_normal_f32(V x)
{
    ROCRAND_AMD_FP_EXACT
    const float p = sob_uf32(x);
    return R_SQRT2 * roc_f_erfinv(2.0f * p - 1.0f);
}
template<class V>
__device__ __forceinline__ double sob_normal_f64(V x)
{
    return normal_double_1(sob_uf64(x));
}

} // namespace rocrand_amd_engine

