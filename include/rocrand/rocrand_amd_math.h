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

// ---------------------------------------------------------------- double functions
// The double transforms' log, exp, sin and cos, evaluated in double-double arithmetic
// (Dekker's exact sum and product) and rounded once: each result is within
// 2^-66 (relative) of the exact value before that rounding, so it is the correctly rounded
// double except for arguments within that distance of a rounding boundary.  The
// reference's host path takes these steps through the host libm, which rounds them
// correctly but for rare hard cases; evaluated this way the device reproduces its
// intermediate values.  Host and device compile the same code (the host twin of the
// tests checks them against quad precision).
__host__ __device__ __forceinline__ void rr_two_sum(double a, double b, double& s, double& e)
{
    ROCRAND_AMD_FP_EXACT
    s               = a + b;
    const double bb = s - a;
    e               = (a - (s - bb)) + (b - bb);
}
// |a| >= |b| (or a = 0)
__host__ __device__ __forceinline__ void rr_fast_two_sum(double a, double b, double& s, double& e)
{
    ROCRAND_AMD_FP_EXACT
    s = a + b;
    e = b - (s - a);
}
__host__ __device__ __forceinline__ void rr_two_prod(double a, double b, double& p, double& e)
{
    ROCRAND_AMD_FP_EXACT
    p = a * b;
    e = __builtin_fma(a, b, -p);
}
__host__ __device__ __forceinline__ double rr_bits_to_double(uint64_t b)
{
    double x;
    __builtin_memcpy(&x, &b, sizeof(x));
    return x;
}
__host__ __device__ __forceinline__ uint64_t rr_double_to_bits(double x)
{
    uint64_t b;
    __builtin_memcpy(&b, &x, sizeof(b));
    return b;
}

// ln x: x = 2^k m, m in [sqrt(1/2), sqrt(2)), c the centre of m's table interval
// (rocrand_amd_mathtab.h), r = m/c - 1 exact as a double-double (|r| < 2^-8), and
// ln x = k ln 2 - ln(1/c) + log1p(r) with log1p to r^10.
__host__ __device__ __forceinline__ double log_cr(double x)
{
    ROCRAND_AMD_FP_EXACT
    if(!(x > 0.0) || x == __builtin_inf())
        return x == 0.0 ? -__builtin_inf() : (x > 0.0 ? x : __builtin_nan(""));
    int kadj = 0;
    if(x < 0x1p-1022)
    {
        x *= 0x1p54; // subnormal: scale into the normal range
        kadj = -54;
    }
    const uint64_t ix  = rr_double_to_bits(x);
    const uint32_t tmp = (uint32_t)(ix >> 32) - RR_DLOG_OFF;
    const int      ke  = (int)tmp >> 20;
    const double   k   = (double)(ke + kadj);
    const double   m   = rr_bits_to_double(ix - ((uint64_t)(int64_t)ke << 52));
    const double*  e   = rr_dlog_tab[(tmp >> 13) & 127u];
    double         p, pe, rh, rl;
    rr_two_prod(m, e[0], p, pe);
    rr_two_sum(p - 1.0, pe, rh, rl); // p - 1 is exact: p is within 2^-7 of 1
    const double r2  = rh * rh;
    const double r2e = __builtin_fma(rh, rh, -r2);
    double       q   = __builtin_fma(rh, -1.0 / 10, 1.0 / 9);
    q                = __builtin_fma(q, rh, -1.0 / 8);
    q                = __builtin_fma(q, rh, 1.0 / 7);
    q                = __builtin_fma(q, rh, -1.0 / 6);
    q                = __builtin_fma(q, rh, 1.0 / 5);
    q                = __builtin_fma(q, rh, -1.0 / 4);
    q                = __builtin_fma(q, rh, 1.0 / 3);
    // log1p(r) = r - r^2/2 + r^3 q(r); the tail's share of r^2 / 2 and r^3 / 3 folded below
    double a, ae, b, be, c, ce;
    rr_two_sum(k * RR_DLN2_HI, e[1], a, ae);
    rr_two_sum(a, rh, b, be);
    rr_two_sum(b, -0.5 * r2, c, ce);
    const double tail = (rl - 0.5 * r2e - rh * rl) + r2 * (rh * q + rl);
    const double lo   = (ae + be + ce) + (k * RR_DLN2_MI + e[2]) + tail;
    return c + lo;
}

// e^x: x = (64 k + j) ln2/64 + r (|r| <= ln2/128, the product with the head of ln2/64
// exact), e^x = 2^k 2^(j/64) (1 + expm1(r)) with 2^(j/64) a double-double and expm1 to r^8.
__host__ __device__ __forceinline__ double exp_cr(double x)
{
    ROCRAND_AMD_FP_EXACT
    if(x != x)
        return x;
    if(x > 709.782712893384)
        return __builtin_inf();
    if(x < -745.1332191019412)
        return 0.0;
    const double qd = __builtin_rint(x * RR_DEXP_INV_STEP);
    const int    q  = (int)qd;
    double       rh, rl;
    rr_two_sum(__builtin_fma(-qd, RR_DEXP_STEP_HI, x), -qd * RR_DEXP_STEP_MI, rh, rl);
    rl               = __builtin_fma(-qd, RR_DEXP_STEP_LO, rl);
    const double r2  = rh * rh;
    const double r2e = __builtin_fma(rh, rh, -r2);
    double       pp  = __builtin_fma(rh, 1.0 / 40320, 1.0 / 5040);
    pp               = __builtin_fma(pp, rh, 1.0 / 720);
    pp               = __builtin_fma(pp, rh, 1.0 / 120);
    pp               = __builtin_fma(pp, rh, 1.0 / 24);
    pp               = __builtin_fma(pp, rh, 1.0 / 6);
    // expm1(r) = eh + el
    const double  eh = rh;
    const double  el = __builtin_fma(0.5, r2, (rl + (0.5 * r2e + rh * rl)) + r2 * rh * pp);
    const double2 t  = rr_dexp_tab[q & 63];
    double        s1, s1e, s, se;
    rr_two_prod(t.x, eh, s1, s1e);
    rr_fast_two_sum(t.x, s1, s, se);
    const double lo = (se + s1e) + (t.x * el + t.y) + t.y * (eh + el);
    return ldexp(s + lo, q >> 6);
}

// sin t and cos t for t in [0, 2 pi + 2^-48]: t = q pi/64 + r (|r| <= pi/128, the products
// with the two heads of pi/64 exact), then the angle-addition formulas with sin(q pi/64),
// cos(q pi/64) as double-doubles and sin r to r^9, cos r to r^10.  Where sin t (cos t) is
// near 0 the table entry is an exact 0 and the result keeps its relative accuracy.
__host__ __device__ __forceinline__ void sincos_cr(double t, double& sn, double& cs)
{
    ROCRAND_AMD_FP_EXACT
    const double qd = __builtin_rint(t * RR_DSC_INV_STEP);
    const int    q  = (int)qd;
    double       rh, rl;
    rr_two_sum(__builtin_fma(-qd, RR_DSC_STEP_HI, t), -qd * RR_DSC_STEP_MI, rh, rl);
    rl               = __builtin_fma(-qd, RR_DSC_STEP_LO, rl);
    const double r2  = rh * rh;
    const double r2e = __builtin_fma(rh, rh, -r2);
    double       ps  = __builtin_fma(r2, 1.0 / 362880, -1.0 / 5040);
    ps               = __builtin_fma(ps, r2, 1.0 / 120);
    ps               = __builtin_fma(ps, r2, -1.0 / 6);
    double pc        = __builtin_fma(r2, -1.0 / 3628800, 1.0 / 40320);
    pc               = __builtin_fma(pc, r2, -1.0 / 720);
    pc               = __builtin_fma(pc, r2, 1.0 / 24);
    // sin r = rh + sl, cos r = 1 + ch + cl
    const double  sl = __builtin_fma(r2 * rh, ps, __builtin_fma(-0.5 * r2, rl, rl));
    const double  ch = -0.5 * r2;
    const double  cl = __builtin_fma(r2 * r2, pc, -(0.5 * r2e + rh * rl));
    const double* T  = rr_dsc_tab[q];
    const double  S = T[0], Sl = T[1], C = T[2], Cl = T[3];
    double        p1, p1e, p2, p2e, a, ae, b, be;
    // sin t = S cos r + C sin r
    rr_two_prod(C, rh, p1, p1e);
    rr_two_prod(S, ch, p2, p2e);
    rr_two_sum(S, p1, a, ae);
    rr_two_sum(a, p2, b, be);
    sn = b + (((ae + be) + (p1e + p2e)) + ((S * cl + C * sl) + (Sl * (1.0 + ch) + Cl * (rh + sl))));
    // cos t = C cos r - S sin r
    rr_two_prod(-S, rh, p1, p1e);
    rr_two_prod(C, ch, p2, p2e);
    rr_two_sum(C, p1, a, ae);
    rr_two_sum(a, p2, b, be);
    cs = b + (((ae + be) + (p1e + p2e)) + ((C * cl - S * sl) + (Cl * (1.0 + ch) - Sl * (rh + sl))));
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
// [2^-32, 1] and a float angle v in (0, 2 pi], evaluated in double and rounded to float
// once.  qoff = 16 q selects the angle table entry for q = round(128 v / pi), which the
// callers take from the word the angle came from (integer operations).
//   -2 ln u: u's exponent and 6 leading mantissa bits index rr_bm_logtab (a = 2^-k / c,
//     b = 2 ln a), r = u a - 1 (|r| < 2^-7) and -2 ln u = b + r (-2 + r (1 + r (-2/3 + r/2)))
//     (log1p to r^4, relative error below 2^-35; u = 1 gives b = -0 and r = 0, so -0 as the
//     reference's -2.0f * logf(1.0f));
//   sqrt: the hardware double reciprocal square root (relative error below 2^-22) and one
//     coupled Newton step (error below 2^-42); the radius keeps the sign of a zero;
//   sin/cos: v = q P + a with P = pi/128 to 15 bits, a = v - q P exact in float
//     (|a| < 0.0124), sin a to a^3 and cos a to a^4 (relative errors below 2^-32 and 2^-44),
//     then the angle-addition formulas with sin(q P), cos(q P) from rr_bm_sctab.
// The pair before its rounding is within 2^-30 of exact, so each output is within 0.51
// ulp of the exact value, and within 2 ulp of the reference's float pipeline (logf, sqrtf,
// sinf/cosf and the product: four roundings, rocrand_normal.h:51-66).  UNIT rounds
// sin(v) s + 0 instead of sin(v) s: the output of mean + 1 * x at mean +0 (a zero is +0).
template<bool UNIT = false>
__device__ __forceinline__ void box_muller_dd(float u, float v, uint32_t qoff, float& r0, float& r1)
{
    ROCRAND_AMD_FP_EXACT
    const uint32_t loff = (__float_as_uint(u) >> 13) & 0x1fff0u; // 16 * (exponent * 64 + i)
    const double2  lt   = *reinterpret_cast<const double2*>(reinterpret_cast<const char*>(rr_bm_logtab)
                                                         - RR_BM_EMIN * 64u * 16u + loff);
    const double   r    = __builtin_fma((double)u, lt.x, -1.0);
    double         p    = __builtin_fma(r, 0.5, -2.0 / 3);
    p                   = __builtin_fma(p, r, 1.0);
    p                   = __builtin_fma(p, r, -2.0);
    const double   x    = __builtin_fma(p, r, lt.y);
    const double   y    = __builtin_amdgcn_rsq(__builtin_fmax(x, 0x1p-126));
    const double   t    = x * y;
    const double   h    = __builtin_fma(-t, y, 1.0);
    const double   s    = __builtin_fma(t * h, 0.5, t);

    const double2 st = *reinterpret_cast<const double2*>(reinterpret_cast<const char*>(rr_bm_sctab) + qoff);
    const double  a  = (double)__builtin_fmaf(-(float)qoff, RR_BM_ANGLE_STEP * 0.0625f, v);
    const double  a2 = a * a;
    const double  sa = a * __builtin_fma(a2, -1.0 / 6, 1.0);
    const double  ca = __builtin_fma(a2, __builtin_fma(a2, 1.0 / 24, -0.5), 1.0);
    const double  sn = __builtin_fma(st.x, ca, st.y * sa);
    const double  cs = __builtin_fma(st.y, ca, -(st.x * sa));
    if constexpr(UNIT)
    {
        r0 = (float)__builtin_fma(sn, s, 0.0);
        r1 = (float)__builtin_fma(cs, s, 0.0);
    }
    else
    {
        r0 = (float)(sn * s);
        r1 = (float)(cs * s);
    }
}

// rocrand_normal.h:51-66: u = 2^-32 + x 2^-32, angle 2^-32 2pi + y 2^-32 2pi (rounded as
// the reference rounds them), then the pair above; the angle is about 2 pi (y + 1) 2^-32,
// so q = round(y 2^-24).
template<bool UNIT = false>
__device__ __forceinline__ void box_muller(uint32_t x, uint32_t y, float& r0, float& r1)
{
    ROCRAND_AMD_FP_EXACT
    const float u = uniform_f32(x);
    const float v = R_2POW32_INV_2PI + (float)y * R_2POW32_INV_2PI;
    box_muller_dd<UNIT>(u, v, ((y >> 20) + 8u) & ~15u, r0, r1);
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
    const double   s  = __builtin_sqrt(-2.0 * log_cr(u));
    const double   t  = w * R_PI_DBL;
    double         sn, cs;
    sincos_cr(t, sn, cs);
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
// The MRG outputs y in [1, m] scale by about 2^-32 (MRG32k3a) or 2^-31 (MRG31k3p), so
// q = round(y 2^-24) or round(y 2^-23).
template<int MRG, bool UNIT = false>
__device__ __forceinline__ void mrg_box_muller(uint32_t x, uint32_t y, float& r0, float& r1)
{
    ROCRAND_AMD_FP_EXACT
    constexpr int SH = MRG == 32 ? 20 : 19;
    box_muller_dd<UNIT>(mrg_f32<MRG>(x), mrg_f32<MRG>(y) * R_2PI, ((y >> SH) + 8u) & ~15u, r0, r1);
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
    const double s = __builtin_sqrt(-2.0 * log_cr(u));
    double       sn, cs;
    sincos_cr(v * R_PI_DBL, sn, cs);
    r0 = sn * s;
    r1 = cs * s;
}
// rocrand_normal.h:199-221 roc_d_erfinv and :263-268 normal_distribution_double.
__device__ __forceinline__ double roc_d_erfinv(double x)
{
    ROCRAND_AMD_FP_EXACT
    const double sgn = (x < 0.0) ? -1.0 : 1.0;
    x                = (1.0 - x) * (1.0 + x);
    const double lnx = log_cr(x);
    if(isnan(lnx))
        return 1.0;
    if(isinf(lnx))
        return 0.0;
    const double tt1 = 2.0 / (R_PI_DBL * 0.147) + 0.5 * lnx;
    const double tt2 = 1.0 / 0.147 * lnx;
    return sgn * __builtin_sqrt(-tt1 + __builtin_sqrt(tt1 * tt1 - tt2));
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
    const float lnx = x > 0.0f ? logf_tab(x) : -__builtin_inff();
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

