/*
 * orc_gen.c -- distributions and generator objects of the oracle (test infrastructure).
 *
 * The distribution functors restate library/src/rng/distribution/{uniform,normal,log_normal,poisson}.hpp over the
 * scalar math of library/include/rocrand/rocrand_{uniform,normal,log_normal,discrete}.h,
 * host branch (glibc sinf/cosf/logf/expf, no FMA contraction: build with
 * -ffp-contract=off on baseline x86-64, as the reference host path is built).
 * The generate functions restate the kernel bodies over the DEFAULT grid exactly
 * as library/src/rng/system.hpp:140-214 (host_system::launch) runs them.
 */
#include "orc_internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* rocrand_common.h:24-37 */
#define R_2POW16_INV       1.5258789e-05f
#define R_2POW16_INV_2PI   9.58738e-05f
#define R_2POW32_INV       2.3283064e-10f
#define R_2POW32_INV_DBL   2.3283064365386963e-10
#define R_2POW64_INV_DBL   5.4210108624275221700372640043497e-20
#define R_2POW32_INV_2PI   1.46291807e-09f
#define R_2POW53_INV_DBL   1.1102230246251565e-16
#define R_PI_F             3.141592653f
#define R_PI_DBL           3.1415926535897932
#define R_2PI_F            6.2831855f
#define R_SQRT2_F          1.4142135f
/* rocrand_mrg32k3a.h:45-47, rocrand_mrg31k3p.h:33-37 */
#define MRG32_NORM_DBL 2.3283065498378288e-10
#define MRG32_UINT_NORM 1.000000048661607
#define MRG31_NORM_DBL 4.656612875245796923e-10
#define MRG31_UINT_NORM 2.000000001396983862

/* ------------------------------------------------------------------------ */
/* half conversion (IEEE binary16, round to nearest even, as __float2half)  */
/* ------------------------------------------------------------------------ */
unsigned short orc_float2half(float f)
{
    uint32_t x;
    memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t absx = x & 0x7fffffffu;
    if(absx >= 0x7f800000u) /* inf / nan */
        return (unsigned short)(sign | 0x7c00u | (absx > 0x7f800000u ? 0x200u | ((absx >> 13) & 0x3ffu) : 0));
    if(absx >= 0x477ff000u) /* rounds to >= 65520 -> inf */
        return (unsigned short)(sign | 0x7c00u);
    if(absx < 0x38800000u) /* subnormal half or zero */
    {
        if(absx < 0x33000000u) /* < 2^-25: rounds to 0 (2^-25 exactly ties to even 0) */
            return (unsigned short)sign;
        const uint32_t e     = absx >> 23;
        const uint32_t m     = (absx & 0x7fffffu) | 0x800000u;
        const uint32_t shift = 126 - e; /* 14 + (113 - e) */
        uint32_t       r     = m >> shift;
        const uint32_t rem   = m & ((1u << shift) - 1);
        const uint32_t half  = 1u << (shift - 1);
        if(rem > half || (rem == half && (r & 1u)))
            r++;
        return (unsigned short)(sign | r);
    }
    uint32_t       h   = ((absx >> 13) - (112u << 10));
    const uint32_t rem = absx & 0x1fffu;
    if(rem > 0x1000u || (rem == 0x1000u && (h & 1u)))
        h++;
    return (unsigned short)(sign | h);
}

static float half2float(unsigned short h)
{
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    uint32_t       e    = (h >> 10) & 0x1fu;
    uint32_t       m    = h & 0x3ffu;
    uint32_t       x;
    if(e == 0)
    {
        if(m == 0)
            x = sign;
        else
        {
            e = 113;
            while(!(m & 0x400u))
            {
                m <<= 1;
                e--;
            }
            m &= 0x3ffu;
            x = sign | (e << 23) | (m << 13);
        }
    }
    else if(e == 31)
        x = sign | 0x7f800000u | (m << 13);
    else
        x = sign | ((e + 112) << 23) | (m << 13);
    float f;
    memcpy(&f, &x, 4);
    return f;
}

/* ------------------------------------------------------------------------ */
/* scalar transforms (rocrand_uniform.h, rocrand_normal.h)                   */
/* ------------------------------------------------------------------------ */

/* rocrand_uniform.h:63-66 */
static float uni_f32(uint32_t v) { return R_2POW32_INV + ((float)v * R_2POW32_INV); }
/* rocrand_uniform.h:70-73 */
static float uni_f32_64(uint64_t v) { return R_2POW32_INV + (float)(v >> 32) * R_2POW32_INV; }
/* rocrand_uniform.h:100-107 */
static double uni_f64_2(uint32_t v1, uint32_t v2)
{
    const uint64_t v = (uint64_t)v1 | ((uint64_t)(v2 >> 11) << 32);
    return R_2POW53_INV_DBL + ((double)v * R_2POW53_INV_DBL);
}
/* rocrand_uniform.h:109-116 */
static double uni_f64_64(uint64_t v) { return R_2POW53_INV_DBL + ((double)(v >> 11) * R_2POW53_INV_DBL); }
/* rocrand_uniform.h:154-157 */
static unsigned short uni_f16(unsigned short v) { return orc_float2half(R_2POW16_INV + ((float)v * R_2POW16_INV)); }

/* rocrand_uniform.h:161-219 MRG remaps */
static uint32_t mrg_uint(int t, uint32_t v)
{
    return (uint32_t)((double)(v - 1) * (t == RNG_MRG32K3A ? MRG32_UINT_NORM : MRG31_UINT_NORM));
}
static float mrg_f32(int t, uint32_t v) { return (float)((double)v * (t == RNG_MRG32K3A ? MRG32_NORM_DBL : MRG31_NORM_DBL)); }
static double mrg_f64(int t, uint32_t v) { return (double)v * (t == RNG_MRG32K3A ? MRG32_NORM_DBL : MRG31_NORM_DBL); }

/* rocrand_normal.h:51-66 box_muller, host branch */
static void box_muller(uint32_t x, uint32_t y, float* r0, float* r1)
{
    const float u = R_2POW32_INV + ((float)x * R_2POW32_INV);
    const float v = R_2POW32_INV_2PI + ((float)y * R_2POW32_INV_2PI);
    const float s = sqrtf(-2.0f * logf(u));
    *r0 = sinf(v) * s;
    *r1 = cosf(v) * s;
}
/* rocrand_normal.h:76-96 box_muller_double, host branch */
static void box_muller_double(uint32_t a, uint32_t b, uint32_t c, uint32_t d, double* r0, double* r1)
{
    const uint64_t v1 = (uint64_t)a ^ ((uint64_t)b << (53 - 32));
    const double   u  = R_2POW53_INV_DBL + ((double)v1 * R_2POW53_INV_DBL);
    const uint64_t v2 = (uint64_t)c ^ ((uint64_t)d << (53 - 32));
    const double   w  = (R_2POW53_INV_DBL * 2.0) + ((double)v2 * (R_2POW53_INV_DBL * 2.0));
    const double   s  = sqrt(-2.0 * log(u));
    *r0 = sin(w * R_PI_DBL) * s;
    *r1 = cos(w * R_PI_DBL) * s;
}
/* rocrand_normal.h:108-135 box_muller_half, host (no half math) branch */
static void box_muller_half(unsigned short x, unsigned short y, unsigned short* h0, unsigned short* h1)
{
    const float u = R_2POW16_INV + ((float)x * R_2POW16_INV);
    const float v = R_2POW16_INV_2PI + ((float)y * R_2POW16_INV_2PI);
    const float s = sqrtf(-2.0f * logf(u));
    *h0 = orc_float2half(sinf(v) * s);
    *h1 = orc_float2half(cosf(v) * s);
}
/* rocrand_normal.h:138-154 mrg_box_muller */
static void mrg_box_muller(int t, uint32_t x, uint32_t y, float* r0, float* r1)
{
    const float u = mrg_f32(t, x);
    const float v = mrg_f32(t, y) * R_2PI_F;
    const float s = sqrtf(-2.0f * logf(u));
    *r0 = sinf(v) * s;
    *r1 = cosf(v) * s;
}
/* rocrand_normal.h:156-172 mrg_box_muller_double: the uniforms are float-precision. */
static void mrg_box_muller_double(int t, uint32_t x, uint32_t y, double* r0, double* r1)
{
    const double u = (double)mrg_f32(t, x);
    const double v = (double)mrg_f32(t, y) * 2.0;
    const double s = sqrt(-2.0 * log(u));
    *r0 = sin(v * R_PI_DBL) * s;
    *r1 = cos(v * R_PI_DBL) * s;
}
/* rocrand_normal.h:199-221 roc_d_erfinv + :263-268 normal_distribution_double(uint) */
static double roc_d_erfinv(double x)
{
    const double sgn = (x < 0.0) ? -1.0 : 1.0;
    x                = (1.0 - x) * (1.0 + x);
    const double lnx = log(x);
    if(isnan(lnx))
        return 1.0;
    else if(isinf(lnx))
        return 0.0;
    const double tt1 = 2.0 / (R_PI_DBL * 0.147) + 0.5 * lnx;
    const double tt2 = 1.0 / (0.147) * lnx;
    return (sgn * sqrt(-tt1 + sqrt(tt1 * tt1 - tt2)));
}
static double normal_double_1(double p) { return (double)R_SQRT2_F * roc_d_erfinv(2.0 * p - 1.0); }

/* ------------------------------------------------------------------------ */
/* Poisson tables (library/src/rng/distribution/poisson.hpp:110-161,        */
/* discrete.hpp:189-265) and sampling (rocrand_discrete.h:61-90)             */
/* ------------------------------------------------------------------------ */
#define LAMBDA_THRESHOLD_HUGE 4000.0

typedef struct
{
    unsigned size, offset;
    double*  prob;
    unsigned* alias;
} ptable;

static void poisson_probs(double lambda, double** pout, unsigned* size, unsigned* offset)
{
    const size_t capacity   = 2 * (size_t)(16.0 * (2.0 + sqrt(lambda)));
    double*      p          = (double*)calloc(capacity, sizeof(double));
    const double p_epsilon  = 1e-12;
    const double log_lambda = log(lambda);
    const int    left       = (int)((size_t)(long)(int)floor(lambda) - capacity / 2);
    int lo = 0;
    for(int i = (int)(capacity / 2); i >= 0; i--)
    {
        const double x  = left + i;
        const double pp = exp(x * log_lambda - lgamma(x + 1.0) - lambda);
        if(pp < p_epsilon)
        {
            lo = i + 1;
            break;
        }
        p[i] = pp;
    }
    int hi = (int)capacity - 1;
    for(int i = (int)(capacity / 2) + 1; i < (int)capacity; i++)
    {
        const double x  = left + i;
        const double pp = exp(x * log_lambda - lgamma(x + 1.0) - lambda);
        if(pp < p_epsilon)
        {
            hi = i - 1;
            break;
        }
        p[i] = pp;
    }
    for(int i = lo; i <= hi; i++)
        p[i - lo] = p[i];
    *size   = (unsigned)(hi - lo + 1);
    *offset = (unsigned)(left + lo);
    *pout   = p;
}

/* normalize, discrete.hpp:189-200 */
static void normalize(double* p, unsigned size)
{
    double sum = 0.0;
    for(unsigned i = 0; i < size; i++)
        sum += p[i];
    for(unsigned i = 0; i < size; i++)
        p[i] /= sum;
}

/* Vose alias over normalized p (consumed), LIFO stacks, discrete.hpp:212-265 */
static void vose_alias(double* p, unsigned size, double* prob, unsigned* alias)
{
    const double average = 1.0 / size;
    unsigned*    small   = (unsigned*)malloc(size * sizeof(unsigned));
    unsigned*    large   = (unsigned*)malloc(size * sizeof(unsigned));
    unsigned     ns = 0, nl = 0;
    for(unsigned i = 0; i < size; i++)
    {
        if(p[i] >= average)
            large[nl++] = i;
        else
            small[ns++] = i;
    }
    for(unsigned i = 0; i < size; i++)
        alias[i] = 0;
    while(ns > 0 && nl > 0)
    {
        const unsigned less = small[--ns];
        const unsigned more = large[--nl];
        prob[less]  = p[less] * size;
        alias[less] = more;
        p[more]     = (p[more] + p[less]) - average;
        if(p[more] >= average)
            large[nl++] = more;
        else
            small[ns++] = more;
    }
    for(unsigned i = 0; i < ns; i++)
        prob[small[i]] = 1.0;
    for(unsigned i = 0; i < nl; i++)
        prob[large[i]] = 1.0;
    free(small);
    free(large);
}

static void build_alias(double* p, unsigned size, double* prob, unsigned* alias)
{
    normalize(p, size);
    vose_alias(p, size, prob, alias);
}

int orc_discrete_tables(const double* weights, unsigned size, double* prob, unsigned* alias, double* cdf)
{
    /* discrete.hpp:111-143 create(): normalize once, then the alias table from a copy and
     * the running sums (create_cdf, discrete.hpp:277-286) from the normalized weights */
    double* p = (double*)malloc(size * sizeof(double));
    if(p == NULL)
        return -1;
    memcpy(p, weights, size * sizeof(double));
    normalize(p, size);
    double sum = 0.0;
    for(unsigned i = 0; i < size; i++)
    {
        sum += p[i];
        cdf[i] = sum;
    }
    vose_alias(p, size, prob, alias);
    free(p);
    return 0;
}

int orc_poisson_table(double lambda, unsigned int* offset, double* prob, unsigned int* alias)
{
    double*  p;
    unsigned size;
    poisson_probs(lambda, &p, &size, offset);
    build_alias(p, size, prob, alias);
    free(p);
    return (int)size;
}

static unsigned discrete_alias(double x, const ptable* t)
{
    const double   nx  = t->size * x;
    const double   fnx = floor(nx);
    const double   y   = nx - fnx;
    const unsigned i   = (unsigned)fnx;
    return t->offset + (y < t->prob[i] ? i : t->alias[i]);
}

/* ------------------------------------------------------------------------ */
/* distribution functors                                                    */
/* ------------------------------------------------------------------------ */
typedef struct dist dist;
struct dist
{
    unsigned in_w, out_w, esz;
    void (*apply)(const dist*, const uint64_t* in, uint8_t* out);
    int      mrg;            /* MRG engine type for remapping, 0 otherwise */
    float    mf, sf;         /* float mean / stddev */
    double   md, sd;         /* double mean / stddev */
    unsigned short mh, sh;   /* half mean / stddev */
    int      huge;
    double   lambda, sqrt_lambda;
    ptable   pt;
};

#define OUT(T, i) (((T*)out)[i])

static void d_u32(const dist* d, const uint64_t* in, uint8_t* out)
{
    uint32_t v = (uint32_t)in[0];
    if(d->mrg) v = mrg_uint(d->mrg, v);
    OUT(uint32_t, 0) = v;
}
static void d_u8(const dist* d, const uint64_t* in, uint8_t* out)
{
    uint32_t v = (uint32_t)in[0];
    if(d->mrg) v = mrg_uint(d->mrg, v);
    memcpy(out, &v, 4);
}
static void d_u16(const dist* d, const uint64_t* in, uint8_t* out) { d_u8(d, in, out); }
static void d_f32(const dist* d, const uint64_t* in, uint8_t* out)
{
    OUT(float, 0) = d->mrg ? mrg_f32(d->mrg, (uint32_t)in[0]) : uni_f32((uint32_t)in[0]);
}
static void d_f64(const dist* d, const uint64_t* in, uint8_t* out)
{
    OUT(double, 0) = d->mrg ? mrg_f64(d->mrg, (uint32_t)in[0]) : uni_f64_2((uint32_t)in[0], (uint32_t)in[1]);
}
static void d_f16(const dist* d, const uint64_t* in, uint8_t* out)
{
    uint32_t v = (uint32_t)in[0];
    if(d->mrg) v = mrg_uint(d->mrg, v);
    OUT(unsigned short, 0) = uni_f16((unsigned short)v);
    OUT(unsigned short, 1) = uni_f16((unsigned short)(v >> 16));
}
static void d_nrm_f32(const dist* d, const uint64_t* in, uint8_t* out)
{
    float r0, r1;
    if(d->mrg) mrg_box_muller(d->mrg, (uint32_t)in[0], (uint32_t)in[1], &r0, &r1);
    else box_muller((uint32_t)in[0], (uint32_t)in[1], &r0, &r1);
    OUT(float, 0) = d->mf + r0 * d->sf;
    OUT(float, 1) = d->mf + r1 * d->sf;
}
static void d_ln_f32(const dist* d, const uint64_t* in, uint8_t* out)
{
    float r0, r1;
    if(d->mrg) mrg_box_muller(d->mrg, (uint32_t)in[0], (uint32_t)in[1], &r0, &r1);
    else box_muller((uint32_t)in[0], (uint32_t)in[1], &r0, &r1);
    OUT(float, 0) = expf(d->mf + r0 * d->sf);
    OUT(float, 1) = expf(d->mf + r1 * d->sf);
}
static void d_nrm_f64(const dist* d, const uint64_t* in, uint8_t* out)
{
    double r0, r1;
    if(d->mrg) mrg_box_muller_double(d->mrg, (uint32_t)in[0], (uint32_t)in[1], &r0, &r1);
    else box_muller_double((uint32_t)in[0], (uint32_t)in[1], (uint32_t)in[2], (uint32_t)in[3], &r0, &r1);
    OUT(double, 0) = d->md + r0 * d->sd;
    OUT(double, 1) = d->md + r1 * d->sd;
}
static void d_ln_f64(const dist* d, const uint64_t* in, uint8_t* out)
{
    double r0, r1;
    if(d->mrg) mrg_box_muller_double(d->mrg, (uint32_t)in[0], (uint32_t)in[1], &r0, &r1);
    else box_muller_double((uint32_t)in[0], (uint32_t)in[1], (uint32_t)in[2], (uint32_t)in[3], &r0, &r1);
    OUT(double, 0) = exp(d->md + r0 * d->sd);
    OUT(double, 1) = exp(d->md + r1 * d->sd);
}
/* normal.hpp:65-79 (ThreeFry2x32 double, max input width 2): one word, erfinv. */
static void d_nrm_f64_1(const dist* d, const uint64_t* in, uint8_t* out)
{
    const double v = normal_double_1(R_2POW32_INV_DBL + ((double)(uint32_t)in[0] * R_2POW32_INV_DBL));
    OUT(double, 0) = d->md + v * d->sd;
}
static void d_ln_f64_1(const dist* d, const uint64_t* in, uint8_t* out)
{
    const double v = normal_double_1(R_2POW32_INV_DBL + ((double)(uint32_t)in[0] * R_2POW32_INV_DBL));
    OUT(double, 0) = exp(d->md + v * d->sd);
}
static void half_pair(const dist* d, uint32_t v, unsigned short* h0, unsigned short* h1)
{
    if(d->mrg) v = mrg_uint(d->mrg, v);
    box_muller_half((unsigned short)v, (unsigned short)(v >> 16), h0, h1);
}
static void d_nrm_f16(const dist* d, const uint64_t* in, uint8_t* out)
{
    unsigned short h0, h1;
    half_pair(d, (uint32_t)in[0], &h0, &h1);
    const float m = half2float(d->mh), s = half2float(d->sh);
    OUT(unsigned short, 0) = orc_float2half(m + (s * half2float(h0)));
    OUT(unsigned short, 1) = orc_float2half(m + (s * half2float(h1)));
}
static void d_ln_f16(const dist* d, const uint64_t* in, uint8_t* out)
{
    unsigned short h0, h1;
    half_pair(d, (uint32_t)in[0], &h0, &h1);
    const float m = half2float(d->mh), s = half2float(d->sh);
    OUT(unsigned short, 0) = orc_float2half(expf(m + (s * half2float(h0))));
    OUT(unsigned short, 1) = orc_float2half(expf(m + (s * half2float(h1))));
}
static void d_poisson(const dist* d, const uint64_t* in, uint8_t* out)
{
    uint32_t r = (uint32_t)in[0];
    if(d->mrg) r = mrg_uint(d->mrg, r);
    if(d->huge)
    {
        /* poisson.hpp:80-108 poisson_distribution_huge */
        const double nd = normal_double_1(R_2POW32_INV_DBL + ((double)r * R_2POW32_INV_DBL));
        OUT(uint32_t, 0) = (uint32_t)round(d->sqrt_lambda * nd + d->lambda);
        return;
    }
    OUT(uint32_t, 0) = discrete_alias((double)r * R_2POW32_INV_DBL, &d->pt);
}

/* 64-bit-word engines (ThreeFry 2x64 / 4x64), uniform.hpp:113-218, normal.hpp:114-198 */
static void d64_u64(const dist* d, const uint64_t* in, uint8_t* out) { (void)d; OUT(uint64_t, 0) = in[0]; }
static void d64_bytes(const dist* d, const uint64_t* in, uint8_t* out) { (void)d; memcpy(out, &in[0], 8); }
static void d64_f32(const dist* d, const uint64_t* in, uint8_t* out) { (void)d; OUT(float, 0) = uni_f32_64(in[0]); }
static void d64_f64(const dist* d, const uint64_t* in, uint8_t* out) { (void)d; OUT(double, 0) = uni_f64_64(in[0]); }
static void d64_f16(const dist* d, const uint64_t* in, uint8_t* out)
{
    (void)d;
    for(int i = 0; i < 4; i++)
        OUT(unsigned short, i) = uni_f16((unsigned short)(in[0] >> (16 * i)));
}
static void d64_nrm_f32(const dist* d, const uint64_t* in, uint8_t* out)
{
    float r0, r1;
    box_muller((uint32_t)in[0], (uint32_t)(in[0] >> 32), &r0, &r1);
    OUT(float, 0) = d->mf + r0 * d->sf;
    OUT(float, 1) = d->mf + r1 * d->sf;
}
static void d64_ln_f32(const dist* d, const uint64_t* in, uint8_t* out)
{
    float r0, r1;
    box_muller((uint32_t)in[0], (uint32_t)(in[0] >> 32), &r0, &r1);
    OUT(float, 0) = expf(d->mf + r0 * d->sf);
    OUT(float, 1) = expf(d->mf + r1 * d->sf);
}
static void d64_nrm_f64(const dist* d, const uint64_t* in, uint8_t* out)
{
    double r0, r1;
    box_muller_double((uint32_t)in[0], (uint32_t)(in[0] >> 32), (uint32_t)in[1], (uint32_t)(in[1] >> 32), &r0, &r1);
    OUT(double, 0) = d->md + r0 * d->sd;
    OUT(double, 1) = d->md + r1 * d->sd;
}
static void d64_ln_f64(const dist* d, const uint64_t* in, uint8_t* out)
{
    double r0, r1;
    box_muller_double((uint32_t)in[0], (uint32_t)(in[0] >> 32), (uint32_t)in[1], (uint32_t)(in[1] >> 32), &r0, &r1);
    OUT(double, 0) = exp(d->md + r0 * d->sd);
    OUT(double, 1) = exp(d->md + r1 * d->sd);
}
static void d64_half_pair(uint64_t v, unsigned short* h0, unsigned short* h1)
{
    box_muller_half((unsigned short)v, (unsigned short)(v >> 32), h0, h1);
}
static void d64_nrm_f16(const dist* d, const uint64_t* in, uint8_t* out)
{
    unsigned short h0, h1;
    d64_half_pair(in[0], &h0, &h1);
    const float m = half2float(d->mh), s = half2float(d->sh);
    OUT(unsigned short, 0) = orc_float2half(m + (s * half2float(h0)));
    OUT(unsigned short, 1) = orc_float2half(m + (s * half2float(h1)));
}
static void d64_ln_f16(const dist* d, const uint64_t* in, uint8_t* out)
{
    unsigned short h0, h1;
    d64_half_pair(in[0], &h0, &h1);
    const float m = half2float(d->mh), s = half2float(d->sh);
    OUT(unsigned short, 0) = orc_float2half(expf(m + (s * half2float(h0))));
    OUT(unsigned short, 1) = orc_float2half(expf(m + (s * half2float(h1))));
}
static void d64_poisson(const dist* d, const uint64_t* in, uint8_t* out)
{
    if(d->huge)
    {
        const double nd = normal_double_1(uni_f64_64(in[0]));
        OUT(uint32_t, 0) = (uint32_t)round(d->sqrt_lambda * nd + d->lambda);
        return;
    }
    OUT(uint32_t, 0) = discrete_alias((double)in[0] * R_2POW64_INV_DBL, &d->pt);
}

/* ------------------------------------------------------------------------ */
/* generator objects                                                        */
/* ------------------------------------------------------------------------ */
#define S_ENGINES 131072u      /* 512 blocks x 256 threads, config_defaults.hpp:33-73 */
#define CTR_STRIDE 262144u     /* 1024 blocks x 256 threads, config_defaults.hpp:54-108 */

struct orc_gen
{
    int        type;
    uint64_t   seed;
    uint32_t   seed4[4];
    uint64_t   offset;
    unsigned   dims;
    int        initialized;
    ctr_engine ce;
    unsigned   start;
    void*      engines;
    unsigned   S;          /* engines of the interleaved engines (grid threads x blocks) */
    unsigned   ctr_stride; /* grid stride of the counter engines' generate kernel */
    struct orc_ext* ext; /* MTGP32 / MT19937 / Sobol state (orc_ext.c) */
};

int  orc_ext_create(orc_gen* g);
void orc_ext_destroy(orc_gen* g);
int  orc_ext_generate(orc_gen* g, int kind, void* out, size_t n, uint64_t addr, double a, double b);

static int is_counter(int t)
{
    return t == RNG_PHILOX4_32_10 || t == RNG_THREEFRY2_32_20 || t == RNG_THREEFRY2_64_20
           || t == RNG_THREEFRY4_32_20 || t == RNG_THREEFRY4_64_20;
}
static int is_interleaved(int t)
{
    return t == RNG_MRG32K3A || t == RNG_MRG31K3P || t == RNG_XORWOW || t == RNG_LFSR113;
}
static int is_ext(int t)
{
    return t == RNG_MTGP32 || t == RNG_MT19937 || t == RNG_SOBOL32 || t == RNG_SCRAMBLED_SOBOL32
           || t == RNG_SOBOL64 || t == RNG_SCRAMBLED_SOBOL64;
}

orc_gen* orc_create(int rng_type)
{
    if(rng_type == 400)
        rng_type = RNG_XORWOW; /* ROCRAND_RNG_PSEUDO_DEFAULT, rocrand.cpp:115-125 */
    if(rng_type == 500)
        rng_type = RNG_SOBOL32; /* ROCRAND_RNG_QUASI_DEFAULT */
    if(!is_counter(rng_type) && !is_interleaved(rng_type) && !is_ext(rng_type))
        return NULL;
    orc_gen* g    = (orc_gen*)calloc(1, sizeof(orc_gen));
    g->type       = rng_type;
    g->dims       = 1;
    g->S          = S_ENGINES;
    g->ctr_stride = CTR_STRIDE;
    if(rng_type == RNG_MRG32K3A || rng_type == RNG_MRG31K3P)
        g->seed = 12345ull;
    if(rng_type == RNG_LFSR113)
    {
        g->seed4[0] = 2; g->seed4[1] = 8; g->seed4[2] = 16; g->seed4[3] = 128;
    }
    if(is_ext(rng_type) && orc_ext_create(g) != 0)
    {
        free(g);
        return NULL;
    }
    return g;
}

void orc_destroy(orc_gen* g)
{
    if(!g)
        return;
    if(is_ext(g->type))
        orc_ext_destroy(g);
    free(g->engines);
    free(g);
}

/* Accessors for orc_ext.c */
int      orc_gen_type(const orc_gen* g) { return g->type; }
uint64_t orc_gen_seed(const orc_gen* g) { return g->seed; }
uint64_t orc_gen_offset(const orc_gen* g) { return g->offset; }
unsigned orc_gen_dims(const orc_gen* g) { return g->dims; }
struct orc_ext* orc_gen_ext(orc_gen* g) { return g->ext; }
void     orc_gen_set_ext(orc_gen* g, struct orc_ext* e) { g->ext = e; }
int*     orc_gen_initialized(orc_gen* g) { return &g->initialized; }

int orc_set_seed(orc_gen* g, unsigned long long seed)
{
    if(!g) return ST_NOT_CREATED;
    if(g->type == RNG_SOBOL32 || g->type == RNG_SOBOL64 || g->type == RNG_SCRAMBLED_SOBOL32
       || g->type == RNG_SCRAMBLED_SOBOL64)
        return ST_SUCCESS; /* sobol.hpp:581-584: the seed is ignored, nothing is reset */
    if(g->type == RNG_LFSR113)
    {
        /* lfsr113.hpp:240-255: only x and y are raised to their minimum seeds */
        uint32_t s[4] = {(uint32_t)seed, (uint32_t)(seed >> 16), (uint32_t)(seed >> 32), (uint32_t)(seed >> 48)};
        if(s[0] < 2) s[0] += 2;
        if(s[1] < 8) s[1] += 8;
        memcpy(g->seed4, s, sizeof(s));
    }
    else if(g->type == RNG_MRG32K3A || g->type == RNG_MRG31K3P)
        g->seed = seed == 0 ? 12345ull : seed;
    else
        g->seed = seed;
    g->initialized = 0;
    return ST_SUCCESS;
}

int orc_set_seed_uint4(orc_gen* g, const unsigned int seed[4])
{
    if(!g) return ST_NOT_CREATED;
    if(g->type != RNG_LFSR113)
        return ST_TYPE_ERROR;
    uint32_t s[4] = {seed[0], seed[1], seed[2], seed[3]};
    if(s[0] < 2) s[0] += 2;
    if(s[1] < 8) s[1] += 8;
    if(s[2] < 16) s[2] += 16;
    if(s[3] < 128) s[3] += 128;
    memcpy(g->seed4, s, sizeof(s));
    g->initialized = 0;
    return ST_SUCCESS;
}

int orc_set_offset(orc_gen* g, unsigned long long offset)
{
    if(!g) return ST_NOT_CREATED;
    if(g->type == RNG_MTGP32 || g->type == RNG_MT19937)
        return ST_TYPE_ERROR;
    g->offset      = offset;
    g->initialized = 0;
    return ST_SUCCESS;
}

int orc_set_dimensions(orc_gen* g, unsigned int dims)
{
    if(!g) return ST_NOT_CREATED;
    if(!(g->type == RNG_SOBOL32 || g->type == RNG_SOBOL64 || g->type == RNG_SCRAMBLED_SOBOL32
         || g->type == RNG_SCRAMBLED_SOBOL64))
        return ST_TYPE_ERROR;
    if(dims < 1 || dims > 20000)
        return ST_OUT_OF_RANGE;
    g->dims        = dims;
    g->initialized = 0;
    return ST_SUCCESS;
}

int orc_set_ordering(orc_gen* g, int ordering)
{
    (void)ordering; /* the grid is set by orc_set_grid (the DEFAULT grid unless changed) */
    if(!g) return ST_NOT_CREATED;
    return ST_SUCCESS;
}

/* The launch grid the reference's generate kernels would run with under a tuned
 * (DYNAMIC-ordering) configuration (config_types.hpp:314-336): only threads x blocks
 * matters to the values -- the engine count of the interleaved engines (mrg.hpp:284-337,
 * get_least_common_grid_size) and the grid stride of the counter engines'
 * kernels (philox4x32_10.hpp:106-214).  Resets the generator like set_offset. */
int orc_set_grid(orc_gen* g, unsigned threads, unsigned blocks)
{
    if(!g) return ST_NOT_CREATED;
    const unsigned long long size = (unsigned long long)threads * blocks;
    if(size == 0 || size % 256 != 0 || size > (1u << 22))
        return ST_OUT_OF_RANGE;
    if(is_counter(g->type))
        g->ctr_stride = (unsigned)size;
    else if(is_interleaved(g->type))
        g->S = (unsigned)size;
    else
        return ST_TYPE_ERROR;
    g->initialized = 0;
    return ST_SUCCESS;
}

/* Build the functor for (engine family, kind). Returns 0 on success, else a status. */
static int make_dist(const orc_gen* g, int kind, double a, double b, dist* d)
{
    memset(d, 0, sizeof(*d));
    const int w64 = g->type == RNG_THREEFRY2_64_20 || g->type == RNG_THREEFRY4_64_20;
    d->mrg = (g->type == RNG_MRG32K3A || g->type == RNG_MRG31K3P) ? g->type : 0;
    d->mf = (float)a; d->sf = (float)b; d->md = a; d->sd = b;
    d->mh = orc_float2half((float)a); d->sh = orc_float2half((float)b);
    if(kind == ORC_POISSON)
    {
        if(a <= 0.0)
            return ST_OUT_OF_RANGE;
        d->lambda = a;
        d->sqrt_lambda = sqrt(a);
        if(a > LAMBDA_THRESHOLD_HUGE)
            d->huge = 1;
        else
        {
            double* p;
            poisson_probs(a, &p, &d->pt.size, &d->pt.offset);
            d->pt.prob  = (double*)calloc(d->pt.size, sizeof(double));
            d->pt.alias = (unsigned*)calloc(d->pt.size, sizeof(unsigned));
            build_alias(p, d->pt.size, d->pt.prob, d->pt.alias);
            free(p);
        }
    }
#define SET(iw, ow, sz, fn) do { d->in_w = iw; d->out_w = ow; d->esz = sz; d->apply = fn; } while(0)
    if(!w64)
    {
        switch(kind)
        {
            case ORC_U8: SET(1, 4, 1, d_u8); break;
            case ORC_U16: SET(1, 2, 2, d_u16); break;
            case ORC_U32: SET(1, 1, 4, d_u32); break;
            case ORC_U64: return ST_TYPE_ERROR;
            case ORC_UNIFORM_F32: SET(1, 1, 4, d_f32); break;
            case ORC_UNIFORM_F64: if(d->mrg) SET(1, 1, 8, d_f64); else SET(2, 1, 8, d_f64); break;
            case ORC_UNIFORM_F16: SET(1, 2, 2, d_f16); break;
            case ORC_NORMAL_F32: SET(2, 2, 4, d_nrm_f32); break;
            case ORC_LOGNORMAL_F32: SET(2, 2, 4, d_ln_f32); break;
            case ORC_NORMAL_F64:
                if(d->mrg) SET(2, 2, 8, d_nrm_f64);
                else if(g->type == RNG_THREEFRY2_32_20) SET(1, 1, 8, d_nrm_f64_1);
                else SET(4, 2, 8, d_nrm_f64);
                break;
            case ORC_LOGNORMAL_F64:
                if(d->mrg) SET(2, 2, 8, d_ln_f64);
                else if(g->type == RNG_THREEFRY2_32_20) SET(1, 1, 8, d_ln_f64_1);
                else SET(4, 2, 8, d_ln_f64);
                break;
            case ORC_NORMAL_F16: SET(1, 2, 2, d_nrm_f16); break;
            case ORC_LOGNORMAL_F16: SET(1, 2, 2, d_ln_f16); break;
            case ORC_POISSON: SET(1, 1, 4, d_poisson); break;
            default: return ST_TYPE_ERROR;
        }
    }
    else
    {
        switch(kind)
        {
            case ORC_U8: SET(1, 8, 1, d64_bytes); break;
            case ORC_U16: SET(1, 4, 2, d64_bytes); break;
            case ORC_U32: SET(1, 2, 4, d64_bytes); break;
            case ORC_U64: SET(1, 1, 8, d64_u64); break;
            case ORC_UNIFORM_F32: SET(1, 1, 4, d64_f32); break;
            case ORC_UNIFORM_F64: SET(1, 1, 8, d64_f64); break;
            case ORC_UNIFORM_F16: SET(1, 4, 2, d64_f16); break;
            case ORC_NORMAL_F32: SET(1, 2, 4, d64_nrm_f32); break;
            case ORC_LOGNORMAL_F32: SET(1, 2, 4, d64_ln_f32); break;
            case ORC_NORMAL_F64: SET(2, 2, 8, d64_nrm_f64); break;
            case ORC_LOGNORMAL_F64: SET(2, 2, 8, d64_ln_f64); break;
            case ORC_NORMAL_F16: SET(1, 2, 2, d64_nrm_f16); break;
            case ORC_LOGNORMAL_F16: SET(1, 2, 2, d64_ln_f16); break;
            case ORC_POISSON: SET(1, 1, 4, d64_poisson); break;
            default: return ST_TYPE_ERROR;
        }
    }
#undef SET
    return 0;
}

static void free_dist(dist* d)
{
    free(d->pt.prob);
    free(d->pt.alias);
}

/* Distributions over 32-bit engine words for MTGP32 / MT19937 (orc_ext.c, orc_mt19937.c):
   the same uniform/normal/log-normal/Poisson functors as Philox (mtgp32.hpp:375-470). */
int orc_w32_dist(int kind, double a, double b, void** handle, unsigned* in_w, unsigned* out_w, unsigned* esz)
{
    orc_gen tmp;
    memset(&tmp, 0, sizeof(tmp));
    tmp.type = RNG_MTGP32;
    dist* d  = (dist*)calloc(1, sizeof(dist));
    const int st = make_dist(&tmp, kind, a, b, d);
    if(st != 0)
    {
        free_dist(d);
        free(d);
        return st;
    }
    *handle = d;
    *in_w   = d->in_w;
    *out_w  = d->out_w;
    *esz    = d->esz;
    return 0;
}
void orc_w32_apply(void* handle, const uint64_t* in, uint8_t* out)
{
    const dist* d = (const dist*)handle;
    d->apply(d, in, out);
}
void orc_w32_free(void* handle)
{
    free_dist((dist*)handle);
    free(handle);
}

/* CDF table of the Poisson pmf (discrete.hpp:189-200 normalize, :276-286 running sum), used by
   the quasi-random generators (sobol.hpp:500-503 select DISCRETE_METHOD_CDF). */
int orc_poisson_cdf(double lambda, unsigned* size, unsigned* offset, double** cdf)
{
    double* p;
    poisson_probs(lambda, &p, size, offset);
    double sum = 0.0;
    for(unsigned i = 0; i < *size; i++)
        sum += p[i];
    for(unsigned i = 0; i < *size; i++)
        p[i] /= sum;
    double run = 0.0;
    for(unsigned i = 0; i < *size; i++)
    {
        run += p[i];
        p[i] = run;
    }
    *cdf = p;
    return 0;
}

/* philox4x32_10.hpp:106-214 generate_philox and threefry.hpp:86-198 generate_threefry. */
static void gen_counter(orc_gen* g, const dist* d, uint8_t* out, size_t n, uint64_t addr)
{
    const unsigned dim = (unsigned)g->ce.dim;
    const unsigned in_w = d->in_w, out_w = d->out_w, esz = d->esz;
    const unsigned opt = dim / in_w, full = opt * out_w;
    const size_t   mis = (full - (addr / esz) % full) % full;
    const unsigned head = (unsigned)(n < mis ? n : mis);
    const unsigned tail = (unsigned)((n - head) % full);
    const size_t   vec_n = (n - head) / full;
    const int      philox = g->type == RNG_PHILOX4_32_10;
    uint8_t        tmp[64];
    const uint32_t stride = g->ctr_stride;
    for(uint32_t t = 0; t < stride; t++)
    {
        const int active = (t < vec_n) || (t == 0 && head > 0) || (t == vec_n % stride && tail > 0);
        if(!active)
            continue;
        ctr_engine     e = g->ce;
        const uint32_t eoff = philox ? 4u * t + (t == 0 ? 0u : head)
                                     : dim * t + (t == 0 ? 0u : head / out_w * in_w);
        ce_discard(&e, eoff);
        if(t == 0 && head > 0)
        {
            for(unsigned s = 0; s < opt; s++)
            {
                if(s * out_w >= head)
                    break;
                uint64_t in[4];
                for(unsigned i = 0; i < in_w; i++)
                    in[i] = ce_next(&e);
                d->apply(d, in, tmp);
                for(unsigned o = 0; o < out_w; o++)
                    if(s * out_w + o < head)
                        memcpy(out + (size_t)(s * out_w + o) * esz, tmp + o * esz, esz);
            }
        }
        size_t index = t;
        while(index < vec_n)
        {
            uint64_t v[4];
            ce_next_leap(&e, stride, v);
            for(unsigned s = 0; s < opt; s++)
                d->apply(d, v + s * in_w, out + (mis + index * full + (size_t)s * out_w) * esz);
            index += stride;
        }
        if(index == vec_n && tail > 0)
        {
            for(unsigned s = 0; s < opt; s++)
            {
                if(s * out_w >= tail)
                    break;
                uint64_t in[4];
                for(unsigned i = 0; i < in_w; i++)
                    in[i] = ce_next(&e);
                d->apply(d, in, tmp);
                for(unsigned o = 0; o < out_w; o++)
                    if(s * out_w + o < tail)
                        memcpy(out + (n - tail + s * out_w + o) * esz, tmp + o * esz, esz);
            }
        }
    }
    /* philox4x32_10.hpp:350-354: host engine advances by ceil(n/out_w)*in_w words. */
    ce_discard(&g->ce, (uint64_t)((n + out_w - 1) / out_w) * in_w);
}

static uint32_t inter_next(int type, void* engines, unsigned eid)
{
    switch(type)
    {
        case RNG_MRG32K3A:
        case RNG_MRG31K3P: return mrg_next(&((mrg_engine*)engines)[eid]);
        case RNG_XORWOW: return xorwow_next(&((xorwow_engine*)engines)[eid]);
        default: return lfsr_next(&((lfsr_engine*)engines)[eid]);
    }
}

/* mrg.hpp:45-63 init_engines_mrg, xorwow.hpp:45-61, lfsr113.hpp:45-62.
 * Engine e = seed state jumped by e subsequences, then by offset/S + [e < start] steps.
 * The subsequence jump is applied incrementally (state_e = J state_{e-1}, J the
 * one-subsequence matrix), which is the same GF(2)/mod-m linear map as the
 * reference's per-engine digit expansion. */
static void init_interleaved(orc_gen* g)
{
    const unsigned   S = g->S;
    const uint64_t   base_off = g->offset / S;
    g->start = (unsigned)(g->offset % S);
    free(g->engines);
    if(g->type == RNG_MRG32K3A || g->type == RNG_MRG31K3P)
    {
        mrg_engine* es = (mrg_engine*)malloc(S * sizeof(mrg_engine));
        mrg_engine  cur;
        mrg_seed_only(&cur, g->type, g->seed);
        for(unsigned e = 0; e < S; e++)
        {
            if(e > 0)
                mrg_discard_subsequence(&cur, 1);
            es[e] = cur;
            mrg_discard(&es[e], base_off + (e < g->start ? 1 : 0));
        }
        g->engines = es;
    }
    else if(g->type == RNG_XORWOW)
    {
        xorwow_engine* es = (xorwow_engine*)malloc(S * sizeof(xorwow_engine));
        xorwow_engine  cur;
        xorwow_seed_only(&cur, g->seed);
        const uint32_t* J = orc_gf2_matrix(RNG_XORWOW, 1, 0);
        for(unsigned e = 0; e < S; e++)
        {
            if(e > 0)
                xorwow_apply(&cur, J);
            es[e] = cur;
            xorwow_discard(&es[e], base_off + (e < g->start ? 1 : 0));
        }
        g->engines = es;
    }
    else
    {
        lfsr_engine* es = (lfsr_engine*)malloc(S * sizeof(lfsr_engine));
        lfsr_engine  cur;
        lfsr_seed_only(&cur, g->seed4);
        const uint32_t* J = orc_gf2_matrix(RNG_LFSR113, 1, 0);
        /* lfsr113.hpp:345-360 passes offset/S through an unsigned int. */
        const uint32_t off32 = (uint32_t)base_off;
        for(unsigned e = 0; e < S; e++)
        {
            if(e > 0)
                lfsr_apply(&cur, J);
            es[e] = cur;
            lfsr_discard(&es[e], (uint32_t)(off32 + (e < g->start ? 1u : 0u)));
        }
        g->engines = es;
    }
    g->initialized = 1;
}

/* mrg.hpp:65-157 generate_mrg (same body: xorwow.hpp:63-161, lfsr113.hpp:66-160). */
static void gen_interleaved(orc_gen* g, const dist* d, uint8_t* out, size_t n, uint64_t addr)
{
    const unsigned S = g->S;
    const unsigned in_w = d->in_w, out_w = d->out_w, esz = d->esz;
    const size_t   mis = (out_w - (addr / esz) % out_w) % out_w;
    const unsigned head = (unsigned)(n < mis ? n : mis);
    const unsigned tail = (unsigned)((n - head) % out_w);
    const size_t   vec_n = (n - head) / out_w;
    uint8_t        tmp[64];
    for(unsigned id = 0; id < S; id++)
    {
        const unsigned eid = (id + g->start) % S;
        size_t         index = id;
        uint64_t       in[4];
        while(index < vec_n)
        {
            for(unsigned i = 0; i < in_w; i++)
                in[i] = inter_next(g->type, g->engines, eid);
            d->apply(d, in, out + (mis + index * out_w) * esz);
            index += S;
        }
        if(out_w > 1 && index == vec_n)
        {
            if(head > 0)
            {
                for(unsigned i = 0; i < in_w; i++)
                    in[i] = inter_next(g->type, g->engines, eid);
                d->apply(d, in, tmp);
                for(unsigned o = 0; o < out_w; o++)
                    if(o < head)
                        memcpy(out + (size_t)o * esz, tmp + o * esz, esz);
            }
            if(tail > 0)
            {
                for(unsigned i = 0; i < in_w; i++)
                    in[i] = inter_next(g->type, g->engines, eid);
                d->apply(d, in, tmp);
                for(unsigned o = 0; o < out_w; o++)
                    if(o < tail)
                        memcpy(out + (n - tail + o) * esz, tmp + o * esz, esz);
            }
        }
    }
    /* mrg.hpp:385-389 */
    g->start = (unsigned)((g->start + (n + out_w - 1) / out_w) % S);
}

int orc_generate(orc_gen* g, int kind, void* out, size_t n, uint64_t addr, double a, double b)
{
    if(!g)
        return ST_NOT_CREATED;
    if(is_ext(g->type))
        return orc_ext_generate(g, kind, out, n, addr, a, b);
    dist d;
    int  st = make_dist(g, kind, a, b, &d);
    if(st != 0)
        return st;
    if(is_counter(g->type))
    {
        if(!g->initialized)
        {
            ce_init(&g->ce, g->type, g->seed, 0, g->offset);
            g->initialized = 1;
        }
        gen_counter(g, &d, (uint8_t*)out, n, addr);
    }
    else
    {
        if(!g->initialized)
            init_interleaved(g);
        gen_interleaved(g, &d, (uint8_t*)out, n, addr);
    }
    free_dist(&d);
    return ST_SUCCESS;
}
