/* Host twin of the float transcendentals of rocrand_amd/csrc/rr_dist.hpp (logf_tab,
 * sincosf_acc, expf_acc): the same double operations in the same order, with the tables
 * of rocrand_amd_mathtab.h (written to mathtab_twin.h by the test).  Compiled with gcc
 * -ffp-contract=off; fma() is the correctly rounded fused multiply-add, as the device's
 * v_fma_f64.
 *
 * usage: fastmath_twin STRIDE
 * Prints, for every STRIDE-th float of each domain, the number of results that differ
 * from the correctly rounded one (the long double libm result rounded to float):
 *   log  [2^-32, 1]   sin, cos  (0, 7]   exp  (-105, 90)
 *
 * usage: fastmath_twin bm N [mrg]
 * The Box-Muller pair of box_muller_dd (one double expression rounded once) over N
 * pseudo-random word pairs (32-bit words, or MRG32k3a values in [1, m1] with "mrg"):
 *   bm N  COUNT_BEYOND_2ULP_OF_HOST  MAX_ULP_TO_HOST  MAX_ERR_MILLIULP  HOST_MAX_ERR_MILLIULP
 * where the host is the reference's float pipeline over this libm (sqrtf(-2 logf(u)),
 * sinf(v) s, cosf(v) s) and the error is against the long double evaluation.  The
 * device's double reciprocal square root (within 2^-22 relative) is emulated by a
 * correctly rounded one: one Newton step leaves the radius insensitive to the seed's
 * last bits.  The angle's table entry comes from the word as on the device (q =
 * round(y 2^-24) for 32-bit words, round(y 2^-24) too for MRG32k3a values).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mathtab_twin.h"

static const double LN2        = 6.93147180559945286227e-01;

static inline uint32_t fbits(float x) { uint32_t b; memcpy(&b, &x, 4); return b; }
static inline float bitsf(uint32_t b) { float x; memcpy(&x, &b, 4); return x; }

static float logf_tab(float u)
{
    const uint32_t ix = fbits(u), tmp = ix - LOGTAB_OFF;
    const uint32_t i = (tmp >> (23 - LOGTAB_BITS)) & ((1u << LOGTAB_BITS) - 1u);
    const int k = (int32_t)tmp >> 23;
    const float z = bitsf(ix - (tmp & 0xff800000u));
    const double r = fma((double)z, LOGTAB[i][0], -1.0), r2 = r * r;
    double p = fma(r, -1.0 / 6, 1.0 / 5);
    p = fma(p, r, -0.25);
    p = fma(p, r, 1.0 / 3);
    p = fma(p, r, -0.5);
    const double hi = fma((double)k, LN2, LOGTAB[i][1]);
    return (float)(hi + fma(p, r2, r));
}

static void sincosf_acc(float xf, float* s, float* c)
{
    const double x = xf, q = rint(x * SC_INV_STEP);
    const int k = (int)q & ((1 << SCTAB_BITS) - 1);
    double r = fma(-q, SC_STEP_HI, x);
    r = fma(-q, SC_STEP_LO, r);
    const double r2 = r * r;
    double ps = fma(r2, -1.0 / 5040, 1.0 / 120);
    ps = fma(ps, r2, -1.0 / 6);
    const double sr = fma(r * r2, ps, r);
    double pc = fma(r2, -1.0 / 720, 1.0 / 24);
    pc = fma(pc, r2, -0.5);
    const double cr = fma(r2, pc, 1.0);
    *s = (float)fma(SCTAB[k][0], cr, SCTAB[k][1] * sr);
    *c = (float)fma(SCTAB[k][1], cr, -(SCTAB[k][0] * sr));
}

static float expf_acc(float xf)
{
    const double x = xf, xc = x < -110.0 ? -110.0 : (x > 100.0 ? 100.0 : x);
    const double kd = rint(xc * EXP_INV_STEP);
    const int k = (int)kd;
    double r = fma(-kd, EXP_STEP_HI, xc);
    r = fma(-kd, EXP_STEP_LO, r);
    double p = fma(r, 1.0 / 120, 1.0 / 24);
    p = fma(p, r, 1.0 / 6);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = p * r;
    const double t = EXPTAB[k & ((1 << EXPTAB_BITS) - 1)];
    const float y = (float)ldexp(fma(t, p, t), k >> EXPTAB_BITS);
    return xf != xf ? xf : y;
}

/* box_muller_dd: the pair in double, rounded once; qoff = 16 q (the angle table entry) */
static void box_muller_dd(float u, float v, uint32_t qoff, float* r0, float* r1)
{
    const uint32_t loff = (fbits(u) >> 13) & 0x1fff0u;
    const double* lt = BMLOGTAB[loff / 16 - BM_EMIN * 64];
    const double r = fma((double)u, lt[0], -1.0);
    double p = fma(r, 0.5, -2.0 / 3);
    p = fma(p, r, 1.0);
    p = fma(p, r, -2.0);
    const double x = fma(p, r, lt[1]);
    const double y = 1.0 / sqrt(x > 0x1p-126 ? x : 0x1p-126);
    const double t = x * y, h = fma(-t, y, 1.0), s = fma(t * h, 0.5, t);
    const double* st = BMSCTAB[qoff / 16];
    const double a = (double)fmaf(-(float)qoff, BM_ANGLE_STEP * 0.0625f, v);
    const double a2 = a * a;
    const double sa = a * fma(a2, -1.0 / 6, 1.0);
    const double ca = fma(a2, fma(a2, 1.0 / 24, -0.5), 1.0);
    const double sn = fma(st[0], ca, st[1] * sa);
    const double cs = fma(st[1], ca, -(st[0] * sa));
    *r0 = (float)(sn * s);
    *r1 = (float)(cs * s);
}

static int64_t ford(float x)
{
    const int32_t i = (int32_t)fbits(x);
    return i < 0 ? (int64_t)INT32_MIN - i : i;
}

/* |got - want| in ulps of want's binade, against a long double reference */
static double ulp_err(float got, long double want)
{
    const float w = (float)want;
    const float sp = nextafterf(fabsf(w), INFINITY) - fabsf(w);
    return (double)(fabsl((long double)got - want) / sp);
}

static int box_muller_sample(long n, int mrg)
{
    uint64_t st = 0x2545F4914F6CDD1Dull;
    long beyond = 0;
    int64_t maxd = 0;
    double maxe = 0.0, maxh = 0.0;
    for(long i = 0; i < n + 2; i++)
    {
        uint32_t w[2];
        for(int j = 0; j < 2; j++)
        {
            st ^= st << 13; st ^= st >> 7; st ^= st << 17;
            w[j] = (uint32_t)(st >> 11);
            if(mrg)
                w[j] = 1u + (uint32_t)((st >> 8) % 4294967087ull);
        }
        if(i == n) /* u = 1: the radius is -0 */
            w[0] = mrg ? 4294967087u : 0xffffffffu;
        if(i == n + 1) /* the smallest u */
            w[0] = mrg ? 1u : 0u;
        float u, v;
        if(mrg)
        {
            u = (float)((double)w[0] * 2.3283065498378288e-10);
            v = (float)((double)w[1] * 2.3283065498378288e-10) * 6.2831855f;
        }
        else
        {
            u = 2.3283064e-10f + (float)w[0] * 2.3283064e-10f;
            v = 1.46291807e-09f + (float)w[1] * 1.46291807e-09f;
        }
        float d0, d1;
        box_muller_dd(u, v, ((w[1] >> 20) + 8u) & ~15u, &d0, &d1);
        const float hs = sqrtf(-2.0f * logf(u));
        const float h[2] = {sinf(v) * hs, cosf(v) * hs};
        const long double S = sqrtl(-2.0L * logl((long double)u));
        const long double e[2] = {sinl((long double)v) * S, cosl((long double)v) * S};
        const float d[2] = {d0, d1};
        for(int j = 0; j < 2; j++)
        {
            if(signbit(d[j]) != signbit(h[j]))
                beyond++;
            int64_t dd = ford(d[j]) - ford(h[j]);
            dd = dd < 0 ? -dd : dd;
            beyond += dd > 2;
            maxd = dd > maxd ? dd : maxd;
            if(e[j] != 0.0L)
            {
                const double er = ulp_err(d[j], e[j]), eh = ulp_err(h[j], e[j]);
                maxe = er > maxe ? er : maxe;
                maxh = eh > maxh ? eh : maxh;
            }
        }
    }
    printf("bm %ld %ld %ld %ld %ld\n", 2 * (n + 2), beyond, (long)maxd, (long)(maxe * 1000), (long)(maxh * 1000));
    return 0;
}

int main(int argc, char** argv)
{
    if(argc > 2 && strcmp(argv[1], "bm") == 0)
        return box_muller_sample(atol(argv[2]), argc > 3 && strcmp(argv[3], "mrg") == 0);
    const int64_t stride = argc > 1 ? atoll(argv[1]) : 1;
    long nlog = 0, blog = 0, ntrig = 0, bsin = 0, bcos = 0, nexp = 0, bexp = 0;
    for(int64_t b = fbits(ldexpf(1.0f, -32)); b <= fbits(1.0f); b += stride)
    {
        const float x = bitsf((uint32_t)b);
        nlog++;
        blog += logf_tab(x) != (float)logl((long double)x);
    }
    for(int64_t b = 1; b <= fbits(7.0f); b += stride)
    {
        const float x = bitsf((uint32_t)b);
        float s, c;
        sincosf_acc(x, &s, &c);
        ntrig++;
        bsin += s != (float)sinl((long double)x);
        bcos += c != (float)cosl((long double)x);
    }
    for(int64_t b = 0; b <= 0xffffffffLL; b += stride)
    {
        const float x = bitsf((uint32_t)b);
        if(!(x > -105.0f && x < 90.0f))
            continue;
        nexp++;
        bexp += expf_acc(x) != (float)expl((long double)x);
    }
    printf("log %ld %ld\nsin %ld %ld\ncos %ld %ld\nexp %ld %ld\n", nlog, blog, ntrig, bsin, ntrig, bcos, nexp,
           bexp);
    /* special values the transforms can produce */
    printf("special %d %d %d %d\n", logf_tab(1.0f) == 0.0f, expf_acc(-200.0f) == 0.0f, isinf(expf_acc(200.0f)),
           isnan(expf_acc(NAN)));
    return 0;
}
