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

int main(int argc, char** argv)
{
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
