/* Host twin of the double functions of include/rocrand/rocrand_amd_math.h (log_cr, exp_cr,
 * sincos_cr): the same source compiled for the host, checked against quad precision
 * (libquadmath).  The device compiles the same code; fma is the correctly rounded fused
 * multiply-add on both.
 *
 * usage: dmath_twin N
 * For each function, N pseudo-random arguments of the ranges the transforms use plus a
 * wider range; prints per function the count of results that are not the correctly
 * rounded double of the quad-precision value and the largest error in ulps:
 *   NAME COUNT MISROUNDED MAX_ERR_MICROULP
 * then special values: log 1, log 0, log of the smallest subnormal, exp 0, exp above the
 * overflow threshold, exp below the underflow threshold, sin pi, cos pi, sin 2pi.
 */
#include <rocrand/rocrand_amd_math.h>

#if !defined(__HIP_DEVICE_COMPILE__)
#include <math.h>
#include <quadmath.h>
#include <stdio.h>
#include <stdlib.h>

using namespace rocrand_amd_engine;

static uint64_t g_state = 0x9E3779B97F4A7C15ull;
static uint64_t next64()
{
    g_state ^= g_state << 13;
    g_state ^= g_state >> 7;
    g_state ^= g_state << 17;
    return g_state;
}
static double unit53() { return (double)(next64() >> 11) * 0x1p-53; }

struct tally
{
    const char* name;
    long        n = 0, bad = 0;
    double      worst = 0.0;
    void add(double got, __float128 exact)
    {
        n++;
        const double e  = (double)exact;
        bad += got != e;
        const double sp = nextafter(fabs(e), INFINITY) - fabs(e);
        const double u  = (double)(fabsq((__float128)got - exact) / (__float128)sp);
        worst           = u > worst ? u : worst;
    }
    void print() const { printf("%s %ld %ld %ld\n", name, n, bad, (long)(worst * 1e6)); }
};

int main(int argc, char** argv)
{
    const long n = argc > 1 ? atol(argv[1]) : 100000;
    tally      lg{"log"}, lw{"log_wide"}, ex{"exp"}, ew{"exp_wide"}, sn{"sin"}, cs{"cos"};
    for(long i = 0; i < n; i++)
    {
        // Box-Muller uniform 2^-53 + k 2^-53 (rocrand_normal.h:76-96)
        const double u = 0x1p-53 + (double)(next64() >> 11) * 0x1p-53;
        lg.add(log_cr(u), logq((__float128)u));
        // any positive normal double down to 2^-1000 (the inverse error function's argument)
        const double x = ldexp(1.0 + unit53(), -(int)(next64() % 1000));
        lw.add(log_cr(x), logq((__float128)x));
        // log-normal exponents: mean + stddev * v
        const double z = (unit53() - 0.5) * 80.0;
        ex.add(exp_cr(z), expq((__float128)z));
        const double zw = -708.0 + unit53() * (709.7 + 708.0);
        ew.add(exp_cr(zw), expq((__float128)zw));
        // Box-Muller angle w pi, w = 2^-52 + k 2^-52
        const double w = 0x1p-52 + (double)(next64() >> 11) * 0x1p-52;
        const double t = w * 3.1415926535897932;
        double       s, c;
        sincos_cr(t, s, c);
        sn.add(s, sinq((__float128)t));
        cs.add(c, cosq((__float128)t));
    }
    lg.print();
    lw.print();
    ex.print();
    ew.print();
    sn.print();
    cs.print();
    double s1, c1, s2, c2;
    sincos_cr(3.1415926535897932, s1, c1);
    sincos_cr(2.0 * 3.1415926535897932, s2, c2);
    printf("special %d %d %d %d %d %d %d %d %d\n", log_cr(1.0) == 0.0, log_cr(0.0) == -INFINITY,
           log_cr(0x1p-1074) == (double)logq((__float128)0x1p-1074), exp_cr(0.0) == 1.0, isinf(exp_cr(710.0)),
           exp_cr(-746.0) == 0.0, s1 == (double)sinq((__float128)3.1415926535897932), c1 == -1.0,
           s2 == (double)sinq((__float128)(2.0 * 3.1415926535897932)));
    return 0;
}
#endif
