"""Write include/rocrand/rocrand_amd_mathtab.h: the tables of the float-argument logarithm
(Box-Muller radius), sine and cosine (Box-Muller angle) and exponential (log-normal
outputs) evaluated in double.

Logarithm, 64 entries: a positive normal float u is split as u = 2^k z with z in
[0x3f330000, 0x3fb30000) (about [0.699, 1.398)), so that k = 0 around u = 1.  Entry i
covers the z whose bit pattern minus 0x3f330000 has bits 17..22 equal to i; it holds
invc = 1/c for the centre c of that interval (rounded to double) and logc = -ln(invc)
rounded to double, with invc = 1 and logc = 0 exactly for the interval that contains 1.
Then ln u = k ln 2 + logc + log1p(z invc - 1), |z invc - 1| < 2^-7.

Sine and cosine, 64 entries: sin(k pi/32), cos(k pi/32) rounded to double (exact 0 and +-1
at multiples of pi/2); with x = q pi/32 + r, |r| <= pi/64, k = q mod 64,
sin x = sin(a_k) cos r + cos(a_k) sin r and cos x = cos(a_k) cos r - sin(a_k) sin r.  The
reduction constant pi/32 is split into a head with its low 8 bits clear (exact products
with q < 2^7) and the rounded remainder.

Exponential, 32 entries: 2^(j/32) rounded to double, j = 0..31; with x = (32 e + j) ln2/32
+ r, |r| <= ln2/64, exp x = 2^e 2^(j/32) exp(r).  The reduction constant ln2/32 is split
into a head with its low 16 bits clear (exact products with |32 e + j| < 2^16) and the
rounded remainder.

Box-Muller pair (evaluated in double and rounded once), logarithm, 33 x 64 entries: a
float u in [2^-32, 1] with bit pattern b has biased exponent e = b >> 23 in [95, 127] and
interval i = bits 17..22 of b (z = u 2^(127-e) in [1 + i/64, 1 + (i+1)/64)).  Entry
(e - 95) * 64 + i holds a = 2^(127-e) / c, c the centre of the interval, and
b = -2 ln(u / (a u)) = -2 ((e - 127) ln 2 + ln c) for the rounded a, so that
-2 ln u = b - 2 log1p(u a - 1), |u a - 1| < 2^-7.  The interval [1, 1 + 1/64) of e = 127
holds only u = 1: its entry is a = 1, b = -0, which gives -2 ln 1 = -0 (the reference's
-2.0f * logf(1.0f)).  The last interval below 1, u in [1 - 2^-7, 1), takes c = 1 (a = 1,
b = 0) instead of its centre: r = u - 1 is then exact and small with ln u, so -2 ln u
keeps its relative accuracy where it approaches 0.

Box-Muller sine and cosine, 257 entries: angles q P, q = 0..256, P = pi/128 rounded to a
float of 15 significant bits (q P and v - q P are exact in float for the angles v the
transform sees, |v - q P| < 0.0124); entry q holds sin(q P), cos(q P) rounded to double.

Double-precision logarithm, exponential and sine/cosine (the double transforms' steps,
each evaluated in double-double and rounded once):
  log, 128 entries: x = 2^k m with the high word of m in [0x3fe6a09e, 0x3ff6a09e)
    (m in [sqrt(1/2), sqrt(2))); entry i covers the m whose high word minus 0x3fe6a09e has
    bits 13..19 equal to i and holds inv = 1/c for the centre c of that interval (rounded
    to double; exactly 1 for the interval that contains 1) and -ln(inv) as a double-double.
    Then ln x = k ln 2 - ln(inv) + log1p(m inv - 1), |m inv - 1| < 2^-8.
  exp, 64 entries: 2^(j/64) as a double-double; x = (64 k + j) ln2/64 + r, |r| <= ln2/128.
  sin/cos, 129 entries: sin(q pi/64) and cos(q pi/64), q = 0..128, as double-doubles (exact
    0 and +-1 at multiples of pi/2); t = q pi/64 + r, |r| <= pi/128.
The reduction constants ln 2, ln 2/64 and pi/64 are split into heads with enough low bits
clear that their products with the integer multiples the arguments need are exact.

    python rocrand_amd/data/make_math_tables.py
"""
import os
import struct

import mpmath

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(os.path.dirname(os.path.dirname(HERE)), "include", "rocrand", "rocrand_amd_mathtab.h")
OFF = 0x3F330000
N_LOG = 64
N_EXP = 32
N_SC = 64


def f32(bits):
    return struct.unpack("<f", struct.pack("<I", bits))[0]


def log_entries():
    one = 0x3F800000
    out = []
    for i in range(N_LOG):
        lo_bits = OFF + (i << 17)
        hi_bits = OFF + ((i + 1) << 17)
        if lo_bits <= one < hi_bits:
            out.append((1.0, 0.0))
            continue
        c = (mpmath.mpf(f32(lo_bits)) + mpmath.mpf(f32(hi_bits))) / 2
        invc = float(1 / c)
        logc = float(-mpmath.log(mpmath.mpf(invc)))
        out.append((invc, logc))
    return out


def exp_entries():
    return [float(mpmath.power(2, mpmath.mpf(j) / N_EXP)) for j in range(N_EXP)]


def sincos_entries():
    out = []
    for k in range(N_SC):
        if k % (N_SC // 4) == 0:  # multiples of pi/2: exact values
            q = k // (N_SC // 4)
            out.append(([0.0, 1.0, 0.0, -1.0][q], [1.0, 0.0, -1.0, 0.0][q]))
            continue
        a = k * mpmath.pi / (N_SC // 2)
        out.append((float(mpmath.sin(a)), float(mpmath.cos(a))))
    return out


BM_EMIN = 95  # biased exponent of 2^-32
BM_EMAX = 127


def bm_log_entries():
    out = []
    for e in range(BM_EMIN, BM_EMAX + 1):
        scale = mpmath.power(2, BM_EMAX - e)
        for i in range(64):
            if e == BM_EMAX and i == 0:
                out.append((1.0, -0.0))
                continue
            if e == BM_EMAX - 1 and i == 63:
                # u in [1 - 2^-7, 1): c = 1, so r = u - 1 exactly and -2 ln u keeps its
                # relative accuracy as u approaches 1
                out.append((1.0, 0.0))
                continue
            lo = 1 + mpmath.mpf(i) / 64
            c = lo + mpmath.mpf(1) / 128
            a = float(scale / c)
            # -2 ln u = -2 ln(u a) + 2 ln a, exactly, for the rounded a
            b = float(2 * mpmath.log(mpmath.mpf(a)))
            out.append((a, b))
    return out


def bm_angle_step():
    m, e = mpmath.frexp(mpmath.pi / 128)
    return float(mpmath.nint(m * 2**15) * mpmath.power(2, e - 15))


def bm_sincos_entries():
    p = mpmath.mpf(bm_angle_step())
    return [(float(mpmath.sin(q * p)), float(mpmath.cos(q * p))) for q in range(257)]


DLOG_OFF = 0x3FE6A09E


def f64_from_high(hw):
    return struct.unpack("<d", struct.pack("<Q", hw << 32))[0]


def dd(x):
    hi = float(x)
    return hi, float(x - mpmath.mpf(hi))


def dlog_entries():
    out = []
    for i in range(128):
        lo = mpmath.mpf(f64_from_high(DLOG_OFF + i * 0x2000))
        hi = mpmath.mpf(f64_from_high(DLOG_OFF + (i + 1) * 0x2000))
        if lo <= 1 < hi:
            out.append((1.0, 0.0, 0.0))
            continue
        inv = float(2 / (lo + hi))
        t_hi, t_lo = dd(-mpmath.log(mpmath.mpf(inv)))
        out.append((inv, t_hi, t_lo))
    return out


def dexp_entries():
    return [dd(mpmath.power(2, mpmath.mpf(j) / 64)) for j in range(64)]


def dsincos_entries():
    out = []
    for q in range(129):
        if q % 32 == 0:
            k = q // 32
            out.append(([0.0, 1.0, 0.0, -1.0, 0.0][k], 0.0, [1.0, 0.0, -1.0, 0.0, 1.0][k], 0.0))
            continue
        a = q * mpmath.pi / 64
        s_hi, s_lo = dd(mpmath.sin(a))
        c_hi, c_lo = dd(mpmath.cos(a))
        out.append((s_hi, s_lo, c_hi, c_lo))
    return out


def split3(c, clear_bits):
    """c = h + m + l with h and m rounded to 53 - clear_bits significant bits."""
    parts = []
    rest = mpmath.mpf(c)
    for _ in range(2):
        m, e = mpmath.frexp(rest)
        keep = 53 - clear_bits
        v = mpmath.nint(m * 2**keep) * mpmath.power(2, e - keep)
        parts.append(float(v))
        rest = rest - v
    parts.append(float(rest))
    return parts


def split(c, clear_bits):
    head = float(c)
    bits = struct.unpack("<Q", struct.pack("<d", head))[0] & ~((1 << clear_bits) - 1)
    head = struct.unpack("<d", struct.pack("<Q", bits))[0]
    return head, float(c - mpmath.mpf(head))


def ln2_split():
    head, tail = split(mpmath.log(2) / N_EXP, 16)
    return head, tail, float(N_EXP / mpmath.log(2))


def main():
    mpmath.mp.prec = 200
    log_rows = "\n".join(f"    {{{a.hex()}, {b.hex()}}}," for a, b in log_entries())
    exp_rows = "\n".join(f"    {a.hex()}," for a in exp_entries())
    head, tail, inv = ln2_split()
    sc_rows = "\n".join(f"    {{{a.hex()}, {b.hex()}}}," for a, b in sincos_entries())
    sc_head, sc_tail = split(mpmath.pi / (N_SC // 2), 8)
    sc_inv = float((N_SC // 2) / mpmath.pi)
    dlog_rows = "\n".join(f"    {{{a.hex()}, {b.hex()}, {c.hex()}, 0.0}}," for a, b, c in dlog_entries())
    dexp_rows = "\n".join(f"    {{{a.hex()}, {b.hex()}}}," for a, b in dexp_entries())
    dsc_rows = "\n".join(f"    {{{a.hex()}, {b.hex()}, {c.hex()}, {d.hex()}}}," for a, b, c, d in dsincos_entries())
    ln2 = split3(mpmath.log(2), 11)            # k ln2 exact for |k| < 2^11
    l64 = split3(mpmath.log(2) / 64, 17)       # q ln2/64 exact for |q| < 2^17
    p64 = split3(mpmath.pi / 64, 8)            # q pi/64 exact for q < 2^8
    bm_log_rows = "\n".join(f"    {{{a.hex()}, {b.hex()}}}," for a, b in bm_log_entries())
    bm_sc_rows = "\n".join(f"    {{{a.hex()}, {b.hex()}}}," for a, b in bm_sincos_entries())
    bm_p = bm_angle_step()
    text = f"""// rocrand_amd_mathtab.h -- generated by rocrand_amd/data/make_math_tables.py; do not edit.
// Tables of the double-evaluated float logarithm, sine/cosine and exponential (rocrand_amd_math.h).
#pragma once

#include <hip/hip_runtime.h>

// Device code reads the tables from constant memory; host code (the double functions are
// also host functions) from ordinary host arrays with the same values.
#if defined(__HIP_DEVICE_COMPILE__)
#define RR_MATH_TABLE static __constant__
#else
#define RR_MATH_TABLE static
#endif

namespace rocrand_amd_engine
{{
constexpr uint32_t RR_LOGTAB_OFF  = 0x{OFF:08x}u;
constexpr int      RR_LOGTAB_BITS = 6;
// {{1/c, -ln(1/c)}}
RR_MATH_TABLE const double2 rr_logtab[{N_LOG}] = {{
{log_rows}
}};

constexpr int    RR_EXPTAB_BITS  = 5;
constexpr double RR_EXP_INV_STEP = {inv.hex()}; // {N_EXP} / ln 2
constexpr double RR_EXP_STEP_HI  = {head.hex()}; // ln 2 / {N_EXP}, low 16 bits clear
constexpr double RR_EXP_STEP_LO  = {tail.hex()}; // ln 2 / {N_EXP} - RR_EXP_STEP_HI
// 2^(j/{N_EXP})
RR_MATH_TABLE const double rr_exptab[{N_EXP}] = {{
{exp_rows}
}};

constexpr int    RR_SCTAB_BITS  = 6;
constexpr double RR_SC_INV_STEP = {sc_inv.hex()}; // {N_SC // 2} / pi
constexpr double RR_SC_STEP_HI  = {sc_head.hex()}; // pi / {N_SC // 2}, low 8 bits clear
constexpr double RR_SC_STEP_LO  = {sc_tail.hex()}; // pi / {N_SC // 2} - RR_SC_STEP_HI
// {{sin(k pi/{N_SC // 2}), cos(k pi/{N_SC // 2})}}
RR_MATH_TABLE const double2 rr_sctab[{N_SC}] = {{
{sc_rows}
}};

// Box-Muller pair: {{2^(127-e)/c, 2 ln(2^(127-e)/c)}} for biased exponents e = {BM_EMIN}..{BM_EMAX}
constexpr uint32_t RR_BM_EMIN = {BM_EMIN}u;
RR_MATH_TABLE const double2 rr_bm_logtab[{(BM_EMAX - BM_EMIN + 1) * 64}] = {{
{bm_log_rows}
}};
constexpr float RR_BM_ANGLE_STEP = {bm_p.hex()}f; // pi/128 to 15 bits
// {{sin(q RR_BM_ANGLE_STEP), cos(q RR_BM_ANGLE_STEP)}}, q = 0..256
RR_MATH_TABLE const double2 rr_bm_sctab[257] = {{
{bm_sc_rows}
}};

// Double-precision log: {{1/c, -ln(1/c) (head, tail), 0}} for m in [sqrt(1/2), sqrt(2))
constexpr uint32_t RR_DLOG_OFF = 0x{DLOG_OFF:08x}u;
RR_MATH_TABLE const double rr_dlog_tab[128][4] = {{
{dlog_rows}
}};
constexpr double RR_DLN2_HI = {ln2[0].hex()};
constexpr double RR_DLN2_MI = {ln2[1].hex()};
constexpr double RR_DLN2_LO = {ln2[2].hex()};
// Double-precision exp: 2^(j/64) (head, tail)
RR_MATH_TABLE const double2 rr_dexp_tab[64] = {{
{dexp_rows}
}};
constexpr double RR_DEXP_INV_STEP = {float(64 / mpmath.log(2)).hex()}; // 64 / ln 2
constexpr double RR_DEXP_STEP_HI  = {l64[0].hex()};
constexpr double RR_DEXP_STEP_MI  = {l64[1].hex()};
constexpr double RR_DEXP_STEP_LO  = {l64[2].hex()};
// Double-precision sin/cos: sin(q pi/64), cos(q pi/64) (heads, tails), q = 0..128
RR_MATH_TABLE const double rr_dsc_tab[129][4] = {{
{dsc_rows}
}};
constexpr double RR_DSC_INV_STEP = {float(64 / mpmath.pi).hex()}; // 64 / pi
constexpr double RR_DSC_STEP_HI  = {p64[0].hex()};
constexpr double RR_DSC_STEP_MI  = {p64[1].hex()};
constexpr double RR_DSC_STEP_LO  = {p64[2].hex()};
}} // namespace rocrand_amd_engine
"""
    with open(OUT, "w") as f:
        f.write(text)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
