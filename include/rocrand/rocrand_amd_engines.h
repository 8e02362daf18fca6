// rocrand_amd_engines.h -- engine arithmetic shared by the library's host logic, its
// gfx950 kernels and the device API (rocrand_kernel.h).
//
// Each engine reproduces the output stream of the reference engine it names
// (file:line citations into the reference tree).  Counter-based engines are
// expressed as pure block functions plus 128-bit word-position arithmetic; the
// stateful engines (MRG32k3a, MRG31k3p, XORWOW, LFSR113) keep their state in
// plain registers and are stored structure-of-arrays in HBM by the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <stdint.h>

// Floating-point expressions in these headers round every operation separately (no
// fused multiply-add contraction), whatever -ffp-contract the including code uses; the
// pragma applies to the function body it opens.
#define ROCRAND_AMD_FP_EXACT _Pragma("clang fp contract(off)")

namespace rocrand_amd_engine
{

// ----------------------------------------------------------------------------
// Philox4x32-10 (rocrand_philox4x32_10.h:58-63 constants, :279-315 rounds).
// ----------------------------------------------------------------------------
constexpr uint32_t PHILOX_M0 = 0xD2511F53u;
constexpr uint32_t PHILOX_M1 = 0xCD9E8D57u;
constexpr uint32_t PHILOX_W0 = 0x9E3779B9u;
constexpr uint32_t PHILOX_W1 = 0xBB67AE85u;

struct w4
{
    uint32_t v[4];
};
struct q4
{
    uint64_t v[4];
};

// a ^ b ^ c in one instruction on gfx950 (v_bitop3_b32, truth table 0x96).
__host__ __device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c)
{
#if defined(__HIP_DEVICE_COMPILE__) && __has_builtin(__builtin_amdgcn_bitop3_b32)
    return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
#else
    return a ^ b ^ c;
#endif
}

__host__ __device__ __forceinline__ void philox_block(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                                      uint32_t k0, uint32_t k1, uint32_t out[4])
{
    uint32_t x = c0, y = c1, z = c2, w = c3;
#pragma unroll
    for(int r = 0; r < 10; r++)
    {
        // One 32x32->64 multiply per lane word: v_mad_u64_u32 on gfx950.
        const uint64_t p0 = (uint64_t)PHILOX_M0 * x;
        const uint64_t p1 = (uint64_t)PHILOX_M1 * z;
        const uint32_t nx = xor3((uint32_t)(p1 >> 32), y, k0);
        const uint32_t nz = xor3((uint32_t)(p0 >> 32), w, k1);
        y                 = (uint32_t)p1;
        w                 = (uint32_t)p0;
        x                 = nx;
        z                 = nz;
        k0 += PHILOX_W0;
        k1 += PHILOX_W1;
    }
    out[0] = x;
    out[1] = y;
    out[2] = z;
    out[3] = w;
}

// ----------------------------------------------------------------------------
// ThreeFry-{2,4}x{32,64}-20 (rocrand_threefry4_impl.h:178-250,
// rocrand_threefry2_impl.h threefry_rounds; rotation tables :60-92).
// ----------------------------------------------------------------------------
template<class W>
__host__ __device__ __forceinline__ W rotl(W x, int d)
{
    constexpr int B = sizeof(W) * 8;
    return (x << (d & (B - 1))) | (x >> ((B - d) & (B - 1)));
}
// A 64-bit rotation as two 32-bit funnel shifts of its halves (one v_alignbit_b32 each
// for the constant amounts of the rounds, where the 64-bit form takes two 64-bit shifts
// and two ORs); the same value as the generic form.
template<>
__host__ __device__ __forceinline__ uint64_t rotl<uint64_t>(uint64_t x, int d)
{
    d &= 63;
    uint32_t hi = (uint32_t)(x >> 32), lo = (uint32_t)x;
    if(d >= 32)
    {
        const uint32_t t = hi;
        hi               = lo;
        lo               = t;
        d -= 32;
    }
    if(d == 0)
        return ((uint64_t)hi << 32) | lo;
    const uint32_t nh = (hi << d) | (lo >> (32 - d));
    const uint32_t nl = (lo << d) | (hi >> (32 - d));
    return ((uint64_t)nh << 32) | nl;
}

template<class W>
__host__ __device__ constexpr W ks_parity()
{
    return sizeof(W) == 4 ? (W)0x1BD11BDAu : (W)0x1BD11BDAA9FC1A22ull;
}

template<class W>
__host__ __device__ __forceinline__ int tf4_rot(int r, int i)
{
    if(sizeof(W) == 4)
    {
        constexpr int R[8][2] = {{10, 26}, {11, 21}, {13, 27}, {23, 5}, {6, 20}, {17, 11}, {25, 10}, {18, 20}};
        return R[r & 7][i];
    }
    constexpr int R[8][2] = {{14, 16}, {52, 57}, {23, 40}, {5, 37}, {25, 33}, {46, 12}, {58, 22}, {32, 32}};
    return R[r & 7][i];
}
template<class W>
__host__ __device__ __forceinline__ int tf2_rot(int r)
{
    if(sizeof(W) == 4)
    {
        constexpr int R[8] = {13, 15, 26, 6, 17, 29, 16, 24};
        return R[r & 7];
    }
    constexpr int R[8] = {16, 42, 12, 31, 16, 32, 24, 21};
    return R[r & 7];
}

template<class W>
__host__ __device__ __forceinline__ void threefry4_block(const W c[4], const W k[4], W out[4])
{
    W ks[5];
    ks[4] = ks_parity<W>();
#pragma unroll
    for(int i = 0; i < 4; i++)
    {
        ks[i] = k[i];
        ks[4] ^= k[i];
    }
    W X0 = c[0] + ks[0], X1 = c[1] + ks[1], X2 = c[2] + ks[2], X3 = c[3] + ks[3];
#pragma unroll
    for(int r = 0; r < 20; r++)
    {
        if((r & 2) == 0)
        {
            X0 += X1;
            X1 = rotl<W>(X1, tf4_rot<W>(r, 0));
            X1 ^= X0;
            X2 += X3;
            X3 = rotl<W>(X3, tf4_rot<W>(r, 1));
            X3 ^= X2;
        }
        else
        {
            X0 += X3;
            X3 = rotl<W>(X3, tf4_rot<W>(r, 0));
            X3 ^= X0;
            X2 += X1;
            X1 = rotl<W>(X1, tf4_rot<W>(r, 1));
            X1 ^= X2;
        }
        if((r & 3) == 3)
        {
            const int inj = r / 4;
            X0 += ks[(1 + inj) % 5];
            X1 += ks[(2 + inj) % 5];
            X2 += ks[(3 + inj) % 5];
            X3 += ks[(4 + inj) % 5];
            X3 += (W)(1 + inj);
        }
    }
    out[0] = X0;
    out[1] = X1;
    out[2] = X2;
    out[3] = X3;
}

template<class W>
__host__ __device__ __forceinline__ void threefry2_block(const W c[2], const W k[2], W out[2])
{
    W ks[3];
    ks[2] = ks_parity<W>() ^ k[0] ^ k[1];
    ks[0] = k[0];
    ks[1] = k[1];
    W X0 = c[0] + ks[0], X1 = c[1] + ks[1];
#pragma unroll
    for(int r = 0; r < 20; r++)
    {
        X0 += X1;
        X1 = rotl<W>(X1, tf2_rot<W>(r));
        X1 ^= X0;
        if((r & 3) == 3)
        {
            const int inj = r / 4;
            X0 += ks[(1 + inj) % 3];
            X1 += ks[(2 + inj) % 3];
            X1 += (W)(1 + inj);
        }
    }
    out[0] = X0;
    out[1] = X1;
}

// ----------------------------------------------------------------------------
// Counter-based engine descriptors.  A "word position" is the index of a word
// in the engine's output stream: counter = pos / DIM, substate = pos % DIM, with
// the counter a DIM-word little-endian integer (128-bit for the 4x32 engines).
// ----------------------------------------------------------------------------
enum class ctr_kind : int
{
    philox4x32   = 0,
    threefry2x32 = 1,
    threefry2x64 = 2,
    threefry4x32 = 3,
    threefry4x64 = 4
};

// Key and base counter of one counter-based engine (kernel argument).
struct ctr_key
{
    uint64_t k[4];
};
struct ctr_pos
{
    uint64_t c[4]; // counter words (32-bit values in the low half for x32 engines)
    uint32_t sub;  // substate within the block
};

template<ctr_kind K>
struct ctr_traits;

template<>
struct ctr_traits<ctr_kind::philox4x32>
{
    using word                = uint32_t;
    static constexpr int DIM  = 4;
    __host__ __device__ __forceinline__ static void block(const word c[4], const ctr_key& k, word out[4])
    {
        philox_block(c[0], c[1], c[2], c[3], (uint32_t)k.k[0], (uint32_t)k.k[1], out);
    }
};
template<>
struct ctr_traits<ctr_kind::threefry4x32>
{
    using word               = uint32_t;
    static constexpr int DIM = 4;
    __host__ __device__ __forceinline__ static void block(const word c[4], const ctr_key& k, word out[4])
    {
        const word kk[4] = {(word)k.k[0], (word)k.k[1], (word)k.k[2], (word)k.k[3]};
        threefry4_block<word>(c, kk, out);
    }
};
template<>
struct ctr_traits<ctr_kind::threefry4x64>
{
    using word               = uint64_t;
    static constexpr int DIM = 4;
    __host__ __device__ __forceinline__ static void block(const word c[4], const ctr_key& k, word out[4])
    {
        threefry4_block<word>(c, k.k, out);
    }
};
template<>
struct ctr_traits<ctr_kind::threefry2x32>
{
    using word               = uint32_t;
    static constexpr int DIM = 2;
    __host__ __device__ __forceinline__ static void block(const word c[2], const ctr_key& k, word out[2])
    {
        const word kk[2] = {(word)k.k[0], (word)k.k[1]};
        threefry2_block<word>(c, kk, out);
    }
};
template<>
struct ctr_traits<ctr_kind::threefry2x64>
{
    using word               = uint64_t;
    static constexpr int DIM = 2;
    __host__ __device__ __forceinline__ static void block(const word c[2], const ctr_key& k, word out[2])
    {
        threefry2_block<word>(c, k.k, out);
    }
};

// counter += v (64-bit), full carry propagation across the DIM words.
template<class W, int DIM>
__host__ __device__ __forceinline__ void ctr_add(W c[DIM], uint64_t v)
{
    if constexpr(sizeof(W) == 4)
    {
        const uint64_t lo = (uint64_t)c[0] + (uint32_t)v;
        c[0]              = (W)lo;
        const uint64_t hi = (uint64_t)c[1] + (uint32_t)(v >> 32) + (lo >> 32);
        c[1]              = (W)hi;
        if constexpr(DIM == 4)
        {
            const uint64_t z = (uint64_t)c[2] + (hi >> 32);
            c[2]             = (W)z;
            c[3] += (W)(z >> 32);
        }
    }
    else
    {
        const W old = c[0];
        c[0] += v;
        const W carry = c[0] < old ? 1 : 0;
        c[1] += carry;
        if constexpr(DIM == 4)
        {
            const W c1 = (c[1] == 0 && carry) ? 1 : 0;
            c[2] += c1;
            c[3] += (c[2] == 0 && c1) ? 1 : 0;
        }
    }
}

// ----------------------------------------------------------------------------
// MRG32k3a (rocrand_mrg32k3a.h:38-47 constants, :160-193 next).
// State order per component: g[0] oldest .. g[2] newest.
// ----------------------------------------------------------------------------
constexpr uint64_t MRG32_M1   = 4294967087ull;
constexpr uint64_t MRG32_M2   = 4294944443ull;
constexpr uint64_t MRG32_A12  = 1403580ull;
constexpr uint64_t MRG32_A13N = 810728ull;
constexpr uint64_t MRG32_A21  = 527612ull;
constexpr uint64_t MRG32_A23N = 1370589ull;
constexpr double   MRG32_NORM_DOUBLE = 2.3283065498378288e-10;
constexpr double   MRG32_UINT_NORM   = 1.000000048661607;

// MRG31k3p (rocrand_mrg31k3p.h:27-37, next()).  State order: [0] newest.
constexpr uint32_t MRG31_M1          = 2147483647u;
constexpr uint32_t MRG31_M2          = 2147462579u;
constexpr double   MRG31_NORM_DOUBLE = 4.656612875245796923e-10;
constexpr double   MRG31_UINT_NORM   = 2.000000001396983862;

struct mrg_state
{
    uint32_t g1[3], g2[3];
};

// Residues for the recurrence's range (p < 2^54), rocrand_mrg32k3a.h:333-363:
// 2^32 = 209 (mod m1) and 2^32 = 22853 (mod m2); one fold (m1) or two (m2) and a
// single conditional subtract leave the canonical residue for that range.
__host__ __device__ __forceinline__ uint32_t mrg32_reduce_m1(uint64_t p)
{
    p = 209ull * (p >> 32) + (p & 0xffffffffull);
    return (uint32_t)(p >= MRG32_M1 ? p - MRG32_M1 : p);
}
__host__ __device__ __forceinline__ uint32_t mrg32_reduce_m2(uint64_t p)
{
    p = 22853ull * (p >> 32) + (p & 0xffffffffull);
    p = 22853ull * (p >> 32) + (p & 0xffffffffull);
    return (uint32_t)(p >= MRG32_M2 ? p - MRG32_M2 : p);
}

// (a mod 2^24) * (b mod 2^24) mod 2^32: one v_mad_u32_u24 together with an addition.
__host__ __device__ __forceinline__ uint32_t umul24(uint32_t a, uint32_t b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul24(a, b);
#else
    return (a & 0xffffffu) * (b & 0xffffffu);
#endif
}

// x mod m for m = 2^32 - c and x < 2^33 - c (so hi = x >> 32 is 0 or 1), in three
// 32-bit operations and no lane masks: v = lo + c·hi cannot wrap and is x - m·hi;
// v + c wraps to v - m exactly when v >= m, so the smaller of v and v + c (mod 2^32)
// is the residue.
__host__ __device__ __forceinline__ uint32_t mrg32_fold_final(uint64_t x, uint32_t c)
{
    const uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
    const uint32_t v  = lo + umul24(hi, c);
    const uint32_t w  = v + c;
    return w < v ? w : v;
}

// p1 - p2 + (p1 <= p2 ? m : 0) (rocrand_mrg32k3a.h:190-192, rocrand_mrg31k3p.h next()) from
// one subtraction's borrow: e = p2 - p1 wraps exactly when p1 > p2, and then 0 - e is
// p1 - p2; otherwise m - e is p1 - p2 + m (m itself when p1 == p2).
__host__ __device__ __forceinline__ uint32_t mrg_combine(uint32_t p1, uint32_t p2, uint32_t m)
{
    uint32_t   e;
    const bool gt = __builtin_sub_overflow(p2, p1, &e);
    return (gt ? 0u : m) - e;
}

__host__ __device__ __forceinline__ uint32_t mrg32k3a_next(mrg_state& s)
{
    // t1 < (a12 + a13n) m1 < 2214308 * 2^32, so one fold by 2^32 = 209 (mod m1) leaves
    // x < 2^32 + 462790164; t2 < (a21 + a23n) m2 < 1898201 * 2^32, so two folds by
    // 2^32 = 22853 (mod m2) leave x < 2^32 + 251384: both within mrg32_fold_final's
    // range, and the same residues as the reference's reductions
    const uint64_t t1 = (uint64_t)(uint32_t)MRG32_A12 * s.g1[1]
                        + (uint64_t)(uint32_t)MRG32_A13N * ((uint32_t)MRG32_M1 - s.g1[0]);
    const uint32_t p1 = mrg32_fold_final(209ull * (t1 >> 32) + (uint32_t)t1, 209u);
    s.g1[0]           = s.g1[1];
    s.g1[1]           = s.g1[2];
    s.g1[2]           = p1;
    const uint64_t t2 = (uint64_t)(uint32_t)MRG32_A21 * s.g2[2]
                        + (uint64_t)(uint32_t)MRG32_A23N * ((uint32_t)MRG32_M2 - s.g2[0]);
    const uint64_t f2 = 22853ull * (t2 >> 32) + (uint32_t)t2;
    const uint32_t p2 = mrg32_fold_final(22853ull * (f2 >> 32) + (uint32_t)f2, 22853u);
    s.g2[0]           = s.g2[1];
    s.g2[1]           = s.g2[2];
    s.g2[2]           = p2;
    return mrg_combine(p1, p2, (uint32_t)MRG32_M1);
}

// Seeded state words {g1[0..2], g2[0..2]} (rocrand_mrg32k3a.h:103-120, :318-363): the
// seeding products wrap at 64 bits and reduce with single folds, reproduced literally.
__host__ __device__ __forceinline__ uint64_t mrg32_seed_mod(uint64_t p, bool m2)
{
    const uint64_t c = m2 ? 22853ull : 209ull;
    const uint64_t m = m2 ? MRG32_M2 : MRG32_M1;
    p                = c * (p >> 32) + (p & 0xffffffffull);
    if(m2)
        p = c * (p >> 32) + (p & 0xffffffffull);
    return p >= m ? p - m : p;
}
__host__ __device__ __forceinline__ uint64_t mrg32_seed_mul(uint32_t i, uint64_t j, bool m2)
{
    const uint64_t hi = i / 131072u;
    const uint64_t lo = i - hi * 131072u;
    return mrg32_seed_mod(mrg32_seed_mod(hi * j, m2) * 131072u + mrg32_seed_mod(lo * j, m2), m2);
}
__host__ __device__ __forceinline__ void mrg32k3a_seed_words(uint64_t s, uint32_t w[6])
{
    const uint32_t x = (uint32_t)s ^ 0x55555555u;
    const uint32_t y = (uint32_t)((s >> 32) ^ 0xAAAAAAAAu);
    w[0]             = (uint32_t)mrg32_seed_mul(x, s, false);
    w[1]             = (uint32_t)mrg32_seed_mul(y, s, false);
    w[2]             = (uint32_t)mrg32_seed_mul(x, s, false);
    w[3]             = (uint32_t)mrg32_seed_mul(y, s, true);
    w[4]             = (uint32_t)mrg32_seed_mul(x, s, true);
    w[5]             = (uint32_t)mrg32_seed_mul(y, s, true);
}
// rocrand_mrg31k3p.h seed(): plain remainders of the wrapped 64-bit products.
__host__ __device__ __forceinline__ void mrg31k3p_seed_words(uint64_t s, uint32_t w[6])
{
    const uint32_t x = (uint32_t)(s ^ 0x55555555u);
    const uint32_t y = (uint32_t)((s >> 32) ^ 0xAAAAAAAAu);
    w[0]             = (uint32_t)(((uint64_t)x * s) % MRG31_M1);
    w[1]             = (uint32_t)(((uint64_t)y * s) % MRG31_M1);
    w[2]             = (uint32_t)(((uint64_t)x * s) % MRG31_M1);
    w[3]             = (uint32_t)(((uint64_t)y * s) % MRG31_M2);
    w[4]             = (uint32_t)(((uint64_t)x * s) % MRG31_M2);
    w[5]             = (uint32_t)(((uint64_t)y * s) % MRG31_M2);
}

// x - m if x >= m, else x (x < 2^32): the smaller of x and x - m (mod 2^32).
__host__ __device__ __forceinline__ uint32_t sub_if_ge(uint32_t x, uint32_t m)
{
    const uint32_t y = x - m;
    return y < x ? y : x;
}

__host__ __device__ __forceinline__ uint32_t mrg31k3p_next(mrg_state& s)
{
    uint32_t tmp = (((s.g1[1] & 511u) << 22) + (s.g1[1] >> 9)) + (((s.g1[2] & 16777215u) << 7) + (s.g1[2] >> 24));
    tmp          = sub_if_ge(tmp, MRG31_M1);
    tmp += s.g1[2];
    tmp     = sub_if_ge(tmp, MRG31_M1);
    s.g1[2] = s.g1[1];
    s.g1[1] = s.g1[0];
    s.g1[0] = tmp;
    tmp     = (((s.g2[0] & 65535u) << 15) + 21069u * (s.g2[0] >> 16));
    tmp     = sub_if_ge(tmp, MRG31_M2);
    tmp += ((s.g2[2] & 65535u) << 15);
    tmp = sub_if_ge(tmp, MRG31_M2);
    tmp += 21069u * (s.g2[2] >> 16);
    tmp = sub_if_ge(tmp, MRG31_M2);
    tmp += s.g2[2];
    tmp = sub_if_ge(tmp, MRG31_M2);
    s.g2[2] = s.g2[1];
    s.g2[1] = s.g2[0];
    s.g2[0] = tmp;
    return mrg_combine(s.g1[0], s.g2[0], MRG31_M1);
}

// x mod (2^K - C) for any 64-bit x: three folds by 2^K = C leave x < 2m for the four
// moduli used here (m1, m2 of MRG32k3a and MRG31k3p), then one conditional subtraction.
template<uint32_t K, uint32_t C>
__host__ __device__ __forceinline__ uint32_t fold_mod(uint64_t x)
{
    constexpr uint64_t mask = (1ull << K) - 1;
    constexpr uint64_t m    = (1ull << K) - C;
    x = (x >> K) * C + (x & mask);
    x = (x >> K) * C + (x & mask);
    x = (x >> K) * C + (x & mask);
    return (uint32_t)(x >= m ? x - m : x);
}

// One fold by 2^K = C (mod 2^K - C): a congruent value below 2^48 for the operands of
// mat3_vec_fold.
template<uint32_t K, uint32_t C>
__host__ __device__ __forceinline__ uint64_t fold_once(uint64_t x)
{
    return (x >> K) * C + (x & ((1ull << K) - 1));
}

// 3x3 matrix (row-major, entries < m) times state vector (any 32-bit words), mod
// m = 2^K - C.  Each row accumulates with one fold between products: an entry below m
// times a 32-bit word plus a folded partial stays below 2^64 for the four moduli used
// here, and fold_mod returns the canonical residue of the last sum.
template<uint32_t K, uint32_t C>
__host__ __device__ __forceinline__ void mat3_vec_fold(const uint32_t* A, uint32_t* s)
{
    uint32_t x[3];
#pragma unroll
    for(int r = 0; r < 3; r++)
    {
        uint64_t acc = fold_once<K, C>((uint64_t)A[3 * r + 0] * s[0]);
        acc          = fold_once<K, C>((uint64_t)A[3 * r + 1] * s[1] + acc);
        x[r]         = fold_mod<K, C>((uint64_t)A[3 * r + 2] * s[2] + acc);
    }
    s[0] = x[0];
    s[1] = x[1];
    s[2] = x[2];
}

// 3x3 matrix (row-major, entries < m) times state vector, mod m.
__host__ __device__ __forceinline__ void mat3_vec_mod(const uint32_t* A, uint32_t* s, uint64_t m)
{
    uint32_t x[3];
#pragma unroll
    for(int r = 0; r < 3; r++)
    {
        uint64_t acc = 0;
#pragma unroll
        for(int c = 0; c < 3; c++)
            acc += ((uint64_t)A[3 * r + c] * s[c]) % m;
        x[r] = (uint32_t)(acc % m);
    }
    s[0] = x[0];
    s[1] = x[1];
    s[2] = x[2];
}

// ----------------------------------------------------------------------------
// XORWOW (rocrand_xorwow.h:100-130 seeding, :165-177 next).
// ----------------------------------------------------------------------------
struct xorwow_state
{
    uint32_t d;
    uint32_t x[5];
};

__host__ __device__ __forceinline__ uint32_t xorwow_next(xorwow_state& s)
{
    const uint32_t t = s.x[0] ^ (s.x[0] >> 2);
    s.x[0]           = s.x[1];
    s.x[1]           = s.x[2];
    s.x[2]           = s.x[3];
    s.x[3]           = s.x[4];
    s.x[4]           = (s.x[4] ^ (s.x[4] << 4)) ^ (t ^ (t << 1));
    s.d += 362437u;
    return s.d + s.x[4];
}

__host__ __device__ __forceinline__ void xorwow_seed(xorwow_state& s, uint64_t seed)
{
    s.x[0] = 123456789u;
    s.x[1] = 362436069u;
    s.x[2] = 521288629u;
    s.x[3] = 88675123u;
    s.x[4] = 5783321u;
    s.d    = 6615241u;
    const uint32_t s0 = (uint32_t)seed ^ 0x2c7f967fu;
    const uint32_t s1 = (uint32_t)(seed >> 32) ^ 0xa03697cbu;
    const uint32_t t0 = 1228688033u * s0;
    const uint32_t t1 = 2073658381u * s1;
    s.x[0] += t0;
    s.x[1] ^= t0;
    s.x[2] += t1;
    s.x[3] ^= t1;
    s.x[4] += t0;
    s.d += t1 + t0;
}

// ----------------------------------------------------------------------------
// LFSR113 (rocrand_lfsr113.h:146-162).
// ----------------------------------------------------------------------------
struct lfsr113_state
{
    uint32_t z[4];
};

__host__ __device__ __forceinline__ void lfsr113_step(uint32_t* z)
{
    uint32_t b;
    b    = (((z[0] << 6) ^ z[0]) >> 13);
    z[0] = (((z[0] & 4294967294u) << 18) ^ b);
    b    = (((z[1] << 2) ^ z[1]) >> 27);
    z[1] = (((z[1] & 4294967288u) << 2) ^ b);
    b    = (((z[2] << 13) ^ z[2]) >> 21);
    z[2] = (((z[2] & 4294967280u) << 7) ^ b);
    b    = (((z[3] << 3) ^ z[3]) >> 12);
    z[3] = (((z[3] & 4294967168u) << 13) ^ b);
}

__host__ __device__ __forceinline__ uint32_t lfsr113_next(lfsr113_state& s)
{
    lfsr113_step(s.z);
    return s.z[0] ^ s.z[1] ^ s.z[2] ^ s.z[3];
}

// GF(2) matrix (row per input bit, N words per row) times state, N in {4, 5}.
template<int N>
__host__ __device__ __forceinline__ void gf2_mat_vec(const uint32_t* __restrict__ m, uint32_t* v)
{
    uint32_t r[N];
#pragma unroll
    for(int k = 0; k < N; k++)
        r[k] = 0;
    // the word loop is unrolled so that v and r stay in registers; the bit loop is not
    // (its rows are read as it goes)
#pragma unroll
    for(int i = 0; i < N; i++)
    {
        const uint32_t vi = v[i];
#pragma unroll 2
        for(int j = 0; j < 32; j++)
        {
            const uint32_t  b   = 0u - ((vi >> j) & 1u);
            const uint32_t* row = m + (i * 32 + j) * N;
#pragma unroll
            for(int k = 0; k < N; k++)
                r[k] ^= b & row[k];
        }
    }
#pragma unroll
    for(int k = 0; k < N; k++)
        v[k] = r[k];
}

} // namespace rocrand_amd_engine
