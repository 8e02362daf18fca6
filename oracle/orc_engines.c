/*
 * orc_engines.c -- CPU restatement of the reference engines (oracle, test infrastructure).
 *
 * Each function names the reference file:line it restates.  Arithmetic is kept
 * literal (same wrap-around, same partial reductions) because bit-exact parity
 * with the reference is the contract, not mathematical nicety.
 */
#include "orc_internal.h"

#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Counter-based engines: Philox4x32-10 and ThreeFry{2,4}x{32,64}-20        */
/* ------------------------------------------------------------------------ */

/* Philox constants, rocrand_philox4x32_10.h:58-63 (Random123 values). */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

/* rocrand_philox4x32_10.h:279-315: ten_rounds / single_round / bumpkey. */
void orc_philox4x32_10_block(const unsigned int ctr[4], const unsigned int key[2], unsigned int out[4])
{
    uint32_t x = ctr[0], y = ctr[1], z = ctr[2], w = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for(int r = 0; r < 10; r++)
    {
        const uint64_t p0 = (uint64_t)PHILOX_M0 * x;
        const uint64_t p1 = (uint64_t)PHILOX_M1 * z;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t nx = hi1 ^ y ^ k0;
        const uint32_t ny = lo1;
        const uint32_t nz = hi0 ^ w ^ k1;
        const uint32_t nw = lo0;
        x = nx; y = ny; z = nz; w = nw;
        k0 += PHILOX_W0;
        k1 += PHILOX_W1;
    }
    out[0] = x; out[1] = y; out[2] = z; out[3] = w;
}

/* Skein key-schedule parity, rocrand_threefry_common.h:31-33. */
#define SKEIN_KS_PARITY32 0x1BD11BDAu
#define SKEIN_KS_PARITY64 0x1BD11BDAA9FC1A22ull

static uint32_t rotl32(uint32_t x, int d) { return (x << (d & 31)) | (x >> ((32 - d) & 31)); }
static uint64_t rotl64(uint64_t x, int d) { return (x << (d & 63)) | (x >> ((64 - d) & 63)); }

/* Rotation tables: rocrand_threefry4_impl.h:60-86, rocrand_threefry2_impl.h:80-92. */
static const int ROT32_4[8][2] = {{10, 26}, {11, 21}, {13, 27}, {23, 5}, {6, 20}, {17, 11}, {25, 10}, {18, 20}};
static const int ROT64_4[8][2] = {{14, 16}, {52, 57}, {23, 40}, {5, 37}, {25, 33}, {46, 12}, {58, 22}, {32, 32}};
static const int ROT32_2[8]    = {13, 15, 26, 6, 17, 29, 16, 24};
static const int ROT64_2[8]    = {16, 42, 12, 31, 16, 32, 24, 21};

/* rocrand_threefry4_impl.h:178-250 (threefry_rounds, 4 words). */
static void tf4_rounds(int w64, const uint64_t c[4], const uint64_t k[4], uint64_t out[4])
{
    const uint64_t M = w64 ? ~0ull : 0xffffffffull;
    uint64_t ks[5];
    ks[4] = w64 ? SKEIN_KS_PARITY64 : SKEIN_KS_PARITY32;
    for(int i = 0; i < 4; i++)
    {
        ks[i] = k[i];
        ks[4] ^= k[i];
    }
    uint64_t X[4];
    for(int i = 0; i < 4; i++)
        X[i] = (c[i] + ks[i]) & M;
    for(unsigned r = 0; r < 20; r++)
    {
        const int r0 = w64 ? ROT64_4[r & 7][0] : ROT32_4[r & 7][0];
        const int r1 = w64 ? ROT64_4[r & 7][1] : ROT32_4[r & 7][1];
#define ROTL(v, d) (w64 ? rotl64((v), (d)) : (uint64_t)rotl32((uint32_t)(v), (d)))
        if((r & 2u) == 0)
        {
            X[0] = (X[0] + X[1]) & M;
            X[1] = ROTL(X[1], r0);
            X[1] ^= X[0];
            X[2] = (X[2] + X[3]) & M;
            X[3] = ROTL(X[3], r1);
            X[3] ^= X[2];
        }
        else
        {
            X[0] = (X[0] + X[3]) & M;
            X[3] = ROTL(X[3], r0);
            X[3] ^= X[0];
            X[2] = (X[2] + X[1]) & M;
            X[1] = ROTL(X[1], r1);
            X[1] ^= X[2];
        }
        if((r & 3u) == 3)
        {
            const unsigned inj = r / 4;
            X[0] = (X[0] + ks[(1 + inj) % 5]) & M;
            X[1] = (X[1] + ks[(2 + inj) % 5]) & M;
            X[2] = (X[2] + ks[(3 + inj) % 5]) & M;
            X[3] = (X[3] + ks[(4 + inj) % 5]) & M;
            X[3] = (X[3] + 1 + inj) & M;
        }
    }
    for(int i = 0; i < 4; i++)
        out[i] = X[i];
}

/* rocrand_threefry2_impl.h:~100-135 (threefry_rounds, 2 words). */
static void tf2_rounds(int w64, const uint64_t c[2], const uint64_t k[2], uint64_t out[2])
{
    const uint64_t M = w64 ? ~0ull : 0xffffffffull;
    uint64_t ks[3];
    ks[2] = w64 ? SKEIN_KS_PARITY64 : SKEIN_KS_PARITY32;
    ks[0] = k[0];
    ks[1] = k[1];
    ks[2] ^= k[0];
    ks[2] ^= k[1];
    uint64_t X0 = (c[0] + ks[0]) & M, X1 = (c[1] + ks[1]) & M;
    for(unsigned r = 0; r < 20; r++)
    {
        X0 = (X0 + X1) & M;
        X1 = w64 ? rotl64(X1, ROT64_2[r & 7]) : (uint64_t)rotl32((uint32_t)X1, ROT32_2[r & 7]);
        X1 ^= X0;
        if((r & 3u) == 3)
        {
            const unsigned inj = r / 4;
            X0 = (X0 + ks[(1 + inj) % 3]) & M;
            X1 = (X1 + ks[(2 + inj) % 3]) & M;
            X1 = (X1 + 1 + inj) & M;
        }
    }
    out[0] = X0;
    out[1] = X1;
}
#undef ROTL

void orc_threefry4x32_20_block(const unsigned int ctr[4], const unsigned int key[4], unsigned int out[4])
{
    uint64_t c[4], k[4], o[4];
    for(int i = 0; i < 4; i++) { c[i] = ctr[i]; k[i] = key[i]; }
    tf4_rounds(0, c, k, o);
    for(int i = 0; i < 4; i++) out[i] = (unsigned int)o[i];
}
void orc_threefry2x32_20_block(const unsigned int ctr[2], const unsigned int key[2], unsigned int out[2])
{
    uint64_t c[2] = {ctr[0], ctr[1]}, k[2] = {key[0], key[1]}, o[2];
    tf2_rounds(0, c, k, o);
    out[0] = (unsigned int)o[0];
    out[1] = (unsigned int)o[1];
}
void orc_threefry4x64_20_block(const unsigned long long ctr[4], const unsigned long long key[4], unsigned long long out[4])
{
    uint64_t c[4], k[4], o[4];
    for(int i = 0; i < 4; i++) { c[i] = ctr[i]; k[i] = key[i]; }
    tf4_rounds(1, c, k, o);
    for(int i = 0; i < 4; i++) out[i] = o[i];
}
void orc_threefry2x64_20_block(const unsigned long long ctr[2], const unsigned long long key[2], unsigned long long out[2])
{
    uint64_t c[2] = {ctr[0], ctr[1]}, k[2] = {key[0], key[1]}, o[2];
    tf2_rounds(1, c, k, o);
    out[0] = o[0];
    out[1] = o[1];
}

static void ce_rounds(const ctr_engine* e, const uint64_t c[4], uint64_t out[4])
{
    if(e->family == CE_PHILOX)
    {
        unsigned int cc[4] = {(unsigned)c[0], (unsigned)c[1], (unsigned)c[2], (unsigned)c[3]};
        unsigned int kk[2] = {(unsigned)e->key[0], (unsigned)e->key[1]};
        unsigned int oo[4];
        orc_philox4x32_10_block(cc, kk, oo);
        for(int i = 0; i < 4; i++) out[i] = oo[i];
    }
    else if(e->dim == 4)
    {
        tf4_rounds(e->w64, c, e->key, out);
    }
    else
    {
        tf2_rounds(e->w64, c, e->key, out);
        out[2] = out[3] = 0;
    }
}

/* rocrand_philox4x32_10.h:233-243 / threefry{2,4}_impl.h discard_state(offset):
 * add a 64-bit word offset to the multi-word counter, literal carry logic. */
static void ce_discard_state(ctr_engine* e, uint64_t offset)
{
    const uint64_t M = e->w64 ? ~0ull : 0xffffffffull;
    const uint64_t lo = e->w64 ? offset : (offset & 0xffffffffull);
    const uint64_t hi = e->w64 ? 0 : (offset >> 32);
    uint64_t old[4];
    memcpy(old, e->ctr, sizeof(old));
    e->ctr[0] = (e->ctr[0] + lo) & M;
    e->ctr[1] = (e->ctr[1] + hi + (e->ctr[0] < old[0] ? 1 : 0)) & M;
    if(e->dim == 4)
    {
        e->ctr[2] = (e->ctr[2] + (e->ctr[1] < old[1] ? 1 : 0)) & M;
        e->ctr[3] = (e->ctr[3] + (e->ctr[2] < old[2] ? 1 : 0)) & M;
    }
}

/* rocrand_philox4x32_10.h:252-260 bump_counter. */
static void ce_bump(const ctr_engine* e, uint64_t c[4])
{
    const uint64_t M = e->w64 ? ~0ull : 0xffffffffull;
    c[0] = (c[0] + 1) & M;
    uint64_t add = c[0] == 0 ? 1 : 0;
    c[1] = (c[1] + add) & M;
    if(e->dim == 4)
    {
        add  = c[1] == 0 ? add : 0;
        c[2] = (c[2] + add) & M;
        add  = c[2] == 0 ? add : 0;
        c[3] = (c[3] + add) & M;
    }
}

/* rocrand_philox4x32_10.h:208-217 discard_impl. */
static void ce_discard_impl(ctr_engine* e, uint64_t offset)
{
    const unsigned d = e->dim;
    e->substate += (unsigned)(offset & (d - 1));
    uint64_t counter_offset = offset / d;
    counter_offset += e->substate < d ? 0 : 1;
    if(e->substate >= d)
        e->substate -= d;
    ce_discard_state(e, counter_offset);
}

void ce_discard(ctr_engine* e, uint64_t offset)
{
    ce_discard_impl(e, offset);
    ce_rounds(e, e->ctr, e->res);
}

/* rocrand_philox4x32_10.h:222-231 discard_subsequence_impl (and threefry variants). */
static void ce_discard_subsequence_impl(ctr_engine* e, uint64_t sub)
{
    const uint64_t M = e->w64 ? ~0ull : 0xffffffffull;
    if(e->dim == 2)
    {
        e->ctr[1] = (e->ctr[1] + sub) & M;
        return;
    }
    const uint64_t lo = e->w64 ? sub : (sub & 0xffffffffull);
    const uint64_t hi = e->w64 ? 0 : (sub >> 32);
    const uint64_t old = e->ctr[2];
    e->ctr[2] = (e->ctr[2] + lo) & M;
    e->ctr[3] = (e->ctr[3] + hi + (e->ctr[2] < old ? 1 : 0)) & M;
}

/* Philox seed (rocrand_philox4x32_10.h:134-141, restart :157-171) and ThreeFry
 * seeds (rocrand_threefry{2x32,2x64,4x32,4x64}_20.h seed(): the 4-word keys take
 * {seed, seed>>16, seed>>32, seed>>48} truncated to the word size). */
void ce_init(ctr_engine* e, int rng_type, uint64_t seed, uint64_t sub, uint64_t offset)
{
    memset(e, 0, sizeof(*e));
    switch(rng_type)
    {
        case RNG_PHILOX4_32_10:
            e->family = CE_PHILOX; e->dim = 4; e->w64 = 0;
            e->key[0] = (uint32_t)seed; e->key[1] = (uint32_t)(seed >> 32);
            ce_discard_subsequence_impl(e, sub);
            ce_discard_impl(e, offset);
            ce_rounds(e, e->ctr, e->res);
            return;
        case RNG_THREEFRY2_32_20:
            e->family = CE_THREEFRY; e->dim = 2; e->w64 = 0;
            e->key[0] = (uint32_t)seed; e->key[1] = (uint32_t)(seed >> 32);
            break;
        case RNG_THREEFRY2_64_20:
            e->family = CE_THREEFRY; e->dim = 2; e->w64 = 1;
            e->key[0] = seed; e->key[1] = seed >> 32;
            break;
        case RNG_THREEFRY4_32_20:
            e->family = CE_THREEFRY; e->dim = 4; e->w64 = 0;
            e->key[0] = (uint32_t)seed; e->key[1] = (uint32_t)(seed >> 16);
            e->key[2] = (uint32_t)(seed >> 32); e->key[3] = (uint32_t)(seed >> 48);
            break;
        case RNG_THREEFRY4_64_20:
            e->family = CE_THREEFRY; e->dim = 4; e->w64 = 1;
            e->key[0] = seed; e->key[1] = seed >> 16; e->key[2] = seed >> 32; e->key[3] = seed >> 48;
            break;
        default: abort();
    }
    ce_discard_subsequence_impl(e, sub);
    ce_discard(e, offset);
    ce_rounds(e, e->ctr, e->res);
}

/* rocrand_philox4x32_10.h:180-195 next(). */
uint64_t ce_next(ctr_engine* e)
{
    const uint64_t ret = e->res[e->substate];
    e->substate++;
    if(e->substate == e->dim)
    {
        e->substate = 0;
        ce_bump(e, e->ctr);
        ce_rounds(e, e->ctr, e->res);
    }
    return ret;
}

/* library/src/rng/philox4x32_10.hpp:88-101 next4_leap / threefry.hpp:64-76 next_leap. */
void ce_next_leap(ctr_engine* e, uint32_t leap, uint64_t out[4])
{
    uint64_t ret[4];
    memcpy(ret, e->res, sizeof(ret));
    if(e->substate > 0)
    {
        uint64_t nc[4], nx[4];
        memcpy(nc, e->ctr, sizeof(nc));
        ce_bump(e, nc);
        ce_rounds(e, nc, nx);
        /* interleave, rocrand_philox4x32_10.h:262-276 */
        uint64_t t[4];
        const unsigned d = e->dim, s = e->substate;
        for(unsigned i = 0; i < d; i++)
            t[i] = (i + s < d) ? ret[i + s] : nx[i + s - d];
        memcpy(ret, t, sizeof(t));
    }
    ce_discard_state(e, leap);
    ce_rounds(e, e->ctr, e->res);
    memcpy(out, ret, sizeof(ret));
}

/* ------------------------------------------------------------------------ */
/* MRG32k3a / MRG31k3p                                                       */
/* ------------------------------------------------------------------------ */

#define MRG32_M1  4294967087ull
#define MRG32_M1C 209ull
#define MRG32_M2  4294944443ull
#define MRG32_M2C 22853ull
#define MRG32_A12 1403580ull
#define MRG32_A13N 810728ull
#define MRG32_A21 527612ull
#define MRG32_A23N 1370589ull

#define MRG31_M1 2147483647u
#define MRG31_M2 2147462579u

/* rocrand_mrg32k3a.h:333-339 mod_m1 (single fold + single subtract, literal). */
static uint64_t mrg32_mod_m1(uint64_t p)
{
    p = MRG32_M1C * (p >> 32) + (p & 0xffffffffull);
    if(p >= MRG32_M1)
        p -= MRG32_M1;
    return p;
}
/* rocrand_mrg32k3a.h:356-363 mod_m2 (two folds + single subtract). */
static uint64_t mrg32_mod_m2(uint64_t p)
{
    p = MRG32_M2C * (p >> 32) + (p & 0xffffffffull);
    p = MRG32_M2C * (p >> 32) + (p & 0xffffffffull);
    if(p >= MRG32_M2)
        p -= MRG32_M2;
    return p;
}
/* rocrand_mrg32k3a.h:318-331 mod_mul_m1 / :341-354 mod_mul_m2: 64-bit wrapping
 * products of a 15/17-bit split of i with the full 64-bit seed. */
static uint64_t mrg32_mod_mul_m1(uint32_t i, uint64_t j)
{
    const uint64_t hi = i / 131072u;
    const uint64_t lo = i - hi * 131072u;
    const uint64_t t1 = mrg32_mod_m1(hi * j) * 131072u;
    const uint64_t t2 = mrg32_mod_m1(lo * j);
    return mrg32_mod_m1(t1 + t2);
}
static uint64_t mrg32_mod_mul_m2(uint32_t i, uint64_t j)
{
    const uint64_t hi = i / 131072u;
    const uint64_t lo = i - hi * 131072u;
    const uint64_t t1 = mrg32_mod_m2(hi * j) * 131072u;
    const uint64_t t2 = mrg32_mod_m2(lo * j);
    return mrg32_mod_m2(t1 + t2);
}

/* rocrand_mrg31k3p.h mod_m1 / mod_m2: plain remainders. */
static uint64_t mrg31_mod_m1(uint64_t p) { return p % MRG31_M1; }
static uint64_t mrg31_mod_m2(uint64_t p) { return p % MRG31_M2; }

/* 3x3 mod-m matrix-vector product in the reference's row-major layout
 * (rocrand_mrg32k3a.h:273-315). Exact math: every table entry < m. */
static void matvec_mod(const uint64_t* A, uint32_t* s, uint64_t m)
{
    uint64_t x[3];
    for(int r = 0; r < 3; r++)
    {
        uint64_t acc = 0;
        for(int c = 0; c < 3; c++)
            acc += (A[3 * r + c] * (uint64_t)s[c]) % m;
        x[r] = acc % m;
    }
    for(int r = 0; r < 3; r++)
        s[r] = (uint32_t)x[r];
}

static void matmul_mod(const uint64_t* A, const uint64_t* B, uint64_t* C, uint64_t m)
{
    uint64_t T[9];
    for(int r = 0; r < 3; r++)
        for(int c = 0; c < 3; c++)
        {
            uint64_t acc = 0;
            for(int k = 0; k < 3; k++)
                acc += (A[3 * r + k] * B[3 * k + c]) % m;
            T[3 * r + c] = acc % m;
        }
    memcpy(C, T, sizeof(T));
}

/* Tables T[i] = X^(2^i) for X in {A, A^(2^sub_log2), A^(2^seq_log2)}, i = 0..63
 * (the layout of rocrand_mrg32k3a_precomputed.h / rocrand_mrg31k3p_precomputed.h,
 * derived here from the one-step matrices by repeated squaring). */
typedef struct
{
    uint64_t t[3][2][64][9]; /* [which: step/sub/seq][component][power][3x3] */
    int      ready;
} mrg_tables;
static mrg_tables g_mrg32, g_mrg31;

static void build_mrg_tables(mrg_tables* tb, const uint64_t A1[9], const uint64_t A2[9],
                             uint64_t m1, uint64_t m2, int sub_log2, int seq_log2)
{
    const uint64_t* A[2] = {A1, A2};
    const uint64_t  m[2] = {m1, m2};
    for(int c = 0; c < 2; c++)
    {
        uint64_t X[9];
        memcpy(X, A[c], sizeof(X));
        for(int i = 0; i < 64; i++)
        {
            memcpy(tb->t[0][c][i], X, sizeof(X));
            matmul_mod(X, X, X, m[c]);
        }
        /* A^(2^sub_log2) */
        memcpy(X, A[c], sizeof(X));
        for(int i = 0; i < sub_log2; i++)
            matmul_mod(X, X, X, m[c]);
        for(int i = 0; i < 64; i++)
        {
            memcpy(tb->t[1][c][i], X, sizeof(X));
            matmul_mod(X, X, X, m[c]);
        }
        memcpy(X, A[c], sizeof(X));
        for(int i = 0; i < seq_log2; i++)
            matmul_mod(X, X, X, m[c]);
        for(int i = 0; i < 64; i++)
        {
            memcpy(tb->t[2][c][i], X, sizeof(X));
            matmul_mod(X, X, X, m[c]);
        }
    }
    tb->ready = 1;
}

static void ensure_mrg_tables(void)
{
    if(!g_mrg32.ready)
    {
        /* One-step matrices of MRG32k3a (L'Ecuyer 1999): x_n = 1403580 x_{n-2} - 810728 x_{n-3}
         * mod m1 and y_n = 527612 y_{n-1} - 1370589 y_{n-3} mod m2, state order (oldest..newest). */
        const uint64_t A1[9] = {0, 1, 0, 0, 0, 1, MRG32_M1 - MRG32_A13N, MRG32_A12, 0};
        const uint64_t A2[9] = {0, 1, 0, 0, 0, 1, MRG32_M2 - MRG32_A23N, 0, MRG32_A21};
        build_mrg_tables(&g_mrg32, A1, A2, MRG32_M1, MRG32_M2, 76, 127);
    }
    if(!g_mrg31.ready)
    {
        /* MRG31k3p (L'Ecuyer-Touzin 2000), state order (newest..oldest):
         * x_n = 2^22 x_{n-2} + (2^7+1) x_{n-3} mod m1, y_n = 2^15 y_{n-1} + (2^15+1) y_{n-3} mod m2. */
        const uint64_t A1[9] = {0, 1ull << 22, (1ull << 7) + 1, 1, 0, 0, 0, 1, 0};
        const uint64_t A2[9] = {1ull << 15, 0, (1ull << 15) + 1, 1, 0, 0, 0, 1, 0};
        build_mrg_tables(&g_mrg31, A1, A2, MRG31_M1, MRG31_M2, 72, 134);
    }
}

int orc_mrg_matrix(int rng_type, int component, int which, int index, unsigned long long out[9])
{
    ensure_mrg_tables();
    const mrg_tables* tb = rng_type == RNG_MRG32K3A ? &g_mrg32 : rng_type == RNG_MRG31K3P ? &g_mrg31 : NULL;
    if(!tb || component < 1 || component > 2 || which < 0 || which > 2 || index < 0 || index > 63)
        return -1;
    memcpy(out, tb->t[which][component - 1][index], 9 * sizeof(uint64_t));
    return 0;
}

static void mrg_apply_bits(mrg_engine* e, int which, uint64_t v)
{
    const mrg_tables* tb = e->type == RNG_MRG32K3A ? &g_mrg32 : &g_mrg31;
    const uint64_t m1 = e->type == RNG_MRG32K3A ? MRG32_M1 : MRG31_M1;
    const uint64_t m2 = e->type == RNG_MRG32K3A ? MRG32_M2 : MRG31_M2;
    int i = 0;
    while(v > 0)
    {
        if(v & 1)
        {
            matvec_mod(tb->t[which][0][i], e->g1, m1);
            matvec_mod(tb->t[which][1][i], e->g2, m2);
        }
        v >>= 1;
        i++;
    }
}

/* rocrand_mrg32k3a.h:103-120 seed(); rocrand_mrg31k3p.h seed() (seed 0 -> 12345). */
void mrg_seed_only(mrg_engine* e, int type, uint64_t seed)
{
    ensure_mrg_tables();
    e->type = type;
    if(seed == 0)
        seed = 12345ull;
    if(type == RNG_MRG32K3A)
    {
        const uint32_t x = (uint32_t)seed ^ 0x55555555u;
        const uint32_t y = (uint32_t)((seed >> 32) ^ 0xAAAAAAAAu);
        e->g1[0] = (uint32_t)mrg32_mod_mul_m1(x, seed);
        e->g1[1] = (uint32_t)mrg32_mod_mul_m1(y, seed);
        e->g1[2] = (uint32_t)mrg32_mod_mul_m1(x, seed);
        e->g2[0] = (uint32_t)mrg32_mod_mul_m2(y, seed);
        e->g2[1] = (uint32_t)mrg32_mod_mul_m2(x, seed);
        e->g2[2] = (uint32_t)mrg32_mod_mul_m2(y, seed);
    }
    else
    {
        const uint32_t x = (uint32_t)(seed ^ 0x55555555u);
        const uint32_t y = (uint32_t)((seed >> 32) ^ 0xAAAAAAAAu);
        e->g1[0] = (uint32_t)mrg31_mod_m1((uint64_t)x * seed);
        e->g1[1] = (uint32_t)mrg31_mod_m1((uint64_t)y * seed);
        e->g1[2] = (uint32_t)mrg31_mod_m1((uint64_t)x * seed);
        e->g2[0] = (uint32_t)mrg31_mod_m2((uint64_t)y * seed);
        e->g2[1] = (uint32_t)mrg31_mod_m2((uint64_t)x * seed);
        e->g2[2] = (uint32_t)mrg31_mod_m2((uint64_t)y * seed);
    }
}

void mrg_discard(mrg_engine* e, uint64_t n) { mrg_apply_bits(e, 0, n); }
void mrg_discard_subsequence(mrg_engine* e, uint64_t n) { mrg_apply_bits(e, 1, n); }

void mrg_init(mrg_engine* e, int type, uint64_t seed, uint64_t sub, uint64_t offset)
{
    mrg_seed_only(e, type, seed);
    mrg_discard_subsequence(e, sub);
    mrg_discard(e, offset);
}

/* rocrand_mrg32k3a.h:160-193 next(); rocrand_mrg31k3p.h next(). */
uint32_t mrg_next(mrg_engine* e)
{
    if(e->type == RNG_MRG32K3A)
    {
        const uint32_t p1 = (uint32_t)mrg32_mod_m1(MRG32_A12 * e->g1[1] + MRG32_A13N * (MRG32_M1 - e->g1[0]));
        e->g1[0] = e->g1[1]; e->g1[1] = e->g1[2]; e->g1[2] = p1;
        const uint32_t p2 = (uint32_t)mrg32_mod_m2(MRG32_A21 * e->g2[2] + MRG32_A23N * (MRG32_M2 - e->g2[0]));
        e->g2[0] = e->g2[1]; e->g2[1] = e->g2[2]; e->g2[2] = p2;
        return (p1 - p2) + (p1 <= p2 ? (uint32_t)MRG32_M1 : 0u);
    }
    uint32_t tmp = (((e->g1[1] & 511u) << 22) + (e->g1[1] >> 9)) + (((e->g1[2] & 16777215u) << 7) + (e->g1[2] >> 24));
    tmp -= (tmp >= MRG31_M1) ? MRG31_M1 : 0;
    tmp += e->g1[2];
    tmp -= (tmp >= MRG31_M1) ? MRG31_M1 : 0;
    e->g1[2] = e->g1[1]; e->g1[1] = e->g1[0]; e->g1[0] = tmp;
    tmp = (((e->g2[0] & 65535u) << 15) + 21069u * (e->g2[0] >> 16));
    tmp -= (tmp >= MRG31_M2) ? MRG31_M2 : 0;
    tmp += ((e->g2[2] & 65535u) << 15);
    tmp -= (tmp >= MRG31_M2) ? MRG31_M2 : 0;
    tmp += 21069u * (e->g2[2] >> 16);
    tmp -= (tmp >= MRG31_M2) ? MRG31_M2 : 0;
    tmp += e->g2[2];
    tmp -= (tmp >= MRG31_M2) ? MRG31_M2 : 0;
    e->g2[2] = e->g2[1]; e->g2[1] = e->g2[0]; e->g2[0] = tmp;
    return e->g1[0] - e->g2[0] + (e->g1[0] <= e->g2[0] ? MRG31_M1 : 0u);
}

/* ------------------------------------------------------------------------ */
/* XORWOW and LFSR113: GF(2) jump matrices                                  */
/* ------------------------------------------------------------------------ */

/* Matrix layout of rocrand_xorwow_precomputed.h / rocrand_lfsr113_precomputed.h:
 * row (word i, bit j) at [(i*32 + j) * N .. +N) holds the image of that input bit. */
typedef struct
{
    uint32_t jump[32][5 * 32 * 5];
    uint32_t seq[32][5 * 32 * 5];
    int      ready;
} gf2_tables;
static gf2_tables g_xorwow, g_lfsr;

static void gf2_matvec(const uint32_t* m, uint32_t* v, int N)
{
    uint32_t r[5] = {0, 0, 0, 0, 0};
    for(int ij = 0; ij < N * 32; ij++)
    {
        const int i = ij / 32, j = ij % 32;
        const uint32_t b = (v[i] & (1u << j)) ? 0xffffffffu : 0u;
        for(int k = 0; k < N; k++)
            r[k] ^= b & m[(i * 32 + j) * N + k];
    }
    memcpy(v, r, N * sizeof(uint32_t));
}

/* C = apply A then B (every operand a power of the same matrix, so order is immaterial). */
static void gf2_matmul(const uint32_t* A, const uint32_t* B, uint32_t* C, int N)
{
    uint32_t T[5 * 32 * 5];
    for(int row = 0; row < N * 32; row++)
    {
        uint32_t v[5];
        memcpy(v, A + row * N, N * sizeof(uint32_t));
        gf2_matvec(B, v, N);
        memcpy(T + row * N, v, N * sizeof(uint32_t));
    }
    memcpy(C, T, N * 32 * N * sizeof(uint32_t));
}

static void xorwow_step_x(uint32_t* x)
{
    /* rocrand_xorwow.h:165-177 without the Weyl counter. */
    const uint32_t t = x[0] ^ (x[0] >> 2);
    x[0] = x[1]; x[1] = x[2]; x[2] = x[3]; x[3] = x[4];
    x[4] = (x[4] ^ (x[4] << 4)) ^ (t ^ (t << 1));
}
static void lfsr_step(uint32_t* z)
{
    /* rocrand_lfsr113.h:146-162 */
    uint32_t b;
    b = (((z[0] << 6) ^ z[0]) >> 13);  z[0] = (((z[0] & 4294967294u) << 18) ^ b);
    b = (((z[1] << 2) ^ z[1]) >> 27);  z[1] = (((z[1] & 4294967288u) << 2) ^ b);
    b = (((z[2] << 13) ^ z[2]) >> 21); z[2] = (((z[2] & 4294967280u) << 7) ^ b);
    b = (((z[3] << 3) ^ z[3]) >> 12);  z[3] = (((z[3] & 4294967168u) << 13) ^ b);
}

/* Same construction as tools/xorwow_precomputed_generator.cpp: jump[k] = M^(4^k),
 * seq[k] = M^(2^seq_log2 * 4^k). */
static void build_gf2(gf2_tables* tb, int N, void (*step)(uint32_t*), int seq_log2)
{
    static uint32_t one[5 * 32 * 5], a[5 * 32 * 5];
    for(int i = 0; i < N; i++)
        for(int j = 0; j < 32; j++)
        {
            uint32_t v[5] = {0, 0, 0, 0, 0};
            v[i] = 1u << j;
            step(v);
            memcpy(one + (i * 32 + j) * N, v, N * sizeof(uint32_t));
        }
    const size_t sz = (size_t)N * 32 * N * sizeof(uint32_t);
    memcpy(a, one, sz);
    for(int k = 0; k < 32; k++)
    {
        memcpy(tb->jump[k], a, sz);
        gf2_matmul(a, a, a, N); /* ^2 */
        gf2_matmul(a, a, a, N); /* ^4 */
    }
    memcpy(a, one, sz);
    for(int s = 0; s < seq_log2; s++)
        gf2_matmul(a, a, a, N);
    for(int k = 0; k < 32; k++)
    {
        memcpy(tb->seq[k], a, sz);
        gf2_matmul(a, a, a, N);
        gf2_matmul(a, a, a, N);
    }
    tb->ready = 1;
}

static void ensure_gf2(void)
{
    if(!g_xorwow.ready)
        build_gf2(&g_xorwow, 5, xorwow_step_x, 67);
    if(!g_lfsr.ready)
        build_gf2(&g_lfsr, 4, lfsr_step, 55);
}

const uint32_t* orc_gf2_matrix(int rng_type, int seq, int k)
{
    ensure_gf2();
    const gf2_tables* tb = rng_type == RNG_XORWOW ? &g_xorwow : &g_lfsr;
    return seq ? tb->seq[k] : tb->jump[k];
}

/* rocrand_xorwow.h:186-200 jump(): base-4 digits, digit d applies matrix k d times. */
static void gf2_jump(uint32_t* v, int N, const uint32_t (*mats)[5 * 32 * 5], uint64_t n)
{
    unsigned mi = 0;
    while(n > 0)
    {
        const unsigned is = (unsigned)(n & 3u);
        for(unsigned i = 0; i < is; i++)
            gf2_matvec(mats[mi], v, N);
        mi++;
        n >>= 2;
    }
}

/* rocrand_xorwow.h:100-130 constructor seeding. */
void xorwow_seed_only(xorwow_engine* e, uint64_t seed)
{
    ensure_gf2();
    e->x[0] = 123456789u; e->x[1] = 362436069u; e->x[2] = 521288629u;
    e->x[3] = 88675123u;  e->x[4] = 5783321u;   e->d = 6615241u;
    const uint32_t s0 = (uint32_t)seed ^ 0x2c7f967fu;
    const uint32_t s1 = (uint32_t)(seed >> 32) ^ 0xa03697cbu;
    const uint32_t t0 = 1228688033u * s0;
    const uint32_t t1 = 2073658381u * s1;
    e->x[0] += t0; e->x[1] ^= t0; e->x[2] += t1; e->x[3] ^= t1; e->x[4] += t0;
    e->d += t1 + t0;
}
void xorwow_discard(xorwow_engine* e, uint64_t n)
{
    gf2_jump(e->x, 5, (const uint32_t(*)[5 * 32 * 5])g_xorwow.jump, n);
    e->d += (uint32_t)n * 362437u;
}
void xorwow_discard_subsequence(xorwow_engine* e, uint64_t n)
{
    gf2_jump(e->x, 5, (const uint32_t(*)[5 * 32 * 5])g_xorwow.seq, n);
}
void xorwow_apply(xorwow_engine* e, const uint32_t* m) { gf2_matvec(m, e->x, 5); }
uint32_t xorwow_next(xorwow_engine* e)
{
    const uint32_t t = e->x[0] ^ (e->x[0] >> 2);
    e->x[0] = e->x[1]; e->x[1] = e->x[2]; e->x[2] = e->x[3]; e->x[3] = e->x[4];
    e->x[4] = (e->x[4] ^ (e->x[4] << 4)) ^ (t ^ (t << 1));
    e->d += 362437u;
    return e->d + e->x[4];
}

void lfsr_seed_only(lfsr_engine* e, const uint32_t seed[4])
{
    ensure_gf2();
    memcpy(e->z, seed, 4 * sizeof(uint32_t));
}
void lfsr_discard(lfsr_engine* e, uint64_t n) { gf2_jump(e->z, 4, (const uint32_t(*)[5 * 32 * 5])g_lfsr.jump, n); }
void lfsr_discard_subsequence(lfsr_engine* e, uint32_t n)
{
    gf2_jump(e->z, 4, (const uint32_t(*)[5 * 32 * 5])g_lfsr.seq, n);
}
void lfsr_apply(lfsr_engine* e, const uint32_t* m) { gf2_matvec(m, e->z, 4); }
uint32_t lfsr_next(lfsr_engine* e)
{
    lfsr_step(e->z);
    return e->z[0] ^ e->z[1] ^ e->z[2] ^ e->z[3];
}

/* ------------------------------------------------------------------------ */
/* Raw engine streams (known-answer and jump checks)                        */
/* ------------------------------------------------------------------------ */
int orc_engine_stream(int rng_type, unsigned long long seed, unsigned long long sub,
                      unsigned long long offset, unsigned int* out, size_t n)
{
    switch(rng_type)
    {
        case RNG_PHILOX4_32_10:
        case RNG_THREEFRY2_32_20:
        case RNG_THREEFRY4_32_20:
        {
            ctr_engine e;
            ce_init(&e, rng_type, seed, sub, offset);
            for(size_t i = 0; i < n; i++)
                out[i] = (unsigned int)ce_next(&e);
            return 0;
        }
        case RNG_MRG32K3A:
        case RNG_MRG31K3P:
        {
            mrg_engine e;
            mrg_init(&e, rng_type, seed, sub, offset);
            for(size_t i = 0; i < n; i++)
                out[i] = mrg_next(&e);
            return 0;
        }
        case RNG_XORWOW:
        {
            xorwow_engine e;
            xorwow_seed_only(&e, seed);
            xorwow_discard_subsequence(&e, sub);
            xorwow_discard(&e, offset);
            for(size_t i = 0; i < n; i++)
                out[i] = xorwow_next(&e);
            return 0;
        }
        case RNG_LFSR113:
        {
            /* seed here is the packed uint4 {x,y,z,w} = 4 x 16-bit fields is not used:
             * callers pass the reference's set_seed(u64) mapping through orc_generate. */
            lfsr_engine e;
            const uint32_t s[4] = {(uint32_t)seed, (uint32_t)(seed >> 16), (uint32_t)(seed >> 32), (uint32_t)(seed >> 48)};
            lfsr_seed_only(&e, s);
            lfsr_discard_subsequence(&e, (uint32_t)sub);
            lfsr_discard(&e, offset);
            for(size_t i = 0; i < n; i++)
                out[i] = lfsr_next(&e);
            return 0;
        }
        default: return -1;
    }
}
