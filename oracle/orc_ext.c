/*
 * orc_ext.c -- Sobol32/64 (+ scrambled) and MTGP32 restatements of the oracle
 * (test infrastructure, see oracle.h).  MT19937 lives in orc_mt19937.c.
 *
 * Sobol: library/src/rng/sobol.hpp:72-261 (kernel body), :489-790 (generator),
 *        rocrand_sobol32.h / rocrand_sobol64.h (engines, Gray-code state),
 *        tools/sobol_utils.hpp:59-104 (direction vectors from Joe-Kuo numbers),
 *        tools/scrambled_sobol{32,64}_direction_vector_generator.cpp (+640000 / +1280000),
 *        tools/scrambled_sobol{32,64}_constants_generator.cpp (std::mt19937 draws).
 * MTGP32: rocrand_mtgp32.h:130-145 (state init), :272-317 (recurrence, tempering),
 *         :362-400 (512 generators, seed ^ (seed >> 32) + i + 1),
 *         library/src/rng/mtgp32.hpp:232-305 (block-per-generator kernel body).
 */
#include "orc_internal.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* shared distribution helpers (definitions in orc_gen.c)                   */
/* ------------------------------------------------------------------------ */
unsigned short orc_float2half(float f);

#define R_2POW16_INV     1.5258789e-05f
#define R_2POW32_INV     2.3283064e-10f
#define R_2POW32_INV_DBL 2.3283064365386963e-10
#define R_2POW64_INV_DBL 5.4210108624275221700372640043497e-20
#define R_2POW53_INV_DBL 1.1102230246251565e-16
#define R_PI_F           3.141592653f
#define R_PI_DBL         3.1415926535897932
#define R_SQRT2_F        1.4142135f

static float half2f(unsigned short h)
{
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu, x;
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
            x = sign | (e << 23) | ((m & 0x3ffu) << 13);
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
/* Sobol tables                                                             */
/* ------------------------------------------------------------------------ */
#define SOBOL_DIM 20000

static uint32_t* g_v32;   /* [SOBOL_DIM][32] */
static uint64_t* g_v64;   /* [SOBOL_DIM][64] */
static uint32_t* g_sv32;  /* scrambled */
static uint64_t* g_sv64;
static uint32_t  g_c32[SOBOL_DIM];
static uint64_t  g_c64[SOBOL_DIM];
static int       g_sobol_ready;

/* tools/sobol_utils.hpp:59-90: Bratley-Fox recurrence from (s, a, m_1..m_s). */
static void sobol_vectors(const uint32_t* jk, size_t words, int bits, void* out)
{
    const int shift = bits - 1;
    size_t    pos   = 0;
    for(int d = 0; d < SOBOL_DIM; d++)
    {
        uint64_t v[64];
        if(d == 0)
        {
            for(int j = 0; j < bits; j++)
                v[j] = 1ull << (shift - j);
        }
        else
        {
            if(pos + 2 > words)
                abort();
            const uint32_t s = jk[pos], a = jk[pos + 1];
            const uint32_t* m = jk + pos + 2;
            pos += 2 + s;
            const uint64_t mask = bits == 64 ? ~0ull : 0xffffffffull;
            for(uint32_t j = 0; j < s && j < (uint32_t)bits; j++)
                v[j] = ((uint64_t)m[j] << (shift - j)) & mask;
            for(int j = (int)s; j < bits; j++)
            {
                uint64_t x = v[j - s] ^ (v[j - s] >> s);
                for(uint32_t k = 1; k < s; k++)
                    if((a >> (s - 1 - k)) & 1u)
                        x ^= v[j - k];
                v[j] = x;
            }
        }
        for(int j = 0; j < bits; j++)
        {
            if(bits == 32)
                ((uint32_t*)out)[(size_t)d * 32 + j] = (uint32_t)v[j];
            else
                ((uint64_t*)out)[(size_t)d * 64 + j] = v[j];
        }
    }
}

/* std::mt19937 with its default seed 5489 (the C++ standard's recurrence). */
typedef struct
{
    uint32_t mt[624];
    int      i;
} mt_std;
static void mt_std_seed(mt_std* g, uint32_t s)
{
    g->mt[0] = s;
    for(int i = 1; i < 624; i++)
        g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (uint32_t)i;
    g->i = 624;
}
static uint32_t mt_std_next(mt_std* g)
{
    if(g->i >= 624)
    {
        for(int k = 0; k < 624; k++)
        {
            const uint32_t y = (g->mt[k] & 0x80000000u) | (g->mt[(k + 1) % 624] & 0x7fffffffu);
            g->mt[k]         = g->mt[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        }
        g->i = 0;
    }
    uint32_t y = g->mt[g->i++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

int orc_sobol_load(const char* path)
{
    FILE* f = fopen(path, "rb");
    if(!f)
        return -1;
    fseek(f, 0, SEEK_END);
    const long bytes = ftell(f);
    fseek(f, 0, SEEK_SET);
    uint32_t* jk = (uint32_t*)malloc((size_t)bytes);
    if(fread(jk, 1, (size_t)bytes, f) != (size_t)bytes)
    {
        fclose(f);
        free(jk);
        return -1;
    }
    fclose(f);
    const size_t words = (size_t)bytes / 4;
    free(g_v32); free(g_v64); free(g_sv32); free(g_sv64);
    g_v32  = (uint32_t*)malloc(sizeof(uint32_t) * SOBOL_DIM * 32);
    g_v64  = (uint64_t*)malloc(sizeof(uint64_t) * SOBOL_DIM * 64);
    g_sv32 = (uint32_t*)malloc(sizeof(uint32_t) * SOBOL_DIM * 32);
    g_sv64 = (uint64_t*)malloc(sizeof(uint64_t) * SOBOL_DIM * 64);
    sobol_vectors(jk, words, 32, g_v32);
    sobol_vectors(jk, words, 64, g_v64);
    free(jk);
    /* scrambled direction vectors: integer add of 640000 / 1280000 (tools/sobol_utils.hpp:107-117) */
    for(size_t k = 0; k < (size_t)SOBOL_DIM * 32; k++)
        g_sv32[k] = g_v32[k] + 640000u;
    for(size_t k = 0; k < (size_t)SOBOL_DIM * 64; k++)
        g_sv64[k] = g_v64[k] + 1280000ull;
    /* scramble constants: default-seeded std::mt19937 through uniform_int_distribution over the
       full range: one draw per 32-bit constant, two (high word first) per 64-bit constant */
    mt_std g;
    mt_std_seed(&g, 5489u);
    for(int k = 0; k < SOBOL_DIM; k++)
        g_c32[k] = mt_std_next(&g);
    mt_std_seed(&g, 5489u);
    for(int k = 0; k < SOBOL_DIM; k++)
    {
        const uint64_t hi = mt_std_next(&g);
        const uint64_t lo = mt_std_next(&g);
        g_c64[k]          = (hi << 32) + lo;
    }
    g_sobol_ready = 1;
    return 0;
}

const unsigned int*       orc_sobol32_vectors(void) { return g_v32; }
const unsigned long long* orc_sobol64_vectors(void) { return (const unsigned long long*)g_v64; }
const unsigned int*       orc_scrambled_sobol32_vectors(void) { return g_sv32; }
const unsigned long long* orc_scrambled_sobol64_vectors(void) { return (const unsigned long long*)g_sv64; }
const unsigned int*       orc_scramble32_constants(void) { return g_c32; }
const unsigned long long* orc_scramble64_constants(void) { return (const unsigned long long*)g_c64; }

/* ------------------------------------------------------------------------ */
/* MTGP32 parameters                                                        */
/* ------------------------------------------------------------------------ */
#define MTGP_N     351
#define MTGP_MASK  1023u
#define MTGP_GENS  512
#define MTGP_BLOCK 256

typedef struct
{
    uint32_t pos, sh1, sh2, tbl[16], tmp[16], mask;
} mtgp_param;
static mtgp_param g_mtgp[MTGP_GENS];
static int        g_mtgp_ready;

int orc_mtgp32_load(const char* path)
{
    FILE* f = fopen(path, "rb");
    if(!f)
        return -1;
    uint32_t w[MTGP_GENS * 36];
    const size_t got = fread(w, 4, MTGP_GENS * 36, f);
    fclose(f);
    if(got != MTGP_GENS * 36)
        return -1;
    for(int i = 0; i < MTGP_GENS; i++)
    {
        const uint32_t* p = w + i * 36;
        g_mtgp[i].pos = p[0];
        g_mtgp[i].sh1 = p[1];
        g_mtgp[i].sh2 = p[2];
        memcpy(g_mtgp[i].tbl, p + 3, 64);
        memcpy(g_mtgp[i].tmp, p + 19, 64);
        g_mtgp[i].mask = p[35];
    }
    g_mtgp_ready = 1;
    return 0;
}

typedef struct
{
    uint32_t status[1024];
    uint32_t offset;
} mtgp_state;

/* rocrand_mtgp32.h:130-145 */
static void mtgp_init_state(uint32_t* array, const mtgp_param* p, uint32_t seed)
{
    const int size        = 11213 / 32 + 1;
    const uint32_t hidden = p->tbl[4] ^ (p->tbl[8] << 16);
    uint32_t tmp          = hidden;
    tmp += tmp >> 16;
    tmp += tmp >> 8;
    memset(array, (int)(tmp & 0xff), sizeof(uint32_t) * size);
    array[0] = seed;
    array[1] = hidden;
    for(int i = 1; i < size; i++)
        array[i] ^= 1812433253u * (array[i - 1] ^ (array[i - 1] >> 30)) + (uint32_t)i;
}

/* rocrand_mtgp32.h:272-317: one lane of a block-step */
static uint32_t mtgp_next_thread(mtgp_state* s, const mtgp_param* p, uint32_t t)
{
    const uint32_t off = s->offset;
    /* every generator uses the first parameter set's mask (rocrand_mtgp32.h:381) */
    uint32_t X = (s->status[(t + off) & MTGP_MASK] & g_mtgp[0].mask) ^ s->status[(t + off + 1) & MTGP_MASK];
    uint32_t Y = s->status[(t + off + p->pos) & MTGP_MASK];
    X ^= X << p->sh1;
    Y = X ^ (Y >> p->sh2);
    const uint32_t r = Y ^ p->tbl[Y & 0x0f];
    s->status[(t + off + MTGP_N) & MTGP_MASK] = r;
    uint32_t T = s->status[(t + off + p->pos - 1) & MTGP_MASK];
    T ^= T >> 16;
    T ^= T >> 8;
    return r ^ p->tmp[T & 0x0f];
}

/* one block-step: 256 values (thread t's value at out[t]); offset advances by 256 */
static void mtgp_step(mtgp_state* s, const mtgp_param* p, uint32_t* out)
{
    for(uint32_t t = 0; t < MTGP_BLOCK; t++)
        out[t] = mtgp_next_thread(s, p, t);
    s->offset = (s->offset + MTGP_BLOCK) & MTGP_MASK;
}

/* ------------------------------------------------------------------------ */
/* extension generators                                                     */
/* ------------------------------------------------------------------------ */
struct orc_ext
{
    int         type;
    mtgp_state* mtgp; /* [512] */
    uint32_t    sobol_offset;
    void*       mt; /* MT19937 (orc_mt19937.c) */
};

int      orc_gen_type(const orc_gen* g);
uint64_t orc_gen_seed(const orc_gen* g);
uint64_t orc_gen_offset(const orc_gen* g);
unsigned orc_gen_dims(const orc_gen* g);
struct orc_ext* orc_gen_ext(orc_gen* g);
void     orc_gen_set_ext(orc_gen* g, struct orc_ext* e);
int*     orc_gen_initialized(orc_gen* g);

/* distribution of a dispatch-time kind on 32-bit MTGP/MT19937 words (orc_gen.c) */
int  orc_w32_dist(int kind, double a, double b, void** handle, unsigned* in_w, unsigned* out_w, unsigned* esz);
void orc_w32_apply(void* handle, const uint64_t* in, uint8_t* out);
void orc_w32_free(void* handle);

void* orc_mt19937_create(void);
void  orc_mt19937_destroy(void* m);
int   orc_mt19937_generate(void* m, uint64_t seed, int* initialized, int kind, void* out, size_t n, uint64_t addr,
                           double a, double b);

int orc_ext_create(orc_gen* g)
{
    struct orc_ext* e = (struct orc_ext*)calloc(1, sizeof(struct orc_ext));
    e->type           = orc_gen_type(g);
    if(e->type == RNG_MTGP32)
        e->mtgp = (mtgp_state*)calloc(MTGP_GENS, sizeof(mtgp_state));
    if(e->type == RNG_MT19937)
        e->mt = orc_mt19937_create();
    orc_gen_set_ext(g, e);
    return 0;
}

void orc_ext_destroy(orc_gen* g)
{
    struct orc_ext* e = orc_gen_ext(g);
    if(!e)
        return;
    free(e->mtgp);
    if(e->mt)
        orc_mt19937_destroy(e->mt);
    free(e);
}

/* ---- MTGP32 generate: mtgp32.hpp:232-305 ---- */
static int mtgp_generate(orc_gen* g, struct orc_ext* e, int kind, uint8_t* out, size_t n, uint64_t addr, double a,
                         double b)
{
    if(!g_mtgp_ready)
        return ST_INTERNAL_ERROR;
    void*    dh;
    unsigned in_w, out_w, esz;
    int      st = orc_w32_dist(kind, a, b, &dh, &in_w, &out_w, &esz);
    if(st)
        return st;
    int* init = orc_gen_initialized(g);
    if(!*init)
    {
        uint64_t seed = orc_gen_seed(g);
        seed          = seed ^ (seed >> 32);
        for(int i = 0; i < MTGP_GENS; i++)
        {
            mtgp_init_state(e->mtgp[i].status, &g_mtgp[i], (uint32_t)seed + (uint32_t)i + 1u);
            e->mtgp[i].offset = 0;
        }
        *init = 1;
    }
    const size_t   mis   = (out_w - (addr / esz) % out_w) % out_w;
    const unsigned head  = (unsigned)(n < mis ? n : mis);
    const unsigned tail  = (unsigned)((n - head) % out_w);
    const size_t   vec_n = (n - head) / out_w;
    const size_t   rem   = vec_n % MTGP_BLOCK;
    const size_t   vec_n_down = vec_n - rem;
    const size_t   vec_n_up   = rem == 0 ? vec_n_down : vec_n_down + MTGP_BLOCK;
    const size_t   stride     = (size_t)MTGP_GENS * MTGP_BLOCK;
    uint8_t*       vec_data   = out + mis * esz;
    uint32_t       words[4][MTGP_BLOCK];
    uint8_t        tmp[64];
    for(int bi = 0; bi < MTGP_GENS; bi++)
    {
        mtgp_state*       s = &e->mtgp[bi];
        const mtgp_param* p = &g_mtgp[bi];
        size_t            index = (size_t)bi * MTGP_BLOCK; /* thread 0's index; thread t adds t */
#define MTGP_GEN()                                   \
    for(unsigned i = 0; i < in_w; i++)               \
        mtgp_step(s, p, words[i]);
        while(index < vec_n_down)
        {
            MTGP_GEN();
            for(unsigned t = 0; t < MTGP_BLOCK; t++)
            {
                uint64_t in[4];
                for(unsigned i = 0; i < in_w; i++)
                    in[i] = words[i][t];
                orc_w32_apply(dh, in, vec_data + (index + t) * out_w * esz);
            }
            index += stride;
        }
        if(index < vec_n_up)
        {
            MTGP_GEN();
            for(unsigned t = 0; t < MTGP_BLOCK; t++)
            {
                if(index + t >= vec_n)
                    continue;
                uint64_t in[4];
                for(unsigned i = 0; i < in_w; i++)
                    in[i] = words[i][t];
                orc_w32_apply(dh, in, vec_data + (index + t) * out_w * esz);
            }
            index += stride;
        }
        if(out_w > 1 && (head > 0 || tail > 0))
        {
            MTGP_GEN();
            for(unsigned t = 0; t < MTGP_BLOCK; t++)
            {
                const size_t idx = index + t;
                if(idx != vec_n_up && idx != vec_n_up + 1)
                    continue;
                uint64_t in[4];
                for(unsigned i = 0; i < in_w; i++)
                    in[i] = words[i][t];
                orc_w32_apply(dh, in, tmp);
                for(unsigned o = 0; o < out_w; o++)
                {
                    if(idx == vec_n_up && o < head)
                        memcpy(out + (size_t)o * esz, tmp + o * esz, esz);
                    if(idx == vec_n_up + 1 && o < tail)
                        memcpy(out + (n - tail + o) * esz, tmp + o * esz, esz);
                }
            }
        }
#undef MTGP_GEN
    }
    orc_w32_free(dh);
    return ST_SUCCESS;
}

/* ---- Sobol generate: sobol.hpp:618-718 with the kernel body :72-261 ---- */
static float roc_f_erfinv(float x)
{
    /* rocrand_normal.h:174-197 */
    const float sgn = (x < 0.0f) ? -1.0f : 1.0f;
    x               = (1.0f - x) * (1.0f + x);
    const float lnx = logf(x);
    if(isnan(lnx))
        return 1.0f;
    else if(isinf(lnx))
        return 0.0f;
    const float tt1 = 2.0f / (R_PI_F * 0.147f) + 0.5f * lnx;
    const float tt2 = 1.0f / (0.147f) * lnx;
    return (sgn * sqrtf(-tt1 + sqrtf(tt1 * tt1 - tt2)));
}
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

/* direction value x (32- or 64-bit) -> output, sobol_*_distribution (uniform.hpp:351-442,
   normal.hpp:303-364, log_normal.hpp, poisson CDF) */
typedef struct
{
    int       kind, bits;
    float     mf, sf;
    double    md, sd;
    unsigned short mh, sh;
    int       huge;
    double    lambda, sqrt_lambda;
    unsigned  size, offset;
    double*   cdf;
} sobol_dist;

static float sobol_uniform_f32(uint64_t x, int bits)
{
    if(bits == 32)
        return R_2POW32_INV + ((float)(uint32_t)x * R_2POW32_INV);
    return R_2POW32_INV + (float)(x >> 32) * R_2POW32_INV;
}
static double sobol_uniform_f64(uint64_t x, int bits)
{
    if(bits == 32)
        return R_2POW32_INV_DBL + ((double)(uint32_t)x * R_2POW32_INV_DBL);
    return R_2POW53_INV_DBL + ((double)(x >> 11) * R_2POW53_INV_DBL);
}
static float sobol_normal_f32(uint64_t x, int bits)
{
    const float p = sobol_uniform_f32(x, bits);
    return R_SQRT2_F * roc_f_erfinv(2.0f * p - 1.0f);
}
static double sobol_normal_f64(uint64_t x, int bits)
{
    const double p = sobol_uniform_f64(x, bits);
    return (double)R_SQRT2_F * roc_d_erfinv(2.0 * p - 1.0);
}
static unsigned discrete_cdf(double x, unsigned size, unsigned offset, const double* cdf)
{
    /* rocrand_discrete.h:111-131 */
    unsigned min = 0, max = size - 1;
    do
    {
        const unsigned center = (min + max) / 2;
        const double   p      = cdf[center];
        if(x > p)
            min = center + 1;
        else
            max = center;
    } while(min != max);
    return offset + min;
}

static void sobol_apply(const sobol_dist* d, uint64_t x, uint8_t* out)
{
    const int b = d->bits;
    switch(d->kind)
    {
        case ORC_U8: *(uint8_t*)out = (uint8_t)(x >> (b - 8)); break;
        case ORC_U16: *(uint16_t*)out = (uint16_t)(x >> (b - 16)); break;
        case ORC_U32: *(uint32_t*)out = (uint32_t)(x >> (b - 32)); break;
        case ORC_U64: *(uint64_t*)out = x; break;
        case ORC_UNIFORM_F32: *(float*)out = sobol_uniform_f32(x, b); break;
        case ORC_UNIFORM_F64: *(double*)out = sobol_uniform_f64(x, b); break;
        case ORC_UNIFORM_F16:
            *(uint16_t*)out = orc_float2half(R_2POW16_INV + ((float)(uint16_t)(x >> (b - 16)) * R_2POW16_INV));
            break;
        case ORC_NORMAL_F32: *(float*)out = d->mf + sobol_normal_f32(x, b) * d->sf; break;
        case ORC_NORMAL_F64: *(double*)out = d->md + sobol_normal_f64(x, b) * d->sd; break;
        case ORC_LOGNORMAL_F32: *(float*)out = expf(d->mf + (d->sf * sobol_normal_f32(x, b))); break;
        case ORC_LOGNORMAL_F64: *(double*)out = exp(d->md + (d->sd * sobol_normal_f64(x, b))); break;
        case ORC_NORMAL_F16:
            *(uint16_t*)out = orc_float2half(half2f(d->mh) + sobol_normal_f32(x, b) * half2f(d->sh));
            break;
        case ORC_LOGNORMAL_F16:
            *(uint16_t*)out = orc_float2half(expf(half2f(d->mh) + (half2f(d->sh) * sobol_normal_f32(x, b))));
            break;
        case ORC_POISSON:
            if(d->huge)
            {
                const double nd = (double)R_SQRT2_F * roc_d_erfinv(2.0 * sobol_uniform_f64(x, b) - 1.0);
                *(uint32_t*)out = (uint32_t)round(d->sqrt_lambda * nd + d->lambda);
            }
            else
            {
                const double r = b == 32 ? (double)(uint32_t)x * R_2POW32_INV_DBL : (double)x * R_2POW64_INV_DBL;
                *(uint32_t*)out = discrete_cdf(r, d->size, d->offset, d->cdf);
            }
            break;
    }
}

int orc_poisson_cdf(double lambda, unsigned* size, unsigned* offset, double** cdf); /* orc_gen.c */

static int sobol_generate(orc_gen* g, struct orc_ext* e, int kind, uint8_t* out, size_t n, double a, double b)
{
    if(!g_sobol_ready)
        return ST_INTERNAL_ERROR;
    const int t    = e->type;
    const int is64 = t == RNG_SOBOL64 || t == RNG_SCRAMBLED_SOBOL64;
    const int scr  = t == RNG_SCRAMBLED_SOBOL32 || t == RNG_SCRAMBLED_SOBOL64;
    if(!is64 && kind == ORC_U64)
        return ST_TYPE_ERROR;
    const unsigned dims = orc_gen_dims(g);
    if(n % dims != 0)
        return ST_LENGTH_NOT_MULTIPLE;
    int* init = orc_gen_initialized(g);
    if(!*init)
    {
        e->sobol_offset = (uint32_t)orc_gen_offset(g);
        *init           = 1;
    }
    sobol_dist d;
    memset(&d, 0, sizeof(d));
    d.kind = kind;
    d.bits = is64 ? 64 : 32;
    d.mf = (float)a; d.sf = (float)b; d.md = a; d.sd = b;
    d.mh = orc_float2half((float)a); d.sh = orc_float2half((float)b);
    if(kind == ORC_POISSON)
    {
        if(a <= 0.0)
            return ST_OUT_OF_RANGE;
        d.lambda = a;
        d.sqrt_lambda = sqrt(a);
        if(a > 4000.0)
            d.huge = 1;
        else
            orc_poisson_cdf(a, &d.size, &d.offset, &d.cdf);
    }
    const size_t esz = (kind == ORC_U8) ? 1 : (kind == ORC_U16 || kind == ORC_UNIFORM_F16 || kind == ORC_NORMAL_F16
                                                || kind == ORC_LOGNORMAL_F16) ? 2
                       : (kind == ORC_U64 || kind == ORC_UNIFORM_F64 || kind == ORC_NORMAL_F64
                          || kind == ORC_LOGNORMAL_F64) ? 8 : 4;
    const size_t size = n / dims;
    for(unsigned dim = 0; dim < dims; dim++)
    {
        /* quirk (sobol.hpp:514): the scrambled generator's per-dimension constant is
           element `dim` of the scrambled direction-vector table */
        uint64_t sc = 0;
        if(scr)
            sc = is64 ? g_sv64[dim] : (uint64_t)g_sv32[dim];
        for(size_t p = 0; p < size; p++)
        {
            const uint64_t i = is64 ? (uint64_t)(uint32_t)(e->sobol_offset + (uint32_t)p) : (uint32_t)(e->sobol_offset + (uint32_t)p);
            const uint64_t gray = i ^ (i >> 1);
            uint64_t       x    = 0;
            for(int k = 0; k < d.bits; k++)
                if((gray >> k) & 1u)
                    x ^= is64 ? (scr ? g_sv64[(size_t)dim * 64 + k] : g_v64[(size_t)dim * 64 + k])
                              : (uint64_t)(scr ? g_sv32[(size_t)dim * 32 + k] : g_v32[(size_t)dim * 32 + k]);
            x ^= sc;
            sobol_apply(&d, x, out + ((size_t)dim * size + p) * esz);
        }
    }
    free(d.cdf);
    e->sobol_offset += (uint32_t)size;
    return ST_SUCCESS;
}

int orc_ext_generate(orc_gen* g, int kind, void* out, size_t n, uint64_t addr, double a, double b)
{
    struct orc_ext* e = orc_gen_ext(g);
    switch(e->type)
    {
        case RNG_MTGP32: return mtgp_generate(g, e, kind, (uint8_t*)out, n, addr, a, b);
        case RNG_MT19937:
            return orc_mt19937_generate(e->mt, orc_gen_seed(g), orc_gen_initialized(g), kind, out, n, addr, a, b);
        default: return sobol_generate(g, e, kind, (uint8_t*)out, n, a, b);
    }
}
