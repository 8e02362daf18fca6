/*
 * orc_mt19937.c -- MT19937 restatement of the oracle (test infrastructure, see oracle.h).
 *
 * Follows library/src/rng/mt19937.hpp and mt19937_octo_engine.hpp:
 *   - engines (:86-274): base state init_genrand((seed >> 32) ^ seed); generator g is the
 *     base state jumped by g * 2^1000 words, applied as two radix-256 digits with the jump
 *     polynomials x^(i*2^1000), x^(256*i*2^1000) mod phi, each evaluated by Horner's rule on
 *     the 624-word sliding window exactly as the jump kernel does (one word step per
 *     coefficient, XOR of the start state where the coefficient is set);
 *   - rounds: every generator twists all 624 words (gen_next_n, octo_engine :259-405) before
 *     a round of outputs; tempering :424-433;
 *   - output map (generate_short/long_mt19937, :308-660): input number q of a round
 *     (q < 65536 * (78 / in_w)) is read by thread t = q mod 65536 (generator t / 8, lane
 *     t mod 8) from its local words (q / 65536) * in_w + i, where lane l's local word 0 is
 *     element {0,113,170,283,340,397,510,567}[l] and local word 1 + 7a + b is element
 *     {1,57,114,171,227,284,341,398,454,511,568}[a] + 7l + b (octo_engine :84-166);
 *   - a call consumes vec_size (+1 when a head or tail exists) inputs from the running
 *     input position; the extra input's outputs fill the head (data[o] = out[o]) and the
 *     tail in reverse (data[n - 1 - k] = out[out_w - tail + k]); changing the input width
 *     first moves the position to the next unused whole input of every engine (:951-964);
 *   - the jump polynomials are read from rocrand_amd/data/mt19937_jumps.bin.
 * Generators are set up lazily: only those whose lanes a call reads are jumped.
 */
#include "orc_internal.h"

#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MT_N 624
#define MT_M 397
#define MT_GENS 8192
#define MT_STRIDE 65536
#define MT_MEXP 19937

static uint32_t* g_jump; /* 510 x 624 words */

int orc_mt19937_load(const char* path)
{
    FILE* f = fopen(path, "rb");
    if(!f)
        return -1;
    uint32_t*    t   = (uint32_t*)malloc(sizeof(uint32_t) * 510 * MT_N);
    const size_t got = fread(t, 4, 510 * MT_N, f);
    fclose(f);
    if(got != 510 * MT_N)
    {
        free(t);
        return -1;
    }
    free(g_jump);
    g_jump = t;
    return 0;
}

static uint32_t comp(uint32_t a, uint32_t b, uint32_t m)
{
    const uint32_t y = (a & 0x80000000u) | (b & 0x7fffffffu);
    return m ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
}

static uint32_t temper(uint32_t y)
{
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

/* mt19937.hpp:130-160 (Horner): state <- p(A) state, one word step per coefficient */
__attribute__((target_clones("avx2", "default"))) static void horner_jump(uint32_t* state, const uint32_t* pf)
{
    uint32_t temp[MT_N];
    memset(temp, 0, sizeof(temp));
    unsigned ptr = 0;
    for(int pfi = MT_MEXP - 1; pfi >= 0; pfi--)
    {
        temp[ptr] = comp(temp[ptr], temp[(ptr + 1) % MT_N], temp[(ptr + MT_M) % MT_N]);
        ptr       = (ptr + 1) % MT_N;
        if((pf[pfi / 32] >> (pfi % 32)) & 1u)
        {
            /* temp[(ptr + k) % n] ^= state[k], as two contiguous runs */
            const unsigned run = MT_N - ptr;
            for(unsigned k = 0; k < run; k++)
                temp[ptr + k] ^= state[k];
            for(unsigned k = run; k < MT_N; k++)
                temp[k - run] ^= state[k];
        }
    }
    for(unsigned k = 0; k < MT_N; k++)
        state[k] = temp[(ptr + k) % MT_N];
}

static void twist(uint32_t* x)
{
    /* standard in-place block recurrence: x[k] <- comp(x[k], x[k+1], x[k+m]) for k = 0..n-1 */
    for(unsigned k = 0; k < MT_N; k++)
        x[k] = comp(x[k], x[(k + 1) % MT_N], x[(k + MT_M) % MT_N]);
}

typedef struct
{
    uint64_t  seed;
    uint32_t* state;  /* [MT_GENS][MT_N], valid where ready[g] */
    uint32_t* rounds; /* twists applied to generator g */
    uint8_t*  ready;
    uint32_t  round;  /* twists of the current round */
    uint32_t  start;  /* next input position within the round (0: the round is used up) */
    uint32_t  prev_in_w;
} mt_gen;

void* orc_mt19937_create(void)
{
    mt_gen* m = (mt_gen*)calloc(1, sizeof(mt_gen));
    m->state  = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)MT_GENS * MT_N);
    m->rounds = (uint32_t*)calloc(MT_GENS, sizeof(uint32_t));
    m->ready  = (uint8_t*)calloc(MT_GENS, 1);
    return m;
}

void orc_mt19937_destroy(void* p)
{
    mt_gen* m = (mt_gen*)p;
    if(!m)
        return;
    free(m->state);
    free(m->rounds);
    free(m->ready);
    free(m);
}

static void setup(mt_gen* m, uint64_t seed)
{
    m->seed = seed;
    memset(m->ready, 0, MT_GENS);
    memset(m->rounds, 0, sizeof(uint32_t) * MT_GENS);
    m->round     = 0;
    m->start     = 0;
    m->prev_in_w = 0;
}

/* jumped start state of generator g (mt19937.hpp:86-274) */
static void jump_generator(mt_gen* m, unsigned g)
{
    uint32_t* s = m->state + (size_t)g * MT_N;
    s[0]        = (uint32_t)((m->seed >> 32) ^ m->seed);
    for(unsigned i = 1; i < MT_N; i++)
        s[i] = 1812433253u * (s[i - 1] ^ (s[i - 1] >> 30)) + i;
    unsigned e = g;
    for(unsigned r = 0; r < 2; r++)
    {
        const unsigned digit = e % 256;
        e /= 256;
        if(digit)
            horner_jump(s, g_jump + (size_t)(r * 255 + digit - 1) * MT_N);
    }
    m->rounds[g] = 0;
    m->ready[g]  = 1;
}

typedef struct
{
    mt_gen*         m;
    const uint8_t*  need;
    unsigned        next;
    pthread_mutex_t lock;
} jump_work;

static void* jump_worker(void* p)
{
    jump_work* w = (jump_work*)p;
    for(;;)
    {
        pthread_mutex_lock(&w->lock);
        unsigned g = w->next;
        while(g < MT_GENS && !(w->need[g] && !w->m->ready[g]))
            g++;
        w->next = g + 1;
        pthread_mutex_unlock(&w->lock);
        if(g >= MT_GENS)
            return NULL;
        jump_generator(w->m, g);
    }
}

/* set up every needed generator (the jumps dominate the oracle's cost: spread them over threads) */
static void prepare(mt_gen* m, const uint8_t* need)
{
    jump_work w;
    w.m    = m;
    w.need = need;
    w.next = 0;
    pthread_mutex_init(&w.lock, NULL);
    enum { T = 8 };
    pthread_t th[T];
    for(int i = 0; i < T; i++)
        pthread_create(&th[i], NULL, jump_worker, &w);
    for(int i = 0; i < T; i++)
        pthread_join(th[i], NULL);
    pthread_mutex_destroy(&w.lock);
}

/* generator g's words for the current round */
static const uint32_t* gen_state(mt_gen* m, unsigned g)
{
    uint32_t* s = m->state + (size_t)g * MT_N;
    if(!m->ready[g])
        jump_generator(m, g);
    while(m->rounds[g] < m->round)
    {
        twist(s);
        m->rounds[g]++;
    }
    return s;
}

static const unsigned special[8] = {0, 113, 170, 283, 340, 397, 510, 567};
static const unsigned offsets[11] = {1, 57, 114, 171, 227, 284, 341, 398, 454, 511, 568};

static unsigned element(unsigned lane, unsigned local)
{
    if(local == 0)
        return special[lane];
    const unsigned a = (local - 1) / 7, b = (local - 1) % 7;
    return offsets[a] + 7 * lane + b;
}

int orc_w32_dist(int kind, double a, double b, void** handle, unsigned* in_w, unsigned* out_w, unsigned* esz);
void orc_w32_apply(void* handle, const uint64_t* in, uint8_t* out);
void orc_w32_free(void* handle);

int orc_mt19937_generate(void* p, uint64_t seed, int* initialized, int kind, void* outp, size_t n, uint64_t addr,
                         double a, double b)
{
    mt_gen* m = (mt_gen*)p;
    if(!g_jump)
        return ST_INTERNAL_ERROR;
    void*    dh;
    unsigned in_w, out_w, esz;
    int      st = orc_w32_dist(kind, a, b, &dh, &in_w, &out_w, &esz);
    if(st)
        return st;
    if(!*initialized)
    {
        setup(m, seed);
        *initialized = 1;
    }
    uint8_t*       out   = (uint8_t*)outp;
    const size_t   mis   = (out_w - (addr / esz) % out_w) % out_w;
    const unsigned head  = (unsigned)(n < mis ? n : mis);
    const unsigned tail  = (unsigned)((n - head) % out_w);
    const size_t   vec_n = (n - head) / out_w;
    const unsigned extra = (head > 0 || tail > 0) ? 1 : 0;
    const unsigned ips   = (MT_N / 8) / in_w;
    const unsigned full  = MT_STRIDE * ips;

    if(m->prev_in_w != in_w && m->start > 0)
    {
        const unsigned used = (m->start + MT_STRIDE - 1) / MT_STRIDE * m->prev_in_w;
        m->start            = (used + in_w - 1) / in_w * MT_STRIDE;
        if(m->start >= full)
            m->start = 0;
    }
    const size_t total = vec_n + extra;
    {
        uint8_t* need = (uint8_t*)calloc(MT_GENS, 1);
        if(total >= MT_STRIDE)
            memset(need, 1, MT_GENS);
        else
            for(size_t v = 0; v < total; v++)
                need[((m->start + v) % full % MT_STRIDE) / 8] = 1;
        prepare(m, need);
        free(need);
    }
    uint8_t tmp[64];
    for(size_t v = 0; v < total; v++)
    {
        if(m->start == 0)
            m->round++; /* a fresh round: every generator twists first */
        const unsigned q = m->start;
        const unsigned j = q / MT_STRIDE, t = q % MT_STRIDE;
        const uint32_t* s = gen_state(m, t / 8);
        uint64_t        in[4];
        for(unsigned i = 0; i < in_w; i++)
            in[i] = temper(s[element(t % 8, j * in_w + i)]);
        if(v < vec_n)
            orc_w32_apply(dh, in, out + (mis + v * out_w) * esz);
        else
        {
            orc_w32_apply(dh, in, tmp);
            for(unsigned o = 0; o < out_w; o++)
            {
                if(o < head)
                    memcpy(out + (size_t)o * esz, tmp + (size_t)o * esz, esz);
                if(o + tail >= out_w) /* o > out_w - tail - 1 */
                    memcpy(out + (n + (out_w - tail - 1) - o) * esz, tmp + (size_t)o * esz, esz);
            }
        }
        m->start = (q + 1) % full;
    }
    m->prev_in_w = in_w;
    orc_w32_free(dh);
    return ST_SUCCESS;
}
