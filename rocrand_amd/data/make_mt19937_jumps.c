/*
 * make_mt19937_jumps.c -- jump polynomials of MT19937 for the generator's engine set-up.
 *
 * Writes mt19937_jumps.bin: 510 polynomials of 624 little-endian uint32 words each,
 * bit j of the polynomial (word j / 32, bit j % 32) being the coefficient of x^j:
 *     x^(i * 2^1000)       mod phi(x),  i = 1..255   (polynomials 0..254)
 *     x^(256 * i * 2^1000) mod phi(x),  i = 1..255   (polynomials 255..509)
 * phi is the minimal polynomial (degree 19937) of the MT19937 word recurrence, found
 * with the Berlekamp-Massey algorithm on one bit of the output sequence.  For a state
 * s (624 consecutive words), p(A) s with p = x^J mod phi is the state J words later
 * (up to the unused low 31 bits of its oldest word).  This is the layout of the
 * reference's rocrand_h_mt19937_jump (rocrand_mt19937_precomputed.h:26-31).
 *
 *     gcc -O2 -o make_mt19937_jumps make_mt19937_jumps.c && ./make_mt19937_jumps [out.bin]
 *     ./make_mt19937_jumps --power N     prints x^N mod phi as 624 words (for checks)
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MEXP 19937
#define NW 312 /* 64-bit words for degree < 19968 */
#define N 624
#define M 397

typedef struct
{
    uint64_t w[2 * NW + 2];
} poly; /* room for products before reduction */

static uint64_t phi[NW + 1]; /* bits 0..19937 */

static int  getbit(const uint64_t* a, int i) { return (int)((a[i >> 6] >> (i & 63)) & 1u); }
static void flipbit(uint64_t* a, int i) { a[i >> 6] ^= 1ull << (i & 63); }

/* one bit of the MT19937 word sequence x_k, k >= 1, from init_genrand(5489) */
static void mt_bits(uint8_t* out, int count)
{
    uint32_t x[N + 1];
    x[0] = 5489u;
    for(int i = 1; i < N; i++)
        x[i] = 1812433253u * (x[i - 1] ^ (x[i - 1] >> 30)) + (uint32_t)i;
    /* sliding window: x[ptr..ptr+623]; emit bit 0 of each new word */
    uint32_t* ring = x;
    int       ptr  = 0;
    for(int k = 0; k < count; k++)
    {
        const uint32_t y  = (ring[ptr] & 0x80000000u) | (ring[(ptr + 1) % N] & 0x7fffffffu);
        const uint32_t nw = ring[(ptr + M) % N] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        ring[ptr]         = nw;
        ptr               = (ptr + 1) % N;
        out[k]            = (uint8_t)(nw & 1u);
    }
}

/* Berlekamp-Massey over GF(2); returns the connection polynomial's degree */
static int berlekamp_massey(const uint8_t* s, int n, uint64_t* C /* NW+1 words */)
{
    const int W = NW + 1;
    uint64_t* B = calloc(W, 8);
    uint64_t* T = calloc(W, 8);
    uint64_t* R = calloc(W, 8); /* R bit i = s[n - i] */
    memset(C, 0, W * 8);
    C[0] = 1;
    B[0] = 1;
    int L = 0, m = 1;
    for(int k = 0; k < n; k++)
    {
        /* shift R left by one and insert s[k] */
        for(int i = W - 1; i > 0; i--)
            R[i] = (R[i] << 1) | (R[i - 1] >> 63);
        R[0] = (R[0] << 1) | s[k];
        uint64_t acc = 0;
        for(int i = 0; i < W; i++)
            acc ^= C[i] & R[i];
        const int d = __builtin_parityll(acc);
        if(!d)
        {
            m++;
            continue;
        }
        memcpy(T, C, W * 8);
        /* C ^= B << m */
        const int ws = m >> 6, bs = m & 63;
        for(int i = W - 1; i >= ws; i--)
        {
            uint64_t v = B[i - ws] << bs;
            if(bs && i - ws - 1 >= 0)
                v |= B[i - ws - 1] >> (64 - bs);
            C[i] ^= v;
        }
        if(2 * L <= k)
        {
            L = k + 1 - L;
            memcpy(B, T, W * 8);
            m = 1;
        }
        else
            m++;
    }
    free(B);
    free(T);
    free(R);
    return L;
}

/* a mod phi, a has bits up to `top` */
static void reduce(uint64_t* a, int top)
{
    for(int i = top; i >= MEXP; i--)
    {
        if(!getbit(a, i))
            continue;
        const int sh = i - MEXP, ws = sh >> 6, bs = sh & 63;
        for(int j = 0; j <= NW; j++)
        {
            const uint64_t v = phi[j];
            if(!v)
                continue;
            a[j + ws] ^= v << bs;
            if(bs)
                a[j + ws + 1] ^= v >> (64 - bs);
        }
    }
}

static void mulmod(const uint64_t* a, const uint64_t* b, uint64_t* out)
{
    static uint64_t sh[64][NW + 2];
    uint64_t        r[2 * NW + 4];
    memset(r, 0, sizeof(r));
    for(int s = 0; s < 64; s++)
    {
        sh[s][0] = b[0] << s;
        for(int j = 1; j < NW + 1; j++)
            sh[s][j] = (b[j] << s) | (s ? b[j - 1] >> (64 - s) : 0);
        sh[s][NW + 1] = s ? b[NW] >> (64 - s) : 0;
    }
    for(int i = 0; i < NW + 1; i++)
    {
        uint64_t w = a[i];
        while(w)
        {
            const int s = __builtin_ctzll(w);
            w &= w - 1;
            for(int j = 0; j < NW + 2; j++)
                r[i + j] ^= sh[s][j];
        }
    }
    reduce(r, 2 * MEXP);
    memcpy(out, r, (NW + 1) * 8);
}

static void sqrmod(const uint64_t* a, uint64_t* out)
{
    uint64_t r[2 * NW + 4];
    memset(r, 0, sizeof(r));
    for(int i = 0; i < MEXP; i++)
        if(getbit(a, i))
            flipbit(r, 2 * i);
    reduce(r, 2 * MEXP);
    memcpy(out, r, (NW + 1) * 8);
}

static void x_pow(unsigned long long e, uint64_t* out)
{
    uint64_t r[NW + 1], t[NW + 1];
    memset(r, 0, sizeof(r));
    r[0] = 1;
    for(int b = 63; b >= 0; b--)
    {
        sqrmod(r, t);
        memcpy(r, t, sizeof(r));
        if((e >> b) & 1u)
        {
            /* multiply by x */
            uint64_t s[NW + 2];
            memset(s, 0, sizeof(s));
            for(int j = 0; j < NW + 1; j++)
            {
                s[j] |= r[j] << 1;
                s[j + 1] |= r[j] >> 63;
            }
            reduce(s, MEXP);
            memcpy(r, s, sizeof(r));
        }
    }
    memcpy(out, r, sizeof(r));
}

static void emit(FILE* f, const uint64_t* p)
{
    uint32_t w[N];
    for(int j = 0; j < N; j++)
        w[j] = (uint32_t)(p[j >> 1] >> (32 * (j & 1)));
    fwrite(w, 4, N, f);
}

int main(int argc, char** argv)
{
    const int nbits = 2 * MEXP + 100;
    uint8_t*  s     = malloc(nbits);
    mt_bits(s, nbits);
    uint64_t C[NW + 1];
    const int L = berlekamp_massey(s, nbits, C);
    free(s);
    if(L != MEXP)
    {
        fprintf(stderr, "unexpected linear complexity %d\n", L);
        return 1;
    }
    /* characteristic polynomial: reciprocal of the connection polynomial */
    memset(phi, 0, sizeof(phi));
    for(int i = 0; i <= MEXP; i++)
        if(getbit(C, i))
            flipbit(phi, MEXP - i);

    if(argc == 3 && strcmp(argv[1], "--power") == 0)
    {
        uint64_t p[NW + 1];
        x_pow(strtoull(argv[2], NULL, 10), p);
        for(int j = 0; j < N; j++)
            printf("%u\n", (uint32_t)(p[j >> 1] >> (32 * (j & 1))));
        return 0;
    }

    const char* path = argc > 1 ? argv[1] : "mt19937_jumps.bin";
    FILE*       f    = fopen(path, "wb");
    if(!f)
        return 1;
    uint64_t base[NW + 1], q[NW + 1], cur[NW + 1], t[NW + 1];
    /* x^(2^1000) */
    memset(base, 0, sizeof(base));
    base[0] = 2; /* x */
    for(int i = 0; i < 1000; i++)
    {
        sqrmod(base, t);
        memcpy(base, t, sizeof(t));
    }
    memcpy(cur, base, sizeof(cur));
    for(int i = 1; i < 256; i++)
    {
        emit(f, cur);
        mulmod(cur, base, t);
        memcpy(cur, t, sizeof(t));
    }
    /* cur = base^256 = x^(256 * 2^1000) */
    memcpy(q, cur, sizeof(q));
    for(int i = 1; i < 256; i++)
    {
        emit(f, cur);
        mulmod(cur, q, t);
        memcpy(cur, t, sizeof(t));
    }
    fclose(f);
    fprintf(stderr, "wrote %s\n", path);
    return 0;
}
