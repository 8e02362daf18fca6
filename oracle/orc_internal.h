/* orc_internal.h -- shared types of the oracle (test infrastructure, see oracle.h). */
#ifndef ROCRAND_ORACLE_INTERNAL_H_
#define ROCRAND_ORACLE_INTERNAL_H_

#include "oracle.h"

#include <stdint.h>

/* rocrand_rng_type values, library/include/rocrand/rocrand.h:77-101 */
enum
{
    RNG_XORWOW          = 401,
    RNG_MRG32K3A        = 402,
    RNG_MTGP32          = 403,
    RNG_PHILOX4_32_10   = 404,
    RNG_MRG31K3P        = 405,
    RNG_LFSR113         = 406,
    RNG_MT19937         = 407,
    RNG_THREEFRY2_32_20 = 408,
    RNG_THREEFRY2_64_20 = 409,
    RNG_THREEFRY4_32_20 = 410,
    RNG_THREEFRY4_64_20 = 411,
    RNG_SOBOL32         = 501,
    RNG_SCRAMBLED_SOBOL32 = 502,
    RNG_SOBOL64         = 504,
    RNG_SCRAMBLED_SOBOL64 = 505
};

/* rocrand_status values, rocrand.h:59-72 */
enum
{
    ST_SUCCESS             = 0,
    ST_NOT_CREATED         = 101,
    ST_ALLOCATION_FAILED   = 102,
    ST_TYPE_ERROR          = 103,
    ST_OUT_OF_RANGE        = 104,
    ST_LENGTH_NOT_MULTIPLE = 105,
    ST_INTERNAL_ERROR      = 108
};

enum
{
    CE_PHILOX   = 1,
    CE_THREEFRY = 2
};

/* Counter-based engine state (Philox / ThreeFry); 32-bit words stored in u64 slots. */
typedef struct
{
    int      family, dim, w64;
    uint64_t ctr[4], key[4], res[4];
    unsigned substate;
} ctr_engine;

void     ce_init(ctr_engine* e, int rng_type, uint64_t seed, uint64_t sub, uint64_t offset);
void     ce_discard(ctr_engine* e, uint64_t offset);
uint64_t ce_next(ctr_engine* e);
void     ce_next_leap(ctr_engine* e, uint32_t leap, uint64_t out[4]);

typedef struct
{
    int      type;
    uint32_t g1[3], g2[3];
} mrg_engine;

void     mrg_seed_only(mrg_engine* e, int type, uint64_t seed);
void     mrg_init(mrg_engine* e, int type, uint64_t seed, uint64_t sub, uint64_t offset);
void     mrg_discard(mrg_engine* e, uint64_t n);
void     mrg_discard_subsequence(mrg_engine* e, uint64_t n);
uint32_t mrg_next(mrg_engine* e);

typedef struct
{
    uint32_t d;
    uint32_t x[5];
} xorwow_engine;

void     xorwow_seed_only(xorwow_engine* e, uint64_t seed);
void     xorwow_discard(xorwow_engine* e, uint64_t n);
void     xorwow_discard_subsequence(xorwow_engine* e, uint64_t n);
void     xorwow_apply(xorwow_engine* e, const uint32_t* m);
uint32_t xorwow_next(xorwow_engine* e);

typedef struct
{
    uint32_t z[4];
} lfsr_engine;

void     lfsr_seed_only(lfsr_engine* e, const uint32_t seed[4]);
void     lfsr_discard(lfsr_engine* e, uint64_t n);
void     lfsr_discard_subsequence(lfsr_engine* e, uint32_t n);
void     lfsr_apply(lfsr_engine* e, const uint32_t* m);
uint32_t lfsr_next(lfsr_engine* e);

/* GF(2) jump matrix k of XORWOW / LFSR113 (seq=1: subsequence matrices). */
const uint32_t* orc_gf2_matrix(int rng_type, int seq, int k);

#endif
