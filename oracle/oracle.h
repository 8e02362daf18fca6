/*
 * oracle.h -- CPU restatement of the reference rocRAND host-API generation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * implementation in rocrand_amd/.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product never links or calls it.
 *
 * Every generator restates the reference's kernel bodies exactly as the
 * reference host system runs them (library/src/rng/system.hpp:140-214: loop over
 * blocks, then threads, with the DEFAULT-ordering grid of
 * library/src/rng/config/config_defaults.hpp:33-108).  Status codes are the
 * rocrand_status values of library/include/rocrand/rocrand.h:59-72.
 *
 * Pinning (what checks each part, see tests/test_oracle_*.py):
 *   Philox4x32-10, ThreeFry{2,4}x{32,64}-20 : Random123 known-answer vectors,
 *                                             and oracle/_ref (reference sources).
 *   MRG32k3a/MRG31k3p/XORWOW/LFSR113 jumps  : brute force (k x next() == jump(k)),
 *                                             published RngStreams A^(2^76)/A^(2^127),
 *                                             and oracle/_ref.
 *   Index maps, head/tail, continuity       : oracle/_ref on misaligned/odd sizes,
 *                                             offset/continuity invariants.
 *   Uniform transforms                      : closed form (exact power-of-two scale).
 *   Normal / log-normal                     : oracle/_ref (same glibc), FP64 Box-Muller.
 *   Poisson tables                          : oracle/_ref, pmf closed form, moments.
 *   MTGP32 / MT19937 / Sobol                : oracle/_ref, std::mt19937, scipy Sobol.
 */
#ifndef ROCRAND_ORACLE_H_
#define ROCRAND_ORACLE_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Output kinds understood by orc_generate (one per rocrand_generate_* entry point). */
enum orc_kind
{
    ORC_U8 = 0,
    ORC_U16,
    ORC_U32,
    ORC_U64,
    ORC_UNIFORM_F32,
    ORC_UNIFORM_F64,
    ORC_UNIFORM_F16,
    ORC_NORMAL_F32,
    ORC_NORMAL_F64,
    ORC_NORMAL_F16,
    ORC_LOGNORMAL_F32,
    ORC_LOGNORMAL_F64,
    ORC_LOGNORMAL_F16,
    ORC_POISSON,
    ORC_KIND_COUNT
};

typedef struct orc_gen orc_gen;

/* rng_type uses the rocrand_rng_type values (400..505). NULL on unsupported type. */
orc_gen* orc_create(int rng_type);
void     orc_destroy(orc_gen* g);
int      orc_set_seed(orc_gen* g, unsigned long long seed);
int      orc_set_seed_uint4(orc_gen* g, const unsigned int seed[4]);
int      orc_set_offset(orc_gen* g, unsigned long long offset);
int      orc_set_dimensions(orc_gen* g, unsigned int dimensions);
int      orc_set_ordering(orc_gen* g, int ordering);
int      orc_set_grid(orc_gen* g, unsigned threads, unsigned blocks);

/*
 * Generate n values of `kind` into host buffer `out`.  `addr` is the address
 * the output would have on the device: the reference derives its head/tail
 * split from the pointer's alignment, so parity tests pass the real device
 * pointer here.  a/b are mean/stddev (normal, log-normal) or lambda (poisson, a).
 * Returns a rocrand_status value.
 */
int orc_generate(orc_gen* g, int kind, void* out, size_t n, uint64_t addr, double a, double b);

/* Single-engine access used by the known-answer tests. */
void orc_philox4x32_10_block(const unsigned int ctr[4], const unsigned int key[2], unsigned int out[4]);
void orc_threefry4x32_20_block(const unsigned int ctr[4], const unsigned int key[4], unsigned int out[4]);
void orc_threefry2x32_20_block(const unsigned int ctr[2], const unsigned int key[2], unsigned int out[2]);
void orc_threefry4x64_20_block(const unsigned long long ctr[4], const unsigned long long key[4], unsigned long long out[4]);
void orc_threefry2x64_20_block(const unsigned long long ctr[2], const unsigned long long key[2], unsigned long long out[2]);

/* Raw engine streams: engine(seed, subsequence, offset) then n x next(). */
int orc_engine_stream(int rng_type, unsigned long long seed, unsigned long long subsequence,
                      unsigned long long offset, unsigned int* out, size_t n);

/* Jump tables (for cross-checks): writes 9 words of A^(2^(log2)) for component c (1 or 2). */
int orc_mrg_matrix(int rng_type, int component, int which, int index, unsigned long long out[9]);

/* Poisson alias table for lambda: returns size, writes offset/prob/alias (caller buffers >= 4096). */
int orc_poisson_table(double lambda, unsigned int* offset, double* prob, unsigned int* alias);
/* Tables of rocrand_create_discrete_distribution for `size` weights: alias pair and CDF. */
int orc_discrete_tables(const double* weights, unsigned size, double* prob, unsigned* alias, double* cdf);

/* float -> half, round to nearest even (the host __float2half). */
unsigned short orc_float2half(float f);

/* Direction vectors / scramble constants (computed from the Joe-Kuo data file). */
int orc_sobol_load(const char* joe_kuo_path);
const unsigned int*       orc_sobol32_vectors(void);
const unsigned long long* orc_sobol64_vectors(void);
const unsigned int*       orc_scrambled_sobol32_vectors(void);
const unsigned long long* orc_scrambled_sobol64_vectors(void);
const unsigned int*       orc_scramble32_constants(void);
const unsigned long long* orc_scramble64_constants(void);

/* MTGP32 parameter sets (512 x {pos, sh1, sh2, tbl[16], tmp_tbl[16], mask}). */
int orc_mtgp32_load(const char* params_path);

/* MT19937 jump polynomials (510 x 624 words, see rocrand_amd/data/make_mt19937_jumps.c). */
int orc_mt19937_load(const char* jumps_path);

#ifdef __cplusplus
}
#endif

#endif
