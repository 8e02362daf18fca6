/*
 * rocrand.h -- C ABI of the MI355X-native rocRAND bulk-generation library.
 *
 * Drop-in for the reference header library/include/rocrand/rocrand.h: the
 * opaque handle, every enum value and every entry point keep the reference's
 * names, types and status semantics, so code and bindings written against the
 * reference (hipRAND's AMD backend, rocrand.hpp, the Fortran bind(C) module,
 * the Python ctypes wrapper) link against librocrand.so built here unchanged.
 * Each declaration cites the reference declaration it replaces.
 */
#ifndef ROCRAND_H_
#define ROCRAND_H_

#include "rocrand_discrete_types.h"

#include <hip/hip_fp16.h>
#include <hip/hip_runtime_api.h>
#include <hip/hip_vector_types.h>

#include <stddef.h>

#ifndef ROCRANDAPI
    #define ROCRANDAPI __attribute__((visibility("default")))
#endif

/* Version: 3.2.0 feature level of the reference (rocrand_version.h.in). */
#define ROCRAND_VERSION_MAJOR 3
#define ROCRAND_VERSION_MINOR 2
#define ROCRAND_VERSION_PATCH 0
#define ROCRAND_VERSION (ROCRAND_VERSION_MAJOR * 100000 + ROCRAND_VERSION_MINOR * 100 + ROCRAND_VERSION_PATCH)

/* rocrand.h:41 */
typedef struct rocrand_generator_base_type* rocrand_generator;
/* rocrand.h:46 */
typedef __half half;

#define ROCRAND_DEFAULT_MAX_BLOCK_SIZE 256

#if defined(__cplusplus)
extern "C" {
#endif

/* rocrand.h:59-72 */
typedef enum rocrand_status
{
    ROCRAND_STATUS_SUCCESS                   = 0,
    ROCRAND_STATUS_VERSION_MISMATCH          = 100,
    ROCRAND_STATUS_NOT_CREATED               = 101,
    ROCRAND_STATUS_ALLOCATION_FAILED         = 102,
    ROCRAND_STATUS_TYPE_ERROR                = 103,
    ROCRAND_STATUS_OUT_OF_RANGE              = 104,
    ROCRAND_STATUS_LENGTH_NOT_MULTIPLE       = 105,
    ROCRAND_STATUS_DOUBLE_PRECISION_REQUIRED = 106,
    ROCRAND_STATUS_LAUNCH_FAILURE            = 107,
    ROCRAND_STATUS_INTERNAL_ERROR            = 108
} rocrand_status;

/* rocrand.h:77-101 */
typedef enum rocrand_rng_type
{
    ROCRAND_RNG_PSEUDO_DEFAULT          = 400,
    ROCRAND_RNG_PSEUDO_XORWOW           = 401,
    ROCRAND_RNG_PSEUDO_MRG32K3A         = 402,
    ROCRAND_RNG_PSEUDO_MTGP32           = 403,
    ROCRAND_RNG_PSEUDO_PHILOX4_32_10    = 404,
    ROCRAND_RNG_PSEUDO_MRG31K3P         = 405,
    ROCRAND_RNG_PSEUDO_LFSR113          = 406,
    ROCRAND_RNG_PSEUDO_MT19937          = 407,
    ROCRAND_RNG_PSEUDO_THREEFRY2_32_20  = 408,
    ROCRAND_RNG_PSEUDO_THREEFRY2_64_20  = 409,
    ROCRAND_RNG_PSEUDO_THREEFRY4_32_20  = 410,
    ROCRAND_RNG_PSEUDO_THREEFRY4_64_20  = 411,
    ROCRAND_RNG_QUASI_DEFAULT           = 500,
    ROCRAND_RNG_QUASI_SOBOL32           = 501,
    ROCRAND_RNG_QUASI_SCRAMBLED_SOBOL32 = 502,
    ROCRAND_RNG_QUASI_SOBOL64           = 504,
    ROCRAND_RNG_QUASI_SCRAMBLED_SOBOL64 = 505
} rocrand_rng_type;

/* rocrand.h:106-115 */
typedef enum rocrand_ordering
{
    ROCRAND_ORDERING_PSEUDO_BEST    = 100,
    ROCRAND_ORDERING_PSEUDO_DEFAULT = 101,
    ROCRAND_ORDERING_PSEUDO_SEEDED  = 102,
    ROCRAND_ORDERING_PSEUDO_LEGACY  = 103,
    ROCRAND_ORDERING_PSEUDO_DYNAMIC = 104,
    ROCRAND_ORDERING_QUASI_DEFAULT  = 201
} rocrand_ordering;

/* rocrand.h:120-126 */
typedef enum rocrand_direction_vector_set
{
    ROCRAND_DIRECTION_VECTORS_32_JOEKUO6           = 101,
    ROCRAND_SCRAMBLED_DIRECTION_VECTORS_32_JOEKUO6 = 102,
    ROCRAND_DIRECTION_VECTORS_64_JOEKUO6           = 103,
    ROCRAND_SCRAMBLED_DIRECTION_VECTORS_64_JOEKUO6 = 104
} rocrand_direction_vector_set;

/* rocrand.h:163-164: device generator (output in device memory, async on its stream). */
rocrand_status ROCRANDAPI rocrand_create_generator(rocrand_generator* generator, rocrand_rng_type rng_type);
/* rocrand.h:185-186: host generator (output in host memory, ordered on its stream). */
rocrand_status ROCRANDAPI rocrand_create_generator_host(rocrand_generator* generator, rocrand_rng_type rng_type);
/* rocrand.h:206-207: host generator that returns when the output is written. */
rocrand_status ROCRANDAPI rocrand_create_generator_host_blocking(rocrand_generator* generator,
                                                                 rocrand_rng_type   rng_type);
/* rocrand.h:221 */
rocrand_status ROCRANDAPI rocrand_destroy_generator(rocrand_generator generator);

/* rocrand.h:241-243 */
rocrand_status ROCRANDAPI rocrand_generate(rocrand_generator generator, unsigned int* output_data, size_t n);
/* rocrand.h:264-266 */
rocrand_status ROCRANDAPI rocrand_generate_long_long(rocrand_generator       generator,
                                                     unsigned long long int* output_data,
                                                     size_t                  n);
/* rocrand.h:286-288 */
rocrand_status ROCRANDAPI rocrand_generate_char(rocrand_generator generator, unsigned char* output_data, size_t n);
/* rocrand.h:308-310 */
rocrand_status ROCRANDAPI rocrand_generate_short(rocrand_generator generator, unsigned short* output_data, size_t n);
/* rocrand.h:330-332 */
rocrand_status ROCRANDAPI rocrand_generate_uniform(rocrand_generator generator, float* output_data, size_t n);
/* rocrand.h:352-354 */
rocrand_status ROCRANDAPI rocrand_generate_uniform_double(rocrand_generator generator, double* output_data, size_t n);
/* rocrand.h:374-376 */
rocrand_status ROCRANDAPI rocrand_generate_uniform_half(rocrand_generator generator, half* output_data, size_t n);
/* rocrand.h:398-401 */
rocrand_status ROCRANDAPI rocrand_generate_normal(rocrand_generator generator, float* output_data, size_t n,
                                                  float mean, float stddev);
/* rocrand.h:423-426 */
rocrand_status ROCRANDAPI rocrand_generate_normal_double(rocrand_generator generator, double* output_data,
                                                         size_t n, double mean, double stddev);
/* rocrand.h:448-451 */
rocrand_status ROCRANDAPI rocrand_generate_normal_half(rocrand_generator generator, half* output_data, size_t n,
                                                       half mean, half stddev);
/* rocrand.h:473-476 */
rocrand_status ROCRANDAPI rocrand_generate_log_normal(rocrand_generator generator, float* output_data, size_t n,
                                                      float mean, float stddev);
/* rocrand.h:498-501 */
rocrand_status ROCRANDAPI rocrand_generate_log_normal_double(rocrand_generator generator, double* output_data,
                                                             size_t n, double mean, double stddev);
/* rocrand.h:523-526 */
rocrand_status ROCRANDAPI rocrand_generate_log_normal_half(rocrand_generator generator, half* output_data,
                                                           size_t n, half mean, half stddev);
/* rocrand.h:545-548 */
rocrand_status ROCRANDAPI rocrand_generate_poisson(rocrand_generator generator, unsigned int* output_data,
                                                   size_t n, double lambda);

/* rocrand.h:561 */
rocrand_status ROCRANDAPI rocrand_initialize_generator(rocrand_generator generator);
/* rocrand.h:575 */
rocrand_status ROCRANDAPI rocrand_set_stream(rocrand_generator generator, hipStream_t stream);
/* rocrand.h:596 */
rocrand_status ROCRANDAPI rocrand_set_seed(rocrand_generator generator, unsigned long long seed);
/* rocrand.h:618 */
rocrand_status ROCRANDAPI rocrand_set_seed_uint4(rocrand_generator generator, uint4 seed);
/* rocrand.h:642 */
rocrand_status ROCRANDAPI rocrand_set_offset(rocrand_generator generator, unsigned long long offset);
/* rocrand.h:668 */
rocrand_status ROCRANDAPI rocrand_set_ordering(rocrand_generator generator, rocrand_ordering order);
/* rocrand.h:688-690 */
rocrand_status ROCRANDAPI rocrand_set_quasi_random_generator_dimensions(rocrand_generator generator,
                                                                        unsigned int      dimensions);
/* rocrand.h:700 */
rocrand_status ROCRANDAPI rocrand_get_version(int* version);

/* rocrand.h:717-718 */
rocrand_status ROCRANDAPI rocrand_create_poisson_distribution(double                         lambda,
                                                              rocrand_discrete_distribution* discrete_distribution);
/* rocrand.h:744-747 */
rocrand_status ROCRANDAPI rocrand_create_discrete_distribution(const double*                  probabilities,
                                                               unsigned int                   size,
                                                               unsigned int                   offset,
                                                               rocrand_discrete_distribution* discrete_distribution);
/* rocrand.h:759 */
rocrand_status ROCRANDAPI rocrand_destroy_discrete_distribution(rocrand_discrete_distribution discrete_distribution);

/* rocrand.h:775-776 */
rocrand_status ROCRANDAPI rocrand_get_direction_vectors32(const unsigned int**         vectors,
                                                          rocrand_direction_vector_set set);
/* rocrand.h:792-793 */
rocrand_status ROCRANDAPI rocrand_get_direction_vectors64(const unsigned long long**   vectors,
                                                          rocrand_direction_vector_set set);
/* rocrand.h:807 */
rocrand_status ROCRANDAPI rocrand_get_scramble_constants32(const unsigned int** constants);
/* rocrand.h:821 */
rocrand_status ROCRANDAPI rocrand_get_scramble_constants64(const unsigned long long** constants);

#if defined(__cplusplus)
}
#endif

#endif /* ROCRAND_H_ */
