/*
 * rocrand_mtgp32_11213.h -- the 512 MTGP32 parameter sets for Mersenne exponent 11213
 * (the MTGP authors' dynamic-creator output, as in the reference's
 * library/include/rocrand/rocrand_mtgp32_11213.h:63), for rocrand_make_state_mtgp32 and
 * rocrand_make_constant.  The array lives in librocrand.so (link with -lrocrand).
 */
#ifndef ROCRAND_MTGP32_11213_H_
#define ROCRAND_MTGP32_11213_H_

#include "rocrand_kernel.h"

extern "C" const mtgp32_params_fast_t mtgp32dc_params_fast_11213[MTGP_BN_MAX];

#endif // ROCRAND_MTGP32_11213_H_
