/* The reference's device API (library/include/rocrand/rocrand_kernel.h), compiled
 * host-only from the reference tree and driven by tests/devapi/program.h: its host
 * branch run over the same per-thread call sequences as the GPU parity tests.  Linked
 * into oracle/_ref/libref.so by build_ref.sh.  Test infrastructure only.
 *
 * As in ref_harness.cpp, __HIP_PLATFORM_AMD__ is undefined once the HIP headers are in:
 * it selects `result.data[...]`, a member current HIP vector types no longer have, over
 * the reference's portable `(&result.x)[...]` branch (the same arithmetic). */
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstring>
#undef __HIP_PLATFORM_AMD__

#include <rocrand/rocrand_kernel.h>

#include "../tests/devapi/program.h"

extern "C" int ref_devapi_run(int engine, int fn, unsigned long long seed, const unsigned long long* sub,
                              const unsigned long long* off, int n, int calls, double a, double b,
                              unsigned int dsize, unsigned int doffset, const double* prob,
                              const unsigned int* alias, const double* cdf, const unsigned int* v32,
                              const unsigned long long* v64, unsigned int dims, unsigned long long* out)
{
    if(n < 0 || calls < 0 || ((fn == devapi_test::DISCRETE || fn == devapi_test::DISCRETE4) && dsize == 0))
        return -2;
    rocrand_discrete_distribution_st d;
    d.size        = dsize;
    d.offset      = doffset;
    d.alias       = const_cast<unsigned int*>(alias);
    d.probability = const_cast<double*>(prob);
    d.cdf         = const_cast<double*>(cdf);
    const devapi_test::sobol_tables tables = {v32, v64, dims == 0 ? 1u : dims};
    if(engine >= devapi_test::SOBOL32 && (dims == 0 || !v32 || !v64))
        return -2;
    return devapi_test::run_host(engine, fn, seed, sub, off, n, calls, a, b, &d, tables, (uint64_t*)out);
}

// Bytes of the reference's state object after rocrand_init, as 32-bit words.
extern "C" int ref_devapi_state(int engine, unsigned long long seed, const unsigned long long* sub,
                                const unsigned long long* off, int n, int words, unsigned int* out)
{
    int size = -1;
    const devapi_test::sobol_tables none = {nullptr, nullptr, 1};
    if(engine >= devapi_test::SOBOL32)
        return -1;
    devapi_test::with_engine(engine,
                             [&](auto* tag)
                             {
                                 using S = std::remove_pointer_t<decltype(tag)>;
                                 for(int t = 0; t < n; t++)
                                     devapi_test::state_words<S>(seed, sub[t], off[t], none, out + (size_t)t * words,
                                                                 words);
                                 size = (int)sizeof(S);
                             });
    return size;
}
