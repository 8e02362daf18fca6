/* Host branch of include/rocrand/rocrand_kernel.h driven by program.h: the same calls as
 * the GPU kernels of devapi_gpu.hip, run on the CPU with the host libm. */
#include <rocrand/rocrand_kernel.h>

#include "program.h"

extern "C" int devapi_host_run(int engine, int fn, unsigned long long seed, const unsigned long long* sub,
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

// State bytes after rocrand_init for thread t = (seed, sub[t], off[t]), `words` 32-bit
// words per thread; returns sizeof(state) or -1.
extern "C" int devapi_host_state(int engine, unsigned long long seed, const unsigned long long* sub,
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
