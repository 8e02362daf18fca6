/* Device branch of include/rocrand/rocrand_kernel.h driven by program.h: one GPU thread
 * per call sequence, outputs copied back to the caller's host buffer. */
#include <rocrand/rocrand_kernel.h>

#include "program.h"

#include <hip/hip_runtime.h>

namespace
{

constexpr int BLOCK = 128;

template<class S, int FN>
__global__ __launch_bounds__(BLOCK) void devapi_kernel(unsigned long long seed, const unsigned long long* sub,
                                                       const unsigned long long* off, int n, int calls, double a,
                                                       double b, rocrand_discrete_distribution dist,
                                                       devapi_test::sobol_tables tables, uint64_t* out)
{
    const int t = blockIdx.x * BLOCK + threadIdx.x;
    if(t >= n)
        return;
    devapi_test::run_thread<S, FN>(seed, sub[t], off[t], calls, a, b, dist, tables, out + (size_t)t * calls * 4);
}

struct device_buffers
{
    void* p[9] = {};
    ~device_buffers()
    {
        for(void* q : p)
            if(q)
                (void)hipFree(q);
    }
};

} // namespace

extern "C" int devapi_gpu_run(int engine, int fn, unsigned long long seed, const unsigned long long* sub,
                              const unsigned long long* off, int n, int calls, double a, double b,
                              unsigned int dsize, unsigned int doffset, const double* prob,
                              const unsigned int* alias, const double* cdf, const unsigned int* v32,
                              const unsigned long long* v64, unsigned int dims, unsigned long long* out)
{
    const bool discrete = fn == devapi_test::DISCRETE || fn == devapi_test::DISCRETE4;
    const bool sobol    = engine >= devapi_test::SOBOL32;
    if(n <= 0 || calls <= 0 || (size_t)n * calls > (1u << 26)
       || (discrete && (dsize == 0 || !prob || !alias || !cdf)) || (sobol && (dims == 0 || !v32 || !v64)))
        return -2;
    bool known = false;
    devapi_test::with_engine(engine,
                             [&](auto* tag)
                             {
                                 using S = std::remove_pointer_t<decltype(tag)>;
                                 known   = devapi_test::supported<S>(fn);
                             });
    if(!known)
        return -1;

    device_buffers     buf;
    const size_t       out_bytes = (size_t)n * calls * 4 * sizeof(uint64_t);
    const unsigned int dn        = discrete ? dsize : 1;
    const unsigned int nd        = sobol ? dims : 1;
    if(hipMalloc(&buf.p[0], n * sizeof(unsigned long long)) != hipSuccess
       || hipMalloc(&buf.p[1], n * sizeof(unsigned long long)) != hipSuccess
       || hipMalloc(&buf.p[2], out_bytes) != hipSuccess || hipMalloc(&buf.p[3], dn * sizeof(double)) != hipSuccess
       || hipMalloc(&buf.p[4], dn * sizeof(unsigned int)) != hipSuccess
       || hipMalloc(&buf.p[5], sizeof(rocrand_discrete_distribution_st)) != hipSuccess
       || hipMalloc(&buf.p[6], dn * sizeof(double)) != hipSuccess
       || hipMalloc(&buf.p[7], (size_t)nd * 32 * sizeof(unsigned int)) != hipSuccess
       || hipMalloc(&buf.p[8], (size_t)nd * 64 * sizeof(unsigned long long)) != hipSuccess)
        return -3;
    rocrand_discrete_distribution_st d;
    d.size        = dn;
    d.offset      = doffset;
    d.alias       = (unsigned int*)buf.p[4];
    d.probability = (double*)buf.p[3];
    d.cdf         = (double*)buf.p[6];
    if(hipMemcpy(buf.p[0], sub, n * sizeof(unsigned long long), hipMemcpyHostToDevice) != hipSuccess
       || hipMemcpy(buf.p[1], off, n * sizeof(unsigned long long), hipMemcpyHostToDevice) != hipSuccess
       || hipMemset(buf.p[2], 0, out_bytes) != hipSuccess
       || hipMemcpy(buf.p[5], &d, sizeof(d), hipMemcpyHostToDevice) != hipSuccess)
        return -3;
    if(discrete
       && (hipMemcpy(buf.p[3], prob, dn * sizeof(double), hipMemcpyHostToDevice) != hipSuccess
           || hipMemcpy(buf.p[4], alias, dn * sizeof(unsigned int), hipMemcpyHostToDevice) != hipSuccess
           || hipMemcpy(buf.p[6], cdf, dn * sizeof(double), hipMemcpyHostToDevice) != hipSuccess))
        return -3;
    if(sobol
       && (hipMemcpy(buf.p[7], v32, (size_t)nd * 32 * sizeof(unsigned int), hipMemcpyHostToDevice) != hipSuccess
           || hipMemcpy(buf.p[8], v64, (size_t)nd * 64 * sizeof(unsigned long long), hipMemcpyHostToDevice)
                  != hipSuccess))
        return -3;
    const devapi_test::sobol_tables tables = {(const unsigned int*)buf.p[7], (const unsigned long long*)buf.p[8], nd};

    const unsigned int blocks = (unsigned int)((n + BLOCK - 1) / BLOCK);
    devapi_test::with_engine(
        engine,
        [&](auto* tag)
        {
            using S = std::remove_pointer_t<decltype(tag)>;
            devapi_test::with_fn(fn,
                                 [&](auto fc)
                                 {
                                     constexpr int FN = decltype(fc)::value;
                                     if constexpr(devapi_test::supported<S>(FN))
                                         hipLaunchKernelGGL((devapi_kernel<S, FN>), dim3(blocks), dim3(BLOCK), 0, 0,
                                                            seed, (const unsigned long long*)buf.p[0],
                                                            (const unsigned long long*)buf.p[1], n, calls, a, b,
                                                            (rocrand_discrete_distribution)buf.p[5], tables,
                                                            (uint64_t*)buf.p[2]);
                                 });
        });
    if(hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess)
        return -4;
    if(hipMemcpy(out, buf.p[2], out_bytes, hipMemcpyDeviceToHost) != hipSuccess)
        return -3;
    return 0;
}

// The same kernels over a discrete distribution that already lives on the device (one
// made by rocrand_create_discrete_distribution / rocrand_create_poisson_distribution).
extern "C" int devapi_gpu_run_handle(int engine, int fn, unsigned long long seed, const unsigned long long* sub,
                                     const unsigned long long* off, int n, int calls,
                                     rocrand_discrete_distribution dist, const unsigned int* v32,
                                     const unsigned long long* v64, unsigned int dims, unsigned long long* out)
{
    const bool sobol = engine >= devapi_test::SOBOL32;
    if(n <= 0 || calls <= 0 || (size_t)n * calls > (1u << 26) || dist == nullptr
       || (fn != devapi_test::DISCRETE && fn != devapi_test::DISCRETE4) || (sobol && (dims == 0 || !v32 || !v64)))
        return -2;
    bool known = false;
    devapi_test::with_engine(engine,
                             [&](auto* tag)
                             {
                                 using S = std::remove_pointer_t<decltype(tag)>;
                                 known   = devapi_test::supported<S>(fn);
                             });
    if(!known)
        return -1;
    device_buffers     buf;
    const size_t       out_bytes = (size_t)n * calls * 4 * sizeof(uint64_t);
    const unsigned int nd        = sobol ? dims : 1;
    if(hipMalloc(&buf.p[0], n * sizeof(unsigned long long)) != hipSuccess
       || hipMalloc(&buf.p[1], n * sizeof(unsigned long long)) != hipSuccess
       || hipMalloc(&buf.p[2], out_bytes) != hipSuccess
       || hipMalloc(&buf.p[7], (size_t)nd * 32 * sizeof(unsigned int)) != hipSuccess
       || hipMalloc(&buf.p[8], (size_t)nd * 64 * sizeof(unsigned long long)) != hipSuccess)
        return -3;
    if(hipMemcpy(buf.p[0], sub, n * sizeof(unsigned long long), hipMemcpyHostToDevice) != hipSuccess
       || hipMemcpy(buf.p[1], off, n * sizeof(unsigned long long), hipMemcpyHostToDevice) != hipSuccess
       || hipMemset(buf.p[2], 0, out_bytes) != hipSuccess)
        return -3;
    if(sobol
       && (hipMemcpy(buf.p[7], v32, (size_t)nd * 32 * sizeof(unsigned int), hipMemcpyHostToDevice) != hipSuccess
           || hipMemcpy(buf.p[8], v64, (size_t)nd * 64 * sizeof(unsigned long long), hipMemcpyHostToDevice)
                  != hipSuccess))
        return -3;
    const devapi_test::sobol_tables tables = {(const unsigned int*)buf.p[7], (const unsigned long long*)buf.p[8], nd};
    const unsigned int              blocks = (unsigned int)((n + BLOCK - 1) / BLOCK);
    devapi_test::with_engine(
        engine,
        [&](auto* tag)
        {
            using S = std::remove_pointer_t<decltype(tag)>;
            devapi_test::with_fn(fn,
                                 [&](auto fc)
                                 {
                                     constexpr int FN = decltype(fc)::value;
                                     if constexpr(devapi_test::supported<S>(FN))
                                         hipLaunchKernelGGL((devapi_kernel<S, FN>), dim3(blocks), dim3(BLOCK), 0, 0,
                                                            seed, (const unsigned long long*)buf.p[0],
                                                            (const unsigned long long*)buf.p[1], n, calls, 0.0, 0.0,
                                                            dist, tables, (uint64_t*)buf.p[2]);
                                 });
        });
    if(hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess)
        return -4;
    if(hipMemcpy(out, buf.p[2], out_bytes, hipMemcpyDeviceToHost) != hipSuccess)
        return -3;
    return 0;
}

// ---------------------------------------------------------------- MTGP32
namespace
{

template<int FN>
__global__ __launch_bounds__(256) void devapi_mtgp32_kernel(rocrand_state_mtgp32* states, mtgp32_params* params,
                                                            int calls, double a, double b,
                                                            rocrand_discrete_distribution dist,
                                                            devapi_test::sobol_tables tables, uint64_t* out)
{
    __shared__ rocrand_state_mtgp32 state;
    rocrand_mtgp32_block_copy(&states[blockIdx.x], &state);
    if(params != nullptr)
    {
        if(threadIdx.x == 0)
            rocrand_mtgp32_set_params(&state, params);
        __syncthreads();
    }
    uint64_t* o = out + ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * calls * 4;
    for(int k = 0; k < calls; k++)
        devapi_test::call<rocrand_state_mtgp32, FN>(&state, k, a, b, dist, o + (size_t)k * 4);
    (void)tables;
}

} // namespace

// MTGP32 states 0 .. blocks-1 from `params` (host array of mtgp32_fast_params), one block of
// `threads` threads per state; with use_constant the kernels also reload the tables from a
// rocrand_make_constant copy.  out: [blocks][threads][calls][4] slots.
extern "C" int devapi_gpu_mtgp32(int fn, unsigned long long seed, int blocks, int threads, int calls, double a,
                                 double b, unsigned int dsize, const double* prob, const unsigned int* alias,
                                 const void* params, int use_constant, unsigned long long* out)
{
    if(blocks <= 0 || blocks > MTGP_BN_MAX || threads <= 0 || threads > 256 || calls <= 0 || calls > 4096
       || !params || dsize == 0 || !prob || !alias)
        return -2;
    if(!devapi_test::supported<rocrand_state_mtgp32>(fn))
        return -1;
    device_buffers     buf;
    const size_t       out_bytes = (size_t)blocks * threads * calls * 4 * sizeof(uint64_t);
    if(hipMalloc(&buf.p[0], sizeof(rocrand_state_mtgp32) * blocks) != hipSuccess
       || hipMalloc(&buf.p[1], sizeof(mtgp32_params)) != hipSuccess || hipMalloc(&buf.p[2], out_bytes) != hipSuccess
       || hipMalloc(&buf.p[3], dsize * sizeof(double)) != hipSuccess
       || hipMalloc(&buf.p[4], dsize * sizeof(unsigned int)) != hipSuccess
       || hipMalloc(&buf.p[5], sizeof(rocrand_discrete_distribution_st)) != hipSuccess)
        return -3;
    const mtgp32_fast_params* fp = static_cast<const mtgp32_fast_params*>(params);
    if(rocrand_make_state_mtgp32((rocrand_state_mtgp32*)buf.p[0], fp, blocks, seed) != ROCRAND_STATUS_SUCCESS
       || rocrand_make_constant(fp, (mtgp32_params*)buf.p[1]) != ROCRAND_STATUS_SUCCESS)
        return -3;
    rocrand_discrete_distribution_st d;
    d.size        = dsize;
    d.offset      = 0;
    d.alias       = (unsigned int*)buf.p[4];
    d.probability = (double*)buf.p[3];
    d.cdf         = nullptr;
    if(hipMemcpy(buf.p[3], prob, dsize * sizeof(double), hipMemcpyHostToDevice) != hipSuccess
       || hipMemcpy(buf.p[4], alias, dsize * sizeof(unsigned int), hipMemcpyHostToDevice) != hipSuccess
       || hipMemcpy(buf.p[5], &d, sizeof(d), hipMemcpyHostToDevice) != hipSuccess
       || hipMemset(buf.p[2], 0, out_bytes) != hipSuccess)
        return -3;
    const devapi_test::sobol_tables none = {nullptr, nullptr, 1};
    devapi_test::with_fn(fn,
                         [&](auto fc)
                         {
                             constexpr int FN = decltype(fc)::value;
                             if constexpr(devapi_test::supported<rocrand_state_mtgp32>(FN))
                                 hipLaunchKernelGGL((devapi_mtgp32_kernel<FN>), dim3(blocks), dim3(threads), 0, 0,
                                                    (rocrand_state_mtgp32*)buf.p[0],
                                                    use_constant ? (mtgp32_params*)buf.p[1] : nullptr, calls, a, b,
                                                    (rocrand_discrete_distribution)buf.p[5], none,
                                                    (uint64_t*)buf.p[2]);
                         });
    if(hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess)
        return -4;
    if(hipMemcpy(out, buf.p[2], out_bytes, hipMemcpyDeviceToHost) != hipSuccess)
        return -3;
    return 0;
}

// ---------------------------------------------------------------- state bytes after init
namespace
{
template<class S>
__global__ __launch_bounds__(BLOCK) void devapi_state_kernel(unsigned long long seed, const unsigned long long* sub,
                                                             const unsigned long long* off, int n, int words,
                                                             uint32_t* out)
{
    const int t = blockIdx.x * BLOCK + threadIdx.x;
    if(t >= n)
        return;
    const devapi_test::sobol_tables none = {nullptr, nullptr, 1};
    devapi_test::state_words<S>(seed, sub[t], off[t], none, out + (size_t)t * words, words);
}
} // namespace

extern "C" int devapi_gpu_state(int engine, unsigned long long seed, const unsigned long long* sub,
                                const unsigned long long* off, int n, int words, unsigned int* out)
{
    if(n <= 0 || n > (1 << 20) || words <= 0 || words > 64 || engine >= devapi_test::SOBOL32)
        return -2;
    device_buffers buf;
    const size_t   out_bytes = (size_t)n * words * sizeof(uint32_t);
    if(hipMalloc(&buf.p[0], n * sizeof(unsigned long long)) != hipSuccess
       || hipMalloc(&buf.p[1], n * sizeof(unsigned long long)) != hipSuccess
       || hipMalloc(&buf.p[2], out_bytes) != hipSuccess)
        return -3;
    if(hipMemcpy(buf.p[0], sub, n * sizeof(unsigned long long), hipMemcpyHostToDevice) != hipSuccess
       || hipMemcpy(buf.p[1], off, n * sizeof(unsigned long long), hipMemcpyHostToDevice) != hipSuccess)
        return -3;
    int                size   = -1;
    const unsigned int blocks = (unsigned int)((n + BLOCK - 1) / BLOCK);
    devapi_test::with_engine(engine,
                             [&](auto* tag)
                             {
                                 using S = std::remove_pointer_t<decltype(tag)>;
                                 hipLaunchKernelGGL((devapi_state_kernel<S>), dim3(blocks), dim3(BLOCK), 0, 0, seed,
                                                    (const unsigned long long*)buf.p[0],
                                                    (const unsigned long long*)buf.p[1], n, words, (uint32_t*)buf.p[2]);
                                 size = (int)sizeof(S);
                             });
    if(hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess)
        return -4;
    if(hipMemcpy(out, buf.p[2], out_bytes, hipMemcpyDeviceToHost) != hipSuccess)
        return -3;
    return size;
}

// ---------------------------------------------------------------- throughput of user kernels
// A user kernel as the device API is used: one state per thread (rocrand_init with the
// thread as subsequence), then a grid-stride loop of generate calls writing coalesced
// 16-byte vectors (4 values per thread per iteration).  kind: 0 raw words, 1 uniform f32,
// 2 normal f32, 3 no generation (the same loop storing a constant: its own ceiling).
namespace
{

template<class S>
__global__ __launch_bounds__(256) void bench_init_kernel(S* states, unsigned long long seed)
{
    const unsigned int                 t    = blockIdx.x * 256 + threadIdx.x;
    const devapi_test::sobol_tables    none = {nullptr, nullptr, 1};
    S                                  s;
    devapi_test::init(&s, seed, t, 0, none);
    states[t] = s;
}

template<class S, int KIND>
__device__ __forceinline__ uint4 bench4(S* s)
{
    if constexpr(KIND == 3)
    {
        // the loop's own ceiling: a value that costs nothing to make
        return uint4{threadIdx.x, blockIdx.x, 0u, 1u};
    }
    else if constexpr(KIND == 0)
    {
        if constexpr(devapi_test::has_rocrand4<S>::value)
            return rocrand4(s);
        else
        {
            const unsigned int a = rocrand(s), b = rocrand(s), c = rocrand(s), d = rocrand(s);
            return uint4{a, b, c, d};
        }
    }
    float4 f;
    if constexpr(KIND == 1)
    {
        if constexpr(devapi_test::has_uniform4<S>::value)
            f = rocrand_uniform4(s);
        else
        {
            f.x = rocrand_uniform(s);
            f.y = rocrand_uniform(s);
            f.z = rocrand_uniform(s);
            f.w = rocrand_uniform(s);
        }
    }
    else
    {
        if constexpr(devapi_test::has_normal4<S>::value)
            f = rocrand_normal4(s);
        else
        {
            const float2 a = rocrand_normal2(s), b = rocrand_normal2(s);
            f              = float4{a.x, a.y, b.x, b.y};
        }
    }
    return uint4{__float_as_uint(f.x), __float_as_uint(f.y), __float_as_uint(f.z), __float_as_uint(f.w)};
}

template<class S, int KIND>
__global__ __launch_bounds__(256) void bench_gen_kernel(S* states, uint4* out, unsigned int iters)
{
    const unsigned int t       = blockIdx.x * 256 + threadIdx.x;
    const unsigned int threads = gridDim.x * 256;
    S                  s       = states[t];
    for(unsigned int i = 0; i < iters; i++)
        out[(size_t)i * threads + t] = bench4<S, KIND>(&s);
    states[t] = s;
}

} // namespace

// Times `reps` generate launches of 2^28 values (after 20 untimed) and one init launch;
// writes ms per generate launch and ms of the init into ms[0], ms[1].
extern "C" int devapi_gpu_bench(int engine, int kind, int reps, float* ms)
{
    if(kind < 0 || kind > 3 || reps <= 0 || reps > 1000
       || !(engine == devapi_test::PHILOX || engine == devapi_test::MRG32K3A || engine == devapi_test::MRG31K3P
            || engine == devapi_test::XORWOW || engine == devapi_test::LFSR113 || engine == devapi_test::TF4X32))
        return -2;
    constexpr unsigned int blocks = 2048, threads = blocks * 256;
    constexpr size_t       n4     = (size_t)1 << 26; // 2^28 values as uint4
    const unsigned int     iters  = (unsigned int)(n4 / threads);
    device_buffers         buf;
    if(hipMalloc(&buf.p[0], n4 * sizeof(uint4)) != hipSuccess || hipMalloc(&buf.p[1], (size_t)threads * 64) != hipSuccess)
        return -3;
    hipEvent_t e0, e1;
    if(hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        return -3;
    int  rc  = 0;
    auto run = [&](auto* tag)
        {
            using S = std::remove_pointer_t<decltype(tag)>;
            static_assert(sizeof(S) <= 64, "state slot");
            S* st = (S*)buf.p[1];
            (void)hipEventRecord(e0, 0);
            hipLaunchKernelGGL((bench_init_kernel<S>), dim3(blocks), dim3(256), 0, 0, st, 12345ull);
            (void)hipEventRecord(e1, 0);
            if(hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms[1], e0, e1) != hipSuccess)
                rc = -4;
            auto gen = [&]()
            {
                if(kind == 0)
                    hipLaunchKernelGGL((bench_gen_kernel<S, 0>), dim3(blocks), dim3(256), 0, 0, st, (uint4*)buf.p[0],
                                       iters);
                else if(kind == 1)
                    hipLaunchKernelGGL((bench_gen_kernel<S, 1>), dim3(blocks), dim3(256), 0, 0, st, (uint4*)buf.p[0],
                                       iters);
                else if(kind == 3)
                    hipLaunchKernelGGL((bench_gen_kernel<S, 3>), dim3(blocks), dim3(256), 0, 0, st, (uint4*)buf.p[0],
                                       iters);
                else
                    hipLaunchKernelGGL((bench_gen_kernel<S, 2>), dim3(blocks), dim3(256), 0, 0, st, (uint4*)buf.p[0],
                                       iters);
            };
            for(int r = 0; r < 20; r++)
                gen();
            (void)hipEventRecord(e0, 0);
            for(int r = 0; r < reps; r++)
                gen();
            (void)hipEventRecord(e1, 0);
            if(hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms[0], e0, e1) != hipSuccess)
                rc = -4;
            ms[0] /= reps;
        };
    switch(engine)
    {
        case devapi_test::PHILOX: run((rocrand_state_philox4x32_10*)nullptr); break;
        case devapi_test::MRG32K3A: run((rocrand_state_mrg32k3a*)nullptr); break;
        case devapi_test::MRG31K3P: run((rocrand_state_mrg31k3p*)nullptr); break;
        case devapi_test::XORWOW: run((rocrand_state_xorwow*)nullptr); break;
        case devapi_test::LFSR113: run((rocrand_state_lfsr113*)nullptr); break;
        default: run((rocrand_state_threefry4x32_20*)nullptr); break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if(hipGetLastError() != hipSuccess)
        return -4;
    return rc;
}
