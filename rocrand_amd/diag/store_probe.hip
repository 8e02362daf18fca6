// store_probe.hip -- HBM store ceilings on gfx950 for the write patterns the generators use.
//
// Fills a 2 GiB device buffer with a constant (no arithmetic) and reports GB/s, for 4-, 8-
// and 16-byte stores per lane, plain and non-temporal, in three patterns:
//   linear   one store per lane, the grid covering the buffer (the counter-based kernels)
//   stride   L lanes, lane l storing l, l + L, l + 2L, ... (a grid-stride loop)
//   chunked  S = 131072 engines x J = 8 chunks, lane (j, e) storing e + (j C + k) S for
//            k < C (the interleaved MRG32k3a kernel's layout and schedule)
// Every index is below the element count by construction (checked in each loop).
// Build and run: make -C rocrand_amd/diag && rocrand_amd/diag/store_probe
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#define CK(x)                                                                          \
    do                                                                                 \
    {                                                                                  \
        hipError_t e_ = (x);                                                           \
        if(e_ != hipSuccess)                                                           \
        {                                                                              \
            std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
            return 1;                                                                  \
        }                                                                              \
    } while(0)

using u2 = unsigned int __attribute__((ext_vector_type(2)));
using u4 = unsigned int __attribute__((ext_vector_type(4)));

template<class T>
__device__ __forceinline__ T fill_value()
{
    T v;
    __builtin_memset(&v, 0x3f, sizeof(T));
    return v;
}

template<class T, bool NT>
__device__ __forceinline__ void put(T* p, size_t i)
{
    if constexpr(NT)
        __builtin_nontemporal_store(fill_value<T>(), p + i);
    else
        p[i] = fill_value<T>();
}

template<class T, bool NT>
__global__ __launch_bounds__(256) void linear_kernel(T* p, size_t n)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if(i < n)
        put<T, NT>(p, i);
}

template<class T, bool NT>
__global__ __launch_bounds__(256) void stride_kernel(T* p, size_t n, uint32_t lanes)
{
    for(size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += lanes)
        put<T, NT>(p, i);
}

template<class T, bool NT>
__global__ __launch_bounds__(256) void chunked_kernel(T* p, size_t n, uint32_t S, uint32_t C)
{
    const uint32_t bpc = S / 256;
    const uint32_t j   = blockIdx.x / bpc;
    const uint32_t e   = (blockIdx.x % bpc) * 256 + threadIdx.x;
    for(uint32_t k = 0; k < C; k++)
    {
        const size_t i = e + ((size_t)j * C + k) * S;
        if(i < n)
            put<T, NT>(p, i);
    }
}

template<class T, bool NT>
int run(void* buf, size_t bytes, hipEvent_t a, hipEvent_t b)
{
    const size_t   n = bytes / sizeof(T);
    const uint32_t S = 131072, J = 8, C = (uint32_t)((n / S + J - 1) / J);
    for(int mode = 0; mode < 4; mode++)
    {
        const uint32_t lanes = mode == 1 ? (1u << 20) : (1u << 18);
        auto           go    = [&]() {
            if(mode == 0)
                hipLaunchKernelGGL((linear_kernel<T, NT>), dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, 0,
                                   (T*)buf, n);
            else if(mode <= 2)
                hipLaunchKernelGGL((stride_kernel<T, NT>), dim3(lanes / 256), dim3(256), 0, 0, (T*)buf, n, lanes);
            else
                hipLaunchKernelGGL((chunked_kernel<T, NT>), dim3(S / 256 * J), dim3(256), 0, 0, (T*)buf, n, S, C);
        };
        for(int i = 0; i < 50; i++)
            go();
        CK(hipGetLastError());
        CK(hipDeviceSynchronize());
        CK(hipEventRecord(a));
        for(int i = 0; i < 50; i++)
            go();
        CK(hipEventRecord(b));
        CK(hipEventSynchronize(b));
        float ms = 0;
        CK(hipEventElapsedTime(&ms, a, b));
        ms /= 50;
        static const char* names[] = {"linear", "stride 2^20 lanes", "stride 2^18 lanes", "chunked S=2^17 J=8"};
        std::printf("| %2zu B | %s | %-18s | %7.1f | %5.0f |\n", sizeof(T), NT ? "nt   " : "plain", names[mode],
                    ms * 1e3, (double)bytes / ms / 1e6);
        std::fflush(stdout);
    }
    return 0;
}

int main()
{
    const size_t bytes = 2ull << 30;
    void*        buf   = nullptr;
    hipEvent_t   a, b;
    CK(hipMalloc(&buf, bytes));
    CK(hipEventCreate(&a));
    CK(hipEventCreate(&b));
    std::printf("| store | policy | pattern | us per 2 GiB | GB/s |\n|---|---|---|---|---|\n");
    int rc = run<uint32_t, false>(buf, bytes, a, b) || run<uint32_t, true>(buf, bytes, a, b)
             || run<u2, false>(buf, bytes, a, b) || run<u2, true>(buf, bytes, a, b)
             || run<u4, false>(buf, bytes, a, b) || run<u4, true>(buf, bytes, a, b);
    CK(hipFree(buf));
    return rc;
}
