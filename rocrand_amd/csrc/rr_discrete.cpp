// rr_discrete.cpp -- discrete-distribution entry points of the C ABI
// (library/src/rocrand.cpp:462-588): a device-resident
// rocrand_discrete_distribution_st whose alias, probability and CDF tables are
// device arrays, for samplers running in user kernels.
#include "rr_generator.hpp"

#include <rocrand.h>

#include <hip/hip_runtime.h>

#include <vector>

namespace
{

rocrand_status make_device_distribution(const std::vector<double>& p, unsigned int size, unsigned int offset,
                                        rocrand_discrete_distribution* out)
{
    std::vector<double>       prob, cdf;
    std::vector<unsigned int> alias;
    rr::alias_table_from(p, size, prob, alias, &cdf);
    rocrand_discrete_distribution_st h{};
    h.size   = size;
    h.offset = offset;
    rocrand_discrete_distribution handle = nullptr;
    // Every failure path releases whatever was allocated before it.
    auto fail = [&](rocrand_status status) {
        (void)hipFree(h.probability);
        (void)hipFree(h.alias);
        (void)hipFree(h.cdf);
        (void)hipFree(handle);
        return status;
    };
    if(hipMalloc(&h.probability, size * sizeof(double)) != hipSuccess
       || hipMalloc(&h.alias, size * sizeof(unsigned int)) != hipSuccess
       || hipMalloc(&h.cdf, size * sizeof(double)) != hipSuccess
       || hipMalloc(&handle, sizeof(rocrand_discrete_distribution_st)) != hipSuccess)
        return fail(ROCRAND_STATUS_ALLOCATION_FAILED);
    if(hipMemcpy(h.probability, prob.data(), size * sizeof(double), hipMemcpyHostToDevice) != hipSuccess
       || hipMemcpy(h.alias, alias.data(), size * sizeof(unsigned int), hipMemcpyHostToDevice) != hipSuccess
       || hipMemcpy(h.cdf, cdf.data(), size * sizeof(double), hipMemcpyHostToDevice) != hipSuccess
       || hipMemcpy(handle, &h, sizeof(h), hipMemcpyHostToDevice) != hipSuccess)
        return fail(ROCRAND_STATUS_INTERNAL_ERROR);
    *out = handle;
    return ROCRAND_STATUS_SUCCESS;
}

} // namespace

extern "C" {

rocrand_status ROCRANDAPI rocrand_create_poisson_distribution(double                         lambda,
                                                              rocrand_discrete_distribution* discrete_distribution)
{
    if(discrete_distribution == nullptr || lambda <= 0.0)
        return ROCRAND_STATUS_OUT_OF_RANGE;
    unsigned int              size, offset;
    const std::vector<double> pmf = rr::poisson_pmf(lambda, size, offset);
    return make_device_distribution(pmf, size, offset, discrete_distribution);
}

rocrand_status ROCRANDAPI rocrand_create_discrete_distribution(const double*                  probabilities,
                                                               unsigned int                   size,
                                                               unsigned int                   offset,
                                                               rocrand_discrete_distribution* discrete_distribution)
{
    if(discrete_distribution == nullptr || size == 0 || probabilities == nullptr)
        return ROCRAND_STATUS_OUT_OF_RANGE;
    return make_device_distribution(std::vector<double>(probabilities, probabilities + size), size, offset,
                                    discrete_distribution);
}

rocrand_status ROCRANDAPI rocrand_destroy_discrete_distribution(rocrand_discrete_distribution discrete_distribution)
{
    if(discrete_distribution == nullptr)
        return ROCRAND_STATUS_OUT_OF_RANGE;
    rocrand_discrete_distribution_st h{};
    if(hipMemcpy(&h, discrete_distribution, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess)
        return ROCRAND_STATUS_INTERNAL_ERROR;
    if(hipFree(h.probability) != hipSuccess || hipFree(h.alias) != hipSuccess || hipFree(h.cdf) != hipSuccess
       || hipFree(discrete_distribution) != hipSuccess)
        return ROCRAND_STATUS_INTERNAL_ERROR;
    return ROCRAND_STATUS_SUCCESS;
}

} // extern "C"
