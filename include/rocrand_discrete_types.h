/*
 * rocrand_discrete_types.h -- discrete distribution descriptor of the C ABI.
 * Same layout as the reference library/include/rocrand/rocrand_discrete_types.h:25-41
 * (a device-resident alias/probability/CDF table consumed by device samplers).
 */
#ifndef ROCRAND_DISCRETE_TYPES_H_
#define ROCRAND_DISCRETE_TYPES_H_

struct rocrand_discrete_distribution_st
{
    unsigned int  size;
    unsigned int  offset;
    unsigned int* alias;
    double*       probability;
    double*       cdf;
};

typedef struct rocrand_discrete_distribution_st* rocrand_discrete_distribution;

#endif /* ROCRAND_DISCRETE_TYPES_H_ */
