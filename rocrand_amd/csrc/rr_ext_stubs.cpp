// rr_ext_stubs.cpp -- placeholders for engines not built yet (MTGP32, MT19937, Sobol).
#include "rr_generator.hpp"

namespace rr
{
rocrand_generator make_mtgp32_generator() { return nullptr; }
rocrand_generator make_mt19937_generator() { return nullptr; }
rocrand_generator make_sobol_generator(rocrand_rng_type) { return nullptr; }
} // namespace rr

extern "C" {
rocrand_status ROCRANDAPI rocrand_get_direction_vectors32(const unsigned int**, rocrand_direction_vector_set)
{
    return ROCRAND_STATUS_OUT_OF_RANGE;
}
rocrand_status ROCRANDAPI rocrand_get_direction_vectors64(const unsigned long long**, rocrand_direction_vector_set)
{
    return ROCRAND_STATUS_OUT_OF_RANGE;
}
rocrand_status ROCRANDAPI rocrand_get_scramble_constants32(const unsigned int**) { return ROCRAND_STATUS_OUT_OF_RANGE; }
rocrand_status ROCRANDAPI rocrand_get_scramble_constants64(const unsigned long long**) { return ROCRAND_STATUS_OUT_OF_RANGE; }
}
