// rr_quasi.cpp -- Sobol32 / Sobol64 (plain and scrambled) generators, their direction
// vectors and scramble constants, and the rocrand_get_* table entry points.
//
// Tables (reference layout, [dimension][bit], 20000 dimensions):
//   direction vectors  : Bratley-Fox recurrence over the Joe-Kuo direction numbers
//                        (tools/sobol_utils.hpp:59-104); dimension 0 is van der Corput;
//   scrambled vectors  : the plain vectors plus 640000 (32-bit) / 1280000 (64-bit)
//                        (tools/scrambled_sobol{32,64}_direction_vector_generator.cpp);
//   scramble constants : draws of a default-seeded std::mt19937, one per 32-bit constant
//                        and two (high word first) per 64-bit constant
//                        (tools/scrambled_sobol{32,64}_constants_generator.cpp).
// Generator semantics follow library/src/rng/sobol.hpp:489-790.
#include "rr_generator.hpp"

#include <rocrand.h>

#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <random>
#include <string.h>
#include <vector>

extern "C" const unsigned char rr_joe_kuo_begin[];
extern "C" const unsigned char rr_joe_kuo_end[];

using namespace rr;

namespace
{

constexpr unsigned SOBOL_DIMS = 20000;

struct sobol_host_tables
{
    std::vector<uint32_t>           v32, sv32, c32;
    std::vector<unsigned long long> v64, sv64, c64;
};

template<class V>
void direction_vectors(const uint32_t* jk, size_t words, std::vector<V>& out)
{
    constexpr int B = sizeof(V) * 8;
    out.assign((size_t)SOBOL_DIMS * B, 0);
    size_t pos = 0;
    for(unsigned d = 0; d < SOBOL_DIMS; d++)
    {
        V* v = out.data() + (size_t)d * B;
        if(d == 0)
        {
            for(int j = 0; j < B; j++)
                v[j] = V(1) << (B - 1 - j);
            continue;
        }
        if(pos + 2 > words)
            throw ROCRAND_STATUS_INTERNAL_ERROR;
        const uint32_t  s = jk[pos], a = jk[pos + 1];
        const uint32_t* m = jk + pos + 2;
        pos += 2 + s;
        for(uint32_t j = 0; j < s && j < (uint32_t)B; j++)
            v[j] = (V)m[j] << (B - 1 - j);
        for(int j = (int)s; j < B; j++)
        {
            V x = v[j - s] ^ (v[j - s] >> s);
            for(uint32_t k = 1; k < s; k++)
                if((a >> (s - 1 - k)) & 1u)
                    x ^= v[j - k];
            v[j] = x;
        }
    }
}

const sobol_host_tables& host_tables()
{
    static sobol_host_tables t;
    static std::once_flag    once;
    std::call_once(once, [] {
        const uint32_t* jk    = reinterpret_cast<const uint32_t*>(rr_joe_kuo_begin);
        const size_t    words = (size_t)(rr_joe_kuo_end - rr_joe_kuo_begin) / 4;
        direction_vectors(jk, words, t.v32);
        direction_vectors(jk, words, t.v64);
        t.sv32 = t.v32;
        for(auto& x : t.sv32)
            x += 640000u;
        t.sv64 = t.v64;
        for(auto& x : t.sv64)
            x += 1280000ull;
        std::mt19937 g32;
        t.c32.resize(SOBOL_DIMS);
        for(auto& x : t.c32)
            x = (uint32_t)g32();
        std::mt19937 g64;
        t.c64.resize(SOBOL_DIMS);
        for(auto& x : t.c64)
        {
            const unsigned long long hi = (uint32_t)g64();
            const unsigned long long lo = (uint32_t)g64();
            x                           = (hi << 32) | lo;
        }
    });
    return t;
}

struct device_copy
{
    int                 device = -1;
    sobol_device_tables t{};
};

} // namespace

namespace rr
{

hipError_t get_sobol_device_tables(sobol_device_tables& out)
{
    static std::mutex               mtx;
    static std::vector<device_copy> copies;
    int                             dev = 0;
    hipError_t                      e   = hipGetDevice(&dev);
    if(e != hipSuccess)
        return e;
    std::lock_guard<std::mutex> lock(mtx);
    for(const auto& c : copies)
        if(c.device == dev)
        {
            out = c.t;
            return hipSuccess;
        }
    const sobol_host_tables& h = host_tables();
    device_copy              c;
    c.device = dev;
    void* p[4]{};
    const size_t bytes[4] = {h.v32.size() * 4, h.sv32.size() * 4, h.v64.size() * 8, h.sv64.size() * 8};
    const void*  src[4]   = {h.v32.data(), h.sv32.data(), h.v64.data(), h.sv64.data()};
    for(int i = 0; i < 4; i++)
    {
        if((e = hipMalloc(&p[i], bytes[i])) != hipSuccess)
            return e;
        if((e = hipMemcpy(p[i], src[i], bytes[i], hipMemcpyHostToDevice)) != hipSuccess)
            return e;
    }
    c.t.v32  = static_cast<const uint32_t*>(p[0]);
    c.t.sv32 = static_cast<const uint32_t*>(p[1]);
    c.t.v64  = static_cast<const uint64_t*>(p[2]);
    c.t.sv64 = static_cast<const uint64_t*>(p[3]);
    copies.push_back(c);
    out = c.t;
    return hipSuccess;
}

} // namespace rr

namespace
{

// sobol.hpp:489-790
class sobol_generator final : public rocrand_generator_base_type
{
public:
    explicit sobol_generator(rocrand_rng_type t) : rocrand_generator_base_type(t)
    {
        m_is64      = t == ROCRAND_RNG_QUASI_SOBOL64 || t == ROCRAND_RNG_QUASI_SCRAMBLED_SOBOL64;
        m_scrambled = t == ROCRAND_RNG_QUASI_SCRAMBLED_SOBOL32 || t == ROCRAND_RNG_QUASI_SCRAMBLED_SOBOL64;
        order       = ROCRAND_ORDERING_QUASI_DEFAULT;
        m_poisson.cdf_method = true;
    }

    void set_seed(unsigned long long) override {} // sobol.hpp:581-584: no seed, no reset

    rocrand_status set_order(rocrand_ordering o) override
    {
        if(o != ROCRAND_ORDERING_QUASI_DEFAULT) // is_ordering_quasi
            return ROCRAND_STATUS_OUT_OF_RANGE;
        order = o;
        reset();
        return ROCRAND_STATUS_SUCCESS;
    }

    rocrand_status set_dimensions(unsigned int d) override
    {
        if(d < 1 || d > SOBOL_DIMS)
            return ROCRAND_STATUS_OUT_OF_RANGE;
        m_dims = d;
        reset();
        return ROCRAND_STATUS_SUCCESS;
    }

    rocrand_status init() override
    {
        if(m_initialized)
            return ROCRAND_STATUS_SUCCESS;
        // the device copy of the direction vectors is made here, outside any capture
        sobol_device_tables t;
        if(get_sobol_device_tables(t) != hipSuccess)
            return ROCRAND_STATUS_ALLOCATION_FAILED;
        m_current     = (uint32_t)offset; // the point index is an unsigned int
        m_initialized = true;
        return ROCRAND_STATUS_SUCCESS;
    }

    rocrand_status shape(out_kind k, dist_shape& s) const override
    {
        static constexpr int esz[] = {1, 2, 4, 8, 4, 8, 2, 4, 8, 2, 4, 8, 2, 4};
        if(k == out_kind::u64 && !m_is64)
            return ROCRAND_STATUS_TYPE_ERROR;
        s = dist_shape{1, 1, esz[(int)k]};
        return ROCRAND_STATUS_SUCCESS;
    }

protected:
    void reset() override { m_initialized = false; }

    rocrand_status generate_device(out_kind k, void* data, size_t n, const gen_args& args) override
    {
        dist_shape     s;
        rocrand_status st = shape(k, s);
        if(st != ROCRAND_STATUS_SUCCESS)
            return st;
        sobol_call c{};
        // generate_poisson resolves its table before the length check (sobol.hpp:743-757)
        if((st = make_params(k, args, c.params)) != ROCRAND_STATUS_SUCCESS)
            return st;
        if(n % m_dims != 0)
            return ROCRAND_STATUS_LENGTH_NOT_MULTIPLE;
        if((st = init()) != ROCRAND_STATUS_SUCCESS)
            return st;
        if(data == nullptr)
            return ROCRAND_STATUS_SUCCESS;
        sobol_device_tables t;
        if(get_sobol_device_tables(t) != hipSuccess)
            return ROCRAND_STATUS_ALLOCATION_FAILED;
        const size_t size = n / m_dims;
        if(m_is64)
            c.vectors = m_scrambled ? (const void*)t.sv64 : (const void*)t.v64;
        else
            c.vectors = m_scrambled ? (const void*)t.sv32 : (const void*)t.v32;
        // The scrambled generator's per-dimension constant is entry `dimension` of its
        // direction-vector table: the generator constructor stores the vector table
        // in its scramble-constant pointer (sobol.hpp:514).
        c.scramble = m_scrambled ? c.vectors : nullptr;
        c.data     = data;
        c.size     = size;
        c.offset   = m_current;
        c.dims     = m_dims;
        if(launch_sobol(m_is64, k, c, stream) != hipSuccess)
            return ROCRAND_STATUS_LAUNCH_FAILURE;
        m_current += (uint32_t)size;
        return ROCRAND_STATUS_SUCCESS;
    }

private:
    bool     m_is64, m_scrambled;
    unsigned m_dims        = 1;
    bool     m_initialized = false;
    uint32_t m_current     = 0;
};

} // namespace

namespace rr
{
rocrand_generator make_sobol_generator(rocrand_rng_type t) { return new sobol_generator(t); }
} // namespace rr

extern "C" {

// rocrand.cpp:589-629
rocrand_status ROCRANDAPI rocrand_get_direction_vectors32(const unsigned int** vectors, rocrand_direction_vector_set set)
{
    const sobol_host_tables& t = host_tables();
    switch(set)
    {
        case ROCRAND_DIRECTION_VECTORS_32_JOEKUO6: *vectors = t.v32.data(); return ROCRAND_STATUS_SUCCESS;
        case ROCRAND_SCRAMBLED_DIRECTION_VECTORS_32_JOEKUO6: *vectors = t.sv32.data(); return ROCRAND_STATUS_SUCCESS;
        default: return ROCRAND_STATUS_OUT_OF_RANGE;
    }
}

rocrand_status ROCRANDAPI rocrand_get_direction_vectors64(const unsigned long long** vectors,
                                                          rocrand_direction_vector_set set)
{
    const sobol_host_tables& t = host_tables();
    switch(set)
    {
        case ROCRAND_DIRECTION_VECTORS_64_JOEKUO6: *vectors = t.v64.data(); return ROCRAND_STATUS_SUCCESS;
        case ROCRAND_SCRAMBLED_DIRECTION_VECTORS_64_JOEKUO6: *vectors = t.sv64.data(); return ROCRAND_STATUS_SUCCESS;
        default: return ROCRAND_STATUS_OUT_OF_RANGE;
    }
}

rocrand_status ROCRANDAPI rocrand_get_scramble_constants32(const unsigned int** constants)
{
    *constants = host_tables().c32.data();
    return ROCRAND_STATUS_SUCCESS;
}

rocrand_status ROCRANDAPI rocrand_get_scramble_constants64(const unsigned long long** constants)
{
    *constants = host_tables().c64.data();
    return ROCRAND_STATUS_SUCCESS;
}

} // extern "C"
