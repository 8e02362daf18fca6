// ref_harness.cpp -- drives the REFERENCE rocRAND kernel bodies on the CPU (test infrastructure).
//
// Built by oracle/build_ref.sh from the reference sources where they lie
// (library/include, library/src), host-only, into oracle/_ref/libref.so.  It is a
// second, independent checker for the oracle restatement: the engines, the
// distribution functors and the generate_* kernel bodies executed here are the
// reference's own code; this file only supplies the generator bookkeeping that
// library/src/rng/{philox4x32_10,threefry,mrg,xorwow,lfsr113}.hpp perform around
// them (engine init, launch over the DEFAULT grid the way
// library/src/rng/system.hpp:140-214 host_system::launch does, and the
// post-call engine advance).  The public host generator cannot run here: its
// allocations call the HIP runtime, which needs a GPU.
//
// The reference selects `m_state.result.data[...]` under __HIP_PLATFORM_AMD__, a
// member current HIP vector types no longer have; undefining the macro after the
// HIP headers are in selects the reference's own portable `(&result.x)[...]`
// branch, which is the same arithmetic.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <cstring>
#undef __HIP_PLATFORM_AMD__

#include "rng/lfsr113.hpp"
#include "rng/mrg.hpp"
#include "rng/philox4x32_10.hpp"
#include "rng/threefry.hpp"
#include "rng/xorwow.hpp"

#include <cstdint>
#include <cstdlib>
#include <memory>
#include <vector>

using namespace rocrand_impl::host;

namespace
{

enum kind_t
{
    K_U8 = 0,
    K_U16,
    K_U32,
    K_U64,
    K_UF32,
    K_UF64,
    K_UF16,
    K_NF32,
    K_NF64,
    K_NF16,
    K_LF32,
    K_LF64,
    K_LF16,
    K_POISSON
};

template<class Kernel, class... Args>
void launch(unsigned blocks, unsigned threads, Kernel k, Args... args)
{
    for(unsigned b = 0; b < blocks; b++)
        for(unsigned t = 0; t < threads; t++)
            k(dim3(b), dim3(t), dim3(blocks), dim3(threads), args...);
}

// Scratch buffer whose address has the requested residue modulo 64 bytes, so the
// kernel body's pointer-alignment head/tail split matches the device pointer's.
struct aligned_out
{
    std::vector<unsigned char> buf;
    unsigned char*             p;
    aligned_out(size_t bytes, uint64_t addr)
    {
        buf.resize(bytes + 256);
        uintptr_t base = reinterpret_cast<uintptr_t>(buf.data());
        uintptr_t want = (base + 128) & ~uintptr_t(63);
        p              = reinterpret_cast<unsigned char*>(want + (addr & 63));
    }
};

struct poisson_holder
{
    rocrand_discrete_distribution_st st{};
    std::vector<double>              prob;
    std::vector<unsigned int>        alias;
    explicit poisson_holder(double lambda)
    {
        unsigned int size, offset;
        std::vector<double> p = calculate_poisson_probabilities(lambda, size, offset);
        using F               = discrete_distribution_factory<DISCRETE_METHOD_ALIAS, true>;
        F::normalize(p, size);
        prob.resize(size);
        alias.resize(size);
        F::create_alias_table(p, size, prob.begin(), alias.begin());
        st.size        = size;
        st.offset      = offset;
        st.probability = prob.data();
        st.alias       = alias.data();
    }
};

struct gen_base
{
    int                type;
    unsigned long long seed   = 0;
    uint4              seed4  = {2, 8, 16, 128};
    unsigned long long offset = 0;
    bool               initialized = false;
    virtual ~gen_base() = default;
    virtual int generate(int kind, void* out, size_t n, uint64_t addr, double a, double b) = 0;
};

// ---------------------------------------------------------------- counter-based
template<class Engine, bool IsPhilox>
struct scalar_of
{
    using type = typename Engine::scalar_type;
};
template<class Engine>
struct scalar_of<Engine, true>
{
    using type = unsigned int;
};

template<class Engine, rocrand_rng_type Type, bool IsPhilox>
struct counter_gen : gen_base
{
    using scalar = typename scalar_of<Engine, IsPhilox>::type;
    Engine engine;

    template<class T, class D>
    int run(T* out, size_t n, uint64_t addr, D dist)
    {
        if(!initialized)
        {
            engine      = Engine(seed, 0, offset);
            initialized = true;
        }
        aligned_out tmp(n * sizeof(T), addr);
        T*          data = reinterpret_cast<T*>(tmp.p);
        // config_defaults.hpp:54-108: <<<1024, 256>>> for Philox and every ThreeFry
        if constexpr(IsPhilox)
            launch(1024, 256, generate_philox<T, D>, engine, data, n, dist);
        else
            launch(1024, 256, generate_threefry<Engine, T, D>, engine, data, n, dist);
        std::memcpy(out, data, n * sizeof(T));
        engine.discard((n + D::output_width - 1) / D::output_width * D::input_width);
        return 0;
    }

    int generate(int kind, void* out, size_t n, uint64_t addr, double a, double b) override
    {
        constexpr bool w64 = sizeof(scalar) == 8;
        switch(kind)
        {
            case K_U8: return run((unsigned char*)out, n, addr, uniform_distribution<unsigned char, scalar>());
            case K_U16: return run((unsigned short*)out, n, addr, uniform_distribution<unsigned short, scalar>());
            case K_U32: return run((unsigned int*)out, n, addr, uniform_distribution<unsigned int, scalar>());
            case K_U64:
                if constexpr(w64)
                    return run((unsigned long long*)out, n, addr, uniform_distribution<unsigned long long, scalar>());
                else
                    return ROCRAND_STATUS_TYPE_ERROR;
            case K_UF32: return run((float*)out, n, addr, uniform_distribution<float, scalar>());
            case K_UF64: return run((double*)out, n, addr, uniform_distribution<double, scalar>());
            case K_UF16: return run((__half*)out, n, addr, uniform_distribution<__half, scalar>());
            case K_NF32:
                return run((float*)out, n, addr,
                           normal_distribution<float, scalar, normal_distribution_max_input_width<Type, float>>(float(a), float(b)));
            case K_NF64:
                return run((double*)out, n, addr,
                           normal_distribution<double, scalar, normal_distribution_max_input_width<Type, double>>(a, b));
            case K_NF16:
                return run((__half*)out, n, addr,
                           normal_distribution<__half, scalar, normal_distribution_max_input_width<Type, __half>>(
                               __float2half(float(a)), __float2half(float(b))));
            case K_LF32:
                return run((float*)out, n, addr,
                           log_normal_distribution<float, scalar, log_normal_distribution_max_input_width<Type, float>>(float(a), float(b)));
            case K_LF64:
                return run((double*)out, n, addr,
                           log_normal_distribution<double, scalar, log_normal_distribution_max_input_width<Type, double>>(a, b));
            case K_LF16:
                return run((__half*)out, n, addr,
                           log_normal_distribution<__half, scalar, log_normal_distribution_max_input_width<Type, __half>>(
                               __float2half(float(a)), __float2half(float(b))));
            case K_POISSON:
            {
                if(a <= 0.0)
                    return ROCRAND_STATUS_OUT_OF_RANGE;
                if(a > rocrand_device::detail::lambda_threshold_huge)
                    return run((unsigned int*)out, n, addr, poisson_distribution_huge(a));
                poisson_holder h(a);
                return run((unsigned int*)out, n, addr, poisson_distribution<DISCRETE_METHOD_ALIAS>(h.st));
            }
        }
        return ROCRAND_STATUS_TYPE_ERROR;
    }
};

// ------------------------------------------------------------ interleaved engines
template<class Engine, rocrand_rng_type Type>
struct interleaved_gen : gen_base
{
    static constexpr unsigned S = 131072; // 512 x 256, config_defaults.hpp:33-73
    std::vector<Engine>       engines;
    unsigned                  start = 0;
    using CP                  = default_config_provider<Type>;

    void init()
    {
        start = static_cast<unsigned>(offset % S);
        engines.assign(S, Engine());
        if constexpr(Type == ROCRAND_RNG_PSEUDO_XORWOW)
            launch(S / 256, 256, init_xorwow_engines, engines.data(), start, S, seed, offset / S);
        else if constexpr(Type == ROCRAND_RNG_PSEUDO_LFSR113)
            launch(S / 256, 256, init_lfsr113_engines, engines.data(), start, S, seed4,
                   static_cast<unsigned>(offset / S));
        else
            launch(S / 256, 256, init_engines_mrg<Engine>, engines.data(), start, S, seed, offset / S);
        initialized = true;
    }

    template<class T, class D>
    int run(T* out, size_t n, uint64_t addr, D dist)
    {
        if(!initialized)
            init();
        aligned_out tmp(n * sizeof(T), addr);
        T*          data = reinterpret_cast<T*>(tmp.p);
        if constexpr(Type == ROCRAND_RNG_PSEUDO_XORWOW)
            launch(512, 256, generate_xorwow<CP, false, T, D>, engines.data(), start, data, n, dist);
        else if constexpr(Type == ROCRAND_RNG_PSEUDO_LFSR113)
            launch(512, 256, generate_lfsr113<CP, false, T, D>, engines.data(), start, data, n, dist);
        else
            launch(512, 256, generate_mrg<CP, false, Engine, T, D>, engines.data(), start, data, n, dist);
        std::memcpy(out, data, n * sizeof(T));
        start = static_cast<unsigned>((start + (n + D::output_width - 1) / D::output_width) % S);
        return 0;
    }

    int generate(int kind, void* out, size_t n, uint64_t addr, double a, double b) override
    {
        constexpr bool mrg = Type == ROCRAND_RNG_PSEUDO_MRG32K3A || Type == ROCRAND_RNG_PSEUDO_MRG31K3P;
        if constexpr(mrg)
        {
            switch(kind)
            {
                case K_U8: return run((unsigned char*)out, n, addr, mrg_engine_uniform_distribution<unsigned char, Engine>());
                case K_U16: return run((unsigned short*)out, n, addr, mrg_engine_uniform_distribution<unsigned short, Engine>());
                case K_U32: return run((unsigned int*)out, n, addr, mrg_engine_uniform_distribution<unsigned int, Engine>());
                case K_UF32: return run((float*)out, n, addr, mrg_engine_uniform_distribution<float, Engine>());
                case K_UF64: return run((double*)out, n, addr, mrg_engine_uniform_distribution<double, Engine>());
                case K_UF16: return run((__half*)out, n, addr, mrg_engine_uniform_distribution<__half, Engine>());
                case K_NF32: return run((float*)out, n, addr, mrg_engine_normal_distribution<float, Engine>(float(a), float(b)));
                case K_NF64: return run((double*)out, n, addr, mrg_engine_normal_distribution<double, Engine>(a, b));
                case K_NF16:
                    return run((__half*)out, n, addr,
                               mrg_engine_normal_distribution<__half, Engine>(__float2half(float(a)), __float2half(float(b))));
                case K_LF32: return run((float*)out, n, addr, mrg_engine_log_normal_distribution<float, Engine>(float(a), float(b)));
                case K_LF64: return run((double*)out, n, addr, mrg_engine_log_normal_distribution<double, Engine>(a, b));
                case K_LF16:
                    return run((__half*)out, n, addr,
                               mrg_engine_log_normal_distribution<__half, Engine>(__float2half(float(a)), __float2half(float(b))));
                case K_POISSON:
                {
                    if(a <= 0.0)
                        return ROCRAND_STATUS_OUT_OF_RANGE;
                    if(a > rocrand_device::detail::lambda_threshold_huge)
                        return run((unsigned int*)out, n, addr,
                                   mrg_engine_poisson_distribution<Engine, poisson_distribution_huge>(poisson_distribution_huge(a)));
                    poisson_holder h(a);
                    using PD = poisson_distribution<DISCRETE_METHOD_ALIAS>;
                    return run((unsigned int*)out, n, addr, mrg_engine_poisson_distribution<Engine, PD>(PD(h.st)));
                }
            }
            return ROCRAND_STATUS_TYPE_ERROR;
        }
        else
        {
            switch(kind)
            {
                case K_U8: return run((unsigned char*)out, n, addr, uniform_distribution<unsigned char>());
                case K_U16: return run((unsigned short*)out, n, addr, uniform_distribution<unsigned short>());
                case K_U32: return run((unsigned int*)out, n, addr, uniform_distribution<unsigned int>());
                case K_UF32: return run((float*)out, n, addr, uniform_distribution<float>());
                case K_UF64: return run((double*)out, n, addr, uniform_distribution<double>());
                case K_UF16: return run((__half*)out, n, addr, uniform_distribution<__half>());
                case K_NF32: return run((float*)out, n, addr, normal_distribution<float>(float(a), float(b)));
                case K_NF64: return run((double*)out, n, addr, normal_distribution<double>(a, b));
                case K_NF16:
                    return run((__half*)out, n, addr,
                               normal_distribution<__half>(__float2half(float(a)), __float2half(float(b))));
                case K_LF32: return run((float*)out, n, addr, log_normal_distribution<float>(float(a), float(b)));
                case K_LF64: return run((double*)out, n, addr, log_normal_distribution<double>(a, b));
                case K_LF16:
                    return run((__half*)out, n, addr,
                               log_normal_distribution<__half>(__float2half(float(a)), __float2half(float(b))));
                case K_POISSON:
                {
                    if(a <= 0.0)
                        return ROCRAND_STATUS_OUT_OF_RANGE;
                    if(a > rocrand_device::detail::lambda_threshold_huge)
                        return run((unsigned int*)out, n, addr, poisson_distribution_huge(a));
                    poisson_holder h(a);
                    return run((unsigned int*)out, n, addr, poisson_distribution<DISCRETE_METHOD_ALIAS>(h.st));
                }
            }
            return ROCRAND_STATUS_TYPE_ERROR;
        }
    }
};

} // namespace

extern "C" {

void* ref_create(int type)
{
    gen_base* g = nullptr;
    switch(type)
    {
        case ROCRAND_RNG_PSEUDO_PHILOX4_32_10:
            g = new counter_gen<philox4x32_10_device_engine, ROCRAND_RNG_PSEUDO_PHILOX4_32_10, true>();
            break;
        case ROCRAND_RNG_PSEUDO_THREEFRY2_32_20:
            g = new counter_gen<threefry_device_engine<rocrand_device::threefry2x32_20_engine>,
                                ROCRAND_RNG_PSEUDO_THREEFRY2_32_20, false>();
            break;
        case ROCRAND_RNG_PSEUDO_THREEFRY2_64_20:
            g = new counter_gen<threefry_device_engine<rocrand_device::threefry2x64_20_engine>,
                                ROCRAND_RNG_PSEUDO_THREEFRY2_64_20, false>();
            break;
        case ROCRAND_RNG_PSEUDO_THREEFRY4_32_20:
            g = new counter_gen<threefry_device_engine<rocrand_device::threefry4x32_20_engine>,
                                ROCRAND_RNG_PSEUDO_THREEFRY4_32_20, false>();
            break;
        case ROCRAND_RNG_PSEUDO_THREEFRY4_64_20:
            g = new counter_gen<threefry_device_engine<rocrand_device::threefry4x64_20_engine>,
                                ROCRAND_RNG_PSEUDO_THREEFRY4_64_20, false>();
            break;
        case ROCRAND_RNG_PSEUDO_MRG32K3A:
            g = new interleaved_gen<rocrand_device::mrg32k3a_engine, ROCRAND_RNG_PSEUDO_MRG32K3A>();
            g->seed = 12345;
            break;
        case ROCRAND_RNG_PSEUDO_MRG31K3P:
            g = new interleaved_gen<rocrand_device::mrg31k3p_engine, ROCRAND_RNG_PSEUDO_MRG31K3P>();
            g->seed = 12345;
            break;
        case ROCRAND_RNG_PSEUDO_XORWOW:
            g = new interleaved_gen<xorwow_device_engine, ROCRAND_RNG_PSEUDO_XORWOW>();
            break;
        case ROCRAND_RNG_PSEUDO_LFSR113:
            g = new interleaved_gen<lfsr113_device_engine, ROCRAND_RNG_PSEUDO_LFSR113>();
            break;
        default: return nullptr;
    }
    g->type = type;
    return g;
}

void ref_destroy(void* g) { delete static_cast<gen_base*>(g); }

int ref_set_seed(void* p, unsigned long long seed)
{
    auto* g = static_cast<gen_base*>(p);
    if(g->type == ROCRAND_RNG_PSEUDO_LFSR113)
    {
        // lfsr113.hpp:240-255
        uint4 s = {static_cast<unsigned int>(seed), static_cast<unsigned int>(seed >> 16),
                   static_cast<unsigned int>(seed >> 32), static_cast<unsigned int>(seed >> 48)};
        if(s.x < ROCRAND_LFSR113_DEFAULT_SEED_X)
            s.x += ROCRAND_LFSR113_DEFAULT_SEED_X;
        if(s.y < ROCRAND_LFSR113_DEFAULT_SEED_Y)
            s.y += ROCRAND_LFSR113_DEFAULT_SEED_Y;
        g->seed4 = s;
    }
    else if(g->type == ROCRAND_RNG_PSEUDO_MRG32K3A || g->type == ROCRAND_RNG_PSEUDO_MRG31K3P)
        g->seed = seed == 0 ? 12345ull : seed; // mrg.hpp:240-244
    else
        g->seed = seed;
    g->initialized = false;
    return 0;
}

int ref_set_seed_uint4(void* p, const unsigned int* s4)
{
    auto* g = static_cast<gen_base*>(p);
    if(g->type != ROCRAND_RNG_PSEUDO_LFSR113)
        return ROCRAND_STATUS_TYPE_ERROR;
    uint4 s = {s4[0], s4[1], s4[2], s4[3]}; // lfsr113.hpp:257-272
    if(s.x < ROCRAND_LFSR113_DEFAULT_SEED_X) s.x += ROCRAND_LFSR113_DEFAULT_SEED_X;
    if(s.y < ROCRAND_LFSR113_DEFAULT_SEED_Y) s.y += ROCRAND_LFSR113_DEFAULT_SEED_Y;
    if(s.z < ROCRAND_LFSR113_DEFAULT_SEED_Z) s.z += ROCRAND_LFSR113_DEFAULT_SEED_Z;
    if(s.w < ROCRAND_LFSR113_DEFAULT_SEED_W) s.w += ROCRAND_LFSR113_DEFAULT_SEED_W;
    g->seed4       = s;
    g->initialized = false;
    return 0;
}

int ref_set_offset(void* p, unsigned long long offset)
{
    auto* g        = static_cast<gen_base*>(p);
    g->offset      = offset;
    g->initialized = false;
    return 0;
}

int ref_generate(void* p, int kind, void* out, size_t n, uint64_t addr, double a, double b)
{
    return static_cast<gen_base*>(p)->generate(kind, out, n, addr, a, b);
}

// Raw reference engine stream: Engine(seed, subsequence, offset) then n x next().
int ref_engine_stream(int type, unsigned long long seed, unsigned long long sub,
                      unsigned long long offset, unsigned int* out, size_t n)
{
    switch(type)
    {
        case ROCRAND_RNG_PSEUDO_PHILOX4_32_10:
        {
            rocrand_device::philox4x32_10_engine e(seed, sub, offset);
            for(size_t i = 0; i < n; i++) out[i] = e.next();
            return 0;
        }
        case ROCRAND_RNG_PSEUDO_THREEFRY4_32_20:
        {
            rocrand_device::threefry4x32_20_engine e(seed, sub, offset);
            for(size_t i = 0; i < n; i++) out[i] = e.next();
            return 0;
        }
        case ROCRAND_RNG_PSEUDO_THREEFRY2_32_20:
        {
            rocrand_device::threefry2x32_20_engine e(seed, sub, offset);
            for(size_t i = 0; i < n; i++) out[i] = e.next();
            return 0;
        }
        case ROCRAND_RNG_PSEUDO_MRG32K3A:
        {
            rocrand_device::mrg32k3a_engine e(seed, sub, offset);
            for(size_t i = 0; i < n; i++) out[i] = e.next();
            return 0;
        }
        case ROCRAND_RNG_PSEUDO_MRG31K3P:
        {
            rocrand_device::mrg31k3p_engine e(seed, sub, offset);
            for(size_t i = 0; i < n; i++) out[i] = e.next();
            return 0;
        }
        case ROCRAND_RNG_PSEUDO_XORWOW:
        {
            rocrand_device::xorwow_engine e(seed, sub, offset);
            for(size_t i = 0; i < n; i++) out[i] = e.next();
            return 0;
        }
        case ROCRAND_RNG_PSEUDO_LFSR113:
        {
            uint4 s = {static_cast<unsigned int>(seed), static_cast<unsigned int>(seed >> 16),
                       static_cast<unsigned int>(seed >> 32), static_cast<unsigned int>(seed >> 48)};
            rocrand_device::lfsr113_engine e(s, static_cast<unsigned int>(sub), offset);
            for(size_t i = 0; i < n; i++) out[i] = e.next();
            return 0;
        }
    }
    return -1;
}

} // extern "C"
