// ref_harness.cpp -- drives the REFERENCE rocRAND kernel bodies on the CPU (test infrastructure).
//
// Built by oracle/build_ref.sh from the reference sources where they lie
// (library/include, library/src), host-only, into oracle/_ref/libref.so.  It is a
// second, independent checker for the oracle restatement: the engines, the
// distribution functors and the generate_* kernel bodies executed here are the
// reference's own code; this file only supplies the generator bookkeeping that
// library/src/rng/{philox4x32_10,threefry,mrg,xorwow,lfsr113,mtgp32,sobol}.hpp perform around
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
#include "rng/mt19937.hpp"
#include "rng/mtgp32.hpp"
#include "rng/philox4x32_10.hpp"
#include "rng/sobol.hpp"
#include "rng/threefry.hpp"
#include "rng/xorwow.hpp"

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <memory>
#include <thread>
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

// Blocks split over host threads: for kernels whose blocks write disjoint data.
template<class Kernel, class... Args>
void launch_parallel(unsigned blocks, unsigned threads, Kernel k, Args... args)
{
    const unsigned           T = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for(unsigned w = 0; w < T; w++)
        pool.emplace_back([=] {
            for(unsigned b = w; b < blocks; b += T)
                for(unsigned t = 0; t < threads; t++)
                    k(dim3(b), dim3(t), dim3(blocks), dim3(threads), args...);
        });
    for(auto& th : pool)
        th.join();
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
    // threads x blocks of the generate kernels' grid (0: the DEFAULT config); the kernels
    // run with 256-thread blocks, so only the product matters
    unsigned           grid = 0;
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
        const unsigned blocks = grid ? grid / 256 : 1024;
        if constexpr(IsPhilox)
            launch(blocks, 256, generate_philox<T, D>, engine, data, n, dist);
        else
            launch(blocks, 256, generate_threefry<Engine, T, D>, engine, data, n, dist);
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
    static constexpr unsigned S_DEFAULT = 131072; // 512 x 256, config_defaults.hpp:33-73
    std::vector<Engine>       engines;
    unsigned                  start = 0;
    unsigned                  S     = S_DEFAULT;
    using CP                  = default_config_provider<Type>;

    void init()
    {
        S     = grid ? grid : S_DEFAULT;
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
            launch(S / 256, 256, generate_xorwow<CP, false, T, D>, engines.data(), start, data, n, dist);
        else if constexpr(Type == ROCRAND_RNG_PSEUDO_LFSR113)
            launch(S / 256, 256, generate_lfsr113<CP, false, T, D>, engines.data(), start, data, n, dist);
        else
            launch(S / 256, 256, generate_mrg<CP, false, Engine, T, D>, engines.data(), start, data, n, dist);
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

// The generic 32-bit word distributions (uniform.hpp, normal.hpp, log_normal.hpp, poisson.hpp
// with the alias method) as used by the MTGP32 and MT19937 generators.
template<class G>
int dispatch_w32(G& g, int kind, void* out, size_t n, uint64_t addr, double a, double b)
{
    switch(kind)
    {
        case K_U8: return g.run((unsigned char*)out, n, addr, uniform_distribution<unsigned char>());
        case K_U16: return g.run((unsigned short*)out, n, addr, uniform_distribution<unsigned short>());
        case K_U32: return g.run((unsigned int*)out, n, addr, uniform_distribution<unsigned int>());
        case K_UF32: return g.run((float*)out, n, addr, uniform_distribution<float>());
        case K_UF64: return g.run((double*)out, n, addr, uniform_distribution<double>());
        case K_UF16: return g.run((__half*)out, n, addr, uniform_distribution<__half>());
        case K_NF32: return g.run((float*)out, n, addr, normal_distribution<float>(float(a), float(b)));
        case K_NF64: return g.run((double*)out, n, addr, normal_distribution<double>(a, b));
        case K_NF16:
            return g.run((__half*)out, n, addr, normal_distribution<__half>(__float2half(float(a)), __float2half(float(b))));
        case K_LF32: return g.run((float*)out, n, addr, log_normal_distribution<float>(float(a), float(b)));
        case K_LF64: return g.run((double*)out, n, addr, log_normal_distribution<double>(a, b));
        case K_LF16:
            return g.run((__half*)out, n, addr,
                         log_normal_distribution<__half>(__float2half(float(a)), __float2half(float(b))));
        case K_POISSON:
        {
            if(a <= 0.0)
                return ROCRAND_STATUS_OUT_OF_RANGE;
            if(a > rocrand_device::detail::lambda_threshold_huge)
                return g.run((unsigned int*)out, n, addr, poisson_distribution_huge(a));
            poisson_holder h(a);
            return g.run((unsigned int*)out, n, addr, poisson_distribution<DISCRETE_METHOD_ALIAS>(h.st));
        }
    }
    return ROCRAND_STATUS_TYPE_ERROR;
}

// ------------------------------------------------------------------- MTGP32
// mtgp32.hpp:429-476 init (rocrand_make_state_mtgp32's host loop, rocrand_mtgp32.h:362-390,
// without its hipMemcpy) and :478-521 generate: 512 blocks, one host thread per block.
struct mtgp_gen : gen_base
{
    static constexpr unsigned G = 512; // config_defaults.hpp:45-50
    std::vector<mtgp32_device_engine> engines;
    using CP = default_config_provider<ROCRAND_RNG_PSEUDO_MTGP32>;

    void init()
    {
        engines.assign(G, mtgp32_device_engine());
        unsigned long long s = seed ^ (seed >> 32);
        for(unsigned i = 0; i < G; i++)
        {
            auto& e = engines[i];
            rocrand_device::rocrand_mtgp32_init_state(&e.m_state.status[0], &mtgp32dc_params_fast_11213[i],
                                                      (unsigned int)s + i + 1);
            e.m_state.offset = 0;
            e.m_state.id     = i;
            e.pos_tbl        = mtgp32dc_params_fast_11213[i].pos;
            e.sh1_tbl        = mtgp32dc_params_fast_11213[i].sh1;
            e.sh2_tbl        = mtgp32dc_params_fast_11213[i].sh2;
            e.mask           = mtgp32dc_params_fast_11213[0].mask;
            for(int j = 0; j < MTGP_TS; j++)
            {
                e.param_tbl[j]         = mtgp32dc_params_fast_11213[i].tbl[j];
                e.temper_tbl[j]        = mtgp32dc_params_fast_11213[i].tmp_tbl[j];
                e.single_temper_tbl[j] = mtgp32dc_params_fast_11213[i].flt_tmp_tbl[j];
            }
        }
        initialized = true;
    }

    template<class T, class D>
    int run(T* out, size_t n, uint64_t addr, D dist)
    {
        if(!initialized)
            init();
        aligned_out tmp(n * sizeof(T), addr);
        T*          data = reinterpret_cast<T*>(tmp.p);
        launch(G, 1, generate_mtgp<CP, false, T, D>, engines.data(), data, n, dist);
        std::memcpy(out, data, n * sizeof(T));
        return 0;
    }

    int generate(int kind, void* out, size_t n, uint64_t addr, double a, double b) override
    {
        return dispatch_w32(*this, kind, out, n, addr, a, b);
    }
};

// ------------------------------------------------------------------ MT19937
// mt19937.hpp:793-908 init (jump_ahead_mt19937, init_engines_mt19937) and :911-1042
// generate (short/long kernel choice and the start-input bookkeeping), over the default
// grid 256 x 256 (config_defaults.hpp:76-80: 8192 generators of 8 threads).  The jump
// kernel's host form costs ~10 ms per generator (spread over host threads here); only the
// first `limit` generators are set up (the others stay zero): callers keep every touched
// thread below 8*limit.
// The host form of the kernels shares one head/tail flag among eight emulated lanes
// (mt19937.hpp:355-431), so only aligned calls of whole vectors are comparable.
struct mt_gen : gen_base
{
    static constexpr unsigned G = 8192, STRIDE = 65536, NN = mt19937_constants::n;
    using CP = default_config_provider<ROCRAND_RNG_PSEUDO_MT19937>;
    unsigned              limit = 64;
    std::vector<unsigned> engines;
    unsigned              start_input = 0, prev_in_w = 0;

    void init()
    {
        std::vector<unsigned> base(size_t(G) * NN, 0u);
        launch_parallel(limit, 128, jump_ahead_mt19937<128, CP, false>, base.data(), seed, rocrand_h_mt19937_jump);
        engines.assign(size_t(NN / 8) * STRIDE, 0u);
        launch(256, 256, init_engines_mt19937<CP, false>, engines.data(), (const unsigned*)base.data());
        start_input = 0;
        prev_in_w   = 0;
        initialized = true;
    }

    template<class T, class D>
    int run(T* out, size_t size, uint64_t addr, D dist)
    {
        if(!initialized)
            init();
        constexpr unsigned in_w = D::input_width, out_w = D::output_width;
        constexpr unsigned ips  = (NN / 8) / in_w;
        const unsigned     full_stride = STRIDE * ips;
        using vec_type     = rocrand_impl::aligned_vec_type<T, out_w>;
        aligned_out tmp(size * sizeof(T), addr);
        T*          data = reinterpret_cast<T*>(tmp.p);
        const uintptr_t    up    = reinterpret_cast<uintptr_t>(data);
        const size_t       mis   = (out_w - up / sizeof(T) % out_w) % out_w;
        const unsigned int head  = static_cast<unsigned>(std::min<size_t>(size, mis));
        const unsigned int tail  = static_cast<unsigned>((size - head) % out_w);
        const size_t       vec_size = (size - head) / out_w;
        const unsigned     extra = (head > 0 || tail > 0) ? 1 : 0;
        vec_type*          vec_data = reinterpret_cast<vec_type*>(data + mis);
        if(prev_in_w != in_w && start_input > 0)
        {
            const unsigned used = (start_input + STRIDE - 1) / STRIDE * prev_in_w;
            start_input         = (used + in_w - 1) / in_w * STRIDE;
            if(start_input >= full_stride)
                start_input = 0;
        }
        if(start_input > 0 && start_input + vec_size + extra <= full_stride)
            launch(256, 256, generate_short_mt19937<CP, false, T, vec_type, D>, engines.data(), start_input, data,
                   size, vec_data, vec_size, head, tail, dist);
        else
            launch(256, 256, generate_long_mt19937<CP, false, T, vec_type, D>, engines.data(), start_input, data,
                   size, vec_data, vec_size, head, tail, dist);
        std::memcpy(out, data, size * sizeof(T));
        start_input = static_cast<unsigned>((start_input + vec_size + extra) % full_stride);
        prev_in_w   = in_w;
        return 0;
    }

    int generate(int kind, void* out, size_t n, uint64_t addr, double a, double b) override
    {
        return dispatch_w32(*this, kind, out, n, addr, a, b);
    }
};

// -------------------------------------------------------------------- Sobol
// 32-bit direction vectors come from the reference's own tables (library/src/*.S + .bin);
// the 64-bit tables are absent from the reference tree and are handed in by the caller
// (ref_sobol64_set_vectors).
std::vector<unsigned long long> g_sobol64_vectors, g_scrambled_sobol64_vectors;

struct cdf_holder
{
    rocrand_discrete_distribution_st st{};
    std::vector<double>              cdf;
    explicit cdf_holder(double lambda)
    {
        unsigned int        size, offset;
        std::vector<double> p = calculate_poisson_probabilities(lambda, size, offset);
        using F               = discrete_distribution_factory<DISCRETE_METHOD_CDF, true>;
        F::normalize(p, size);
        cdf.resize(size);
        F::create_cdf(p, size, cdf.begin());
        st.size   = size;
        st.offset = offset;
        st.cdf    = cdf.data();
    }
};

template<bool Is64, bool Scrambled>
struct sobol_gen : gen_base
{
    using constant_type = std::conditional_t<Is64, unsigned long long, unsigned int>;
    using engine_type   = sobol_device_engine_t<Is64, Scrambled, false>;
    unsigned dims           = 1;
    unsigned current_offset = 0;

    static const constant_type* vectors()
    {
        if constexpr(Is64)
            return Scrambled ? g_scrambled_sobol64_vectors.data() : g_sobol64_vectors.data();
        else
            return Scrambled ? rocrand_h_scrambled_sobol32_direction_vectors
                             : rocrand_h_sobol32_direction_vectors;
    }

    // sobol.hpp:618-718
    template<class T, class D>
    int run(T* out, size_t data_size, uint64_t addr, D dist)
    {
        if constexpr(!Is64 && std::is_same_v<T, unsigned long long>)
            return ROCRAND_STATUS_TYPE_ERROR;
        constexpr unsigned opt = sizeof(T) >= sizeof(int) ? 1 : sizeof(int) / sizeof(T);
        if(data_size % dims != 0)
            return ROCRAND_STATUS_LENGTH_NOT_MULTIPLE;
        if(!initialized)
        {
            current_offset = static_cast<unsigned int>(offset);
            initialized    = true;
        }
        const constant_type* v = vectors();
        if(Is64 && v == nullptr)
            return ROCRAND_STATUS_INTERNAL_ERROR;
        // quirk of the generator's constructor (sobol.hpp:514): the scramble-constant
        // pointer receives the direction-vector table
        const constant_type* sc = v;
        aligned_out        tmp(data_size * sizeof(T), addr);
        T*                 data = reinterpret_cast<T*>(tmp.p);
        const size_t       size = data_size / dims;
        constexpr uint32_t threads = 256, max_blocks = 4096, opb = threads * opt;
        const uint32_t blocks = std::min(max_blocks, static_cast<uint32_t>((size + opb - 1) / opb));
        uint32_t       bx     = 1;
        while(bx < (blocks + dims - 1) / dims)
            bx *= 2;
        for(uint32_t by = 0; by < dims; by++)
            for(uint32_t x = 0; x < bx; x++)
                for(uint32_t t = 0; t < threads; t++)
                    generate_sobol_host<opt, Scrambled, engine_type, constant_type, T, D>(
                        dim3(x, by), dim3(t), dim3(bx, dims), dim3(threads), data, size, v, sc,
                        current_offset, dist);
        std::memcpy(out, data, data_size * sizeof(T));
        current_offset += size;
        return 0;
    }

    int generate(int kind, void* out, size_t n, uint64_t addr, double a, double b) override
    {
        switch(kind)
        {
            case K_U8: return run((unsigned char*)out, n, addr, sobol_uniform_distribution<unsigned char>());
            case K_U16: return run((unsigned short*)out, n, addr, sobol_uniform_distribution<unsigned short>());
            case K_U32: return run((unsigned int*)out, n, addr, sobol_uniform_distribution<unsigned int>());
            case K_U64:
                if constexpr(Is64)
                    return run((unsigned long long*)out, n, addr, sobol_uniform_distribution<unsigned long long>());
                else
                    return ROCRAND_STATUS_TYPE_ERROR;
            case K_UF32: return run((float*)out, n, addr, sobol_uniform_distribution<float>());
            case K_UF64: return run((double*)out, n, addr, sobol_uniform_distribution<double>());
            case K_UF16: return run((__half*)out, n, addr, sobol_uniform_distribution<__half>());
            case K_NF32: return run((float*)out, n, addr, sobol_normal_distribution<float>(float(a), float(b)));
            case K_NF64: return run((double*)out, n, addr, sobol_normal_distribution<double>(a, b));
            case K_NF16:
                return run((__half*)out, n, addr,
                           sobol_normal_distribution<__half>(__float2half(float(a)), __float2half(float(b))));
            case K_LF32: return run((float*)out, n, addr, sobol_log_normal_distribution<float>(float(a), float(b)));
            case K_LF64: return run((double*)out, n, addr, sobol_log_normal_distribution<double>(a, b));
            case K_LF16:
                return run((__half*)out, n, addr,
                           sobol_log_normal_distribution<__half>(__float2half(float(a)), __float2half(float(b))));
            case K_POISSON:
            {
                if(a <= 0.0)
                    return ROCRAND_STATUS_OUT_OF_RANGE;
                if(a > rocrand_device::detail::lambda_threshold_huge)
                    return run((unsigned int*)out, n, addr, poisson_distribution_huge(a));
                cdf_holder h(a);
                return run((unsigned int*)out, n, addr, poisson_distribution<DISCRETE_METHOD_CDF>(h.st));
            }
        }
        return ROCRAND_STATUS_TYPE_ERROR;
    }
};

} // namespace

extern "C" {

static bool is_sobol(int t)
{
    return t == ROCRAND_RNG_QUASI_SOBOL32 || t == ROCRAND_RNG_QUASI_SCRAMBLED_SOBOL32
           || t == ROCRAND_RNG_QUASI_SOBOL64 || t == ROCRAND_RNG_QUASI_SCRAMBLED_SOBOL64;
}


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
        case ROCRAND_RNG_PSEUDO_MTGP32: g = new mtgp_gen(); break;
        case ROCRAND_RNG_PSEUDO_MT19937: g = new mt_gen(); break;
        case ROCRAND_RNG_QUASI_SOBOL32: g = new sobol_gen<false, false>(); break;
        case ROCRAND_RNG_QUASI_SCRAMBLED_SOBOL32: g = new sobol_gen<false, true>(); break;
        case ROCRAND_RNG_QUASI_SOBOL64: g = new sobol_gen<true, false>(); break;
        case ROCRAND_RNG_QUASI_SCRAMBLED_SOBOL64: g = new sobol_gen<true, true>(); break;
        default: return nullptr;
    }
    g->type = type;
    return g;
}

void ref_destroy(void* g) { delete static_cast<gen_base*>(g); }

int ref_set_seed(void* p, unsigned long long seed)
{
    auto* g = static_cast<gen_base*>(p);
    if(is_sobol(g->type))
        return 0; // sobol.hpp:581-584: ignored
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

int ref_set_dimensions(void* p, unsigned int dims)
{
    auto* g = static_cast<gen_base*>(p);
    if(!is_sobol(g->type))
        return ROCRAND_STATUS_TYPE_ERROR;
    if(dims < 1 || dims > 20000) // sobol.hpp:573-585
        return ROCRAND_STATUS_OUT_OF_RANGE;
    switch(g->type)
    {
        case ROCRAND_RNG_QUASI_SOBOL32: static_cast<sobol_gen<false, false>*>(g)->dims = dims; break;
        case ROCRAND_RNG_QUASI_SCRAMBLED_SOBOL32: static_cast<sobol_gen<false, true>*>(g)->dims = dims; break;
        case ROCRAND_RNG_QUASI_SOBOL64: static_cast<sobol_gen<true, false>*>(g)->dims = dims; break;
        default: static_cast<sobol_gen<true, true>*>(g)->dims = dims; break;
    }
    g->initialized = false;
    return 0;
}

// The 64-bit direction-vector tables (20000 x 64 words each), which the reference tree lacks.
void ref_sobol64_set_vectors(const unsigned long long* plain, const unsigned long long* scrambled)
{
    g_sobol64_vectors.assign(plain, plain + 1280000);
    g_scrambled_sobol64_vectors.assign(scrambled, scrambled + 1280000);
}

const unsigned int* ref_sobol32_vectors() { return rocrand_h_sobol32_direction_vectors; }
const unsigned int* ref_mt19937_jump() { return rocrand_h_mt19937_jump; }

// MT19937: number of generators whose engines the next initialisation sets up.
int ref_mt19937_set_limit(void* p, unsigned int limit)
{
    auto* g = static_cast<gen_base*>(p);
    if(g->type != ROCRAND_RNG_PSEUDO_MT19937 || limit < 1 || limit > 8192)
        return ROCRAND_STATUS_OUT_OF_RANGE;
    static_cast<mt_gen*>(g)->limit = limit;
    g->initialized                 = false;
    return 0;
}
const unsigned int* ref_scrambled_sobol32_vectors() { return rocrand_h_scrambled_sobol32_direction_vectors; }
const unsigned int* ref_scramble32_constants() { return h_scrambled_sobol32_constants; }
const unsigned long long* ref_scramble64_constants() { return h_scrambled_sobol64_constants; }

// A tuned launch grid (threads x blocks, a multiple of 256) for the counter and
// interleaved engines, as a DYNAMIC-ordering config would select (config_types.hpp:314-336).
int ref_set_grid(void* p, unsigned threads, unsigned blocks)
{
    auto*                    g    = static_cast<gen_base*>(p);
    const unsigned long long size = (unsigned long long)threads * blocks;
    if(size == 0 || size % 256 != 0 || size > (1u << 22))
        return ROCRAND_STATUS_OUT_OF_RANGE;
    switch(g->type)
    {
        case ROCRAND_RNG_PSEUDO_MTGP32:
        case ROCRAND_RNG_PSEUDO_MT19937:
        case ROCRAND_RNG_QUASI_SOBOL32:
        case ROCRAND_RNG_QUASI_SOBOL64:
        case ROCRAND_RNG_QUASI_SCRAMBLED_SOBOL32:
        case ROCRAND_RNG_QUASI_SCRAMBLED_SOBOL64: return ROCRAND_STATUS_TYPE_ERROR;
        default: break;
    }
    g->grid        = (unsigned)size;
    g->initialized = false;
    return 0;
}

int ref_set_offset(void* p, unsigned long long offset)
{
    auto* g        = static_cast<gen_base*>(p);
    if(g->type == ROCRAND_RNG_PSEUDO_MTGP32 || g->type == ROCRAND_RNG_PSEUDO_MT19937)
        return ROCRAND_STATUS_TYPE_ERROR;
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

// The tables rocrand_create_discrete_distribution builds for `size` user weights
// (rocrand.cpp:506-533): the reference's own factory, host-side instance
// (discrete.hpp:111-143 create: normalize, create_alias_table, create_cdf).
// Returns 0, or -1 when the factory reports an error.
int ref_discrete_tables(const double* weights, unsigned int size, unsigned int offset,
                        double* prob, unsigned int* alias, double* cdf)
{
    using F = discrete_distribution_factory<DISCRETE_METHOD_UNIVERSAL, true>;
    rocrand_discrete_distribution_st st{};
    if(F::create(std::vector<double>(weights, weights + size), size, offset, st)
       != ROCRAND_STATUS_SUCCESS)
        return -1;
    std::copy(st.probability, st.probability + size, prob);
    std::copy(st.alias, st.alias + size, alias);
    std::copy(st.cdf, st.cdf + size, cdf);
    F::deallocate(st);
    return 0;
}

// The tables rocrand_create_poisson_distribution builds for `lambda` (rocrand.cpp:462-500:
// calculate_poisson_probabilities, poisson.hpp:110-161, then the same factory).  Returns the
// table size and writes its offset; writes the tables only when the size is at most `cap`.
unsigned int ref_poisson_tables(double lambda, unsigned int cap, unsigned int* offset,
                                double* prob, unsigned int* alias, double* cdf)
{
    unsigned int              size, off;
    const std::vector<double> p = calculate_poisson_probabilities(lambda, size, off);
    *offset                     = off;
    if(size > cap)
        return size;
    return ref_discrete_tables(p.data(), size, off, prob, alias, cdf) == 0 ? size : 0;
}

} // extern "C"
