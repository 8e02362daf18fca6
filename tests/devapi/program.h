/* program.h -- the per-thread call sequence of the device-API parity tests.
 *
 * Written against the device API's public names only (rocrand_init, rocrand,
 * rocrand_uniform, ..., skipahead), so the same sequence compiles three ways:
 *   - device code over include/rocrand/rocrand_kernel.h  (tests/devapi/devapi_gpu.hip)
 *   - host code over include/rocrand/rocrand_kernel.h    (tests/devapi/devapi_host.cpp)
 *   - host code over the reference's rocrand_kernel.h    (oracle/ref_devapi.cpp)
 *
 * Thread t initialises its state with (seed, sub[t], off[t]) (a Sobol state: dimension
 * sub[t] mod dims of the given direction vectors, offset off[t]) and makes `calls` calls of
 * one function; call k writes its result components into out[k * 4 + c] as 64-bit slots
 * (32-bit integers and floats as their bit patterns, zero-extended).  Functions a state
 * type does not have are reported by supported() and never instantiated.
 */
#pragma once

#include <stdint.h>
#include <string.h>
#include <type_traits>
#include <utility>

namespace devapi_test
{

enum engine_id
{
    PHILOX = 0,
    MRG32K3A,
    MRG31K3P,
    XORWOW,
    LFSR113,
    TF2X32,
    TF2X64,
    TF4X32,
    TF4X64,
    SOBOL32,
    SOBOL64,
    SSOBOL32, // scrambled
    SSOBOL64,
    ENGINE_COUNT
};

enum fn_id
{
    WORD = 0, // rocrand
    WORDV,    // rocrand2 / rocrand4
    UNIFORM,
    UNIFORM2,
    UNIFORM4,
    UNIFORM_D,
    UNIFORM_D2,
    UNIFORM_D4,
    NORMAL,
    NORMAL2,
    NORMAL4,
    NORMAL_D,
    NORMAL_D2,
    NORMAL_D4,
    LOGN, // mean a, stddev b
    LOGN2,
    LOGN4,
    LOGN_D,
    LOGN_D2,
    LOGN_D4,
    POISSON, // lambda a
    POISSON4,
    DISCRETE,
    DISCRETE4,
    MIXED, // normal, uniform, normal_double, log_normal(a, b), normal: the Box-Muller caches
    SKIP,  // skipahead / skipahead_subsequence / skipahead_sequence, each followed by rocrand
    FN_COUNT
};

template<engine_id E>
struct state_of;
template<>
struct state_of<PHILOX>
{
    using type = rocrand_state_philox4x32_10;
};
template<>
struct state_of<MRG32K3A>
{
    using type = rocrand_state_mrg32k3a;
};
template<>
struct state_of<MRG31K3P>
{
    using type = rocrand_state_mrg31k3p;
};
template<>
struct state_of<XORWOW>
{
    using type = rocrand_state_xorwow;
};
template<>
struct state_of<LFSR113>
{
    using type = rocrand_state_lfsr113;
};
template<>
struct state_of<TF2X32>
{
    using type = rocrand_state_threefry2x32_20;
};
template<>
struct state_of<TF2X64>
{
    using type = rocrand_state_threefry2x64_20;
};
template<>
struct state_of<TF4X32>
{
    using type = rocrand_state_threefry4x32_20;
};
template<>
struct state_of<TF4X64>
{
    using type = rocrand_state_threefry4x64_20;
};
template<>
struct state_of<SOBOL32>
{
    using type = rocrand_state_sobol32;
};
template<>
struct state_of<SOBOL64>
{
    using type = rocrand_state_sobol64;
};
template<>
struct state_of<SSOBOL32>
{
    using type = rocrand_state_scrambled_sobol32;
};
template<>
struct state_of<SSOBOL64>
{
    using type = rocrand_state_scrambled_sobol64;
};

// Direction vectors of the Sobol states: `dims` dimensions of 32 (or 64) words.
struct sobol_tables
{
    const unsigned int*       v32;
    const unsigned long long* v64;
    unsigned int              dims;
};

// ---- which calls a state type has
#define DEVAPI_HAS(NAME, EXPR)                                                                               \
    template<class S, class = void>                                                                          \
    struct NAME : std::false_type                                                                            \
    {};                                                                                                      \
    template<class S>                                                                                        \
    struct NAME<S, std::void_t<decltype(EXPR)>> : std::true_type                                             \
    {};
DEVAPI_HAS(has_rocrand2, rocrand2(std::declval<S*>()))
DEVAPI_HAS(has_rocrand4, rocrand4(std::declval<S*>()))
DEVAPI_HAS(has_uniform2, rocrand_uniform2(std::declval<S*>()))
DEVAPI_HAS(has_uniform4, rocrand_uniform4(std::declval<S*>()))
DEVAPI_HAS(has_uniform_d2, rocrand_uniform_double2(std::declval<S*>()))
DEVAPI_HAS(has_uniform_d4, rocrand_uniform_double4(std::declval<S*>()))
DEVAPI_HAS(has_normal2, rocrand_normal2(std::declval<S*>()))
DEVAPI_HAS(has_normal_d2, rocrand_normal_double2(std::declval<S*>()))
DEVAPI_HAS(has_logn2, rocrand_log_normal2(std::declval<S*>(), 0.0f, 1.0f))
DEVAPI_HAS(has_logn_d2, rocrand_log_normal_double2(std::declval<S*>(), 0.0, 1.0))
DEVAPI_HAS(has_normal4, rocrand_normal4(std::declval<S*>()))
DEVAPI_HAS(has_normal_d4, rocrand_normal_double4(std::declval<S*>()))
DEVAPI_HAS(has_logn4, rocrand_log_normal4(std::declval<S*>(), 0.0f, 1.0f))
DEVAPI_HAS(has_logn_d4, rocrand_log_normal_double4(std::declval<S*>(), 0.0, 1.0))
DEVAPI_HAS(has_poisson4, rocrand_poisson4(std::declval<S*>(), 1.0))
DEVAPI_HAS(has_discrete4, rocrand_discrete4(std::declval<S*>(), std::declval<rocrand_discrete_distribution>()))
DEVAPI_HAS(has_skipahead, skipahead(1ull, std::declval<S*>()))
DEVAPI_HAS(has_skip_subsequence, skipahead_subsequence(1u, std::declval<S*>()))
#undef DEVAPI_HAS

template<class S>
constexpr bool supported(int fn)
{
    switch(fn)
    {
        case WORDV: return has_rocrand2<S>::value || has_rocrand4<S>::value;
        case UNIFORM2: return has_uniform2<S>::value;
        case UNIFORM4: return has_uniform4<S>::value;
        case UNIFORM_D2: return has_uniform_d2<S>::value;
        case UNIFORM_D4: return has_uniform_d4<S>::value;
        case NORMAL2: return has_normal2<S>::value;
        case NORMAL_D2: return has_normal_d2<S>::value;
        case LOGN2: return has_logn2<S>::value;
        case LOGN_D2: return has_logn_d2<S>::value;
        case NORMAL4: return has_normal4<S>::value;
        case NORMAL_D4: return has_normal_d4<S>::value;
        case LOGN4: return has_logn4<S>::value;
        case LOGN_D4: return has_logn_d4<S>::value;
        case POISSON4: return has_poisson4<S>::value;
        case DISCRETE4: return has_discrete4<S>::value;
        case SKIP: return has_skipahead<S>::value;
        default: return fn >= 0 && fn < FN_COUNT;
    }
}

__host__ __device__ inline uint64_t slot(unsigned int v) { return v; }
__host__ __device__ inline uint64_t slot(unsigned long long v) { return v; }
__host__ __device__ inline uint64_t slot(float v)
{
    uint32_t b;
    memcpy(&b, &v, 4);
    return b;
}
__host__ __device__ inline uint64_t slot(double v)
{
    uint64_t b;
    memcpy(&b, &v, 8);
    return b;
}
template<class V>
__host__ __device__ inline void put2(uint64_t* o, const V& v)
{
    o[0] = slot(v.x);
    o[1] = slot(v.y);
}
template<class V>
__host__ __device__ inline void put4(uint64_t* o, const V& v)
{
    o[0] = slot(v.x);
    o[1] = slot(v.y);
    o[2] = slot(v.z);
    o[3] = slot(v.w);
}

// Sobol: dimension sub mod dims; the scramble constant derives from the seed and the dimension
__host__ __device__ inline void init(rocrand_state_sobol32* s, unsigned long long, unsigned long long sub,
                                     unsigned long long off, const sobol_tables& t)
{
    rocrand_init(t.v32 + (sub % t.dims) * 32, (unsigned int)off, s);
}
__host__ __device__ inline void init(rocrand_state_sobol64* s, unsigned long long, unsigned long long sub,
                                     unsigned long long off, const sobol_tables& t)
{
    rocrand_init(t.v64 + (sub % t.dims) * 64, (unsigned int)off, s);
}
__host__ __device__ inline void init(rocrand_state_scrambled_sobol32* s, unsigned long long seed,
                                     unsigned long long sub, unsigned long long off, const sobol_tables& t)
{
    rocrand_init(t.v32 + (sub % t.dims) * 32, (unsigned int)(seed ^ (sub * 0x9E3779B97F4A7C15ull)), (unsigned int)off,
                 s);
}
__host__ __device__ inline void init(rocrand_state_scrambled_sobol64* s, unsigned long long seed,
                                     unsigned long long sub, unsigned long long off, const sobol_tables& t)
{
    rocrand_init(t.v64 + (sub % t.dims) * 64, seed ^ (sub * 0x9E3779B97F4A7C15ull), (unsigned int)off, s);
}
__host__ __device__ inline void init(rocrand_state_lfsr113* s, unsigned long long seed, unsigned long long sub,
                                     unsigned long long off, const sobol_tables&)
{
    // the seed words must exceed 1, 7, 15 and 127 for the full period
    const uint4 seed4 = {(unsigned int)seed | 2u, (unsigned int)(seed >> 32) | 8u,
                         ((unsigned int)seed ^ 0x9e3779b9u) | 16u, ((unsigned int)(seed >> 32) + 12345u) | 128u};
    rocrand_init(seed4, (unsigned int)sub, off, s);
}
template<class S>
__host__ __device__ inline void init(S* s, unsigned long long seed, unsigned long long sub, unsigned long long off,
                                     const sobol_tables&)
{
    rocrand_init(seed, sub, off, s);
}

template<class S, int FN>
__host__ __device__ inline void call(S* s, int k, double a, double b, rocrand_discrete_distribution dist,
                                     uint64_t* o)
{
    if constexpr(FN == WORD)
        o[0] = slot(rocrand(s));
    else if constexpr(FN == WORDV)
    {
        if constexpr(has_rocrand4<S>::value)
            put4(o, rocrand4(s));
        else
            put2(o, rocrand2(s));
    }
    else if constexpr(FN == UNIFORM)
        o[0] = slot(rocrand_uniform(s));
    else if constexpr(FN == UNIFORM2)
        put2(o, rocrand_uniform2(s));
    else if constexpr(FN == UNIFORM4)
        put4(o, rocrand_uniform4(s));
    else if constexpr(FN == UNIFORM_D)
        o[0] = slot(rocrand_uniform_double(s));
    else if constexpr(FN == UNIFORM_D2)
        put2(o, rocrand_uniform_double2(s));
    else if constexpr(FN == UNIFORM_D4)
        put4(o, rocrand_uniform_double4(s));
    else if constexpr(FN == NORMAL)
        o[0] = slot(rocrand_normal(s));
    else if constexpr(FN == NORMAL2)
        put2(o, rocrand_normal2(s));
    else if constexpr(FN == NORMAL4)
        put4(o, rocrand_normal4(s));
    else if constexpr(FN == NORMAL_D)
        o[0] = slot(rocrand_normal_double(s));
    else if constexpr(FN == NORMAL_D2)
        put2(o, rocrand_normal_double2(s));
    else if constexpr(FN == NORMAL_D4)
        put4(o, rocrand_normal_double4(s));
    else if constexpr(FN == LOGN)
        o[0] = slot(rocrand_log_normal(s, a, b));
    else if constexpr(FN == LOGN2)
        put2(o, rocrand_log_normal2(s, (float)a, (float)b));
    else if constexpr(FN == LOGN4)
        put4(o, rocrand_log_normal4(s, (float)a, (float)b));
    else if constexpr(FN == LOGN_D)
        o[0] = slot(rocrand_log_normal_double(s, a, b));
    else if constexpr(FN == LOGN_D2)
        put2(o, rocrand_log_normal_double2(s, a, b));
    else if constexpr(FN == LOGN_D4)
        put4(o, rocrand_log_normal_double4(s, a, b));
    else if constexpr(FN == POISSON)
        o[0] = slot(rocrand_poisson(s, a));
    else if constexpr(FN == POISSON4)
        put4(o, rocrand_poisson4(s, a));
    else if constexpr(FN == DISCRETE)
        o[0] = slot(rocrand_discrete(s, dist));
    else if constexpr(FN == DISCRETE4)
        put4(o, rocrand_discrete4(s, dist));
    else if constexpr(FN == MIXED)
    {
        switch(k % 5)
        {
            case 0: o[0] = slot(rocrand_normal(s)); break;
            case 1: o[0] = slot(rocrand_uniform(s)); break;
            case 2: o[0] = slot(rocrand_normal_double(s)); break;
            case 3: o[0] = slot(rocrand_log_normal(s, a, b)); break;
            default: o[0] = slot(rocrand_normal(s)); break;
        }
    }
    else if constexpr(FN == SKIP)
    {
        if constexpr(has_skip_subsequence<S>::value)
        {
            switch(k % 3)
            {
                case 0: skipahead(((unsigned long long)(k + 1) << 37) + 1000003ull * (unsigned long long)k, s); break;
                case 1: skipahead_subsequence((unsigned int)(k + 1), s); break;
                default: skipahead_sequence((unsigned int)(k + 1), s); break;
            }
        }
        else
            skipahead(((unsigned long long)(k + 1) << 37) + 1000003ull * (unsigned long long)k, s);
        o[0] = slot(rocrand(s));
    }
}

// One thread's whole sequence.
template<class S, int FN>
__host__ __device__ inline void run_thread(unsigned long long seed, unsigned long long sub, unsigned long long off,
                                           int calls, double a, double b, rocrand_discrete_distribution dist,
                                           const sobol_tables& tables, uint64_t* out)
{
    S s;
    init(&s, seed, sub, off, tables);
    for(int k = 0; k < calls; k++)
        call<S, FN>(&s, k, a, b, dist, out + (size_t)k * 4);
}

// The raw bytes of a state after init(seed, sub, off), as 32-bit words (zero-padded).
template<class S>
__host__ __device__ inline void state_words(unsigned long long seed, unsigned long long sub, unsigned long long off,
                                            const sobol_tables& tables, uint32_t* out, int words)
{
    S s;
    memset(&s, 0, sizeof(S));
    init(&s, seed, sub, off, tables);
    const int n = (int)((sizeof(S) + 3) / 4);
    uint32_t  w[(sizeof(S) + 3) / 4] = {};
    memcpy(w, &s, sizeof(S));
    for(int i = 0; i < words; i++)
        out[i] = i < n ? w[i] : 0u;
}

// Calls fn(std::integral_constant<int, FN>) for the runtime fn id.
template<class F>
inline bool with_fn(int fn, F&& f)
{
    switch(fn)
    {
#define DEVAPI_FN(ID)                                                                                        \
    case ID: f(std::integral_constant<int, ID>{}); return true;
        DEVAPI_FN(WORD)
        DEVAPI_FN(WORDV)
        DEVAPI_FN(UNIFORM)
        DEVAPI_FN(UNIFORM2)
        DEVAPI_FN(UNIFORM4)
        DEVAPI_FN(UNIFORM_D)
        DEVAPI_FN(UNIFORM_D2)
        DEVAPI_FN(UNIFORM_D4)
        DEVAPI_FN(NORMAL)
        DEVAPI_FN(NORMAL2)
        DEVAPI_FN(NORMAL4)
        DEVAPI_FN(NORMAL_D)
        DEVAPI_FN(NORMAL_D2)
        DEVAPI_FN(NORMAL_D4)
        DEVAPI_FN(LOGN)
        DEVAPI_FN(LOGN2)
        DEVAPI_FN(LOGN4)
        DEVAPI_FN(LOGN_D)
        DEVAPI_FN(LOGN_D2)
        DEVAPI_FN(LOGN_D4)
        DEVAPI_FN(POISSON)
        DEVAPI_FN(POISSON4)
        DEVAPI_FN(DISCRETE)
        DEVAPI_FN(DISCRETE4)
        DEVAPI_FN(MIXED)
        DEVAPI_FN(SKIP)
#undef DEVAPI_FN
        default: return false;
    }
}

// Calls f(state_of<E>::type*) for the runtime engine id.
template<class F>
inline bool with_engine(int engine, F&& f)
{
    switch(engine)
    {
        case PHILOX: f((state_of<PHILOX>::type*)nullptr); return true;
        case MRG32K3A: f((state_of<MRG32K3A>::type*)nullptr); return true;
        case MRG31K3P: f((state_of<MRG31K3P>::type*)nullptr); return true;
        case XORWOW: f((state_of<XORWOW>::type*)nullptr); return true;
        case LFSR113: f((state_of<LFSR113>::type*)nullptr); return true;
        case TF2X32: f((state_of<TF2X32>::type*)nullptr); return true;
        case TF2X64: f((state_of<TF2X64>::type*)nullptr); return true;
        case TF4X32: f((state_of<TF4X32>::type*)nullptr); return true;
        case TF4X64: f((state_of<TF4X64>::type*)nullptr); return true;
        case SOBOL32: f((state_of<SOBOL32>::type*)nullptr); return true;
        case SOBOL64: f((state_of<SOBOL64>::type*)nullptr); return true;
        case SSOBOL32: f((state_of<SSOBOL32>::type*)nullptr); return true;
        case SSOBOL64: f((state_of<SSOBOL64>::type*)nullptr); return true;
        default: return false;
    }
}

// The host run of every thread: returns -1 for an unknown or unsupported pair.
inline int run_host(int engine, int fn, unsigned long long seed, const unsigned long long* sub,
                    const unsigned long long* off, int n, int calls, double a, double b,
                    rocrand_discrete_distribution dist, const sobol_tables& tables, uint64_t* out)
{
    int rc = -1;
    with_engine(engine,
                [&](auto* tag)
                {
                    using S = std::remove_pointer_t<decltype(tag)>;
                    if(!supported<S>(fn))
                        return;
                    with_fn(fn,
                            [&](auto fc)
                            {
                                constexpr int FN = decltype(fc)::value;
                                if constexpr(supported<S>(FN))
                                {
                                    for(int t = 0; t < n; t++)
                                        run_thread<S, FN>(seed, sub[t], off[t], calls, a, b, dist, tables,
                                                          out + (size_t)t * calls * 4);
                                    rc = 0;
                                }
                            });
                });
    return rc;
}

} // namespace devapi_test
