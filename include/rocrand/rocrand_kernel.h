/*
 * rocrand_kernel.h -- device API: per-thread generator states for user kernels.
 *
 * Same names, state types and call semantics as the reference's device API
 * (library/include/rocrand/rocrand_kernel.h and the engine / distribution headers it
 * includes), for Philox4x32-10, MRG32k3a, MRG31k3p, XORWOW, LFSR113, ThreeFry
 * 2x32/2x64/4x32/4x64-20, Sobol32/64 (plain and scrambled) and MTGP32: rocrand_init, rocrand / rocrand2 / rocrand4, skipahead,
 * skipahead_subsequence, skipahead_sequence, rocrand_uniform*, rocrand_normal*,
 * rocrand_log_normal*, rocrand_poisson* and rocrand_discrete*.  A state initialised with
 * (seed, subsequence, offset) produces the words of the reference state initialised the
 * same way; the engine arithmetic and the float transforms are the library's own
 * (rocrand_amd_engines.h, rocrand_amd_math.h), so a thread's normals equal the bulk
 * generator's for the same words.  Jump matrices: rocrand_amd_jumptables.h.
 *
 * Sobol32/64 and their scrambled forms take the caller's direction vectors and scramble
 * constants (rocrand_get_direction_vectors32/64, rocrand_get_scramble_constants32/64).
 * MTGP32: block-shared states made by rocrand_make_state_mtgp32 from the parameter sets of
 * rocrand_mtgp32_11213.h (the array is exported by librocrand.so).
 */
#ifndef ROCRAND_KERNEL_H_
#define ROCRAND_KERNEL_H_

#include "rocrand_amd_engines.h"
#include "rocrand_amd_jumptables.h"
#include "rocrand_amd_math.h"
#include "../rocrand.h"
#include "../rocrand_discrete_types.h"

#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ROCRAND_PHILOX4x32_DEFAULT_SEED 0xdeadbeefdeadbeefULL
#define ROCRAND_MRG32K3A_DEFAULT_SEED 12345ULL
#define ROCRAND_MRG31K3P_DEFAULT_SEED 12345ULL
#define ROCRAND_XORWOW_DEFAULT_SEED 0ULL
#define ROCRAND_LFSR113_DEFAULT_SEED_X 2
#define ROCRAND_LFSR113_DEFAULT_SEED_Y 8
#define ROCRAND_LFSR113_DEFAULT_SEED_Z 16
#define ROCRAND_LFSR113_DEFAULT_SEED_W 128

// MTGP32 (rocrand_mtgp32.h:70-90): exponent 11213, a 351-word recurrence kept in a
// 1024-word ring, 256 threads per block step, up to 512 parameter sets.
#define MTGP_MEXP 11213
#define MTGP_N 351
#define MTGP_FLOOR_2P 256
#define MTGP_CEIL_2P 512
#define MTGP_TN MTGP_FLOOR_2P
#define MTGP_LS (MTGP_TN * 3)
#define MTGP_BN_MAX 512
#define MTGP_TS 16
#define MTGP_STATE 1024
#define MTGP_MASK 1023

// One parameter set as the MTGP dynamic creator writes it (rocrand_mtgp32.h:92-107).
struct mtgp32_params_fast_t
{
    int           mexp;
    int           pos;
    int           sh1;
    int           sh2;
    uint32_t      tbl[16];
    uint32_t      tmp_tbl[16];
    uint32_t      flt_tmp_tbl[16];
    uint32_t      mask;
    unsigned char poly_sha1[21];
};


namespace rocrand_device
{
namespace amd_detail
{

namespace eng = ::rocrand_amd_engine;

// Jump tables: one initializer each, a host copy and a device copy.
static const uint32_t h_mrg32k3a_step[]               = ROCRAND_AMD_JT_MRG32K3A_STEP;
static const uint32_t h_mrg32k3a_sub[]                = ROCRAND_AMD_JT_MRG32K3A_SUB;
static const uint32_t h_mrg32k3a_seq[]                = ROCRAND_AMD_JT_MRG32K3A_SEQ;
static const uint32_t h_mrg31k3p_step[]               = ROCRAND_AMD_JT_MRG31K3P_STEP;
static const uint32_t h_mrg31k3p_sub[]                = ROCRAND_AMD_JT_MRG31K3P_SUB;
static const uint32_t h_mrg31k3p_seq[]                = ROCRAND_AMD_JT_MRG31K3P_SEQ;
static const uint32_t h_xorwow_step[]                 = ROCRAND_AMD_JT_XORWOW_STEP;
static const uint32_t h_xorwow_sub[]                  = ROCRAND_AMD_JT_XORWOW_SUB;
static const uint32_t h_lfsr113_step[]                = ROCRAND_AMD_JT_LFSR113_STEP;
static const uint32_t h_lfsr113_sub[]                 = ROCRAND_AMD_JT_LFSR113_SUB;
static __constant__ const uint32_t d_mrg32k3a_step[] = ROCRAND_AMD_JT_MRG32K3A_STEP;
static __constant__ const uint32_t d_mrg32k3a_sub[]  = ROCRAND_AMD_JT_MRG32K3A_SUB;
static __constant__ const uint32_t d_mrg32k3a_seq[]  = ROCRAND_AMD_JT_MRG32K3A_SEQ;
static __constant__ const uint32_t d_mrg31k3p_step[] = ROCRAND_AMD_JT_MRG31K3P_STEP;
static __constant__ const uint32_t d_mrg31k3p_sub[]  = ROCRAND_AMD_JT_MRG31K3P_SUB;
static __constant__ const uint32_t d_mrg31k3p_seq[]  = ROCRAND_AMD_JT_MRG31K3P_SEQ;
static __constant__ const uint32_t d_xorwow_step[]   = ROCRAND_AMD_JT_XORWOW_STEP;
static __constant__ const uint32_t d_xorwow_sub[]    = ROCRAND_AMD_JT_XORWOW_SUB;
static __constant__ const uint32_t d_lfsr113_step[]  = ROCRAND_AMD_JT_LFSR113_STEP;
static __constant__ const uint32_t d_lfsr113_sub[]   = ROCRAND_AMD_JT_LFSR113_SUB;

#if defined(__HIP_DEVICE_COMPILE__)
    #define ROCRAND_AMD_TABLE(name) ::rocrand_device::amd_detail::d_##name
#else
    #define ROCRAND_AMD_TABLE(name) ::rocrand_device::amd_detail::h_##name
#endif

// A^v from a table of A^(2^(first + i)), i < 64, for both MRG components.
template<int MRG>
__host__ __device__ __forceinline__ void mrg_jump(const uint32_t* tab, unsigned long long v, uint32_t* g1,
                                                  uint32_t* g2)
{
    ROCRAND_AMD_FP_EXACT
    for(int i = 0; v > 0; v >>= 1, i++)
        if(v & 1)
        {
            if constexpr(MRG == 32)
            {
                eng::mat3_vec_fold<32, 209>(tab + (i * 2 + 0) * 9, g1);
                eng::mat3_vec_fold<32, 22853>(tab + (i * 2 + 1) * 9, g2);
            }
            else
            {
                eng::mat3_vec_fold<31, 1>(tab + (i * 2 + 0) * 9, g1);
                eng::mat3_vec_fold<31, 21069>(tab + (i * 2 + 1) * 9, g2);
            }
        }
}

// M^v over GF(2) from a table of M^(c 4^k): base-4 digits, matrix k applied digit times.
template<int N>
__host__ __device__ __forceinline__ void gf2_jump(const uint32_t* mats, unsigned long long v, uint32_t* x)
{
    for(int k = 0; v > 0; v >>= 2, k++)
    {
        const unsigned int d = (unsigned int)(v & 3u);
        for(unsigned int i = 0; i < d; i++)
            eng::gf2_mat_vec<N>(mats + (size_t)k * 32 * N * N, x);
    }
}

// 4 x 32-bit and 4 x 64-bit counter vectors as arrays.
template<class V>
struct lanes;
template<>
struct lanes<uint4>
{
    using word = unsigned int;
    static constexpr int n = 4;
    __host__ __device__ static word get(const uint4& v, unsigned int i)
    {
        return i == 0 ? v.x : i == 1 ? v.y : i == 2 ? v.z : v.w;
    }
};
template<>
struct lanes<ulonglong4>
{
    using word = unsigned long long;
    static constexpr int n = 4;
    __host__ __device__ static word get(const ulonglong4& v, unsigned int i)
    {
        return i == 0 ? v.x : i == 1 ? v.y : i == 2 ? v.z : v.w;
    }
};
template<>
struct lanes<uint2>
{
    using word = unsigned int;
    static constexpr int n = 2;
    __host__ __device__ static word get(const uint2& v, unsigned int i) { return i == 0 ? v.x : v.y; }
};
template<>
struct lanes<ulonglong2>
{
    using word = unsigned long long;
    static constexpr int n = 2;
    __host__ __device__ static word get(const ulonglong2& v, unsigned int i) { return i == 0 ? v.x : v.y; }
};

// ---- counter arithmetic of the 4-lane engines (rocrand_philox4x32_10.h:206-260,
//      rocrand_threefry4_impl.h:203-262).  `Word` 32 bit: 64-bit offsets split into two
//      words; 64 bit: the offset is one word.
template<class V>
__host__ __device__ __forceinline__ void add_counter4(V& c, unsigned long long offset)
{
    using W         = typename lanes<V>::word;
    const W   lo    = (W)offset;
    const W   hi    = sizeof(W) == 4 ? (W)(offset >> 32) : (W)0;
    const V   old   = c;
    c.x += lo;
    c.y += hi + (c.x < old.x ? 1 : 0);
    c.z += (c.y < old.y ? 1 : 0);
    c.w += (c.z < old.z ? 1 : 0);
}
template<class V>
__host__ __device__ __forceinline__ void add_subsequence4(V& c, unsigned long long subsequence)
{
    using W       = typename lanes<V>::word;
    const W lo    = (W)subsequence;
    const W hi    = sizeof(W) == 4 ? (W)(subsequence >> 32) : (W)0;
    const W old_z = c.z;
    c.z += lo;
    c.w += hi + (c.z < old_z ? 1 : 0);
}
template<class V>
__host__ __device__ __forceinline__ void bump4(V& c)
{
    c.x++;
    unsigned int add = c.x == 0 ? 1 : 0;
    c.y += add;
    add = c.y == 0 ? add : 0;
    c.z += add;
    add = c.z == 0 ? add : 0;
    c.w += add;
}
template<class V>
__host__ __device__ __forceinline__ V interleave4(const V& prev, const V& next, unsigned int substate)
{
    switch(substate)
    {
        case 1: return V{prev.y, prev.z, prev.w, next.x};
        case 2: return V{prev.z, prev.w, next.x, next.y};
        case 3: return V{prev.w, next.x, next.y, next.z};
        default: return prev;
    }
}

} // namespace amd_detail

// ---------------------------------------------------------------- 4-lane counter engines
// Philox keeps a Box-Muller cache in its state (rocrand_philox4x32_10.h:91-104); the
// ThreeFry states hold counter, key, result and lane only (rocrand_threefry4_impl.h).
struct philox_state_words
{
    uint4        counter;
    uint4        result;
    uint2        key;
    unsigned int substate;
    unsigned int boxmuller_float_state;
    unsigned int boxmuller_double_state;
    float        boxmuller_float;
    double       boxmuller_double;
};
template<class V>
struct threefry_state_words
{
    V            counter;
    V            key;
    V            result;
    unsigned int substate;
};

// Shared mechanics of Philox4x32-10 and ThreeFry4x32/4x64-20: a word position is a
// counter block plus a lane (substate); Derived::rounds maps (counter, key) to the block.
template<class Derived, class State, class V, class K>
class counter4_engine
{
public:
    using state_type = State;
    using word       = typename amd_detail::lanes<V>::word;

    __host__ __device__ void discard(unsigned long long offset)
    {
        discard_impl(offset);
        m_state.result = Derived::rounds(m_state.counter, m_state.key);
    }
    __host__ __device__ void discard_subsequence(unsigned long long subsequence)
    {
        amd_detail::add_subsequence4(m_state.counter, subsequence);
        m_state.result = Derived::rounds(m_state.counter, m_state.key);
    }
    __host__ __device__ word operator()() { return next(); }
    __host__ __device__ word next()
    {
        const word ret = amd_detail::lanes<V>::get(m_state.result, m_state.substate);
        m_state.substate++;
        if(m_state.substate == 4)
        {
            m_state.substate = 0;
            amd_detail::bump4(m_state.counter);
            m_state.result = Derived::rounds(m_state.counter, m_state.key);
        }
        return ret;
    }
    __host__ __device__ V next4()
    {
        const V ret = m_state.result;
        amd_detail::bump4(m_state.counter);
        m_state.result = Derived::rounds(m_state.counter, m_state.key);
        return amd_detail::interleave4(ret, m_state.result, m_state.substate);
    }

    state_type m_state;

protected:
    // counter = (0, 0, subsequence), then `offset` words on; result = the current block
    __host__ __device__ void start(const K& key, unsigned long long subsequence, unsigned long long offset)
    {
        m_state.counter  = V{0, 0, 0, 0};
        m_state.result   = V{0, 0, 0, 0};
        m_state.key      = key;
        m_state.substate = 0;
        amd_detail::add_subsequence4(m_state.counter, subsequence);
        discard(offset);
    }
    __host__ __device__ void discard_impl(unsigned long long offset)
    {
        m_state.substate += offset & 3;
        unsigned long long blocks = offset / 4;
        blocks += m_state.substate < 4 ? 0 : 1;
        m_state.substate -= m_state.substate < 4 ? 0 : 4;
        amd_detail::add_counter4(m_state.counter, blocks);
    }
};

class philox4x32_10_engine : public counter4_engine<philox4x32_10_engine, philox_state_words, uint4, uint2>
{
public:
    __host__ __device__ philox4x32_10_engine() { seed(ROCRAND_PHILOX4x32_DEFAULT_SEED, 0, 0); }
    __host__ __device__ philox4x32_10_engine(unsigned long long seed_value, unsigned long long subsequence,
                                             unsigned long long offset)
    {
        seed(seed_value, subsequence, offset);
    }
    // rocrand_philox4x32_10.h:134-186: key = seed; subsequence = high 64 counter bits.
    __host__ __device__ void seed(unsigned long long seed_value, unsigned long long subsequence,
                                  unsigned long long offset)
    {
        m_state.boxmuller_float_state  = 0;
        m_state.boxmuller_double_state = 0;
        start(uint2{(unsigned int)seed_value, (unsigned int)(seed_value >> 32)}, subsequence, offset);
    }
    __host__ __device__ static uint4 rounds(const uint4& c, const uint2& k)
    {
        uint32_t o[4];
        ::rocrand_amd_engine::philox_block(c.x, c.y, c.z, c.w, k.x, k.y, o);
        return uint4{o[0], o[1], o[2], o[3]};
    }
};

// ThreeFry4: rocrand_threefry4x32_20.h / 4x64_20.h seed(): key words are the seed and
// its right shifts by 16, 32 and 48 bits (truncated to 32-bit words for 4x32).
template<class V>
class threefry4_engine : public counter4_engine<threefry4_engine<V>, threefry_state_words<V>, V, V>
{
    using base = counter4_engine<threefry4_engine<V>, threefry_state_words<V>, V, V>;
    using word = typename base::word;

public:
    __host__ __device__ threefry4_engine(unsigned long long seed_value = 0, unsigned long long subsequence = 0,
                                         unsigned long long offset = 0)
    {
        seed(seed_value, subsequence, offset);
    }
    __host__ __device__ void seed(unsigned long long seed_value = 0, unsigned long long subsequence = 0,
                                  unsigned long long offset = 0)
    {
        base::start(V{(word)seed_value, (word)(seed_value >> 16), (word)(seed_value >> 32),
                      (word)(seed_value >> 48)},
                    subsequence, offset);
    }
    __host__ __device__ static V rounds(const V& c, const V& k)
    {
        const word cc[4] = {c.x, c.y, c.z, c.w};
        const word kk[4] = {k.x, k.y, k.z, k.w};
        word       o[4];
        ::rocrand_amd_engine::threefry4_block<word>(cc, kk, o);
        return V{o[0], o[1], o[2], o[3]};
    }
};
typedef threefry4_engine<uint4>      threefry4x32_20_engine;
typedef threefry4_engine<ulonglong4> threefry4x64_20_engine;

// ---------------------------------------------------------------- 2-lane ThreeFry
// rocrand_threefry2_impl.h: a subsequence adds to the high counter word.
template<class V>
class threefry2_engine
{
public:
    using state_type = threefry_state_words<V>;
    using word       = typename amd_detail::lanes<V>::word;

    __host__ __device__ threefry2_engine(unsigned long long seed_value = 0, unsigned long long subsequence = 0,
                                         unsigned long long offset = 0)
    {
        seed(seed_value, subsequence, offset);
    }
    // rocrand_threefry2x32_20.h / 2x64_20.h seed(): key = {seed, seed >> 32}.
    __host__ __device__ void seed(unsigned long long seed_value = 0, unsigned long long subsequence = 0,
                                  unsigned long long offset = 0)
    {
        m_state.counter  = V{0, 0};
        m_state.result   = V{0, 0};
        m_state.substate = 0;
        m_state.key      = V{(word)seed_value, (word)(seed_value >> 32)};
        m_state.counter.y += (word)subsequence;
        discard(offset);
    }
    __host__ __device__ void discard(unsigned long long offset)
    {
        m_state.substate += offset & 1;
        unsigned long long blocks = offset / 2;
        blocks += m_state.substate < 2 ? 0 : 1;
        m_state.substate -= m_state.substate < 2 ? 0 : 2;
        const word lo  = (word)blocks;
        const word hi  = sizeof(word) == 4 ? (word)(blocks >> 32) : (word)0;
        const word old = m_state.counter.x;
        m_state.counter.x += lo;
        m_state.counter.y += hi + (m_state.counter.x < old ? 1 : 0);
        m_state.result = rounds(m_state.counter, m_state.key);
    }
    __host__ __device__ void discard_subsequence(unsigned long long subsequence)
    {
        m_state.counter.y += (word)subsequence;
        m_state.result = rounds(m_state.counter, m_state.key);
    }
    __host__ __device__ word operator()() { return next(); }
    __host__ __device__ word next()
    {
        const word ret = amd_detail::lanes<V>::get(m_state.result, m_state.substate);
        m_state.substate++;
        if(m_state.substate == 2)
        {
            m_state.substate = 0;
            bump();
            m_state.result = rounds(m_state.counter, m_state.key);
        }
        return ret;
    }
    __host__ __device__ V next2()
    {
        const V ret = m_state.result;
        bump();
        m_state.result = rounds(m_state.counter, m_state.key);
        return m_state.substate == 0 ? ret : V{ret.y, m_state.result.x};
    }
    __host__ __device__ static V rounds(const V& c, const V& k)
    {
        const word cc[2] = {c.x, c.y};
        const word kk[2] = {k.x, k.y};
        word       o[2];
        ::rocrand_amd_engine::threefry2_block<word>(cc, kk, o);
        return V{o[0], o[1]};
    }

    state_type m_state;

private:
    __host__ __device__ void bump()
    {
        m_state.counter.x++;
        m_state.counter.y += m_state.counter.x == 0 ? 1 : 0;
    }
};
typedef threefry2_engine<uint2>      threefry2x32_20_engine;
typedef threefry2_engine<ulonglong2> threefry2x64_20_engine;

// ---------------------------------------------------------------- MRG32k3a / MRG31k3p
struct mrg_state_words
{
    unsigned int g1[3];
    unsigned int g2[3];
    unsigned int boxmuller_float_state;
    unsigned int boxmuller_double_state;
    float        boxmuller_float;
    double       boxmuller_double;
};

// rocrand_mrg32k3a.h / rocrand_mrg31k3p.h: seed 0 means the default seed; a subsequence
// is 2^76 (2^72) steps, a sequence 2^127 (2^134); values lie in [1, m1].
template<int MRG>
class mrg_engine
{
public:
    using state_type = mrg_state_words;

    __host__ __device__ mrg_engine() { seed(12345ull, 0, 0); }
    __host__ __device__ mrg_engine(unsigned long long seed_value, unsigned long long subsequence,
                                   unsigned long long offset)
    {
        seed(seed_value, subsequence, offset);
    }
    __host__ __device__ void seed(unsigned long long seed_value, unsigned long long subsequence,
                                  unsigned long long offset)
    {
        if(seed_value == 0)
            seed_value = 12345ull;
        uint32_t w[6];
        if constexpr(MRG == 32)
            ::rocrand_amd_engine::mrg32k3a_seed_words(seed_value, w);
        else
            ::rocrand_amd_engine::mrg31k3p_seed_words(seed_value, w);
        for(int j = 0; j < 3; j++)
        {
            m_state.g1[j] = w[j];
            m_state.g2[j] = w[3 + j];
        }
        m_state.boxmuller_float_state  = 0;
        m_state.boxmuller_double_state = 0;
        discard_subsequence(subsequence);
        discard(offset);
    }
    __host__ __device__ void discard(unsigned long long offset)
    {
        if constexpr(MRG == 32)
            amd_detail::mrg_jump<32>(ROCRAND_AMD_TABLE(mrg32k3a_step), offset, m_state.g1, m_state.g2);
        else
            amd_detail::mrg_jump<31>(ROCRAND_AMD_TABLE(mrg31k3p_step), offset, m_state.g1, m_state.g2);
    }
    __host__ __device__ void discard_subsequence(unsigned long long subsequence)
    {
        if constexpr(MRG == 32)
            amd_detail::mrg_jump<32>(ROCRAND_AMD_TABLE(mrg32k3a_sub), subsequence, m_state.g1, m_state.g2);
        else
            amd_detail::mrg_jump<31>(ROCRAND_AMD_TABLE(mrg31k3p_sub), subsequence, m_state.g1, m_state.g2);
    }
    __host__ __device__ void discard_sequence(unsigned long long sequence)
    {
        if constexpr(MRG == 32)
            amd_detail::mrg_jump<32>(ROCRAND_AMD_TABLE(mrg32k3a_seq), sequence, m_state.g1, m_state.g2);
        else
            amd_detail::mrg_jump<31>(ROCRAND_AMD_TABLE(mrg31k3p_seq), sequence, m_state.g1, m_state.g2);
    }
    __host__ __device__ unsigned int operator()() { return next(); }
    __host__ __device__ unsigned int next()
    {
        ::rocrand_amd_engine::mrg_state s;
        for(int j = 0; j < 3; j++)
        {
            s.g1[j] = m_state.g1[j];
            s.g2[j] = m_state.g2[j];
        }
        const unsigned int v = MRG == 32 ? ::rocrand_amd_engine::mrg32k3a_next(s)
                                         : ::rocrand_amd_engine::mrg31k3p_next(s);
        for(int j = 0; j < 3; j++)
        {
            m_state.g1[j] = s.g1[j];
            m_state.g2[j] = s.g2[j];
        }
        return v;
    }

    state_type m_state;
};
typedef mrg_engine<32> mrg32k3a_engine;
typedef mrg_engine<31> mrg31k3p_engine;

// ---------------------------------------------------------------- XORWOW
struct xorwow_state_words
{
    unsigned int d;
    unsigned int boxmuller_float_state;
    unsigned int boxmuller_double_state;
    float        boxmuller_float;
    double       boxmuller_double;
    unsigned int x[5];
};

// rocrand_xorwow.h:85-160: a subsequence is 2^67 steps (the Weyl counter d is unchanged
// by it: 2^67 * 362437 = 0 mod 2^32).
class xorwow_engine
{
public:
    using state_type = xorwow_state_words;

    __host__ __device__ xorwow_engine() : xorwow_engine(ROCRAND_XORWOW_DEFAULT_SEED, 0, 0) {}
    __host__ __device__ xorwow_engine(unsigned long long seed_value, unsigned long long subsequence,
                                      unsigned long long offset)
    {
        ::rocrand_amd_engine::xorwow_state s;
        ::rocrand_amd_engine::xorwow_seed(s, seed_value);
        m_state.d = s.d;
        for(int j = 0; j < 5; j++)
            m_state.x[j] = s.x[j];
        discard_subsequence(subsequence);
        discard(offset);
        m_state.boxmuller_float_state  = 0;
        m_state.boxmuller_double_state = 0;
    }
    __host__ __device__ void discard(unsigned long long offset)
    {
        amd_detail::gf2_jump<5>(ROCRAND_AMD_TABLE(xorwow_step), offset, m_state.x);
        m_state.d += (unsigned int)offset * 362437u;
    }
    __host__ __device__ void discard_subsequence(unsigned long long subsequence)
    {
        amd_detail::gf2_jump<5>(ROCRAND_AMD_TABLE(xorwow_sub), subsequence, m_state.x);
    }
    __host__ __device__ unsigned int operator()() { return next(); }
    __host__ __device__ unsigned int next()
    {
        ::rocrand_amd_engine::xorwow_state s;
        s.d = m_state.d;
        for(int j = 0; j < 5; j++)
            s.x[j] = m_state.x[j];
        const unsigned int v = ::rocrand_amd_engine::xorwow_next(s);
        m_state.d            = s.d;
        for(int j = 0; j < 5; j++)
            m_state.x[j] = s.x[j];
        return v;
    }

    state_type m_state;
};

// ---------------------------------------------------------------- LFSR113
struct lfsr113_state_words
{
    uint4 z;
    uint4 subsequence; // the seed the subsequence jumps start from
};

// rocrand_lfsr113.h:78-175: the seed words are used as given; a subsequence is 2^55 steps
// (an unsigned int count).
class lfsr113_engine
{
public:
    using state_type = lfsr113_state_words;

    __host__ __device__ lfsr113_engine(uint4 seed_value = uint4{ROCRAND_LFSR113_DEFAULT_SEED_X,
                                                                ROCRAND_LFSR113_DEFAULT_SEED_Y,
                                                                ROCRAND_LFSR113_DEFAULT_SEED_Z,
                                                                ROCRAND_LFSR113_DEFAULT_SEED_W},
                                       unsigned int subsequence = 0, unsigned long long offset = 0)
    {
        seed(seed_value, subsequence, offset);
    }
    __host__ __device__ void seed(uint4 seed_value, unsigned long long subsequence, unsigned long long offset = 0)
    {
        m_state.subsequence = seed_value;
        m_state.z           = seed_value;
        discard_subsequence((unsigned int)subsequence);
        discard(offset);
    }
    __host__ __device__ void discard() { next(); }
    __host__ __device__ void discard(unsigned long long offset)
    {
        uint32_t z[4] = {m_state.z.x, m_state.z.y, m_state.z.z, m_state.z.w};
        amd_detail::gf2_jump<4>(ROCRAND_AMD_TABLE(lfsr113_step), offset, z);
        m_state.z = uint4{z[0], z[1], z[2], z[3]};
    }
    __host__ __device__ void discard_subsequence(unsigned int subsequence)
    {
        uint32_t z[4] = {m_state.z.x, m_state.z.y, m_state.z.z, m_state.z.w};
        amd_detail::gf2_jump<4>(ROCRAND_AMD_TABLE(lfsr113_sub), subsequence, z);
        m_state.z = uint4{z[0], z[1], z[2], z[3]};
    }
    __host__ __device__ unsigned int operator()() { return next(); }
    __host__ __device__ unsigned int next()
    {
        ::rocrand_amd_engine::lfsr113_state s;
        s.z[0]               = m_state.z.x;
        s.z[1]               = m_state.z.y;
        s.z[2]               = m_state.z.z;
        s.z[3]               = m_state.z.w;
        const unsigned int v = ::rocrand_amd_engine::lfsr113_next(s);
        m_state.z            = uint4{s.z[0], s.z[1], s.z[2], s.z[3]};
        return v;
    }

    state_type m_state;
};

// ---------------------------------------------------------------- Sobol (quasi-random)
// rocrand_sobol32.h / rocrand_sobol64.h: the state keeps a copy of the dimension's
// direction vectors, the point index i and the current value d (the Gray-code XOR of the
// vectors selected by i ^ (i >> 1)); each step XORs in the vector of i's lowest zero bit.
template<class W, int BITS>
struct sobol_state_words
{
    W d;
    W i;
    W vectors[BITS];
};

template<class W, int BITS>
class sobol_engine
{
public:
    using state_type = sobol_state_words<W, BITS>;

    __host__ __device__ sobol_engine() : m_state() {}
    __host__ __device__ sobol_engine(const W* vectors, W offset)
    {
        m_state.d = 0;
        m_state.i = 0;
        for(int k = 0; k < BITS; k++)
            m_state.vectors[k] = vectors[k];
        discard(offset);
    }
    // point i + offset: recomputed from its Gray code
    __host__ __device__ void discard(W offset)
    {
        m_state.i += offset;
        const W g = m_state.i ^ (m_state.i >> 1);
        W       d = 0;
        for(int k = 0; k < BITS; k++)
            d ^= ((g >> k) & 1) ? m_state.vectors[k] : W(0);
        m_state.d = d;
    }
    __host__ __device__ void discard()
    {
        m_state.d ^= m_state.vectors[lowest_zero_bit(m_state.i)];
        m_state.i++;
    }
    // i += stride for a power of two stride, i a multiple of it (the lanes of a
    // strided loop): the Gray codes of i and i + stride differ in bits log2(stride) - 1
    // and the lowest zero bit of i | (stride - 1)
    __host__ __device__ void discard_stride(W stride)
    {
        if(stride == 1)
        {
            discard();
            return;
        }
        m_state.d ^= m_state.vectors[lowest_zero_bit(~stride) - 1];
        m_state.d ^= m_state.vectors[lowest_zero_bit(m_state.i | (stride - 1))];
        m_state.i += stride;
    }
    __host__ __device__ W operator()() { return next(); }
    __host__ __device__ W next()
    {
        const W p = m_state.d;
        discard();
        return p;
    }
    __host__ __device__ W current() const { return m_state.d; }

    state_type m_state;

private:
    // index of the lowest zero bit; 0 for an all-ones value (as the reference's device
    // branch, whose ffs(~x) is 0 there)
    __host__ __device__ static unsigned int lowest_zero_bit(W x)
    {
        const W y = ~x;
        if(y == 0)
            return 0;
        if constexpr(BITS == 32)
            return (unsigned int)__builtin_ctz((unsigned int)y);
        else
            return (unsigned int)__builtin_ctzll((unsigned long long)y);
    }
};
typedef sobol_engine<unsigned int, 32>       sobol32_engine;
typedef sobol_engine<unsigned long long, 64> sobol64_engine;

// rocrand_scrambled_sobol32.h / 64.h: the Sobol value XOR a per-dimension constant.
template<class W, int BITS>
class scrambled_sobol_engine
{
public:
    __host__ __device__ scrambled_sobol_engine() : m_engine(), scramble_constant() {}
    __host__ __device__ scrambled_sobol_engine(const W* vectors, W scramble, W offset)
        : m_engine(vectors, 0), scramble_constant(scramble)
    {
        m_engine.discard(offset);
    }
    __host__ __device__ void discard(W offset) { m_engine.discard(offset); }
    __host__ __device__ void discard() { m_engine.discard(); }
    __host__ __device__ void discard_stride(W stride) { m_engine.discard_stride(stride); }
    __host__ __device__ W    operator()() { return next(); }
    __host__ __device__ W    next() { return m_engine.next() ^ scramble_constant; }
    __host__ __device__ W    current() const { return m_engine.current() ^ scramble_constant; }

private:
    sobol_engine<W, BITS> m_engine;
    W                     scramble_constant;
};
typedef scrambled_sobol_engine<unsigned int, 32>       scrambled_sobol32_engine;
typedef scrambled_sobol_engine<unsigned long long, 64> scrambled_sobol64_engine;

// ---------------------------------------------------------------- MTGP32
// Parameter tables of up to MTGP_BN_MAX states in device memory (rocrand_make_constant).
struct mtgp32_params
{
    unsigned int pos_tbl[MTGP_BN_MAX];
    unsigned int param_tbl[MTGP_BN_MAX][MTGP_TS];
    unsigned int temper_tbl[MTGP_BN_MAX][MTGP_TS];
    unsigned int single_temper_tbl[MTGP_BN_MAX][MTGP_TS];
    unsigned int sh1_tbl[MTGP_BN_MAX];
    unsigned int sh2_tbl[MTGP_BN_MAX];
    unsigned int mask[1];
};
typedef mtgp32_params_fast_t mtgp32_fast_params;

// A ring of the recurrence: status[(offset + i) & MTGP_MASK] is word i of the current
// window; id is the parameter set (block) it was made for.
struct mtgp32_state
{
    int          offset;
    int          id;
    unsigned int status[MTGP_STATE];
};

// MTGP's seeding (rocrand_mtgp32.h:130-145): the first mexp/32 + 1 words from the seed,
// the parameter set's hidden seed and Knuth's 1812433253 recurrence over a byte fill.
inline void rocrand_mtgp32_init_state(unsigned int array[], const mtgp32_fast_params* para, unsigned int seed)
{
    const int          size   = para->mexp / 32 + 1;
    const unsigned int hidden = para->tbl[4] ^ (para->tbl[8] << 16);
    unsigned int       fill   = hidden;
    fill += fill >> 16;
    fill += fill >> 8;
    memset(array, (int)(fill & 0xff), sizeof(unsigned int) * size);
    array[0] = seed;
    array[1] = hidden;
    for(int i = 1; i < size; i++)
        array[i] ^= 1812433253u * (array[i - 1] ^ (array[i - 1] >> 30)) + (unsigned int)i;
}

// One MTGP32 generator shared by the threads of a block: each call of next() by every
// thread of the block advances the ring by blockDim.x words, thread t taking word t of
// the step (rocrand_mtgp32.h:147-357).  The state must be visible to the whole block
// (LDS, or one global copy per block) and every thread must make the same calls.
class mtgp32_engine
{
public:
    __host__ __device__ mtgp32_engine() {}
    __host__ __device__ mtgp32_engine(const mtgp32_state& state, const mtgp32_params* params, int bid)
    {
        m_state = state;
        load_params(params, bid);
    }
    // block-cooperative copy on the device (ends with a barrier), plain copy on the host
    __host__ __device__ void copy(const mtgp32_engine* src)
    {
#if defined(__HIP_DEVICE_COMPILE__)
        const unsigned int t = threadIdx.x;
        for(unsigned int i = t; i < MTGP_STATE; i += blockDim.x)
            m_state.status[i] = src->m_state.status[i];
        if(t == 0)
        {
            m_state.offset = src->m_state.offset;
            m_state.id     = src->m_state.id;
            pos_tbl        = src->pos_tbl;
            sh1_tbl        = src->sh1_tbl;
            sh2_tbl        = src->sh2_tbl;
            mask           = src->mask;
        }
        if(t < MTGP_TS)
        {
            param_tbl[t]         = src->param_tbl[t];
            temper_tbl[t]        = src->temper_tbl[t];
            single_temper_tbl[t] = src->single_temper_tbl[t];
        }
        __syncthreads();
#else
        *this = *src;
#endif
    }
    __host__ __device__ void set_params(mtgp32_params* params) { load_params(params, m_state.id); }
    __host__ __device__ unsigned int operator()() { return next(); }
    // 32-bit output; the host has no block and returns 0 (as the reference)
    __host__ __device__ unsigned int next()
    {
#if defined(__HIP_DEVICE_COMPILE__)
        const unsigned int t = threadIdx.x;
        const unsigned int r = recur(t);
        const unsigned int o = r ^ temper_tbl[fold(m_state.status[(t + m_state.offset + pos_tbl - 1) & MTGP_MASK])];
        advance();
        return o;
#else
        return 0;
#endif
    }
    // the float-tempered output: bits of a float in [1, 2) shifted right by 9
    __host__ __device__ unsigned int next_single()
    {
#if defined(__HIP_DEVICE_COMPILE__)
        const unsigned int t = threadIdx.x;
        const unsigned int r = recur(t);
        const unsigned int o
            = (r >> 9) ^ single_temper_tbl[fold(m_state.status[(t + m_state.offset + pos_tbl - 1) & MTGP_MASK])];
        advance();
        return o;
#else
        return 0;
#endif
    }

    mtgp32_state m_state;
    unsigned int pos_tbl;
    unsigned int param_tbl[MTGP_TS];
    unsigned int temper_tbl[MTGP_TS];
    unsigned int sh1_tbl;
    unsigned int sh2_tbl;
    unsigned int single_temper_tbl[MTGP_TS];
    unsigned int mask;

private:
    __host__ __device__ void load_params(const mtgp32_params* params, int bid)
    {
        pos_tbl = params->pos_tbl[bid];
        sh1_tbl = params->sh1_tbl[bid];
        sh2_tbl = params->sh2_tbl[bid];
        mask    = params->mask[0];
        for(int j = 0; j < MTGP_TS; j++)
        {
            param_tbl[j]         = params->param_tbl[bid][j];
            temper_tbl[j]        = params->temper_tbl[bid][j];
            single_temper_tbl[j] = params->single_temper_tbl[bid][j];
        }
    }
    // word MTGP_N of thread t's window, stored back into the ring.  A step reads ring
    // words offset .. offset + pos + blockDim.x - 1 and writes offset + MTGP_N .., so with
    // pos + blockDim.x <= MTGP_N (every parameter set here, blocks of up to 256) the
    // threads of a step never read a word another thread writes.
    __host__ __device__ unsigned int recur(unsigned int t)
    {
        const unsigned int o = m_state.offset;
        unsigned int       x = (m_state.status[(t + o) & MTGP_MASK] & mask) ^ m_state.status[(t + o + 1) & MTGP_MASK];
        x ^= x << sh1_tbl;
        const unsigned int y = x ^ (m_state.status[(t + o + pos_tbl) & MTGP_MASK] >> sh2_tbl);
        const unsigned int r = y ^ param_tbl[y & 0x0f];
        m_state.status[(t + o + MTGP_N) & MTGP_MASK] = r;
        return r;
    }
    __host__ __device__ static unsigned int fold(unsigned int v)
    {
        v ^= v >> 16;
        v ^= v >> 8;
        return v & 0x0f;
    }
#if defined(__HIP_DEVICE_COMPILE__)
    __device__ void advance()
    {
        __syncthreads();
        if(threadIdx.x == 0)
            m_state.offset = (m_state.offset + (int)blockDim.x) & MTGP_MASK;
        __syncthreads();
    }
#endif
};

} // namespace rocrand_device

typedef rocrand_device::sobol32_engine           rocrand_state_sobol32;
typedef rocrand_device::sobol64_engine           rocrand_state_sobol64;
typedef rocrand_device::scrambled_sobol32_engine rocrand_state_scrambled_sobol32;
typedef rocrand_device::scrambled_sobol64_engine rocrand_state_scrambled_sobol64;
typedef rocrand_device::mtgp32_engine            rocrand_state_mtgp32;
typedef rocrand_device::mtgp32_state             mtgp32_state;
typedef rocrand_device::mtgp32_fast_params       mtgp32_fast_params;
typedef rocrand_device::mtgp32_params            mtgp32_params;

// n states for blocks 0 .. n-1 from parameter sets 0 .. n-1, copied to `state` (device
// memory): seed ^ (seed >> 32) plus block + 1 seeds each (rocrand_mtgp32.h:362-398);
// every state masks with the first set's mask.
__host__ inline rocrand_status rocrand_make_state_mtgp32(rocrand_state_mtgp32*    state,
                                                         const mtgp32_fast_params params[], int n,
                                                         unsigned long long seed)
{
    if(n < 0 || n > MTGP_BN_MAX)
        return ROCRAND_STATUS_OUT_OF_RANGE;
    rocrand_state_mtgp32* h = (rocrand_state_mtgp32*)malloc(sizeof(rocrand_state_mtgp32) * (n > 0 ? n : 1));
    if(h == NULL)
        return ROCRAND_STATUS_ALLOCATION_FAILED;
    seed ^= seed >> 32;
    for(int i = 0; i < n; i++)
    {
        rocrand_device::rocrand_mtgp32_init_state(h[i].m_state.status, &params[i], (unsigned int)seed + i + 1);
        h[i].m_state.offset = 0;
        h[i].m_state.id     = i;
        h[i].pos_tbl        = params[i].pos;
        h[i].sh1_tbl        = params[i].sh1;
        h[i].sh2_tbl        = params[i].sh2;
        h[i].mask           = params[0].mask;
        for(int j = 0; j < MTGP_TS; j++)
        {
            h[i].param_tbl[j]         = params[i].tbl[j];
            h[i].temper_tbl[j]        = params[i].tmp_tbl[j];
            h[i].single_temper_tbl[j] = params[i].flt_tmp_tbl[j];
        }
    }
    const hipError_t e = hipMemcpy(state, h, sizeof(rocrand_state_mtgp32) * n, hipMemcpyDefault);
    free(h);
    return e == hipSuccess ? ROCRAND_STATUS_SUCCESS : ROCRAND_STATUS_ALLOCATION_FAILED;
}

// The parameter tables of all MTGP_BN_MAX sets into a device-resident mtgp32_params, for
// rocrand_mtgp32_set_params (rocrand_mtgp32.h:400-470).
__host__ inline rocrand_status rocrand_make_constant(const mtgp32_fast_params params[], mtgp32_params* p)
{
    mtgp32_params* h = (mtgp32_params*)malloc(sizeof(mtgp32_params));
    if(h == NULL)
        return ROCRAND_STATUS_ALLOCATION_FAILED;
    h->mask[0] = params[0].mask;
    for(int i = 0; i < MTGP_BN_MAX; i++)
    {
        h->pos_tbl[i] = params[i].pos;
        h->sh1_tbl[i] = params[i].sh1;
        h->sh2_tbl[i] = params[i].sh2;
        for(int j = 0; j < MTGP_TS; j++)
        {
            h->param_tbl[i][j]         = params[i].tbl[j];
            h->temper_tbl[i][j]        = params[i].tmp_tbl[j];
            h->single_temper_tbl[i][j] = params[i].flt_tmp_tbl[j];
        }
    }
    const hipError_t e = hipMemcpy(p, h, sizeof(mtgp32_params), hipMemcpyHostToDevice);
    free(h);
    return e == hipSuccess ? ROCRAND_STATUS_SUCCESS : ROCRAND_STATUS_ALLOCATION_FAILED;
}
typedef rocrand_device::philox4x32_10_engine   rocrand_state_philox4x32_10;
typedef rocrand_device::mrg32k3a_engine        rocrand_state_mrg32k3a;
typedef rocrand_device::mrg31k3p_engine        rocrand_state_mrg31k3p;
typedef rocrand_device::xorwow_engine          rocrand_state_xorwow;
typedef rocrand_device::lfsr113_engine         rocrand_state_lfsr113;
typedef rocrand_device::threefry2x32_20_engine rocrand_state_threefry2x32_20;
typedef rocrand_device::threefry2x64_20_engine rocrand_state_threefry2x64_20;
typedef rocrand_device::threefry4x32_20_engine rocrand_state_threefry4x32_20;
typedef rocrand_device::threefry4x64_20_engine rocrand_state_threefry4x64_20;

// ---------------------------------------------------------------- init / raw words
#define ROCRAND_AMD_INIT(STATE)                                                                              \
    __host__ __device__ __forceinline__ void rocrand_init(unsigned long long seed,                           \
                                                          unsigned long long subsequence,                    \
                                                          unsigned long long offset, STATE* state)           \
    {                                                                                                        \
        *state = STATE(seed, subsequence, offset);                                                           \
    }
ROCRAND_AMD_INIT(rocrand_state_philox4x32_10)
ROCRAND_AMD_INIT(rocrand_state_mrg32k3a)
ROCRAND_AMD_INIT(rocrand_state_mrg31k3p)
ROCRAND_AMD_INIT(rocrand_state_xorwow)
ROCRAND_AMD_INIT(rocrand_state_threefry2x32_20)
ROCRAND_AMD_INIT(rocrand_state_threefry2x64_20)
ROCRAND_AMD_INIT(rocrand_state_threefry4x32_20)
ROCRAND_AMD_INIT(rocrand_state_threefry4x64_20)
#undef ROCRAND_AMD_INIT

__host__ __device__ __forceinline__ void rocrand_init(const uint4 seed, const unsigned int subsequence,
                                                      rocrand_state_lfsr113* state)
{
    *state = rocrand_state_lfsr113(seed, subsequence);
}
__host__ __device__ __forceinline__ void rocrand_init(const uint4 seed, const unsigned int subsequence,
                                                      const unsigned long long offset, rocrand_state_lfsr113* state)
{
    *state = rocrand_state_lfsr113(seed, subsequence, offset);
}

__host__ __device__ __forceinline__ unsigned int rocrand(rocrand_state_philox4x32_10* state)
{
    return state->next();
}
__host__ __device__ __forceinline__ uint4 rocrand4(rocrand_state_philox4x32_10* state) { return state->next4(); }
// MRG words are remapped from [1, m1] to 32 bits (rocrand_mrg32k3a.h:410-420).
__host__ __device__ __forceinline__ unsigned int rocrand(rocrand_state_mrg32k3a* state)
{
    ROCRAND_AMD_FP_EXACT
    return (unsigned int)((double)(state->next() - 1) * ::rocrand_amd_engine::MRG32_UINT_NORM);
}
__host__ __device__ __forceinline__ unsigned int rocrand(rocrand_state_mrg31k3p* state)
{
    ROCRAND_AMD_FP_EXACT
    return (unsigned int)((double)(state->next() - 1) * ::rocrand_amd_engine::MRG31_UINT_NORM);
}
__host__ __device__ __forceinline__ unsigned int rocrand(rocrand_state_xorwow* state) { return state->next(); }
__host__ __device__ __forceinline__ unsigned int rocrand(rocrand_state_lfsr113* state) { return state->next(); }
__host__ __device__ __forceinline__ unsigned int rocrand(rocrand_state_threefry2x32_20* state)
{
    return state->next();
}
__host__ __device__ __forceinline__ uint2 rocrand2(rocrand_state_threefry2x32_20* state) { return state->next2(); }
__host__ __device__ __forceinline__ unsigned long long rocrand(rocrand_state_threefry2x64_20* state)
{
    return state->next();
}
__host__ __device__ __forceinline__ ulonglong2 rocrand2(rocrand_state_threefry2x64_20* state)
{
    return state->next2();
}
__host__ __device__ __forceinline__ unsigned int rocrand(rocrand_state_threefry4x32_20* state)
{
    return state->next();
}
__host__ __device__ __forceinline__ uint4 rocrand4(rocrand_state_threefry4x32_20* state) { return state->next4(); }
__host__ __device__ __forceinline__ unsigned long long rocrand(rocrand_state_threefry4x64_20* state)
{
    return state->next();
}
__host__ __device__ __forceinline__ ulonglong4 rocrand4(rocrand_state_threefry4x64_20* state)
{
    return state->next4();
}

// ---------------------------------------------------------------- skipahead
#define ROCRAND_AMD_SKIP(STATE, SUBSEQ_T, SEQ_CALL)                                                          \
    __host__ __device__ __forceinline__ void skipahead(unsigned long long offset, STATE* state)              \
    {                                                                                                        \
        state->discard(offset);                                                                              \
    }                                                                                                        \
    __host__ __device__ __forceinline__ void skipahead_subsequence(SUBSEQ_T subsequence, STATE* state)        \
    {                                                                                                        \
        state->discard_subsequence(subsequence);                                                             \
    }                                                                                                        \
    __host__ __device__ __forceinline__ void skipahead_sequence(SUBSEQ_T sequence, STATE* state)              \
    {                                                                                                        \
        state->SEQ_CALL(sequence);                                                                           \
    }
ROCRAND_AMD_SKIP(rocrand_state_philox4x32_10, unsigned long long, discard_subsequence)
ROCRAND_AMD_SKIP(rocrand_state_mrg32k3a, unsigned long long, discard_sequence)
ROCRAND_AMD_SKIP(rocrand_state_mrg31k3p, unsigned long long, discard_sequence)
ROCRAND_AMD_SKIP(rocrand_state_xorwow, unsigned long long, discard_subsequence)
ROCRAND_AMD_SKIP(rocrand_state_lfsr113, unsigned int, discard_subsequence)
#undef ROCRAND_AMD_SKIP

// Sobol: the caller passes the dimension's direction vectors (rocrand_get_direction_vectors32/64)
// and, for the scrambled states, its constant (rocrand_get_scramble_constants32/64).
__host__ __device__ __forceinline__ void rocrand_init(const unsigned int* vectors, const unsigned int offset,
                                                      rocrand_state_sobol32* state)
{
    *state = rocrand_state_sobol32(vectors, offset);
}
__host__ __device__ __forceinline__ void rocrand_init(const unsigned long long* vectors, const unsigned int offset,
                                                      rocrand_state_sobol64* state)
{
    *state = rocrand_state_sobol64(vectors, offset);
}
__host__ __device__ __forceinline__ void rocrand_init(const unsigned int* vectors, const unsigned int scramble_constant,
                                                      const unsigned int               offset,
                                                      rocrand_state_scrambled_sobol32* state)
{
    *state = rocrand_state_scrambled_sobol32(vectors, scramble_constant, offset);
}
__host__ __device__ __forceinline__ void rocrand_init(const unsigned long long*        vectors,
                                                      const unsigned long long         scramble_constant,
                                                      const unsigned int               offset,
                                                      rocrand_state_scrambled_sobol64* state)
{
    *state = rocrand_state_scrambled_sobol64(vectors, scramble_constant, offset);
}
__host__ __device__ __forceinline__ unsigned int rocrand(rocrand_state_sobol32* state) { return state->next(); }
__host__ __device__ __forceinline__ unsigned long long rocrand(rocrand_state_sobol64* state) { return state->next(); }
__host__ __device__ __forceinline__ unsigned int rocrand(rocrand_state_scrambled_sobol32* state)
{
    return state->next();
}
__host__ __device__ __forceinline__ unsigned long long rocrand(rocrand_state_scrambled_sobol64* state)
{
    return state->next();
}
// MTGP32: every thread of the block calls together (the generator steps per block)
__device__ __forceinline__ unsigned int rocrand(rocrand_state_mtgp32* state) { return state->next(); }
// the 32-bit states skip an unsigned int count (rocrand_sobol32.h skipahead -> discard)
__host__ __device__ __forceinline__ void skipahead(unsigned long long offset, rocrand_state_sobol32* state)
{
    state->discard((unsigned int)offset);
}
__host__ __device__ __forceinline__ void skipahead(unsigned long long offset, rocrand_state_scrambled_sobol32* state)
{
    state->discard((unsigned int)offset);
}
__host__ __device__ __forceinline__ void skipahead(unsigned long long offset, rocrand_state_sobol64* state)
{
    state->discard(offset);
}
__host__ __device__ __forceinline__ void skipahead(unsigned long long offset, rocrand_state_scrambled_sobol64* state)
{
    state->discard(offset);
}

// ---------------------------------------------------------------- transforms
// Device: the library's transforms (rocrand_amd_math.h); host: the reference's host
// expressions over the host libm.
namespace rocrand_device
{
namespace amd_detail
{

constexpr float  R2P32F   = 2.3283064e-10f;
constexpr double R2P32D   = 2.3283064365386963e-10;
constexpr double R2P53D   = 1.1102230246251565e-16;
constexpr double R2P64D   = 5.4210108624275221700372640043497e-20;
constexpr float  R2P32F2P = 1.46291807e-09f;
constexpr double RPID     = 3.1415926535897932;
constexpr float  R2PIF    = 6.2831855f;
constexpr float  RSQRT2F  = 1.4142135f;

__host__ __device__ __forceinline__ float u_f32(unsigned int v) { return R2P32F + (float)v * R2P32F; }
__host__ __device__ __forceinline__ float u_f32(unsigned long long v) { return R2P32F + (float)(v >> 32) * R2P32F; }
__host__ __device__ __forceinline__ double u_f64(unsigned int v) { return R2P32D + (double)v * R2P32D; }
__host__ __device__ __forceinline__ double u_f64(unsigned int v1, unsigned int v2)
{
    ROCRAND_AMD_FP_EXACT
    const unsigned long long v = (unsigned long long)v1 | ((unsigned long long)(v2 >> 11) << 32);
    return R2P53D + (double)v * R2P53D;
}
__host__ __device__ __forceinline__ double u_f64(unsigned long long v) { return R2P53D + (double)(v >> 11) * R2P53D; }
template<int MRG>
__host__ __device__ __forceinline__ float mrg_f32(unsigned int v)
{
    return (float)((double)v
                   * (MRG == 32 ? ::rocrand_amd_engine::MRG32_NORM_DOUBLE : ::rocrand_amd_engine::MRG31_NORM_DOUBLE));
}
template<int MRG>
__host__ __device__ __forceinline__ double mrg_f64(unsigned int v)
{
    return (double)v
           * (MRG == 32 ? ::rocrand_amd_engine::MRG32_NORM_DOUBLE : ::rocrand_amd_engine::MRG31_NORM_DOUBLE);
}

// rocrand_normal.h:51-66 (angle 2^-32 2 pi + y 2^-32 2 pi).  LN selects, on the device, the
// variate the log-normals take: the reference's own float steps (logf, sqrtf, sinf/cosf,
// each rounded), which exp does not magnify beyond the log-normal bar; the normals take
// the pair evaluated in double and rounded once.
template<bool LN = false>
__host__ __device__ __forceinline__ float2 box_muller(unsigned int x, unsigned int y)
{
    ROCRAND_AMD_FP_EXACT
#if defined(__HIP_DEVICE_COMPILE__)
    float r0, r1;
    if constexpr(LN)
        ::rocrand_amd_engine::box_muller_stepwise(x, y, r0, r1);
    else
        ::rocrand_amd_engine::box_muller(x, y, r0, r1);
    return float2{r0, r1};
#else
    const float u = u_f32(x);
    const float v = R2P32F2P + (float)y * R2P32F2P;
    const float s = sqrtf(-2.0f * logf(u));
    return float2{sinf(v) * s, cosf(v) * s};
#endif
}
template<bool LN = false>
__host__ __device__ __forceinline__ float2 box_muller(unsigned long long v)
{
    return box_muller<LN>((unsigned int)v, (unsigned int)(v >> 32));
}
// rocrand_normal.h:76-96
__host__ __device__ __forceinline__ double2 box_muller_double(unsigned int a, unsigned int b, unsigned int c,
                                                              unsigned int d)
{
    ROCRAND_AMD_FP_EXACT
#if defined(__HIP_DEVICE_COMPILE__)
    double r0, r1;
    ::rocrand_amd_engine::box_muller_double(a, b, c, d, r0, r1);
    return double2{r0, r1};
#else
    const unsigned long long v1 = (unsigned long long)a ^ ((unsigned long long)b << 21);
    const double             u  = R2P53D + (double)v1 * R2P53D;
    const unsigned long long v2 = (unsigned long long)c ^ ((unsigned long long)d << 21);
    const double             w  = R2P53D * 2.0 + (double)v2 * (R2P53D * 2.0);
    const double             s  = sqrt(-2.0 * log(u));
    return double2{sin(w * RPID) * s, cos(w * RPID) * s};
#endif
}
__host__ __device__ __forceinline__ double2 box_muller_double(uint4 v) { return box_muller_double(v.x, v.y, v.z, v.w); }
__host__ __device__ __forceinline__ double2 box_muller_double(ulonglong2 v)
{
    return box_muller_double((unsigned int)v.x, (unsigned int)(v.x >> 32), (unsigned int)v.y,
                             (unsigned int)(v.y >> 32));
}
// rocrand_normal.h:138-172
template<int MRG, bool LN = false>
__host__ __device__ __forceinline__ float2 mrg_box_muller(unsigned int x, unsigned int y)
{
    ROCRAND_AMD_FP_EXACT
#if defined(__HIP_DEVICE_COMPILE__)
    float r0, r1;
    if constexpr(LN)
        ::rocrand_amd_engine::mrg_box_muller_stepwise<MRG>(x, y, r0, r1);
    else
        ::rocrand_amd_engine::mrg_box_muller<MRG>(x, y, r0, r1);
    return float2{r0, r1};
#else
    const float u = mrg_f32<MRG>(x);
    const float v = mrg_f32<MRG>(y) * R2PIF;
    const float s = sqrtf(-2.0f * logf(u));
    return float2{sinf(v) * s, cosf(v) * s};
#endif
}
template<int MRG>
__host__ __device__ __forceinline__ double2 mrg_box_muller_double(unsigned int x, unsigned int y)
{
    ROCRAND_AMD_FP_EXACT
#if defined(__HIP_DEVICE_COMPILE__)
    double r0, r1;
    ::rocrand_amd_engine::mrg_box_muller_double<MRG>(x, y, r0, r1);
    return double2{r0, r1};
#else
    const double u = (double)mrg_f32<MRG>(x);
    const double v = (double)mrg_f32<MRG>(y) * 2.0;
    const double s = sqrt(-2.0 * log(u));
    return double2{sin(v * RPID) * s, cos(v * RPID) * s};
#endif
}
// rocrand_normal.h:174-268: one word through the inverse error function
__host__ __device__ __forceinline__ float normal_1_host(float p)
{
    ROCRAND_AMD_FP_EXACT
    float z   = 2.0f * p - 1.0f;
    float sgn = z < 0.0f ? -1.0f : 1.0f;
    z         = (1.0f - z) * (1.0f + z);
    float lnx = logf(z);
    if(isnan(lnx))
        return RSQRT2F * 1.0f;
    if(isinf(lnx))
        return RSQRT2F * 0.0f;
    float tt1 = 2.0f / (3.141592653f * 0.147f) + 0.5f * lnx;
    float tt2 = 1.0f / 0.147f * lnx;
    return RSQRT2F * (sgn * sqrtf(-tt1 + sqrtf(tt1 * tt1 - tt2)));
}
__host__ __device__ __forceinline__ float normal_1(unsigned int x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return ::rocrand_amd_engine::sob_normal_f32((uint32_t)x);
#else
    return normal_1_host(u_f32(x));
#endif
}
__host__ __device__ __forceinline__ float normal_1(unsigned long long x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return ::rocrand_amd_engine::sob_normal_f32((uint64_t)x);
#else
    return normal_1_host(u_f32(x));
#endif
}
template<class W>
__host__ __device__ __forceinline__ double normal_double_1(W x)
{
    ROCRAND_AMD_FP_EXACT
#if defined(__HIP_DEVICE_COMPILE__)
    return ::rocrand_amd_engine::normal_double_1(u_f64(x));
#else
    double z   = 2.0 * u_f64(x) - 1.0;
    double sgn = z < 0.0 ? -1.0 : 1.0;
    z          = (1.0 - z) * (1.0 + z);
    double lnx = log(z);
    if(isnan(lnx))
        return (double)RSQRT2F * 1.0;
    if(isinf(lnx))
        return (double)RSQRT2F * 0.0;
    double tt1 = 2.0 / (RPID * 0.147) + 0.5 * lnx;
    double tt2 = 1.0 / 0.147 * lnx;
    return (double)RSQRT2F * (sgn * sqrt(-tt1 + sqrt(tt1 * tt1 - tt2)));
#endif
}
__host__ __device__ __forceinline__ float expf_(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return ::rocrand_amd_engine::expf_acc(x);
#else
    return expf(x);
#endif
}

// rocrand_discrete.h:61-80 (the alias method; gfx9 devices and the host use it for every
// engine).  A 64-bit word within 2^10 of 2^64 converts to x = 1: the index is held at
// size - 1 there, where the reference's read would pass the end of its tables.
__host__ __device__ __forceinline__ unsigned int discrete_alias(double x, const rocrand_discrete_distribution_st& d)
{
    ROCRAND_AMD_FP_EXACT
    const double       nx  = d.size * x;
    const double       fnx = floor(nx);
    const double       y   = nx - fnx;
    unsigned int       i   = (unsigned int)fnx;
    i                      = i < d.size ? i : d.size - 1;
    return d.offset + (y < d.probability[i] ? i : d.alias[i]);
}
// rocrand_discrete.h:109-132: binary search of the CDF (the Sobol states use it so that
// the sample stays monotone in the point)
__host__ __device__ __forceinline__ unsigned int discrete_cdf(double x, const rocrand_discrete_distribution_st& d)
{
    unsigned int lo = 0, hi = d.size - 1;
    while(lo != hi)
    {
        const unsigned int c = (lo + hi) / 2;
        if(x > d.cdf[c])
            lo = c + 1;
        else
            hi = c;
    }
    return d.offset + lo;
}
__host__ __device__ __forceinline__ double draw(unsigned int r) { return (double)r * R2P32D; }
__host__ __device__ __forceinline__ double draw(unsigned long long r) { return (double)r * R2P64D; }

} // namespace amd_detail
} // namespace rocrand_device

// Box-Muller value cache of the states that keep one (rocrand_common.h:157-192
// engine_boxmuller_helper).  The cached double is returned through float there
// (get_double's return type), so the second value of a double pair has float precision.
#define ROCRAND_AMD_BM_FLOAT(state, EXPR2)                                                                   \
    if((state)->m_state.boxmuller_float_state)                                                               \
    {                                                                                                        \
        (state)->m_state.boxmuller_float_state = 0;                                                          \
        return (state)->m_state.boxmuller_float;                                                             \
    }                                                                                                        \
    const float2 r_                         = (EXPR2);                                                       \
    (state)->m_state.boxmuller_float_state = 1;                                                              \
    (state)->m_state.boxmuller_float       = r_.y;                                                           \
    return r_.x;
#define ROCRAND_AMD_BM_DOUBLE(state, EXPR2)                                                                  \
    if((state)->m_state.boxmuller_double_state)                                                              \
    {                                                                                                        \
        (state)->m_state.boxmuller_double_state = 0;                                                         \
        return (double)(float)(state)->m_state.boxmuller_double;                                             \
    }                                                                                                        \
    const double2 r_                         = (EXPR2);                                                      \
    (state)->m_state.boxmuller_double_state = 1;                                                             \
    (state)->m_state.boxmuller_double       = r_.y;                                                          \
    return r_.x;

namespace rocrand_device
{
namespace amd_detail
{
// Box-Muller draws: two words (the MRGs use their raw values in [1, m1]).
template<bool LN = false>
__host__ __device__ __forceinline__ float2 bm2(rocrand_state_philox4x32_10* s)
{
    const unsigned int a = rocrand(s);
    const unsigned int b = rocrand(s);
    return box_muller<LN>(a, b);
}
template<bool LN = false>
__host__ __device__ __forceinline__ float2 bm2(rocrand_state_xorwow* s)
{
    const unsigned int a = rocrand(s);
    const unsigned int b = rocrand(s);
    return box_muller<LN>(a, b);
}
template<bool LN = false, int MRG>
__host__ __device__ __forceinline__ float2 bm2(mrg_engine<MRG>* s)
{
    const unsigned int a = s->next();
    const unsigned int b = s->next();
    return mrg_box_muller<MRG, LN>(a, b);
}
template<int MRG>
__host__ __device__ __forceinline__ double2 bmd2(mrg_engine<MRG>* s)
{
    const unsigned int a = s->next();
    const unsigned int b = s->next();
    return mrg_box_muller_double<MRG>(a, b);
}
template<class S>
__host__ __device__ __forceinline__ double2 bmd2_words(S* s)
{
    const unsigned int a = rocrand(s);
    const unsigned int b = rocrand(s);
    const unsigned int c = rocrand(s);
    const unsigned int d = rocrand(s);
    return box_muller_double(a, b, c, d);
}
} // namespace amd_detail
} // namespace rocrand_device

// ---------------------------------------------------------------- uniform (rocrand_uniform.h)
__host__ __device__ __forceinline__ float rocrand_uniform(rocrand_state_philox4x32_10* state)
{
    return rocrand_device::amd_detail::u_f32(rocrand(state));
}
__host__ __device__ __forceinline__ float2 rocrand_uniform2(rocrand_state_philox4x32_10* state)
{
    const unsigned int a = rocrand(state);
    const unsigned int b = rocrand(state);
    return float2{rocrand_device::amd_detail::u_f32(a), rocrand_device::amd_detail::u_f32(b)};
}
__host__ __device__ __forceinline__ float4 rocrand_uniform4(rocrand_state_philox4x32_10* state)
{
    using namespace rocrand_device::amd_detail;
    const uint4 v = rocrand4(state);
    return float4{u_f32(v.x), u_f32(v.y), u_f32(v.z), u_f32(v.w)};
}
__host__ __device__ __forceinline__ double rocrand_uniform_double(rocrand_state_philox4x32_10* state)
{
    const unsigned int a = rocrand(state);
    const unsigned int b = rocrand(state);
    return rocrand_device::amd_detail::u_f64(a, b);
}
__host__ __device__ __forceinline__ double2 rocrand_uniform_double2(rocrand_state_philox4x32_10* state)
{
    using namespace rocrand_device::amd_detail;
    const uint4 v = rocrand4(state);
    return double2{u_f64(v.x, v.y), u_f64(v.z, v.w)};
}
__host__ __device__ __forceinline__ double4 rocrand_uniform_double4(rocrand_state_philox4x32_10* state)
{
    using namespace rocrand_device::amd_detail;
    const uint4 a = rocrand4(state);
    const uint4 b = rocrand4(state);
    return double4{u_f64(a.x, a.y), u_f64(a.z, a.w), u_f64(b.x, b.y), u_f64(b.z, b.w)};
}
#define ROCRAND_AMD_MRG_UNIFORM(STATE, MRG)                                                                  \
    __host__ __device__ __forceinline__ float rocrand_uniform(STATE* state)                                  \
    {                                                                                                        \
        return rocrand_device::amd_detail::mrg_f32<MRG>(state->next());                                      \
    }                                                                                                        \
    __host__ __device__ __forceinline__ double rocrand_uniform_double(STATE* state)                          \
    {                                                                                                        \
        return rocrand_device::amd_detail::mrg_f64<MRG>(state->next());                                      \
    }
ROCRAND_AMD_MRG_UNIFORM(rocrand_state_mrg32k3a, 32)
ROCRAND_AMD_MRG_UNIFORM(rocrand_state_mrg31k3p, 31)
#undef ROCRAND_AMD_MRG_UNIFORM
__host__ __device__ __forceinline__ float rocrand_uniform(rocrand_state_xorwow* state)
{
    return rocrand_device::amd_detail::u_f32(rocrand(state));
}
__host__ __device__ __forceinline__ double rocrand_uniform_double(rocrand_state_xorwow* state)
{
    const unsigned int a = rocrand(state);
    const unsigned int b = rocrand(state);
    return rocrand_device::amd_detail::u_f64(a, b);
}
// One word per value: LFSR113 and the ThreeFry states.
#define ROCRAND_AMD_ONE_WORD_UNIFORM(STATE)                                                                  \
    __host__ __device__ __forceinline__ float rocrand_uniform(STATE* state)                                  \
    {                                                                                                        \
        return rocrand_device::amd_detail::u_f32(rocrand(state));                                            \
    }                                                                                                        \
    __host__ __device__ __forceinline__ double rocrand_uniform_double(STATE* state)                          \
    {                                                                                                        \
        return rocrand_device::amd_detail::u_f64(rocrand(state));                                            \
    }
ROCRAND_AMD_ONE_WORD_UNIFORM(rocrand_state_lfsr113)
ROCRAND_AMD_ONE_WORD_UNIFORM(rocrand_state_threefry2x32_20)
ROCRAND_AMD_ONE_WORD_UNIFORM(rocrand_state_threefry2x64_20)
ROCRAND_AMD_ONE_WORD_UNIFORM(rocrand_state_threefry4x32_20)
ROCRAND_AMD_ONE_WORD_UNIFORM(rocrand_state_threefry4x64_20)
ROCRAND_AMD_ONE_WORD_UNIFORM(rocrand_state_sobol32)
ROCRAND_AMD_ONE_WORD_UNIFORM(rocrand_state_sobol64)
ROCRAND_AMD_ONE_WORD_UNIFORM(rocrand_state_scrambled_sobol32)
ROCRAND_AMD_ONE_WORD_UNIFORM(rocrand_state_scrambled_sobol64)
#undef ROCRAND_AMD_ONE_WORD_UNIFORM

// ---------------------------------------------------------------- normal (rocrand_normal.h:340-1450)
__host__ __device__ __forceinline__ float rocrand_normal(rocrand_state_philox4x32_10* state)
{
    ROCRAND_AMD_BM_FLOAT(state, rocrand_device::amd_detail::bm2(state))
}
__host__ __device__ __forceinline__ float2 rocrand_normal2(rocrand_state_philox4x32_10* state)
{
    return rocrand_device::amd_detail::bm2(state);
}
__host__ __device__ __forceinline__ float4 rocrand_normal4(rocrand_state_philox4x32_10* state)
{
    using namespace rocrand_device::amd_detail;
    const uint4  v  = rocrand4(state);
    const float2 r1 = box_muller(v.x, v.y);
    const float2 r2 = box_muller(v.z, v.w);
    return float4{r1.x, r1.y, r2.x, r2.y};
}
__host__ __device__ __forceinline__ double rocrand_normal_double(rocrand_state_philox4x32_10* state)
{
    ROCRAND_AMD_BM_DOUBLE(state, rocrand_device::amd_detail::box_muller_double(rocrand4(state)))
}
__host__ __device__ __forceinline__ double2 rocrand_normal_double2(rocrand_state_philox4x32_10* state)
{
    return rocrand_device::amd_detail::box_muller_double(rocrand4(state));
}
__host__ __device__ __forceinline__ double4 rocrand_normal_double4(rocrand_state_philox4x32_10* state)
{
    using namespace rocrand_device::amd_detail;
    const double2 r1 = box_muller_double(rocrand4(state));
    const double2 r2 = box_muller_double(rocrand4(state));
    return double4{r1.x, r1.y, r2.x, r2.y};
}
#define ROCRAND_AMD_MRG_NORMAL(STATE)                                                                        \
    __host__ __device__ __forceinline__ float rocrand_normal(STATE* state)                                   \
    {                                                                                                        \
        ROCRAND_AMD_BM_FLOAT(state, rocrand_device::amd_detail::bm2(state))                                  \
    }                                                                                                        \
    __host__ __device__ __forceinline__ float2 rocrand_normal2(STATE* state)                                 \
    {                                                                                                        \
        return rocrand_device::amd_detail::bm2(state);                                                       \
    }                                                                                                        \
    __host__ __device__ __forceinline__ double rocrand_normal_double(STATE* state)                           \
    {                                                                                                        \
        ROCRAND_AMD_BM_DOUBLE(state, rocrand_device::amd_detail::bmd2(state))                                \
    }                                                                                                        \
    __host__ __device__ __forceinline__ double2 rocrand_normal_double2(STATE* state)                         \
    {                                                                                                        \
        return rocrand_device::amd_detail::bmd2(state);                                                      \
    }
ROCRAND_AMD_MRG_NORMAL(rocrand_state_mrg32k3a)
ROCRAND_AMD_MRG_NORMAL(rocrand_state_mrg31k3p)
#undef ROCRAND_AMD_MRG_NORMAL
__host__ __device__ __forceinline__ float rocrand_normal(rocrand_state_xorwow* state)
{
    ROCRAND_AMD_BM_FLOAT(state, rocrand_device::amd_detail::bm2(state))
}
__host__ __device__ __forceinline__ float2 rocrand_normal2(rocrand_state_xorwow* state)
{
    return rocrand_device::amd_detail::bm2(state);
}
__host__ __device__ __forceinline__ double rocrand_normal_double(rocrand_state_xorwow* state)
{
    ROCRAND_AMD_BM_DOUBLE(state, rocrand_device::amd_detail::bmd2_words(state))
}
__host__ __device__ __forceinline__ double2 rocrand_normal_double2(rocrand_state_xorwow* state)
{
    return rocrand_device::amd_detail::bmd2_words(state);
}
// LFSR113 and ThreeFry: one word through the inverse error function for single values.
#define ROCRAND_AMD_ONE_WORD_NORMAL(STATE)                                                                   \
    __host__ __device__ __forceinline__ float rocrand_normal(STATE* state)                                   \
    {                                                                                                        \
        return rocrand_device::amd_detail::normal_1(rocrand(state));                                         \
    }                                                                                                        \
    __host__ __device__ __forceinline__ double rocrand_normal_double(STATE* state)                           \
    {                                                                                                        \
        return rocrand_device::amd_detail::normal_double_1(rocrand(state));                                  \
    }
ROCRAND_AMD_ONE_WORD_NORMAL(rocrand_state_lfsr113)
ROCRAND_AMD_ONE_WORD_NORMAL(rocrand_state_threefry2x32_20)
ROCRAND_AMD_ONE_WORD_NORMAL(rocrand_state_threefry2x64_20)
ROCRAND_AMD_ONE_WORD_NORMAL(rocrand_state_threefry4x32_20)
ROCRAND_AMD_ONE_WORD_NORMAL(rocrand_state_threefry4x64_20)
ROCRAND_AMD_ONE_WORD_NORMAL(rocrand_state_sobol32)
ROCRAND_AMD_ONE_WORD_NORMAL(rocrand_state_sobol64)
ROCRAND_AMD_ONE_WORD_NORMAL(rocrand_state_scrambled_sobol32)
ROCRAND_AMD_ONE_WORD_NORMAL(rocrand_state_scrambled_sobol64)
#undef ROCRAND_AMD_ONE_WORD_NORMAL
__device__ __forceinline__ float rocrand_uniform(rocrand_state_mtgp32* state)
{
    return rocrand_device::amd_detail::u_f32(rocrand(state));
}
__device__ __forceinline__ double rocrand_uniform_double(rocrand_state_mtgp32* state)
{
    return rocrand_device::amd_detail::u_f64(rocrand(state));
}
__device__ __forceinline__ float rocrand_normal(rocrand_state_mtgp32* state)
{
    return rocrand_device::amd_detail::normal_1(rocrand(state));
}
__device__ __forceinline__ double rocrand_normal_double(rocrand_state_mtgp32* state)
{
    return rocrand_device::amd_detail::normal_double_1(rocrand(state));
}
__host__ __device__ __forceinline__ float2 rocrand_normal2(rocrand_state_lfsr113* state)
{
    const unsigned int a = rocrand(state);
    const unsigned int b = rocrand(state);
    return rocrand_device::amd_detail::box_muller(a, b);
}
__host__ __device__ __forceinline__ double2 rocrand_normal_double2(rocrand_state_lfsr113* state)
{
    return rocrand_device::amd_detail::bmd2_words(state);
}
__host__ __device__ __forceinline__ float2 rocrand_normal2(rocrand_state_threefry2x32_20* state)
{
    const uint2 v = rocrand2(state);
    return rocrand_device::amd_detail::box_muller(v.x, v.y);
}
__host__ __device__ __forceinline__ double2 rocrand_normal_double2(rocrand_state_threefry2x32_20* state)
{
    const uint2 a = rocrand2(state);
    const uint2 b = rocrand2(state);
    return rocrand_device::amd_detail::box_muller_double(a.x, a.y, b.x, b.y);
}
__host__ __device__ __forceinline__ float2 rocrand_normal2(rocrand_state_threefry2x64_20* state)
{
    return rocrand_device::amd_detail::box_muller(rocrand(state));
}
__host__ __device__ __forceinline__ double2 rocrand_normal_double2(rocrand_state_threefry2x64_20* state)
{
    return rocrand_device::amd_detail::box_muller_double(rocrand2(state));
}
__host__ __device__ __forceinline__ float2 rocrand_normal2(rocrand_state_threefry4x32_20* state)
{
    const unsigned int a = rocrand(state);
    const unsigned int b = rocrand(state);
    return rocrand_device::amd_detail::box_muller(a, b);
}
__host__ __device__ __forceinline__ double2 rocrand_normal_double2(rocrand_state_threefry4x32_20* state)
{
    return rocrand_device::amd_detail::box_muller_double(rocrand4(state));
}
// rocrand_normal.h: the two 64-bit words convert to 32 bits (their low halves)
__host__ __device__ __forceinline__ float2 rocrand_normal2(rocrand_state_threefry4x64_20* state)
{
    const unsigned long long a = rocrand(state);
    const unsigned long long b = rocrand(state);
    return rocrand_device::amd_detail::box_muller((unsigned int)a, (unsigned int)b);
}
__host__ __device__ __forceinline__ double2 rocrand_normal_double2(rocrand_state_threefry4x64_20* state)
{
    const unsigned long long a = rocrand(state);
    const unsigned long long b = rocrand(state);
    return rocrand_device::amd_detail::box_muller_double(ulonglong2{a, b});
}

// ---------------------------------------------------------------- log-normal (rocrand_log_normal.h)
namespace rocrand_device
{
namespace amd_detail
{
// The Box-Muller pairs of rocrand_normal2 with the log-normal variate (box_muller<true>),
// and rocrand_normal's single value sharing the state's Box-Muller cache.
__host__ __device__ __forceinline__ float2 normal2_ln(rocrand_state_philox4x32_10* s) { return bm2<true>(s); }
__host__ __device__ __forceinline__ float2 normal2_ln(rocrand_state_xorwow* s) { return bm2<true>(s); }
__host__ __device__ __forceinline__ float2 normal2_ln(rocrand_state_mrg32k3a* s) { return bm2<true>(s); }
__host__ __device__ __forceinline__ float2 normal2_ln(rocrand_state_mrg31k3p* s) { return bm2<true>(s); }
__host__ __device__ __forceinline__ float2 normal2_ln(rocrand_state_lfsr113* s)
{
    const unsigned int a = rocrand(s);
    const unsigned int b = rocrand(s);
    return box_muller<true>(a, b);
}
__host__ __device__ __forceinline__ float2 normal2_ln(rocrand_state_threefry2x32_20* s)
{
    const uint2 v = rocrand2(s);
    return box_muller<true>(v.x, v.y);
}
__host__ __device__ __forceinline__ float2 normal2_ln(rocrand_state_threefry4x32_20* s)
{
    const unsigned int a = rocrand(s);
    const unsigned int b = rocrand(s);
    return box_muller<true>(a, b);
}
__host__ __device__ __forceinline__ float2 normal2_ln(rocrand_state_threefry4x64_20* s)
{
    const unsigned long long a = rocrand(s);
    const unsigned long long b = rocrand(s);
    return box_muller<true>((unsigned int)a, (unsigned int)b);
}
#define ROCRAND_AMD_NORMAL_LN(STATE)                                                                         \
    __host__ __device__ __forceinline__ float normal_ln(STATE* state)                                        \
    {                                                                                                        \
        ROCRAND_AMD_BM_FLOAT(state, normal2_ln(state))                                                       \
    }
ROCRAND_AMD_NORMAL_LN(rocrand_state_philox4x32_10)
ROCRAND_AMD_NORMAL_LN(rocrand_state_mrg32k3a)
ROCRAND_AMD_NORMAL_LN(rocrand_state_mrg31k3p)
ROCRAND_AMD_NORMAL_LN(rocrand_state_xorwow)
#undef ROCRAND_AMD_NORMAL_LN
__host__ __device__ __forceinline__ float ln_f(float mean, float stddev, float r)
{
    ROCRAND_AMD_FP_EXACT
    return expf_(mean + stddev * r);
}
// the ThreeFry single-value forms take double parameters: the sum is a double, then
// converted to float for expf
__host__ __device__ __forceinline__ float ln_fd(double mean, double stddev, float r)
{
    ROCRAND_AMD_FP_EXACT
    return expf_((float)(mean + stddev * (double)r));
}
__host__ __device__ __forceinline__ double ln_d(double mean, double stddev, double r)
{
    ROCRAND_AMD_FP_EXACT
#if defined(__HIP_DEVICE_COMPILE__)
    return ::rocrand_amd_engine::exp_cr(mean + stddev * r);
#else
    return exp(mean + stddev * r);
#endif
}
} // namespace amd_detail
} // namespace rocrand_device

#define ROCRAND_AMD_LN_FROM_NORMAL(STATE)                                                                    \
    __host__ __device__ __forceinline__ float2 rocrand_log_normal2(STATE* state, float mean, float stddev)   \
    {                                                                                                        \
        const float2 r = rocrand_device::amd_detail::normal2_ln(state);                                      \
        return float2{rocrand_device::amd_detail::ln_f(mean, stddev, r.x),                                   \
                      rocrand_device::amd_detail::ln_f(mean, stddev, r.y)};                                  \
    }                                                                                                        \
    __host__ __device__ __forceinline__ double2 rocrand_log_normal_double2(STATE* state, double mean,        \
                                                                           double stddev)                    \
    {                                                                                                        \
        const double2 r = rocrand_normal_double2(state);                                                     \
        return double2{rocrand_device::amd_detail::ln_d(mean, stddev, r.x),                                  \
                       rocrand_device::amd_detail::ln_d(mean, stddev, r.y)};                                 \
    }
// States with a Box-Muller cache: the single value shares it with rocrand_normal.
#define ROCRAND_AMD_LN_CACHED(STATE)                                                                         \
    ROCRAND_AMD_LN_FROM_NORMAL(STATE)                                                                        \
    __host__ __device__ __forceinline__ float rocrand_log_normal(STATE* state, float mean, float stddev)     \
    {                                                                                                        \
        return rocrand_device::amd_detail::ln_f(mean, stddev, rocrand_device::amd_detail::normal_ln(state)); \
    }                                                                                                        \
    __host__ __device__ __forceinline__ double rocrand_log_normal_double(STATE* state, double mean,          \
                                                                         double stddev)                      \
    {                                                                                                        \
        return rocrand_device::amd_detail::ln_d(mean, stddev, rocrand_normal_double(state));                 \
    }
ROCRAND_AMD_LN_CACHED(rocrand_state_philox4x32_10)
ROCRAND_AMD_LN_CACHED(rocrand_state_mrg32k3a)
ROCRAND_AMD_LN_CACHED(rocrand_state_mrg31k3p)
ROCRAND_AMD_LN_CACHED(rocrand_state_xorwow)
#undef ROCRAND_AMD_LN_CACHED
__host__ __device__ __forceinline__ float rocrand_log_normal(rocrand_state_lfsr113* state, float mean, float stddev)
{
    return rocrand_device::amd_detail::ln_f(mean, stddev, rocrand_normal(state));
}
__host__ __device__ __forceinline__ double rocrand_log_normal_double(rocrand_state_lfsr113* state, double mean,
                                                                     double stddev)
{
    return rocrand_device::amd_detail::ln_d(mean, stddev, rocrand_normal_double(state));
}
ROCRAND_AMD_LN_FROM_NORMAL(rocrand_state_lfsr113)
#define ROCRAND_AMD_LN_THREEFRY(STATE)                                                                       \
    ROCRAND_AMD_LN_FROM_NORMAL(STATE)                                                                        \
    __host__ __device__ __forceinline__ float rocrand_log_normal(STATE* state, double mean, double stddev)   \
    {                                                                                                        \
        return rocrand_device::amd_detail::ln_fd(mean, stddev, rocrand_normal(state));                       \
    }                                                                                                        \
    __host__ __device__ __forceinline__ double rocrand_log_normal_double(STATE* state, double mean,          \
                                                                         double stddev)                      \
    {                                                                                                        \
        return rocrand_device::amd_detail::ln_d(mean, stddev, rocrand_normal_double(state));                 \
    }
ROCRAND_AMD_LN_THREEFRY(rocrand_state_threefry2x32_20)
ROCRAND_AMD_LN_THREEFRY(rocrand_state_threefry4x32_20)
ROCRAND_AMD_LN_THREEFRY(rocrand_state_threefry4x64_20)
#undef ROCRAND_AMD_LN_THREEFRY
#undef ROCRAND_AMD_LN_FROM_NORMAL
// ThreeFry2x64: rocrand_normal2 splits one 64-bit word, but rocrand_log_normal2 draws two
// words and uses their low halves (rocrand_log_normal.h, threefry2x64 overloads).
__host__ __device__ __forceinline__ float rocrand_log_normal(rocrand_state_threefry2x64_20* state, double mean,
                                                             double stddev)
{
    return rocrand_device::amd_detail::ln_fd(mean, stddev, rocrand_normal(state));
}
__host__ __device__ __forceinline__ float2 rocrand_log_normal2(rocrand_state_threefry2x64_20* state, float mean,
                                                               float stddev)
{
    using namespace rocrand_device::amd_detail;
    const unsigned long long a = rocrand(state);
    const unsigned long long b = rocrand(state);
    const float2             r = box_muller<true>((unsigned int)a, (unsigned int)b);
    return float2{ln_f(mean, stddev, r.x), ln_f(mean, stddev, r.y)};
}
__host__ __device__ __forceinline__ double rocrand_log_normal_double(rocrand_state_threefry2x64_20* state,
                                                                     double mean, double stddev)
{
    return rocrand_device::amd_detail::ln_d(mean, stddev, rocrand_normal_double(state));
}
__host__ __device__ __forceinline__ double2 rocrand_log_normal_double2(rocrand_state_threefry2x64_20* state,
                                                                       double mean, double stddev)
{
    using namespace rocrand_device::amd_detail;
    const double2 r = rocrand_normal_double2(state);
    return double2{ln_d(mean, stddev, r.x), ln_d(mean, stddev, r.y)};
}
#define ROCRAND_AMD_LN_SOBOL(STATE)                                                                          \
    __host__ __device__ __forceinline__ float rocrand_log_normal(STATE* state, float mean, float stddev)     \
    {                                                                                                        \
        return rocrand_device::amd_detail::ln_f(mean, stddev, rocrand_normal(state));                        \
    }                                                                                                        \
    __host__ __device__ __forceinline__ double rocrand_log_normal_double(STATE* state, double mean,          \
                                                                         double stddev)                      \
    {                                                                                                        \
        return rocrand_device::amd_detail::ln_d(mean, stddev, rocrand_normal_double(state));                 \
    }
ROCRAND_AMD_LN_SOBOL(rocrand_state_sobol32)
ROCRAND_AMD_LN_SOBOL(rocrand_state_sobol64)
ROCRAND_AMD_LN_SOBOL(rocrand_state_scrambled_sobol32)
ROCRAND_AMD_LN_SOBOL(rocrand_state_scrambled_sobol64)
#undef ROCRAND_AMD_LN_SOBOL
__host__ __device__ __forceinline__ float4 rocrand_log_normal4(rocrand_state_philox4x32_10* state, float mean,
                                                               float stddev)
{
    using namespace rocrand_device::amd_detail;
    const uint4  v  = rocrand4(state);
    const float2 r1 = box_muller<true>(v.x, v.y);
    const float2 r2 = box_muller<true>(v.z, v.w);
    return float4{ln_f(mean, stddev, r1.x), ln_f(mean, stddev, r1.y), ln_f(mean, stddev, r2.x),
                  ln_f(mean, stddev, r2.y)};
}
__host__ __device__ __forceinline__ double4 rocrand_log_normal_double4(rocrand_state_philox4x32_10* state,
                                                                       double mean, double stddev)
{
    using namespace rocrand_device::amd_detail;
    const double4 r = rocrand_normal_double4(state);
    return double4{ln_d(mean, stddev, r.x), ln_d(mean, stddev, r.y), ln_d(mean, stddev, r.z),
                   ln_d(mean, stddev, r.w)};
}

// ---------------------------------------------------------------- Poisson (rocrand_poisson.h)
namespace rocrand_device
{
namespace amd_detail
{
// rocrand_poisson.h:48-160: Knuth's product of uniforms below lambda 64, a transformed
// rejection with a Lanczos log-gamma up to 4000, the normal approximation above.
template<class S>
__host__ __device__ __forceinline__ unsigned int poisson_small(S* state, double lambda)
{
    ROCRAND_AMD_FP_EXACT
    const double limit   = exp(-lambda);
    unsigned int k       = 0;
    double       product = 1.0;
    do
    {
        k++;
        product *= rocrand_uniform_double(state);
    } while(product > limit);
    return k - 1;
}
__host__ __device__ __forceinline__ double lgamma_lanczos(double x)
{
    ROCRAND_AMD_FP_EXACT
    // Lanczos approximation, g = 7, 9 terms (the published coefficient set)
    const double z       = x - 1.0;
    const double coef[9] = {0.99999999999980993227684700473478, 676.520368121885098567009190444019,
                            -1259.13921672240287047156078755283, 771.3234287776530788486528258894,
                            -176.61502916214059906584551354,     12.507343278686904814458936853,
                            -0.13857109526572011689554707,       9.984369578019570859563e-6,
                            1.50563273514931155834e-7};
    double sum = 0.0;
    for(int i = 8; i > 0; i--)
        sum += coef[i] / (z + i);
    sum += coef[0];
    return (0.9189385332046727418 + log(sum) - 7) + (z + 0.5) * log((z + 7 + 0.5) / 2.718281828459045090796);
}
template<class S>
__host__ __device__ __forceinline__ unsigned int poisson_large(S* state, double lambda)
{
    ROCRAND_AMD_FP_EXACT
    const double c          = 0.767 - 3.36 / lambda;
    const double beta       = RPID / sqrt(3.0 * lambda);
    const double alpha      = beta * lambda;
    const double k          = log(c) - lambda - log(beta);
    const double log_lambda = log(lambda);
    while(true)
    {
        const double u = rocrand_uniform_double(state);
        const double x = (alpha - log((1.0 - u) / u)) / beta;
        const double n = floor(x + 0.5);
        if(n < 0)
            continue;
        const double v   = rocrand_uniform_double(state);
        const double y   = alpha - beta * x;
        const double t   = 1.0 + exp(y);
        const double lhs = y + log(v / (t * t));
        const double rhs = k + n * log_lambda - lgamma_lanczos(n + 1.0);
        if(lhs <= rhs)
            return (unsigned int)n;
    }
}
template<class S, class R = unsigned int>
__host__ __device__ __forceinline__ R poisson_huge(S* state, double lambda)
{
    ROCRAND_AMD_FP_EXACT
    const double n = rocrand_normal_double(state);
    return (R)round(sqrt(lambda) * n + lambda);
}
template<class S>
__host__ __device__ __forceinline__ unsigned int poisson(S* state, double lambda)
{
    if(lambda < 64.0)
        return poisson_small(state, lambda);
    if(lambda <= 4000.0)
        return poisson_large(state, lambda);
    return poisson_huge(state, lambda);
}
// rocrand_poisson.h:164-216: inversion (Fishman's ITR) in blocks of e^-500 (LFSR113,
// ThreeFry, Sobol; R is the 64-bit count of the Sobol64 states).  The scale of block j is e^-500 while lambda exceeds the block's end
// (500 (j + 1)) and e^(500 j - lambda) for the last block.  The reference compares lambda
// with the first block's end only, so for 500 < lambda < 1000 its second block scales by
// e^-500 again, its CDF tops out at e^(lambda - 1000) < 1 and a draw above that spins
// until the 32-bit count wraps.  Here every block compares with its own end: the same
// values for lambda <= 500 and lambda >= 1000, a terminating inversion in between.
template<class S, class R = unsigned int>
__host__ __device__ __forceinline__ R poisson_inv(S* state, double lambda)
{
    ROCRAND_AMD_FP_EXACT
    if(lambda >= 1000.0)
        return poisson_huge<S, R>(state, lambda);
    double       x = 1.0, y = 1.0, L;
    R            k  = 0;
    int          pw = 0;
    const double u  = rocrand_uniform_double(state) - R2P32D / 2.0;
    const double ex = exp(-500.0);
    do
    {
        L = lambda > pw + 500.0 ? ex : exp((double)(pw - lambda));
        x *= L;
        y *= L;
        pw += 500;
        while(u > y)
        {
            k++;
            x *= (lambda / (double)k);
            y += x;
        }
    } while((double)pw < lambda);
    return k;
}
} // namespace amd_detail
} // namespace rocrand_device

#define ROCRAND_AMD_POISSON(STATE, FN)                                                                       \
    __host__ __device__ __forceinline__ unsigned int rocrand_poisson(STATE* state, double lambda)            \
    {                                                                                                        \
        return rocrand_device::amd_detail::FN(state, lambda);                                                \
    }
ROCRAND_AMD_POISSON(rocrand_state_philox4x32_10, poisson)
ROCRAND_AMD_POISSON(rocrand_state_mrg32k3a, poisson)
ROCRAND_AMD_POISSON(rocrand_state_mrg31k3p, poisson)
ROCRAND_AMD_POISSON(rocrand_state_xorwow, poisson)
ROCRAND_AMD_POISSON(rocrand_state_lfsr113, poisson_inv)
ROCRAND_AMD_POISSON(rocrand_state_threefry2x32_20, poisson_inv)
ROCRAND_AMD_POISSON(rocrand_state_threefry2x64_20, poisson_inv)
ROCRAND_AMD_POISSON(rocrand_state_threefry4x32_20, poisson_inv)
ROCRAND_AMD_POISSON(rocrand_state_threefry4x64_20, poisson_inv)
ROCRAND_AMD_POISSON(rocrand_state_sobol32, poisson_inv)
ROCRAND_AMD_POISSON(rocrand_state_scrambled_sobol32, poisson_inv)
#undef ROCRAND_AMD_POISSON
__host__ __device__ __forceinline__ unsigned long long rocrand_poisson(rocrand_state_sobol64* state, double lambda)
{
    return rocrand_device::amd_detail::poisson_inv<rocrand_state_sobol64, unsigned long long>(state, lambda);
}
__host__ __device__ __forceinline__ unsigned long long rocrand_poisson(rocrand_state_scrambled_sobol64* state,
                                                                      double                           lambda)
{
    return rocrand_device::amd_detail::poisson_inv<rocrand_state_scrambled_sobol64, unsigned long long>(state,
                                                                                                        lambda);
}
__host__ __device__ __forceinline__ uint4 rocrand_poisson4(rocrand_state_philox4x32_10* state, double lambda)
{
    const unsigned int a = rocrand_poisson(state, lambda);
    const unsigned int b = rocrand_poisson(state, lambda);
    const unsigned int c = rocrand_poisson(state, lambda);
    const unsigned int d = rocrand_poisson(state, lambda);
    return uint4{a, b, c, d};
}

// ---------------------------------------------------------------- discrete (rocrand_discrete.h)
#define ROCRAND_AMD_DISCRETE(STATE)                                                                          \
    __host__ __device__ __forceinline__ unsigned int rocrand_discrete(                                       \
        STATE* state, const rocrand_discrete_distribution discrete_distribution)                             \
    {                                                                                                        \
        using namespace rocrand_device::amd_detail;                                                          \
        return discrete_alias(draw(rocrand(state)), *discrete_distribution);                                 \
    }
ROCRAND_AMD_DISCRETE(rocrand_state_philox4x32_10)
ROCRAND_AMD_DISCRETE(rocrand_state_mrg32k3a)
ROCRAND_AMD_DISCRETE(rocrand_state_mrg31k3p)
ROCRAND_AMD_DISCRETE(rocrand_state_xorwow)
ROCRAND_AMD_DISCRETE(rocrand_state_lfsr113)
ROCRAND_AMD_DISCRETE(rocrand_state_threefry2x32_20)
ROCRAND_AMD_DISCRETE(rocrand_state_threefry2x64_20)
ROCRAND_AMD_DISCRETE(rocrand_state_threefry4x32_20)
ROCRAND_AMD_DISCRETE(rocrand_state_threefry4x64_20)
#undef ROCRAND_AMD_DISCRETE
#define ROCRAND_AMD_DISCRETE_CDF(STATE)                                                                      \
    __host__ __device__ __forceinline__ unsigned int rocrand_discrete(                                       \
        STATE* state, const rocrand_discrete_distribution discrete_distribution)                             \
    {                                                                                                        \
        using namespace rocrand_device::amd_detail;                                                          \
        return discrete_cdf(draw(rocrand(state)), *discrete_distribution);                                   \
    }
ROCRAND_AMD_DISCRETE_CDF(rocrand_state_sobol32)
ROCRAND_AMD_DISCRETE_CDF(rocrand_state_sobol64)
ROCRAND_AMD_DISCRETE_CDF(rocrand_state_scrambled_sobol32)
ROCRAND_AMD_DISCRETE_CDF(rocrand_state_scrambled_sobol64)
#undef ROCRAND_AMD_DISCRETE_CDF
__host__ __device__ __forceinline__ uint4 rocrand_discrete4(rocrand_state_philox4x32_10*        state,
                                                            const rocrand_discrete_distribution discrete_distribution)
{
    using namespace rocrand_device::amd_detail;
    const uint4 v = rocrand4(state);
    return uint4{discrete_alias(draw(v.x), *discrete_distribution), discrete_alias(draw(v.y), *discrete_distribution),
                 discrete_alias(draw(v.z), *discrete_distribution),
                 discrete_alias(draw(v.w), *discrete_distribution)};
}

// ---------------------------------------------------------------- MTGP32 calls (device only)
// Every thread of the block makes each call together (the generator steps per block).
__device__ __forceinline__ void rocrand_mtgp32_block_copy(rocrand_state_mtgp32* src, rocrand_state_mtgp32* dest)
{
    dest->copy(src);
}
__device__ __forceinline__ void rocrand_mtgp32_set_params(rocrand_state_mtgp32* state, mtgp32_params* params)
{
    state->set_params(params);
}
__device__ __forceinline__ float2 rocrand_normal2(rocrand_state_mtgp32* state)
{
    const unsigned int a = rocrand(state);
    const unsigned int b = rocrand(state);
    return rocrand_device::amd_detail::box_muller(a, b);
}
__device__ __forceinline__ double2 rocrand_normal_double2(rocrand_state_mtgp32* state)
{
    return rocrand_device::amd_detail::bmd2_words(state);
}
__device__ __forceinline__ float rocrand_log_normal(rocrand_state_mtgp32* state, float mean, float stddev)
{
    return rocrand_device::amd_detail::ln_f(mean, stddev, rocrand_normal(state));
}
__device__ __forceinline__ float2 rocrand_log_normal2(rocrand_state_mtgp32* state, float mean, float stddev)
{
    using namespace rocrand_device::amd_detail;
    const unsigned int a = rocrand(state);
    const unsigned int b = rocrand(state);
    const float2       r = box_muller<true>(a, b);
    return float2{ln_f(mean, stddev, r.x), ln_f(mean, stddev, r.y)};
}
__device__ __forceinline__ double rocrand_log_normal_double(rocrand_state_mtgp32* state, double mean, double stddev)
{
    return rocrand_device::amd_detail::ln_d(mean, stddev, rocrand_normal_double(state));
}
__device__ __forceinline__ double2 rocrand_log_normal_double2(rocrand_state_mtgp32* state, double mean,
                                                              double stddev)
{
    using namespace rocrand_device::amd_detail;
    const double2 r = rocrand_normal_double2(state);
    return double2{ln_d(mean, stddev, r.x), ln_d(mean, stddev, r.y)};
}
__device__ __forceinline__ unsigned int rocrand_poisson(rocrand_state_mtgp32* state, double lambda)
{
    return rocrand_device::amd_detail::poisson_inv(state, lambda);
}
__device__ __forceinline__ unsigned int rocrand_discrete(rocrand_state_mtgp32*               state,
                                                         const rocrand_discrete_distribution discrete_distribution)
{
    using namespace rocrand_device::amd_detail;
    return discrete_alias(draw(rocrand(state)), *discrete_distribution);
}

#undef ROCRAND_AMD_BM_FLOAT
#undef ROCRAND_AMD_BM_DOUBLE


#endif // ROCRAND_KERNEL_H_
