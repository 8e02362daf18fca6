// rr_mt19937.hip -- MT19937 (8192 generators, 2^1000 words apart) on gfx950.
//
// Streams (library/src/rng/mt19937.hpp, mt19937_octo_engine.hpp):
//   * generator g starts from init_genrand((seed >> 32) ^ seed) jumped by g * 2^1000 words,
//     as two radix-256 digits, with the polynomials x^(i*2^1000) and x^(256*i*2^1000)
//     mod phi (jump kernel :86-274);
//   * before each round every generator twists its 624 words (gen_next_n :259-405); input
//     number q of a round (q < 65536 * (78 / in_w)) comes from thread t = q mod 65536
//     (generator t / 8, lane t mod 8), local words (q / 65536) * in_w + i, where lane l's
//     local word 0 is element {0,113,170,283,340,397,510,567}[l] and local word 1 + 7a + b
//     is element {1,57,114,171,227,284,341,398,454,511,568}[a] + 7l + b (:84-166);
//   * a call consumes vec_n inputs (+1 when there is a head or a tail) from the running
//     position; the extra input fills the head (data[o] = out[o]) and, in reverse, the
//     tail (data[n - 1 - k] = out[out_w - tail + k]) (:355-431, :604-629).
//
// Kernels (own schedule; only the map above is fixed):
//   * jump: one 256-thread block per generator evaluates p(A) s by Horner's rule 32
//     coefficients at a time: the accumulator window advances 32 words (32 independent
//     recurrence steps, since a step reads 397 words ahead) and then gains
//     XOR_{set k} A^k s, whose word i is S[i + k] for the start state's sequence S
//     extended by 31 words.  Bit-identical to the one-coefficient-per-step evaluation.
//   * generate: one block per 8 generators keeps their 8 x 624 words in LDS (element-major)
//     for the whole call, in two buffers: a round twists the current words into the other
//     buffer, each thread following chains of words 227 apart (a word's new value needs
//     the new value 227 words before it, kept in a register), one barrier per round;
//     each wavefront then writes 64 consecutive inputs (8 generators x 8 lanes) per local
//     word while the next round's twist may already write the first buffer.
#include "rr_generator.hpp"

#include <rocrand.h>

#include <algorithm>
#include <mutex>
#include <vector>

extern "C" const unsigned char rr_mt19937_jumps_begin[];
extern "C" const unsigned char rr_mt19937_jumps_end[];

namespace rr
{
namespace
{

constexpr unsigned MT_N      = 624;
constexpr unsigned MT_M      = 397;
constexpr unsigned MT_GENS   = 8192; // config_defaults.hpp:76-80: 256 x 256 threads / 8
constexpr unsigned MT_STRIDE = 65536;
constexpr unsigned MT_LOCAL  = MT_N / 8; // 78 words per lane
constexpr unsigned GPB       = 8;        // generators per generate block

__device__ __forceinline__ uint32_t mt_comp(uint32_t a, uint32_t b, uint32_t m)
{
    const uint32_t y = (a & 0x80000000u) | (b & 0x7fffffffu);
    return m ^ (y >> 1) ^ ((0u - (y & 1u)) & 0x9908b0dfu);
}

__device__ __forceinline__ uint32_t mt_temper(uint32_t y)
{
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

__device__ __forceinline__ unsigned wrap(unsigned i) { return i >= MT_N ? i - MT_N : i; }

__global__ __launch_bounds__(256) void mt_jump_kernel(uint32_t* states, uint32_t seedu, const uint32_t* jumps)
{
    __shared__ uint32_t S[MT_N + 32];
    __shared__ uint32_t ring[MT_N];
    __shared__ uint32_t cur[MT_N];
    const unsigned g = blockIdx.x, t = threadIdx.x;
    if(t == 0)
    {
        uint32_t x = seedu;
        cur[0]     = x;
        for(unsigned i = 1; i < MT_N; i++)
        {
            x      = 1812433253u * (x ^ (x >> 30)) + i;
            cur[i] = x;
        }
    }
    __syncthreads();
    unsigned e = g;
    for(unsigned r = 0; r < 2; r++)
    {
        const unsigned digit = e & 255u;
        e >>= 8;
        if(digit == 0)
            continue;
        const uint32_t* pf = jumps + (size_t)(r * 255 + digit - 1) * MT_N;
        for(unsigned i = t; i < MT_N; i += 256)
        {
            S[i]    = cur[i];
            ring[i] = 0;
        }
        __syncthreads();
        if(t < 31)
            S[MT_N + t] = mt_comp(S[t], S[t + 1], S[t + MT_M]);
        __syncthreads();
        unsigned ptr = 0;
        for(int w = MT_N - 1; w >= 0; w--)
        {
            // coefficients 32w .. 32w+31; the top word holds only x^19936
            const unsigned b    = (w == MT_N - 1) ? 1u : 32u;
            const uint32_t bits = pf[w] & (b == 32u ? 0xffffffffu : 1u);
            uint32_t       nw   = 0;
            if(t < b)
                nw = mt_comp(ring[wrap(ptr + t)], ring[wrap(ptr + t + 1)], ring[wrap(ptr + t + MT_M)]);
            __syncthreads();
            if(t < b)
                ring[wrap(ptr + t)] = nw;
            ptr = wrap(ptr + b);
            __syncthreads();
            if(bits)
            {
                for(unsigned i = t; i < MT_N; i += 256)
                {
                    uint32_t acc = 0, m = bits;
                    while(m)
                    {
                        const unsigned k = __builtin_ctz(m);
                        m &= m - 1;
                        acc ^= S[i + k];
                    }
                    ring[wrap(ptr + i)] ^= acc;
                }
                __syncthreads();
            }
        }
        for(unsigned i = t; i < MT_N; i += 256)
            cur[i] = ring[wrap(ptr + i)];
        __syncthreads();
    }
    for(unsigned i = t; i < MT_N; i += 256)
        states[(size_t)g * MT_N + i] = cur[i];
}

struct mt_call
{
    uint32_t*   states; // [MT_GENS][MT_N], words of the current round
    void*       data;
    size_t      n, mis, vec_n;
    uint32_t    head, tail, extra;
    uint32_t    start; // running input position within the current round (0: used up)
    dist_params params;
};

// State element read by lane `lane` as its local word `local` (mt19937_octo_engine.hpp:84-166).
__constant__ unsigned mt_special[8]  = {0, 113, 170, 283, 340, 397, 510, 567};
__constant__ unsigned mt_offsets[11] = {1, 57, 114, 171, 227, 284, 341, 398, 454, 511, 568};

// `local` is wavefront-uniform in the generate kernel (scalar division and table load);
// `special` = mt_special[lane], loaded once per thread.
__device__ __forceinline__ unsigned mt_element(unsigned lane, unsigned special, unsigned local)
{
    if(local == 0)
        return special;
    const unsigned a = (local - 1) / 7, b = (local - 1) - 7 * a;
    return mt_offsets[a] + 7 * lane + b;
}

// LDS layout of the block's generators: element-major, word k of generator g at
// st[k * GPB + g].  A wavefront's output reads (8 generators x 8 lanes, lanes 7 words
// apart) then fall on 64 distinct banks, and the twist's reads are consecutive.
__device__ __forceinline__ unsigned lds_at(unsigned k, unsigned g) { return k * GPB + g; }

// One round of the recurrence on GPB generators: the round's words in `src`, the next
// round's written to `dst` (both element-major, so item it = 8k + g is word k of
// generator g).  Word k of the new round is x'[k] = f(x[k], x[k + 1], m) with
// m = x[k + 397] for k < 227 and m = x'[k - 227] above, and x[624] read as x'[0].
// A thread takes whole chains k0, k0 + 227, k0 + 454 (k0 < 227) of one generator: each
// link's m is the previous link's result, kept in a register, and every other operand
// is an old word, so no thread waits for another's writes; the one exception, x'[0] for
// k = 623, the thread of chain 169 computes itself from old words.  Nothing in `src`
// is written, so a single barrier -- the new words visible to the output reads -- ends
// the round; `src` is next written by the following round's twist, after every thread
// has passed this barrier and so finished its reads of it.
__device__ __forceinline__ void twist_block(const uint32_t* src, uint32_t* dst, unsigned t)
{
    constexpr unsigned W = GPB * MT_N, D = GPB * 227, J = GPB * MT_M;
    for(unsigned it = t; it < D; it += 256)
    {
        const uint32_t a = mt_comp(src[it], src[it + GPB], src[it + J]);
        dst[it]          = a;
        const uint32_t b = mt_comp(src[it + D], src[it + D + GPB], a);
        dst[it + D]      = b;
        const unsigned k3 = it + 2 * D;
        if(k3 < W)
        {
            uint32_t nx;
            if(k3 + GPB < W)
                nx = src[k3 + GPB];
            else
            {
                const unsigned g0 = k3 + GPB - W; // word 0 of this generator
                nx                = mt_comp(src[g0], src[g0 + GPB], src[g0 + J]);
            }
            dst[k3] = mt_comp(src[k3], nx, b);
        }
    }
    __syncthreads();
}

// 4 blocks per CU (the LDS of 2 x 8 generators each fills it): at most 128 VGPRs
template<class D>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void mt_generate_kernel(mt_call c)
{
    using out_t          = typename D::out_t;
    constexpr int IN     = D::in_w;
    constexpr int OUT    = D::out_w;
    constexpr int BYTES  = OUT * (int)sizeof(out_t);
    constexpr unsigned IPS  = MT_LOCAL / IN;
    constexpr uint64_t FULL = (uint64_t)MT_STRIDE * IPS;
    static_assert(BYTES == 4 || BYTES == 8 || BYTES == 16, "vec size");

    // the round's words and the next round's (twist_block), 2 x 39 KiB: 4 blocks per CU
    __shared__ uint32_t buf[2][GPB * MT_N];
    const unsigned t = threadIdx.x, blk = blockIdx.x;
    uint32_t*      gstate = c.states + (size_t)blk * GPB * MT_N;
    uint32_t*      st     = buf[0];
    uint32_t*      nx     = buf[1];
    for(unsigned i = t; i < GPB * MT_N; i += 256)
        st[lds_at(i % MT_N, i / MT_N)] = gstate[i];
    __syncthreads();

    // the wavefront index as a scalar: the local word j, its element offsets and the
    // wavefront's input range below are then uniform (scalar loads and branches)
    const unsigned p = t & 63u, quarter = __builtin_amdgcn_readfirstlane(t >> 6);
    const unsigned gl = p >> 3, lane = p & 7u;
    const uint64_t total = c.vec_n + c.extra;
    const uint64_t qend  = (uint64_t)c.start + total;
    out_t* const   data  = static_cast<out_t*>(c.data);
    out_t* const   vec_data = data + c.mis;
    const unsigned special = mt_special[lane];
    alignas(16) out_t o[OUT];
    uint32_t     in[IN];

    // LDS indices of this thread's inputs, the same every round: local word
    // j = quarter + 4m (m < MW) reads eoff[m][i] for input i (held in registers: every
    // loop over m below is unrolled)
    constexpr unsigned MW = (IPS + 3) / 4;
    unsigned           eoff[MW][IN];
#pragma unroll
    for(unsigned m = 0; m < MW; m++)
#pragma unroll
        for(int i = 0; i < IN; i++)
        {
            const unsigned j = quarter + 4 * m;
            eoff[m][i]       = j < IPS ? lds_at(mt_element(lane, special, j * IN + i), gl) : 0;
        }

    for(uint64_t R = 0; R * FULL < qend; R++)
    {
        if(R > 0 || c.start == 0)
        {
            twist_block(st, nx, t);
            uint32_t* const w = st;
            st                = nx;
            nx                = w;
        }
        const uint64_t R0 = R * FULL;
        if(R0 >= c.start && R0 + FULL <= c.start + c.vec_n)
        {
            // every input of the round is a whole vec: no range checks.  Lane p is the
            // reference's thread blk * 64 + p; its local word j is input
            // R0 + j * 65536 + blk * 64 + p.
            // Up to 8-byte vecs all of the thread's words are read first (an invalid last j
            // reads word 0 and is not stored), so the LDS reads overlap one another; the
            // double-pair transforms need those registers themselves.
            constexpr bool BATCH = BYTES <= 8;
            out_t* const   rbase = vec_data + (R0 - c.start + blk * 64u + p) * OUT;
            uint32_t       w[BATCH ? MW : 1][IN];
            if constexpr(BATCH)
            {
#pragma unroll
                for(unsigned m = 0; m < MW; m++)
#pragma unroll
                    for(int i = 0; i < IN; i++)
                        w[m][i] = st[eoff[m][i]];
            }
#pragma unroll
            for(unsigned m = 0; m < MW; m++)
            {
                const unsigned j = quarter + 4 * m;
                if(j < IPS)
                {
#pragma unroll
                    for(int i = 0; i < IN; i++)
                        in[i] = mt_temper(BATCH ? w[BATCH ? m : 0][i] : st[eoff[m][i]]);
                    D::apply(in, o, c.params);
                    out_t* dst = rbase + (size_t)j * MT_STRIDE * OUT;
                    if constexpr(BYTES == 4)
                        *reinterpret_cast<uint32_t*>(dst) = *reinterpret_cast<const uint32_t*>(o);
                    else if constexpr(BYTES == 8)
                        *reinterpret_cast<uint2*>(dst) = *reinterpret_cast<const uint2*>(o);
                    else
                        *reinterpret_cast<uint4*>(dst) = *reinterpret_cast<const uint4*>(o);
                }
            }
            continue;
        }
        for(unsigned j = quarter; j < IPS; j += 4)
        {
            const uint64_t Q = R0 + (uint64_t)j * MT_STRIDE + blk * 64u + p;
            if(Q < c.start || Q >= qend)
                continue;
            const uint64_t v = Q - c.start;
#pragma unroll
            for(int i = 0; i < IN; i++)
                in[i] = mt_temper(st[lds_at(mt_element(lane, special, j * IN + i), gl)]);
            D::apply(in, o, c.params);
            if(v < c.vec_n)
            {
                out_t* dst = vec_data + v * OUT;
                if constexpr(BYTES == 4)
                    *reinterpret_cast<uint32_t*>(dst) = *reinterpret_cast<const uint32_t*>(o);
                else if constexpr(BYTES == 8)
                    *reinterpret_cast<uint2*>(dst) = *reinterpret_cast<const uint2*>(o);
                else
                    *reinterpret_cast<uint4*>(dst) = *reinterpret_cast<const uint4*>(o);
            }
            else
            {
                for(uint32_t k = 0; k < (uint32_t)OUT; k++)
                {
                    if(k < c.head)
                        data[k] = o[k];
                    if(k + c.tail >= (uint32_t)OUT)
                        data[c.n + (OUT - c.tail - 1) - k] = o[k];
                }
            }
        }
        // no barrier: the next round twists into the buffer these reads do not touch
    }
    for(unsigned i = t; i < GPB * MT_N; i += 256)
        gstate[i] = st[lds_at(i % MT_N, i / MT_N)];
}

template<class D>
hipError_t launch_gen(const mt_call& c, hipStream_t s)
{
    hipLaunchKernelGGL((mt_generate_kernel<D>), dim3(MT_GENS / GPB), dim3(256), 0, s, c);
    return hipGetLastError();
}

hipError_t dispatch(out_kind o, const mt_call& c, hipStream_t s)
{
    constexpr word_family F = word_family::w32;
    switch(o)
    {
        case out_kind::u8: return launch_gen<dist<F, out_kind::u8>>(c, s);
        case out_kind::u16: return launch_gen<dist<F, out_kind::u16>>(c, s);
        case out_kind::u32: return launch_gen<dist<F, out_kind::u32>>(c, s);
        case out_kind::uniform_f32: return launch_gen<dist<F, out_kind::uniform_f32>>(c, s);
        case out_kind::uniform_f64: return launch_gen<dist<F, out_kind::uniform_f64>>(c, s);
        case out_kind::uniform_f16: return launch_gen<dist<F, out_kind::uniform_f16>>(c, s);
        case out_kind::normal_f32: return launch_gen<dist<F, out_kind::normal_f32>>(c, s);
        case out_kind::normal_f64: return launch_gen<dist<F, out_kind::normal_f64>>(c, s);
        case out_kind::normal_f16: return launch_gen<dist<F, out_kind::normal_f16>>(c, s);
        case out_kind::lognormal_f32: return launch_gen<dist<F, out_kind::lognormal_f32>>(c, s);
        case out_kind::lognormal_f64: return launch_gen<dist<F, out_kind::lognormal_f64>>(c, s);
        case out_kind::lognormal_f16: return launch_gen<dist<F, out_kind::lognormal_f16>>(c, s);
        case out_kind::poisson: return launch_gen<dist<F, out_kind::poisson>>(c, s);
        case out_kind::u64: return hipErrorInvalidValue;
    }
    return hipErrorInvalidValue;
}

// The jump polynomials of the current device, uploaded once.
hipError_t device_jumps(const uint32_t** out)
{
    static std::mutex                                  mtx;
    static std::vector<std::pair<int, const uint32_t*>> copies;
    int                                                dev = 0;
    hipError_t                                         e   = hipGetDevice(&dev);
    if(e != hipSuccess)
        return e;
    std::lock_guard<std::mutex> lock(mtx);
    for(const auto& c : copies)
        if(c.first == dev)
        {
            *out = c.second;
            return hipSuccess;
        }
    const size_t bytes = (size_t)(rr_mt19937_jumps_end - rr_mt19937_jumps_begin);
    if(bytes != (size_t)510 * MT_N * 4)
        return hipErrorInvalidImage;
    void* p = nullptr;
    if((e = hipMalloc(&p, bytes)) != hipSuccess)
        return e;
    if((e = hipMemcpy(p, rr_mt19937_jumps_begin, bytes, hipMemcpyHostToDevice)) != hipSuccess)
        return e;
    copies.emplace_back(dev, static_cast<const uint32_t*>(p));
    *out = static_cast<const uint32_t*>(p);
    return hipSuccess;
}

template<word_family F, out_kind K>
constexpr dist_shape mt_shape()
{
    return dist_shape{dist<F, K>::in_w, dist<F, K>::out_w, (int)sizeof(typename dist<F, K>::out_t)};
}

// mt19937.hpp:665-1122
class mt19937_generator final : public rocrand_generator_base_type
{
public:
    mt19937_generator() : rocrand_generator_base_type(ROCRAND_RNG_PSEUDO_MT19937) {}
    ~mt19937_generator() override
    {
        if(m_states)
            (void)hipFree(m_states);
    }

    void set_seed(unsigned long long seed) override
    {
        m_seed = seed;
        reset();
    }

    rocrand_status set_offset(unsigned long long) override { return ROCRAND_STATUS_TYPE_ERROR; }

    rocrand_status set_order(rocrand_ordering o) override
    {
        // mt19937.hpp:765-779: DEFAULT, BEST and LEGACY only
        if(o != ROCRAND_ORDERING_PSEUDO_DEFAULT && o != ROCRAND_ORDERING_PSEUDO_BEST
           && o != ROCRAND_ORDERING_PSEUDO_LEGACY)
            return ROCRAND_STATUS_OUT_OF_RANGE;
        order = o;
        reset();
        return ROCRAND_STATUS_SUCCESS;
    }

    rocrand_status init() override
    {
        if(m_initialized)
            return ROCRAND_STATUS_SUCCESS;
        if(!m_states && hipMalloc(&m_states, sizeof(uint32_t) * MT_GENS * MT_N) != hipSuccess)
        {
            m_states = nullptr;
            return ROCRAND_STATUS_ALLOCATION_FAILED;
        }
        const uint32_t* jumps = nullptr;
        if(device_jumps(&jumps) != hipSuccess)
            return ROCRAND_STATUS_ALLOCATION_FAILED;
        const uint32_t seedu = (uint32_t)((m_seed >> 32) ^ m_seed);
        hipLaunchKernelGGL(mt_jump_kernel, dim3(MT_GENS), dim3(256), 0, stream, static_cast<uint32_t*>(m_states),
                           seedu, jumps);
        if(hipGetLastError() != hipSuccess)
            return ROCRAND_STATUS_LAUNCH_FAILURE;
        m_start       = 0;
        m_prev_in_w   = 0;
        m_initialized = true;
        return ROCRAND_STATUS_SUCCESS;
    }

    rocrand_status shape(out_kind k, dist_shape& s) const override
    {
        constexpr word_family F = word_family::w32;
        switch(k)
        {
            case out_kind::u8: s = mt_shape<F, out_kind::u8>(); break;
            case out_kind::u16: s = mt_shape<F, out_kind::u16>(); break;
            case out_kind::u32: s = mt_shape<F, out_kind::u32>(); break;
            case out_kind::uniform_f32: s = mt_shape<F, out_kind::uniform_f32>(); break;
            case out_kind::uniform_f64: s = mt_shape<F, out_kind::uniform_f64>(); break;
            case out_kind::uniform_f16: s = mt_shape<F, out_kind::uniform_f16>(); break;
            case out_kind::normal_f32: s = mt_shape<F, out_kind::normal_f32>(); break;
            case out_kind::normal_f64: s = mt_shape<F, out_kind::normal_f64>(); break;
            case out_kind::normal_f16: s = mt_shape<F, out_kind::normal_f16>(); break;
            case out_kind::lognormal_f32: s = mt_shape<F, out_kind::lognormal_f32>(); break;
            case out_kind::lognormal_f64: s = mt_shape<F, out_kind::lognormal_f64>(); break;
            case out_kind::lognormal_f16: s = mt_shape<F, out_kind::lognormal_f16>(); break;
            case out_kind::poisson: s = mt_shape<F, out_kind::poisson>(); break;
            default: return ROCRAND_STATUS_TYPE_ERROR;
        }
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
        mt_call c{};
        if((st = make_params(k, args, c.params)) != ROCRAND_STATUS_SUCCESS)
            return st;
        if((st = init()) != ROCRAND_STATUS_SUCCESS)
            return st;
        if(data == nullptr)
            return ROCRAND_STATUS_SUCCESS;
        const unsigned in_w = (unsigned)s.in_w, out_w = (unsigned)s.out_w;
        const unsigned full = MT_STRIDE * (MT_LOCAL / in_w);
        const uintptr_t addr = reinterpret_cast<uintptr_t>(data);
        c.mis   = (out_w - (addr / s.esz) % out_w) % out_w;
        c.head  = (uint32_t)std::min<size_t>(n, c.mis);
        c.tail  = (uint32_t)((n - c.head) % out_w);
        c.vec_n = (n - c.head) / out_w;
        c.extra = (c.head > 0 || c.tail > 0) ? 1u : 0u;
        // mt19937.hpp:951-964: a new input width starts at the next whole unused input
        if(m_prev_in_w != in_w && m_start > 0)
        {
            const unsigned used = (m_start + MT_STRIDE - 1) / MT_STRIDE * m_prev_in_w;
            m_start             = (used + in_w - 1) / in_w * MT_STRIDE;
            if(m_start >= full)
                m_start = 0;
        }
        c.states = static_cast<uint32_t*>(m_states);
        c.data   = data;
        c.n      = n;
        c.start  = m_start;
        if(dispatch(k, c, stream) != hipSuccess)
            return ROCRAND_STATUS_LAUNCH_FAILURE;
        m_start     = (uint32_t)(((uint64_t)m_start + c.vec_n + c.extra) % full);
        m_prev_in_w = in_w;
        return ROCRAND_STATUS_SUCCESS;
    }

private:
    unsigned long long m_seed        = 0;
    bool               m_initialized = false;
    void*              m_states      = nullptr;
    uint32_t           m_start       = 0;
    uint32_t           m_prev_in_w   = 0;
};

} // namespace

rocrand_generator make_mt19937_generator() { return new mt19937_generator(); }

} // namespace rr
