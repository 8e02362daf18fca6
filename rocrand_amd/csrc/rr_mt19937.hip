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
//     for the whole call; a round twists them in three barrier-separated phases ([0,227)
//     reads only old words, [227,454) and [454,624) read the words 227 before them, already
//     new); each wavefront then writes 64 consecutive inputs (8 generators x 8 lanes) per
//     local word.
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

__device__ __forceinline__ unsigned mt_element(unsigned lane, unsigned local)
{
    if(local == 0)
        return mt_special[lane];
    const unsigned a = (local - 1) / 7, b = (local - 1) - 7 * a;
    return mt_offsets[a] + 7 * lane + b;
}

// LDS layout of the block's generators: element-major, word k of generator g at
// st[k * GPB + g].  A wavefront's output reads (8 generators x 8 lanes, lanes 7 words
// apart) then fall on 64 distinct banks, and the twist's reads are consecutive.
__device__ __forceinline__ unsigned lds_at(unsigned k, unsigned g) { return k * GPB + g; }

// One round of the recurrence on GPB generators held in LDS.
__device__ __forceinline__ void twist_block(uint32_t* st, unsigned t)
{
    constexpr unsigned PER = (GPB * 227 + 255) / 256; // items per thread per phase
    uint32_t           v[PER];
    // phase 1: k in [0, 227): old words only
    // phase 2: k in [227, 454): x[k - 227] is new
    // phase 3: k in [454, 624): x[k - 227] new, and k = 623 reads the new x[0]
    const unsigned lo[3] = {0, 227, 454}, len[3] = {227, 227, 170};
#pragma unroll
    for(int ph = 0; ph < 3; ph++)
    {
        const unsigned items = GPB * len[ph];
#pragma unroll
        for(unsigned r = 0; r < PER; r++)
        {
            const unsigned it = t + 256 * r;
            if(it < items)
            {
                const unsigned g = it % GPB, k = lo[ph] + it / GPB;
                v[r] = mt_comp(st[lds_at(k, g)], st[lds_at(wrap(k + 1), g)], st[lds_at(wrap(k + MT_M), g)]);
            }
        }
        __syncthreads();
#pragma unroll
        for(unsigned r = 0; r < PER; r++)
        {
            const unsigned it = t + 256 * r;
            if(it < items)
            {
                const unsigned g = it % GPB, k = lo[ph] + it / GPB;
                st[lds_at(k, g)] = v[r];
            }
        }
        __syncthreads();
    }
}

template<class D>
__global__ __launch_bounds__(256) void mt_generate_kernel(mt_call c)
{
    using out_t          = typename D::out_t;
    constexpr int IN     = D::in_w;
    constexpr int OUT    = D::out_w;
    constexpr int BYTES  = OUT * (int)sizeof(out_t);
    constexpr unsigned IPS  = MT_LOCAL / IN;
    constexpr uint64_t FULL = (uint64_t)MT_STRIDE * IPS;
    static_assert(BYTES == 4 || BYTES == 8 || BYTES == 16, "vec size");

    __shared__ uint32_t st[GPB * MT_N];
    const unsigned t = threadIdx.x, blk = blockIdx.x;
    uint32_t*      gstate = c.states + (size_t)blk * GPB * MT_N;
    for(unsigned i = t; i < GPB * MT_N; i += 256)
        st[lds_at(i % MT_N, i / MT_N)] = gstate[i];
    __syncthreads();

    const unsigned p = t & 63u, quarter = t >> 6;
    const unsigned gl = p >> 3, lane = p & 7u;
    const unsigned tid = blk * 64u + p; // the reference's thread index
    const uint64_t total = c.vec_n + c.extra;
    const uint64_t qend  = (uint64_t)c.start + total;
    out_t* const   data  = static_cast<out_t*>(c.data);
    out_t* const   vec_data = data + c.mis;
    alignas(16) out_t o[OUT];
    uint32_t     in[IN];

    for(uint64_t R = 0; R * FULL < qend; R++)
    {
        if(R > 0 || c.start == 0)
            twist_block(st, t);
        for(unsigned j = quarter; j < IPS; j += 4)
        {
            const uint64_t Q = R * FULL + (uint64_t)j * MT_STRIDE + tid;
            if(Q < c.start || Q >= qend)
                continue;
            const uint64_t v = Q - c.start;
#pragma unroll
            for(int i = 0; i < IN; i++)
                in[i] = mt_temper(st[lds_at(mt_element(lane, j * IN + i), gl)]);
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
        __syncthreads();
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
