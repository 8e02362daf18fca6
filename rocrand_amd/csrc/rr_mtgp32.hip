// rr_mtgp32.hip -- MTGP32 (Mersenne exponent 11213, 512 parameter sets) on gfx950.
//
// Output map (library/src/rng/mtgp32.hpp:232-305): generator b (one 256-thread block)
// advances its 351-word recurrence 256 words per block-step; thread t's word of a
// step feeds vec index b*256 + t + j*(512*256) on the block's j-th use, taking
// in_w consecutive steps per vec.  The head and tail of a misaligned or ragged
// call come from one extra use, by the threads whose index equals vec_n_up and
// vec_n_up + 1.  State per generator (rocrand_mtgp32.h:130-145, 272-317): a 1024-word
// ring, the recurrence parameters and the tempering table of its parameter set;
// every generator masks with the first set's mask (rocrand_mtgp32.h:381).
//
// The ring lives in LDS for the whole call, twice: word i at ring[i] and ring[i + 1024],
// so a step's reads at off + t + {0, 1, pos - 1, pos} (off < 1024, t < 256, pos <= 95)
// need no wrap-around masks and share one address (the two pairs differ by immediate
// offsets from it and from pos - 1).  Every write goes to both copies.  One barrier per
// block-step suffices: a step reads ring[off .. off+350] and writes ring[off+351 ..
// off+606] (indices mod 1024), so reads and writes of one step never alias, and the next
// step's reads see this step's writes after the barrier.
//
// The recurrence table of every parameter set is GF(2)-linear in its 4-bit index
// (tbl[i ^ j] = tbl[i] ^ tbl[j], tbl[0] = 0: MTGP is an F2-linear generator; the host
// checks it for all 512 sets when it builds the states), so its lookup, which sits on
// the step's critical path (ring reads -> new word -> ring write -> barrier), is the XOR
// of the entries at 1, 2, 4 and 8 selected by the index bits: about ten VALU operations
// on block-uniform operands instead of a dependent LDS round trip.  The tempering
// table's lookup only feeds the output; it stays an LDS read, issued as soon as the
// ring reads return so that its latency overlaps the recurrence.
#include "rr_generator.hpp"

#include <rocrand.h>

#include <algorithm>
#include <mutex>
#include <string.h>
#include <vector>

extern "C" const unsigned char rr_mtgp32_params_begin[];
extern "C" const unsigned char rr_mtgp32_params_end[];

namespace rr
{
namespace
{

constexpr unsigned MTGP_GENS  = 512; // config_defaults.hpp:45-50: 512 blocks
constexpr unsigned MTGP_BLOCK = 256;
constexpr unsigned MTGP_N     = 351; // 11213 / 32 + 1
constexpr unsigned MTGP_RING  = 1024;
constexpr unsigned MTGP_MASKI = MTGP_RING - 1;

// One generator's device state: ring, ring offset, parameters.
struct mtgp_state
{
    uint32_t status[MTGP_RING];
    uint32_t offset;
    uint32_t pos, sh1, sh2, mask;
    uint32_t tbl[16], tmp[16];
};

struct mtgp_param
{
    uint32_t pos, sh1, sh2, tbl[16], tmp[16], mask;
};

const mtgp_param* params()
{
    static_assert(sizeof(mtgp_param) == 36 * 4, "layout of mtgp32_11213.bin");
    return reinterpret_cast<const mtgp_param*>(rr_mtgp32_params_begin);
}

struct mtgp_call
{
    mtgp_state* states;
    void*       data;
    size_t      n, mis, vec_n;
    uint32_t    head, tail;
    dist_params params;
};

// tbl[k & 15] of a GF(2)-linear 16-entry table from its entries at 1, 2, 4 and 8
__host__ __device__ __forceinline__ uint32_t mtgp_lin16(uint32_t k, uint32_t e1, uint32_t e2, uint32_t e4,
                                                        uint32_t e8)
{
    return (e1 & (0u - (k & 1u))) ^ (e2 & (0u - ((k >> 1) & 1u))) ^ (e4 & (0u - ((k >> 2) & 1u)))
           ^ (e8 & (0u - ((k >> 3) & 1u)));
}

template<class D>
__global__ __launch_bounds__(MTGP_BLOCK) void mtgp_kernel(mtgp_call c)
{
    using out_t         = typename D::out_t;
    constexpr int IN    = D::in_w;
    constexpr int OUT   = D::out_w;
    constexpr int BYTES = OUT * (int)sizeof(out_t);
    static_assert(BYTES == 2 || BYTES == 4 || BYTES == 8 || BYTES == 16, "vec size");

    __shared__ uint32_t ring[2 * MTGP_RING];
    __shared__ uint32_t ttbl[16];

    const uint32_t b = blockIdx.x, t = threadIdx.x;
    mtgp_state*    g = c.states + b;
    for(uint32_t i = t; i < MTGP_RING; i += MTGP_BLOCK)
    {
        const uint32_t w   = g->status[i];
        ring[i]             = w;
        ring[i + MTGP_RING] = w;
    }
    const uint32_t pos = g->pos, sh1 = g->sh1, sh2 = g->sh2, mask = g->mask;
    if(t < 16)
        ttbl[t] = g->tmp[t];
    // Poisson: the block samples from its copy of the packed alias table in dynamic LDS
    // (a per-lane load from global memory would wait for the block's earlier stores)
    if constexpr(dist_kind<D>::value == out_kind::poisson)
    {
        extern __shared__ uint64_t lds_packed[];
        const uint32_t             size = c.params.huge ? 0u : c.params.pt.size;
        for(uint32_t i = t; i < size; i += MTGP_BLOCK)
            lds_packed[i] = c.params.pt.packed[i];
        c.params.pt.packed = lds_packed;
    }
    const uint32_t p1 = g->tbl[1], p2 = g->tbl[2], p4 = g->tbl[4], p8 = g->tbl[8];
    uint32_t       off = g->offset;
    __syncthreads();

    // one block-step: this thread's tempered word
    auto step = [&]() -> uint32_t {
        const uint32_t* rp = ring + t + off;
        uint32_t        X  = (rp[0] & mask) ^ rp[1];
        uint32_t        Y  = rp[pos];
        uint32_t        T  = rp[pos - 1];
        X ^= X << sh1;
        Y = X ^ (Y >> sh2);
        T ^= T >> 16;
        T ^= T >> 8;
        const uint32_t tt  = ttbl[T & 0x0f];
        const uint32_t r   = Y ^ mtgp_lin16(Y, p1, p2, p4, p8);
        uint32_t* const wp = ring + ((t + off + MTGP_N) & MTGP_MASKI);
        wp[0]              = r;
        wp[MTGP_RING]      = r;
        const uint32_t o   = r ^ tt;
        off = (off + MTGP_BLOCK) & MTGP_MASKI;
        __syncthreads();
        return o;
    };

    out_t* const data     = static_cast<out_t*>(c.data);
    out_t* const vec_data = data + c.mis;
    const size_t stride   = (size_t)MTGP_GENS * MTGP_BLOCK;
    const size_t rem      = c.vec_n % MTGP_BLOCK;
    const size_t down     = c.vec_n - rem;
    const size_t up       = rem == 0 ? down : down + MTGP_BLOCK;
    alignas(16) out_t o[OUT];
    uint32_t     in[IN];
    // the block's vec index before this thread's, block-uniform (a scalar loop counter and
    // store base): every thread takes part in every step
    const size_t   base0 = (size_t)b * MTGP_BLOCK;
    const uint32_t steps = base0 < down ? (uint32_t)((down - base0 + stride - 1) / stride) : 0u;
    out_t*         bptr  = vec_data + base0 * OUT;
    for(uint32_t k = 0; k < steps; k++, bptr += stride * OUT)
    {
#pragma unroll
        for(int i = 0; i < IN; i++)
            in[i] = step();
        D::apply(in, o, c.params);
        out_t* const dst = bptr + t * OUT;
        if constexpr(BYTES == 2)
            *reinterpret_cast<uint16_t*>(dst) = *reinterpret_cast<const uint16_t*>(o);
        else if constexpr(BYTES == 4)
            *reinterpret_cast<uint32_t*>(dst) = *reinterpret_cast<const uint32_t*>(o);
        else if constexpr(BYTES == 8)
            *reinterpret_cast<uint2*>(dst) = *reinterpret_cast<const uint2*>(o);
        else
            *reinterpret_cast<uint4*>(dst) = *reinterpret_cast<const uint4*>(o);
    }
    size_t index = base0 + (size_t)steps * stride + t;
    if(index - t < up)
    {
#pragma unroll
        for(int i = 0; i < IN; i++)
            in[i] = step();
        D::apply(in, o, c.params);
        if(index < c.vec_n)
            for(int j = 0; j < OUT; j++)
                vec_data[index * OUT + j] = o[j];
        index += stride;
    }
    if(OUT > 1 && (c.head > 0 || c.tail > 0))
    {
#pragma unroll
        for(int i = 0; i < IN; i++)
            in[i] = step();
        D::apply(in, o, c.params);
        if(index == up)
            for(uint32_t j = 0; j < (uint32_t)OUT; j++)
                if(j < c.head)
                    data[j] = o[j];
        if(index == up + 1)
            for(uint32_t j = 0; j < (uint32_t)OUT; j++)
                if(j < c.tail)
                    data[c.n - c.tail + j] = o[j];
    }

    for(uint32_t i = t; i < MTGP_RING; i += MTGP_BLOCK)
        g->status[i] = ring[i];
    if(t == 0)
        g->offset = off;
}

template<class D>
hipError_t launch_gen(const mtgp_call& c, hipStream_t s)
{
    const size_t lds = dist_kind<D>::value == out_kind::poisson && !c.params.huge
                           ? c.params.pt.size * sizeof(uint64_t)
                           : 0;
    hipLaunchKernelGGL((mtgp_kernel<D>), dim3(MTGP_GENS), dim3(MTGP_BLOCK), lds, s, c);
    return hipGetLastError();
}

hipError_t dispatch(out_kind o, const mtgp_call& c, hipStream_t s)
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

// mtgp32.hpp:308-617
class mtgp32_generator final : public rocrand_generator_base_type
{
public:
    mtgp32_generator() : rocrand_generator_base_type(ROCRAND_RNG_PSEUDO_MTGP32) {}
    ~mtgp32_generator() override
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

    rocrand_status init() override
    {
        if(m_initialized)
            return ROCRAND_STATUS_SUCCESS;
        if(!m_states && hipMalloc(&m_states, sizeof(mtgp_state) * MTGP_GENS) != hipSuccess)
        {
            m_states = nullptr;
            return ROCRAND_STATUS_ALLOCATION_FAILED;
        }
        // rocrand_make_state_mtgp32 (rocrand_mtgp32.h:362-390) and
        // rocrand_mtgp32_init_state (:130-145)
        std::vector<mtgp_state> h(MTGP_GENS);
        const mtgp_param*       p    = params();
        const uint64_t          s64  = m_seed ^ (m_seed >> 32);
        for(unsigned i = 0; i < MTGP_GENS; i++)
        {
            // the kernel evaluates the recurrence table from its entries at 1, 2, 4 and 8 (the
            // tempering table is linear too and checked with it), and its
            // unmasked ring reads reach off + 255 + pos with off < 1024
            if(p[i].pos == 0 || p[i].pos + MTGP_BLOCK > MTGP_N)
                return ROCRAND_STATUS_INTERNAL_ERROR;
            for(uint32_t k = 0; k < 16; k++)
                if(p[i].tbl[k] != mtgp_lin16(k, p[i].tbl[1], p[i].tbl[2], p[i].tbl[4], p[i].tbl[8])
                   || p[i].tmp[k] != mtgp_lin16(k, p[i].tmp[1], p[i].tmp[2], p[i].tmp[4], p[i].tmp[8]))
                    return ROCRAND_STATUS_INTERNAL_ERROR;
            mtgp_state& g = h[i];
            memset(&g, 0, sizeof(g));
            const uint32_t hidden = p[i].tbl[4] ^ (p[i].tbl[8] << 16);
            uint32_t       tmp    = hidden;
            tmp += tmp >> 16;
            tmp += tmp >> 8;
            memset(g.status, (int)(tmp & 0xff), sizeof(uint32_t) * MTGP_N);
            g.status[0] = (uint32_t)s64 + i + 1;
            g.status[1] = hidden;
            for(unsigned k = 1; k < MTGP_N; k++)
                g.status[k] ^= 1812433253u * (g.status[k - 1] ^ (g.status[k - 1] >> 30)) + k;
            g.offset = 0;
            g.pos    = p[i].pos;
            g.sh1    = p[i].sh1;
            g.sh2    = p[i].sh2;
            g.mask   = p[0].mask;
            memcpy(g.tbl, p[i].tbl, sizeof(g.tbl));
            memcpy(g.tmp, p[i].tmp, sizeof(g.tmp));
        }
        if(hipMemcpyAsync(m_states, h.data(), sizeof(mtgp_state) * MTGP_GENS, hipMemcpyHostToDevice, stream)
               != hipSuccess
           || hipStreamSynchronize(stream) != hipSuccess)
            return ROCRAND_STATUS_ALLOCATION_FAILED;
        m_initialized = true;
        return ROCRAND_STATUS_SUCCESS;
    }

    rocrand_status shape(out_kind k, dist_shape& s) const override
    {
        switch(k)
        {
#define RR_SHAPE(K)                                                                                    \
    case K:                                                                                            \
        s = dist_shape{dist<word_family::w32, K>::in_w, dist<word_family::w32, K>::out_w,              \
                       (int)sizeof(typename dist<word_family::w32, K>::out_t)};                        \
        return ROCRAND_STATUS_SUCCESS;
            RR_SHAPE(out_kind::u8)
            RR_SHAPE(out_kind::u16)
            RR_SHAPE(out_kind::u32)
            RR_SHAPE(out_kind::uniform_f32)
            RR_SHAPE(out_kind::uniform_f64)
            RR_SHAPE(out_kind::uniform_f16)
            RR_SHAPE(out_kind::normal_f32)
            RR_SHAPE(out_kind::normal_f64)
            RR_SHAPE(out_kind::normal_f16)
            RR_SHAPE(out_kind::lognormal_f32)
            RR_SHAPE(out_kind::lognormal_f64)
            RR_SHAPE(out_kind::lognormal_f16)
            RR_SHAPE(out_kind::poisson)
#undef RR_SHAPE
            default: return ROCRAND_STATUS_TYPE_ERROR;
        }
    }

protected:
    void reset() override { m_initialized = false; }

    rocrand_status generate_device(out_kind k, void* data, size_t n, const gen_args& args) override
    {
        dist_shape     s;
        rocrand_status st = shape(k, s);
        if(st != ROCRAND_STATUS_SUCCESS)
            return st;
        mtgp_call c{};
        if((st = make_params(k, args, c.params)) != ROCRAND_STATUS_SUCCESS)
            return st;
        if((st = init()) != ROCRAND_STATUS_SUCCESS)
            return st;
        if(data == nullptr)
            return ROCRAND_STATUS_SUCCESS;
        const uintptr_t addr = reinterpret_cast<uintptr_t>(data);
        c.mis    = (s.out_w - (addr / s.esz) % s.out_w) % s.out_w;
        c.head   = (uint32_t)std::min<size_t>(n, c.mis);
        c.tail   = (uint32_t)((n - c.head) % s.out_w);
        c.vec_n  = (n - c.head) / s.out_w;
        c.states = static_cast<mtgp_state*>(m_states);
        c.data   = data;
        c.n      = n;
        if(dispatch(k, c, stream) != hipSuccess)
            return ROCRAND_STATUS_LAUNCH_FAILURE;
        return ROCRAND_STATUS_SUCCESS;
    }

private:
    unsigned long long m_seed        = 0;
    bool               m_initialized = false;
    void*              m_states      = nullptr;
};

} // namespace

rocrand_generator make_mtgp32_generator() { return new mtgp32_generator(); }

} // namespace rr
