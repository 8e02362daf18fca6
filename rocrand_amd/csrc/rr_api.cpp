// rr_api.cpp -- generator objects and the extern "C" rocrand_* entry points.
//
// Every entry point keeps the reference's argument meaning and status codes
// (library/src/rocrand.cpp:186-630, generator templates in library/src/rng/).
#include "rr_dynamic.hpp"
#include "rr_generator.hpp"

#include <rocrand.h>

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

using namespace rr;

// ============================================================================
// Poisson tables
// ============================================================================
namespace rr
{

// poisson.hpp:110-161: pmf around floor(lambda), truncated where p < 1e-12.
std::vector<double> poisson_pmf(double lambda, unsigned int& size, unsigned int& offset)
{
    const size_t        capacity   = 2 * static_cast<size_t>(16.0 * (2.0 + std::sqrt(lambda)));
    std::vector<double> p(capacity);
    const double        p_epsilon  = 1e-12;
    const double        log_lambda = std::log(lambda);
    const int           left       = static_cast<int>(std::floor(lambda)) - static_cast<int>(capacity / 2);
    int                 lo         = 0;
    for(int i = static_cast<int>(capacity / 2); i >= 0; i--)
    {
        const double x  = left + i;
        const double pp = std::exp(x * log_lambda - std::lgamma(x + 1.0) - lambda);
        if(pp < p_epsilon)
        {
            lo = i + 1;
            break;
        }
        p[i] = pp;
    }
    int hi = static_cast<int>(capacity) - 1;
    for(int i = static_cast<int>(capacity / 2) + 1; i < static_cast<int>(capacity); i++)
    {
        const double x  = left + i;
        const double pp = std::exp(x * log_lambda - std::lgamma(x + 1.0) - lambda);
        if(pp < p_epsilon)
        {
            hi = i - 1;
            break;
        }
        p[i] = pp;
    }
    for(int i = lo; i <= hi; i++)
        p[i - lo] = p[i];
    size   = static_cast<unsigned int>(hi - lo + 1);
    offset = static_cast<unsigned int>(left + lo);
    return p;
}

// discrete.hpp:189-265: normalize, then Vose's alias method with LIFO work lists.
void alias_table_from(const std::vector<double>& p_in, unsigned int size, std::vector<double>& prob,
                      std::vector<unsigned int>& alias, std::vector<double>* cdf)
{
    std::vector<double> p(p_in.begin(), p_in.begin() + size);
    double              sum = 0.0;
    for(unsigned int i = 0; i < size; i++)
        sum += p[i];
    for(unsigned int i = 0; i < size; i++)
        p[i] /= sum;
    if(cdf)
    {
        cdf->assign(size, 0.0);
        double acc = 0.0;
        for(unsigned int i = 0; i < size; i++)
        {
            acc += p[i];
            (*cdf)[i] = acc;
        }
    }
    prob.assign(size, 0.0);
    alias.assign(size, 0u);
    const double              average = 1.0 / size;
    std::vector<unsigned int> small, large;
    small.reserve(size);
    large.reserve(size);
    for(unsigned int i = 0; i < size; i++)
        (p[i] >= average ? large : small).push_back(i);
    while(!small.empty() && !large.empty())
    {
        const unsigned int less = small.back();
        small.pop_back();
        const unsigned int more = large.back();
        large.pop_back();
        prob[less]  = p[less] * size;
        alias[less] = more;
        p[more]     = (p[more] + p[less]) - average;
        (p[more] >= average ? large : small).push_back(more);
    }
    for(unsigned int i : small)
        prob[i] = 1.0;
    for(unsigned int i : large)
        prob[i] = 1.0;
}

unsigned int poisson_table(double lambda, unsigned int& offset, std::vector<double>& prob,
                           std::vector<unsigned int>& alias)
{
    unsigned int        size;
    std::vector<double> p = poisson_pmf(lambda, size, offset);
    alias_table_from(p, size, prob, alias, nullptr);
    return size;
}

std::vector<uint64_t> pack_alias(const std::vector<double>& prob, const std::vector<unsigned int>& alias)
{
    std::vector<uint64_t> out(prob.size());
    for(size_t i = 0; i < prob.size(); i++)
    {
        // prob in [0, 1]: the scaling by 2^32 is exact and the ceiling fits 33 bits
        const uint64_t thr = (uint64_t)std::ceil(std::ldexp(prob[i], 32));
        out[i]             = thr | ((uint64_t)alias[i] << 33);
    }
    return out;
}

// Scratch for the alias probabilities that the CDF method computes and drops.
std::vector<double>& alias_unused_prob()
{
    thread_local std::vector<double> v;
    return v;
}

constexpr double   LAMBDA_HUGE      = 4000.0; // rocrand_poisson.h:52-53
constexpr unsigned POISSON_MAX_SIZE = 4096;   // >= table size at LAMBDA_HUGE (2086)

poisson_cache::~poisson_cache()
{
    if(m_copied)
        (void)hipEventSynchronize(m_copied), (void)hipEventDestroy(m_copied);
    (void)hipFree(m_dprob);
    (void)hipFree(m_dalias);
    (void)hipHostFree(m_hprob);
    (void)hipHostFree(m_halias);
    (void)hipFree(m_dpacked);
    (void)hipHostFree(m_hpacked);
}

rocrand_status poisson_cache::reserve()
{
    if(m_dprob)
        return ROCRAND_STATUS_SUCCESS;
    if(hipMalloc(&m_dprob, POISSON_MAX_SIZE * sizeof(double)) != hipSuccess
       || hipMalloc(&m_dalias, POISSON_MAX_SIZE * sizeof(unsigned int)) != hipSuccess
       || hipHostMalloc(&m_hprob, POISSON_MAX_SIZE * sizeof(double)) != hipSuccess
       || hipHostMalloc(&m_halias, POISSON_MAX_SIZE * sizeof(unsigned int)) != hipSuccess
       || hipMalloc(&m_dpacked, POISSON_MAX_SIZE * sizeof(uint64_t)) != hipSuccess
       || hipHostMalloc(&m_hpacked, POISSON_MAX_SIZE * sizeof(uint64_t)) != hipSuccess
       || hipEventCreateWithFlags(&m_copied, hipEventDisableTiming) != hipSuccess)
        return ROCRAND_STATUS_ALLOCATION_FAILED;
    return ROCRAND_STATUS_SUCCESS;
}

void poisson_cache::fill_staging(void* arg)
{
    const captured_table& t = *static_cast<const captured_table*>(arg);
    memcpy(t.cache->m_hprob, t.prob.data(), t.size * sizeof(double));
    memcpy(t.cache->m_halias, t.alias.data(), t.size * sizeof(unsigned int));
    if(!t.packed.empty())
        memcpy(t.cache->m_hpacked, t.packed.data(), t.size * sizeof(uint64_t));
}

rocrand_status poisson_cache::get(double lambda, hipStream_t stream, dist_params& p)
{
    p.lambda      = lambda;
    p.sqrt_lambda = std::sqrt(lambda);
    p.huge        = lambda > LAMBDA_HUGE ? 1 : 0;
    if(p.huge)
        return ROCRAND_STATUS_SUCCESS;
    rocrand_status st = reserve();
    if(st != ROCRAND_STATUS_SUCCESS)
        return st;
    if(lambda != m_lambda)
    {
        std::vector<double>       prob;
        std::vector<unsigned int> alias;
        if(cdf_method)
        {
            std::vector<double> pmf = poisson_pmf(lambda, m_size, m_offset);
            alias_table_from(pmf, m_size, alias_unused_prob(), alias, &prob);
        }
        else
            m_size = poisson_table(lambda, m_offset, prob, alias);
        if(m_size > POISSON_MAX_SIZE)
            return ROCRAND_STATUS_INTERNAL_ERROR;
        std::vector<uint64_t> packed;
        if(!cdf_method)
            packed = pack_alias(prob, alias);
        hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
        if(hipStreamIsCapturing(stream, &cap) != hipSuccess)
            return ROCRAND_STATUS_INTERNAL_ERROR;
        if(cap == hipStreamCaptureStatusActive)
        {
            // the staging is filled when the graph runs, in stream order before the copies
            auto t    = std::make_unique<captured_table>();
            t->cache  = this;
            t->size   = m_size;
            t->prob   = std::move(prob);
            t->alias  = std::move(alias);
            t->packed = std::move(packed);
            if(hipLaunchHostFunc(stream, &poisson_cache::fill_staging, t.get()) != hipSuccess)
                return ROCRAND_STATUS_INTERNAL_ERROR;
            m_captured.push_back(std::move(t));
        }
        else
        {
            // The staging buffers may still feed the previous table's copy.
            if(hipEventSynchronize(m_copied) != hipSuccess)
                return ROCRAND_STATUS_INTERNAL_ERROR;
            memcpy(m_hprob, prob.data(), m_size * sizeof(double));
            memcpy(m_halias, alias.data(), m_size * sizeof(unsigned int));
            if(!cdf_method)
                memcpy(m_hpacked, packed.data(), m_size * sizeof(uint64_t));
        }
        if(hipMemcpyAsync(m_dprob, m_hprob, m_size * sizeof(double), hipMemcpyHostToDevice, stream) != hipSuccess
           || hipMemcpyAsync(m_dalias, m_halias, m_size * sizeof(unsigned int), hipMemcpyHostToDevice, stream)
                  != hipSuccess
           || (!cdf_method
               && hipMemcpyAsync(m_dpacked, m_hpacked, m_size * sizeof(uint64_t), hipMemcpyHostToDevice, stream)
                      != hipSuccess))
            return ROCRAND_STATUS_INTERNAL_ERROR;
        if(cap != hipStreamCaptureStatusActive && hipEventRecord(m_copied, stream) != hipSuccess)
            return ROCRAND_STATUS_INTERNAL_ERROR;
        m_lambda = lambda;
    }
    p.pt.size   = m_size;
    p.pt.offset = m_offset;
    p.pt.prob   = cdf_method ? nullptr : m_dprob;
    p.pt.alias  = m_dalias;
    p.pt.cdf    = cdf_method ? m_dprob : nullptr;
    p.pt.packed = cdf_method ? nullptr : m_dpacked;
    return ROCRAND_STATUS_SUCCESS;
}

} // namespace rr

// ============================================================================
// Generator base
// ============================================================================
rocrand_generator_base_type::~rocrand_generator_base_type()
{
    if(m_staged)
        (void)hipStreamSynchronize(m_staged_stream); // pending host-function copies read m_pinned
    if(m_staging)
        (void)hipFree(m_staging);
    if(m_pinned)
        (void)hipHostFree(m_pinned);
}

rocrand_status rocrand_generator_base_type::set_seed_uint4(uint4)
{
    return ROCRAND_STATUS_TYPE_ERROR; // generator_type.hpp:245-251
}

rocrand_status rocrand_generator_base_type::set_offset(unsigned long long off)
{
    offset = off;
    reset();
    return ROCRAND_STATUS_SUCCESS;
}

bool rocrand_generator_base_type::pseudo_order_ok(rocrand_ordering o) const
{
    // philox4x32_10.hpp:258-278: DYNAMIC is device-only; SEEDED and QUASI are rejected.
    if(host_mode != 0 && o == ROCRAND_ORDERING_PSEUDO_DYNAMIC)
        return false;
    return o == ROCRAND_ORDERING_PSEUDO_DEFAULT || o == ROCRAND_ORDERING_PSEUDO_DYNAMIC
           || o == ROCRAND_ORDERING_PSEUDO_BEST || o == ROCRAND_ORDERING_PSEUDO_LEGACY;
}

rocrand_status rocrand_generator_base_type::set_order(rocrand_ordering o)
{
    if(!pseudo_order_ok(o))
        return ROCRAND_STATUS_OUT_OF_RANGE;
    order = o;
    reset();
    return ROCRAND_STATUS_SUCCESS;
}

rocrand_status rocrand_generator_base_type::set_dimensions(unsigned int)
{
    return ROCRAND_STATUS_TYPE_ERROR; // generator_type.hpp:283-288
}

rocrand_status rocrand_generator_base_type::set_stream(hipStream_t s)
{
    stream = s;
    return ROCRAND_STATUS_SUCCESS;
}

rocrand_status rocrand_generator_base_type::make_params(out_kind k, const gen_args& args, dist_params& p)
{
    p    = dist_params{};
    p.mf = (float)args.a;
    p.sf = (float)args.b;
    p.md = args.a;
    p.sd = args.b;
    p.mh = args.ha;
    p.sh = args.hb;
    // the standard normal's affine step is the identity (mean +0: -0 + x would differ)
    p.unit = k == out_kind::normal_f32 && __builtin_bit_cast(uint32_t, p.mf) == 0u && p.sf == 1.0f;
    if(k == out_kind::poisson)
        return m_poisson.get(args.a, stream, p);
    return ROCRAND_STATUS_SUCCESS;
}

// Device staging of `bytes` whose address has the host pointer's alignment modulo 128
// bytes (the reference's head/tail split depends on the output address), plus pinned
// staging of the same size for stream-ordered host generators.
rocrand_status rocrand_generator_base_type::stage_host(void* data, size_t bytes, void*& dev)
{
    const size_t need        = bytes + 256;
    const bool   grow_dev    = need > m_staging_bytes;
    const bool   grow_pinned = host_mode == 1 && bytes > m_pinned_bytes;
    if(m_staged && (grow_dev || grow_pinned || m_staged_stream != stream))
    {
        // earlier calls' kernels, copies and host functions still use the buffers
        if(hipStreamSynchronize(m_staged_stream) != hipSuccess)
            return ROCRAND_STATUS_INTERNAL_ERROR;
        m_staged = false;
    }
    if(grow_dev)
    {
        (void)hipFree(m_staging);
        m_staging       = nullptr;
        m_staging_bytes = 0;
        if(hipMalloc(&m_staging, need) != hipSuccess)
            return ROCRAND_STATUS_ALLOCATION_FAILED;
        m_staging_bytes = need;
    }
    if(grow_pinned)
    {
        (void)hipHostFree(m_pinned);
        m_pinned       = nullptr;
        m_pinned_bytes = 0;
        if(hipHostMalloc(&m_pinned, bytes, hipHostMallocDefault) != hipSuccess)
            return ROCRAND_STATUS_ALLOCATION_FAILED;
        m_pinned_bytes = bytes;
    }
    const uintptr_t base = (reinterpret_cast<uintptr_t>(m_staging) + 127) & ~uintptr_t(127);
    dev                  = reinterpret_cast<void*>(base + (reinterpret_cast<uintptr_t>(data) & 127));
    return ROCRAND_STATUS_SUCCESS;
}

namespace
{
// The host-function step of a stream-ordered host generator: the kernel's output has
// reached pinned memory when the stream runs it (library/src/rng/system.hpp:195-207 runs
// the kernel bodies themselves from such a callback).
struct host_copy
{
    void*       dst;
    const void* src;
    size_t      bytes;
    static void run(void* p)
    {
        const host_copy* c = static_cast<const host_copy*>(p);
        std::memcpy(c->dst, c->src, c->bytes);
        delete c;
    }
};
} // namespace

rocrand_status rocrand_generator_base_type::generate(out_kind k, void* data, size_t n, const gen_args& args)
{
    if(host_mode == 0 || data == nullptr)
        return generate_device(k, data, n, args);
    // Host generator: run the device path into staging memory, then copy out.  Blocking
    // generators (rocrand_create_generator_host_blocking) return with the values written;
    // the others (rocrand_create_generator_host) return at once and the values are written
    // when the generator's stream reaches the call, as the reference's host-function launch.
    dist_shape     s;
    rocrand_status st = shape(k, s);
    if(st != ROCRAND_STATUS_SUCCESS)
        return st;
    const size_t bytes = n * s.esz;
    void*        dev   = nullptr;
    st                 = stage_host(data, bytes, dev);
    if(st != ROCRAND_STATUS_SUCCESS)
        return st;
    st = generate_device(k, dev, n, args);
    if(st != ROCRAND_STATUS_SUCCESS)
        return st;
    m_staged        = true;
    m_staged_stream = stream;
    if(bytes == 0)
        return ROCRAND_STATUS_SUCCESS;
    if(host_mode == 2)
    {
        if(hipMemcpyAsync(data, dev, bytes, hipMemcpyDeviceToHost, stream) != hipSuccess
           || hipStreamSynchronize(stream) != hipSuccess)
            return ROCRAND_STATUS_INTERNAL_ERROR;
        return ROCRAND_STATUS_SUCCESS;
    }
    if(hipMemcpyAsync(m_pinned, dev, bytes, hipMemcpyDeviceToHost, stream) != hipSuccess)
        return ROCRAND_STATUS_INTERNAL_ERROR;
    host_copy* job = new(std::nothrow) host_copy{data, m_pinned, bytes};
    if(job == nullptr)
        return ROCRAND_STATUS_ALLOCATION_FAILED;
    if(hipLaunchHostFunc(stream, &host_copy::run, job) != hipSuccess)
    {
        delete job;
        return ROCRAND_STATUS_LAUNCH_FAILURE;
    }
    return ROCRAND_STATUS_SUCCESS;
}

namespace
{

template<word_family F, out_kind K, bool B = false>
constexpr dist_shape shape_of()
{
    return dist_shape{dist<F, K, B>::in_w, dist<F, K, B>::out_w, (int)sizeof(typename dist<F, K, B>::out_t)};
}

template<word_family F>
rocrand_status family_shape(out_kind k, bool tf2x32, dist_shape& s)
{
    switch(k)
    {
        case out_kind::u8: s = shape_of<F, out_kind::u8>(); break;
        case out_kind::u16: s = shape_of<F, out_kind::u16>(); break;
        case out_kind::u32: s = shape_of<F, out_kind::u32>(); break;
        case out_kind::u64:
            if constexpr(F == word_family::w64)
                s = shape_of<F, out_kind::u64>();
            else
                return ROCRAND_STATUS_TYPE_ERROR;
            break;
        case out_kind::uniform_f32: s = shape_of<F, out_kind::uniform_f32>(); break;
        case out_kind::uniform_f64: s = shape_of<F, out_kind::uniform_f64>(); break;
        case out_kind::uniform_f16: s = shape_of<F, out_kind::uniform_f16>(); break;
        case out_kind::normal_f32: s = shape_of<F, out_kind::normal_f32>(); break;
        case out_kind::lognormal_f32: s = shape_of<F, out_kind::lognormal_f32>(); break;
        case out_kind::normal_f16: s = shape_of<F, out_kind::normal_f16>(); break;
        case out_kind::lognormal_f16: s = shape_of<F, out_kind::lognormal_f16>(); break;
        case out_kind::poisson: s = shape_of<F, out_kind::poisson>(); break;
        case out_kind::normal_f64:
            if constexpr(F == word_family::w32)
                s = tf2x32 ? shape_of<F, out_kind::normal_f64, true>() : shape_of<F, out_kind::normal_f64, false>();
            else
                s = shape_of<F, out_kind::normal_f64>();
            break;
        case out_kind::lognormal_f64:
            if constexpr(F == word_family::w32)
                s = tf2x32 ? shape_of<F, out_kind::lognormal_f64, true>()
                           : shape_of<F, out_kind::lognormal_f64, false>();
            else
                s = shape_of<F, out_kind::lognormal_f64>();
            break;
    }
    return ROCRAND_STATUS_SUCCESS;
}

// ============================================================================
// Counter-based generators: Philox4x32-10, ThreeFry{2,4}x{32,64}-20
// (philox4x32_10.hpp:216-432, threefry.hpp:201-494)
// ============================================================================
class counter_generator final : public rocrand_generator_base_type
{
public:
    explicit counter_generator(rocrand_rng_type t) : rocrand_generator_base_type(t)
    {
        switch(t)
        {
            case ROCRAND_RNG_PSEUDO_PHILOX4_32_10: m_kind = ctr_kind::philox4x32; m_dim = 4; m_w64 = false; break;
            case ROCRAND_RNG_PSEUDO_THREEFRY2_32_20: m_kind = ctr_kind::threefry2x32; m_dim = 2; m_w64 = false; break;
            case ROCRAND_RNG_PSEUDO_THREEFRY2_64_20: m_kind = ctr_kind::threefry2x64; m_dim = 2; m_w64 = true; break;
            case ROCRAND_RNG_PSEUDO_THREEFRY4_32_20: m_kind = ctr_kind::threefry4x32; m_dim = 4; m_w64 = false; break;
            default: m_kind = ctr_kind::threefry4x64; m_dim = 4; m_w64 = true; break;
        }
    }

    void set_seed(unsigned long long seed) override
    {
        m_seed = seed;
        reset();
    }

    rocrand_status init() override
    {
        if(m_initialized)
            return ROCRAND_STATUS_SUCCESS;
        // the reference kernel's grid stride under this ordering (config_types.hpp:314-336)
        m_stride = CTR_REF_STRIDE;
        if(order == ROCRAND_ORDERING_PSEUDO_DYNAMIC)
        {
            unsigned       threads, blocks;
            rocrand_status st = dynamic_grid(type, threads, blocks);
            if(st != ROCRAND_STATUS_SUCCESS)
                return st;
            m_stride = threads * blocks;
            if((m_stride & (m_stride - 1)) != 0) // the kernels mask vec indices with it
                return ROCRAND_STATUS_INTERNAL_ERROR;
        }
        // engine(seed, subsequence 0, offset): the stream starts at word `offset`.
        m_pos         = offset;
        m_initialized = true;
        return ROCRAND_STATUS_SUCCESS;
    }

    rocrand_status shape(out_kind k, dist_shape& s) const override
    {
        return m_w64 ? family_shape<word_family::w64>(k, false, s)
                     : family_shape<word_family::w32>(k, m_kind == ctr_kind::threefry2x32, s);
    }

protected:
    void reset() override { m_initialized = false; }

    ctr_key key() const
    {
        ctr_key k{};
        const uint64_t s = m_seed;
        switch(m_kind)
        {
            case ctr_kind::philox4x32:
            case ctr_kind::threefry2x32: k.k[0] = (uint32_t)s; k.k[1] = (uint32_t)(s >> 32); break;
            case ctr_kind::threefry2x64: k.k[0] = s; k.k[1] = s >> 32; break;
            case ctr_kind::threefry4x32:
                k.k[0] = (uint32_t)s; k.k[1] = (uint32_t)(s >> 16); k.k[2] = (uint32_t)(s >> 32); k.k[3] = (uint32_t)(s >> 48);
                break;
            case ctr_kind::threefry4x64: k.k[0] = s; k.k[1] = s >> 16; k.k[2] = s >> 32; k.k[3] = s >> 48; break;
        }
        return k;
    }

    ctr_pos to_pos(unsigned __int128 words) const
    {
        ctr_pos                 p{};
        const unsigned __int128 c = words / (unsigned)m_dim;
        p.sub                     = (uint32_t)(words % (unsigned)m_dim);
        if(m_w64)
        {
            p.c[0] = (uint64_t)c;
            p.c[1] = (uint64_t)(c >> 64);
        }
        else
        {
            for(int i = 0; i < 4; i++)
                p.c[i] = (uint64_t)((c >> (32 * i)) & 0xffffffffu);
        }
        return p;
    }

public:
    ctr_kind kind() const { return m_kind; }

    // Resolves the next generate call (head/tail split, word positions) without
    // launching; c.params is left to the caller.
    void prepare(const dist_shape& s, void* data, size_t n, counter_call& c)
    {
        const unsigned opt  = (unsigned)m_dim / s.in_w;
        const unsigned full = opt * s.out_w;
        const uintptr_t addr = reinterpret_cast<uintptr_t>(data);
        c.mis   = (full - (addr / s.esz) % full) % full;
        c.head  = (uint32_t)std::min<size_t>(n, c.mis);
        c.tail  = (uint32_t)((n - c.head) % full);
        c.vec_n = (n - c.head) / full;
        // Words before vec 0 for every thread but the reference's thread 0 (H), and the
        // words thread 0 actually consumed writing the head (C0).
        const uint64_t H = (m_kind == ctr_kind::philox4x32) ? c.head : (c.head / s.out_w) * s.in_w;
        const uint64_t C0 = (uint64_t)((c.head + s.out_w - 1) / s.out_w) * s.in_w;
        c.key       = key();
        c.pos       = to_pos(m_pos);
        c.vec_base  = to_pos(m_pos + H);
        c.vec_base0 = to_pos(m_pos + C0);
        c.quirk      = (H != C0) ? 1 : 0;
        c.ref_stride = m_stride;
        // the fast kernels align their 64-vec chunks to the output's 128-byte lines
        const uintptr_t vaddr  = addr + (uintptr_t)c.mis * s.esz;
        const unsigned  vbytes = full * s.esz;
        c.lane_shift = 128u % vbytes == 0 ? (uint32_t)((vaddr / vbytes) % (128u / vbytes)) : 0u;
        c.data       = data;
        c.n          = n;
    }

protected:
    rocrand_status generate_device(out_kind k, void* data, size_t n, const gen_args& args) override
    {
        dist_shape     s;
        rocrand_status st = shape(k, s);
        if(st != ROCRAND_STATUS_SUCCESS)
            return st;
        if((st = init()) != ROCRAND_STATUS_SUCCESS)
            return st;
        if(data == nullptr)
            return ROCRAND_STATUS_SUCCESS; // philox4x32_10.hpp:323-326
        counter_call c{};
        if((st = make_params(k, args, c.params)) != ROCRAND_STATUS_SUCCESS)
            return st;
        prepare(s, data, n, c);
        if(n > 0)
        {
            if(launch_counter(m_kind, k, false, c, stream) != hipSuccess)
                return ROCRAND_STATUS_LAUNCH_FAILURE;
        }
        m_pos += (unsigned __int128)((n + s.out_w - 1) / s.out_w) * (unsigned)s.in_w;
        return ROCRAND_STATUS_SUCCESS;
    }

private:
    ctr_kind           m_kind;
    int                m_dim;
    bool               m_w64;
    unsigned long long m_seed        = 0;
    bool               m_initialized = false;
    unsigned __int128  m_pos         = 0;
    uint32_t           m_stride      = CTR_REF_STRIDE;
};

// ============================================================================
// Interleaved stateful generators: MRG32k3a, MRG31k3p, XORWOW, LFSR113
// (mrg.hpp:159-497, xorwow.hpp:164-471, lfsr113.hpp:175-505)
// ============================================================================

class interleaved_generator final : public rocrand_generator_base_type
{
public:
    explicit interleaved_generator(rocrand_rng_type t) : rocrand_generator_base_type(t)
    {
        switch(t)
        {
            case ROCRAND_RNG_PSEUDO_MRG32K3A: m_kind = inter_kind::mrg32k3a; m_seed = 12345; break;
            case ROCRAND_RNG_PSEUDO_MRG31K3P: m_kind = inter_kind::mrg31k3p; m_seed = 12345; break;
            case ROCRAND_RNG_PSEUDO_LFSR113: m_kind = inter_kind::lfsr113; break;
            default: m_kind = inter_kind::xorwow; break;
        }
    }
    ~interleaved_generator() override
    {
        if(m_states)
            (void)hipFree(m_states);
    }

    void set_seed(unsigned long long seed) override
    {
        if(m_kind == inter_kind::lfsr113)
        {
            // lfsr113.hpp:240-255: only x and y are raised to their minimum seeds.
            uint4 s = {(unsigned)seed, (unsigned)(seed >> 16), (unsigned)(seed >> 32), (unsigned)(seed >> 48)};
            if(s.x < 2u)
                s.x += 2u;
            if(s.y < 8u)
                s.y += 8u;
            m_seed4 = s;
        }
        else if(m_kind == inter_kind::mrg32k3a || m_kind == inter_kind::mrg31k3p)
            m_seed = seed == 0 ? 12345ull : seed; // mrg.hpp:240-244
        else
            m_seed = seed;
        reset();
    }

    rocrand_status set_seed_uint4(uint4 s) override
    {
        if(m_kind != inter_kind::lfsr113)
            return ROCRAND_STATUS_TYPE_ERROR;
        // lfsr113.hpp:257-272
        if(s.x < 2u) s.x += 2u;
        if(s.y < 8u) s.y += 8u;
        if(s.z < 16u) s.z += 16u;
        if(s.w < 128u) s.w += 128u;
        m_seed4 = s;
        reset();
        return ROCRAND_STATUS_SUCCESS;
    }

    rocrand_status init() override
    {
        if(m_initialized)
            return ROCRAND_STATUS_SUCCESS;
        // engine count: the reference's grid under this ordering (mrg.hpp:284-337)
        uint32_t engines = INTERLEAVED_ENGINES;
        if(order == ROCRAND_ORDERING_PSEUDO_DYNAMIC)
        {
            unsigned       threads, blocks;
            rocrand_status st = dynamic_grid(type, threads, blocks);
            if(st != ROCRAND_STATUS_SUCCESS)
                return st;
            engines = threads * blocks;
        }
        if(m_states && engines != m_S)
        {
            // earlier calls' kernels on the stream may still read the old states
            if(hipStreamSynchronize(stream) != hipSuccess)
                return ROCRAND_STATUS_INTERNAL_ERROR;
            (void)hipFree(m_states);
            m_states = nullptr;
        }
        m_S = engines;
        if(!m_states)
        {
            // two state buffers: a call reads one and leaves the advanced states in the other
            if(hipMalloc(&m_states, 2 * state_words() * sizeof(uint32_t)) != hipSuccess)
            {
                m_states = nullptr;
                return ROCRAND_STATUS_ALLOCATION_FAILED;
            }
            m_cur = 0;
        }
        m_start = (uint32_t)(offset % m_S);
        uint32_t seeded[6];
        seed_state(seeded);
        if(launch_interleaved_init(m_kind, cur_states(), seeded, m_start, offset / m_S, m_S, stream) != hipSuccess)
            return ROCRAND_STATUS_LAUNCH_FAILURE;
        m_initialized = true;
        return ROCRAND_STATUS_SUCCESS;
    }

    rocrand_status shape(out_kind k, dist_shape& s) const override
    {
        if(m_kind == inter_kind::mrg32k3a)
            return family_shape<word_family::mrg32>(k, false, s);
        if(m_kind == inter_kind::mrg31k3p)
            return family_shape<word_family::mrg31>(k, false, s);
        return family_shape<word_family::w32>(k, false, s);
    }

    inter_kind kind() const { return m_kind; }

    // Seeded state of one engine before any jump, in the kernel's word order.
    void seed_state(uint32_t* w) const
    {
        if(m_kind == inter_kind::mrg32k3a)
            mrg32k3a_seed_words(m_seed, w); // rocrand_mrg32k3a.h:103-120
        else if(m_kind == inter_kind::mrg31k3p)
            mrg31k3p_seed_words(m_seed, w);
        else if(m_kind == inter_kind::xorwow)
        {
            xorwow_state st;
            xorwow_seed(st, m_seed);
            w[0] = st.d;
            for(int j = 0; j < 5; j++)
                w[1 + j] = st.x[j];
        }
        else
        {
            w[0] = m_seed4.x;
            w[1] = m_seed4.y;
            w[2] = m_seed4.z;
            w[3] = m_seed4.w;
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
        if((st = init()) != ROCRAND_STATUS_SUCCESS)
            return st;
        if(data == nullptr)
            return ROCRAND_STATUS_SUCCESS;
        interleaved_call c{};
        if((st = make_params(k, args, c.params)) != ROCRAND_STATUS_SUCCESS)
            return st;
        const uintptr_t addr = reinterpret_cast<uintptr_t>(data);
        c.mis    = (s.out_w - (addr / s.esz) % s.out_w) % s.out_w;
        c.head   = (uint32_t)std::min<size_t>(n, c.mis);
        c.tail   = (uint32_t)((n - c.head) % s.out_w);
        c.vec_n  = (n - c.head) / s.out_w;
        {
            // the kernels shift engines over lanes so that a wavefront's 64 vecs start on a
            // 128-byte line of the output
            const uintptr_t vaddr  = addr + (uintptr_t)c.mis * s.esz;
            const unsigned  vbytes = s.out_w * s.esz;
            c.lane_shift = 128u % vbytes == 0 ? (uint32_t)((vaddr / vbytes) % (128u / vbytes)) : 0u;
        }
        c.states     = cur_states();
        c.states_out = m_states + (m_cur ^ 1u) * state_words();
        c.start      = m_start;
        c.S          = m_S;
        c.data   = data;
        c.n      = n;
        if(launch_interleaved(m_kind, k, c, stream) != hipSuccess)
            return ROCRAND_STATUS_LAUNCH_FAILURE;
        m_cur ^= 1u;
        // mrg.hpp:385-389
        m_start = (uint32_t)((m_start + (n + s.out_w - 1) / s.out_w) % m_S);
        return ROCRAND_STATUS_SUCCESS;
    }

private:
    inter_kind         m_kind;
    unsigned long long m_seed        = 0;
    uint4              m_seed4       = {2, 8, 16, 128};
    bool               m_initialized = false;
    uint32_t*          m_states      = nullptr; // 2 * state_words() words
    uint32_t           m_cur         = 0;       // which half holds the current states
    uint32_t           m_start       = 0;
    uint32_t           m_S           = INTERLEAVED_ENGINES; // engines

    size_t    state_words() const { return (size_t)inter_state_words(m_kind) * m_S; }
    uint32_t* cur_states() const { return m_states + m_cur * state_words(); }
};

} // namespace

namespace rr
{
rocrand_generator make_counter_generator(rocrand_rng_type t) { return new counter_generator(t); }
rocrand_generator make_interleaved_generator(rocrand_rng_type t) { return new interleaved_generator(t); }

rocrand_status prepare_counter_call(rocrand_rng_type t, unsigned long long seed, unsigned long long offset,
                                    out_kind k, void* data, size_t n, counter_call& c, dist_shape& s,
                                    ctr_kind& kind)
{
    counter_generator g(t);
    g.set_seed(seed);
    g.set_offset(offset);
    rocrand_status st = g.shape(k, s);
    if(st != ROCRAND_STATUS_SUCCESS)
        return st;
    g.init();
    g.prepare(s, data, n, c);
    kind = g.kind();
    return ROCRAND_STATUS_SUCCESS;
}

rocrand_status interleaved_seed_state(rocrand_rng_type t, unsigned long long seed, out_kind k, uint32_t w[6],
                                      inter_kind& kind, dist_shape& s)
{
    interleaved_generator g(t);
    g.set_seed(seed);
    rocrand_status st = g.shape(k, s);
    if(st != ROCRAND_STATUS_SUCCESS)
        return st;
    g.seed_state(w);
    kind = g.kind();
    return ROCRAND_STATUS_SUCCESS;
}
} // namespace rr

// ============================================================================
// C ABI
// ============================================================================
namespace
{

rocrand_status create(rocrand_generator* generator, rocrand_rng_type rng_type, int host_mode)
{
    if(generator == nullptr)
        return ROCRAND_STATUS_OUT_OF_RANGE;
    rocrand_generator g = nullptr;
    try
    {
        switch(rng_type)
        {
            case ROCRAND_RNG_PSEUDO_PHILOX4_32_10:
            case ROCRAND_RNG_PSEUDO_THREEFRY2_32_20:
            case ROCRAND_RNG_PSEUDO_THREEFRY2_64_20:
            case ROCRAND_RNG_PSEUDO_THREEFRY4_32_20:
            case ROCRAND_RNG_PSEUDO_THREEFRY4_64_20: g = make_counter_generator(rng_type); break;
            case ROCRAND_RNG_PSEUDO_DEFAULT:
            case ROCRAND_RNG_PSEUDO_XORWOW: g = make_interleaved_generator(ROCRAND_RNG_PSEUDO_XORWOW); break;
            case ROCRAND_RNG_PSEUDO_MRG32K3A:
            case ROCRAND_RNG_PSEUDO_MRG31K3P:
            case ROCRAND_RNG_PSEUDO_LFSR113: g = make_interleaved_generator(rng_type); break;
            case ROCRAND_RNG_PSEUDO_MTGP32: g = make_mtgp32_generator(); break;
            case ROCRAND_RNG_PSEUDO_MT19937: g = make_mt19937_generator(); break;
            case ROCRAND_RNG_QUASI_DEFAULT:
            case ROCRAND_RNG_QUASI_SOBOL32: g = make_sobol_generator(ROCRAND_RNG_QUASI_SOBOL32); break;
            case ROCRAND_RNG_QUASI_SCRAMBLED_SOBOL32:
            case ROCRAND_RNG_QUASI_SOBOL64:
            case ROCRAND_RNG_QUASI_SCRAMBLED_SOBOL64: g = make_sobol_generator(rng_type); break;
            default: return ROCRAND_STATUS_TYPE_ERROR;
        }
    }
    catch(const std::bad_alloc&)
    {
        return ROCRAND_STATUS_INTERNAL_ERROR;
    }
    catch(rocrand_status st)
    {
        return st;
    }
    if(g == nullptr)
        return ROCRAND_STATUS_TYPE_ERROR;
    g->host_mode = host_mode;
    *generator   = g;
    return ROCRAND_STATUS_SUCCESS;
}

template<class F>
rocrand_status guarded(rocrand_generator g, F&& f)
{
    if(g == nullptr)
        return ROCRAND_STATUS_NOT_CREATED;
    try
    {
        return f();
    }
    catch(const std::bad_alloc&)
    {
        return ROCRAND_STATUS_ALLOCATION_FAILED;
    }
    catch(rocrand_status st)
    {
        return st;
    }
}

rocrand_status gen(rocrand_generator g, out_kind k, void* data, size_t n, double a = 0.0, double b = 0.0)
{
    return guarded(g, [&] {
        gen_args args;
        args.a = a;
        args.b = b;
        return g->generate(k, data, n, args);
    });
}

rocrand_status gen_half(rocrand_generator g, out_kind k, half* data, size_t n, half mean, half stddev)
{
    return guarded(g, [&] {
        gen_args args;
        args.a  = (double)__half2float(mean);
        args.b  = (double)__half2float(stddev);
        args.ha = mean;
        args.hb = stddev;
        return g->generate(k, data, n, args);
    });
}

} // namespace

extern "C" {

rocrand_status ROCRANDAPI rocrand_create_generator(rocrand_generator* generator, rocrand_rng_type rng_type)
{
    return create(generator, rng_type, 0);
}

rocrand_status ROCRANDAPI rocrand_create_generator_host(rocrand_generator* generator, rocrand_rng_type rng_type)
{
    return create(generator, rng_type, 1);
}

rocrand_status ROCRANDAPI rocrand_create_generator_host_blocking(rocrand_generator* generator,
                                                                 rocrand_rng_type   rng_type)
{
    return create(generator, rng_type, 2);
}

rocrand_status ROCRANDAPI rocrand_destroy_generator(rocrand_generator generator)
{
    try
    {
        delete generator;
    }
    catch(rocrand_status st)
    {
        return st;
    }
    return ROCRAND_STATUS_SUCCESS;
}

rocrand_status ROCRANDAPI rocrand_generate(rocrand_generator g, unsigned int* d, size_t n)
{
    return gen(g, out_kind::u32, d, n);
}
rocrand_status ROCRANDAPI rocrand_generate_long_long(rocrand_generator g, unsigned long long int* d, size_t n)
{
    return gen(g, out_kind::u64, d, n);
}
rocrand_status ROCRANDAPI rocrand_generate_char(rocrand_generator g, unsigned char* d, size_t n)
{
    return gen(g, out_kind::u8, d, n);
}
rocrand_status ROCRANDAPI rocrand_generate_short(rocrand_generator g, unsigned short* d, size_t n)
{
    return gen(g, out_kind::u16, d, n);
}
rocrand_status ROCRANDAPI rocrand_generate_uniform(rocrand_generator g, float* d, size_t n)
{
    return gen(g, out_kind::uniform_f32, d, n);
}
rocrand_status ROCRANDAPI rocrand_generate_uniform_double(rocrand_generator g, double* d, size_t n)
{
    return gen(g, out_kind::uniform_f64, d, n);
}
rocrand_status ROCRANDAPI rocrand_generate_uniform_half(rocrand_generator g, half* d, size_t n)
{
    return gen(g, out_kind::uniform_f16, d, n);
}
rocrand_status ROCRANDAPI rocrand_generate_normal(rocrand_generator g, float* d, size_t n, float mean, float sd)
{
    return gen(g, out_kind::normal_f32, d, n, mean, sd);
}
rocrand_status ROCRANDAPI rocrand_generate_normal_double(rocrand_generator g, double* d, size_t n, double mean,
                                                         double sd)
{
    return gen(g, out_kind::normal_f64, d, n, mean, sd);
}
rocrand_status ROCRANDAPI rocrand_generate_normal_half(rocrand_generator g, half* d, size_t n, half mean, half sd)
{
    return gen_half(g, out_kind::normal_f16, d, n, mean, sd);
}
rocrand_status ROCRANDAPI rocrand_generate_log_normal(rocrand_generator g, float* d, size_t n, float mean,
                                                      float sd)
{
    return gen(g, out_kind::lognormal_f32, d, n, mean, sd);
}
rocrand_status ROCRANDAPI rocrand_generate_log_normal_double(rocrand_generator g, double* d, size_t n, double mean,
                                                             double sd)
{
    return gen(g, out_kind::lognormal_f64, d, n, mean, sd);
}
rocrand_status ROCRANDAPI rocrand_generate_log_normal_half(rocrand_generator g, half* d, size_t n, half mean,
                                                           half sd)
{
    return gen_half(g, out_kind::lognormal_f16, d, n, mean, sd);
}
rocrand_status ROCRANDAPI rocrand_generate_poisson(rocrand_generator g, unsigned int* d, size_t n, double lambda)
{
    if(g == nullptr)
        return ROCRAND_STATUS_NOT_CREATED;
    if(lambda <= 0.0)
        return ROCRAND_STATUS_OUT_OF_RANGE; // rocrand.cpp:371-374
    return gen(g, out_kind::poisson, d, n, lambda);
}

rocrand_status ROCRANDAPI rocrand_initialize_generator(rocrand_generator g)
{
    // Everything a later generate call would allocate is allocated here, so that calls
    // captured into a HIP graph allocate nothing (test_rocrand_hipgraphs.cpp:13-66).
    return guarded(g, [&] { return g->initialize(); });
}
rocrand_status ROCRANDAPI rocrand_set_stream(rocrand_generator g, hipStream_t stream)
{
    return guarded(g, [&] { return g->set_stream(stream); });
}
rocrand_status ROCRANDAPI rocrand_set_seed(rocrand_generator g, unsigned long long seed)
{
    return guarded(g, [&] {
        g->set_seed(seed);
        return ROCRAND_STATUS_SUCCESS;
    });
}
rocrand_status ROCRANDAPI rocrand_set_seed_uint4(rocrand_generator g, uint4 seed)
{
    return guarded(g, [&] { return g->set_seed_uint4(seed); });
}
rocrand_status ROCRANDAPI rocrand_set_offset(rocrand_generator g, unsigned long long offset)
{
    return guarded(g, [&] { return g->set_offset(offset); });
}
rocrand_status ROCRANDAPI rocrand_set_ordering(rocrand_generator g, rocrand_ordering order)
{
    return guarded(g, [&] { return g->set_order(order); });
}
rocrand_status ROCRANDAPI rocrand_set_quasi_random_generator_dimensions(rocrand_generator g, unsigned int d)
{
    return guarded(g, [&] { return g->set_dimensions(d); });
}
rocrand_status ROCRANDAPI rocrand_get_version(int* version)
{
    if(version == nullptr)
        return ROCRAND_STATUS_OUT_OF_RANGE;
    *version = ROCRAND_VERSION;
    return ROCRAND_STATUS_SUCCESS;
}

} // extern "C"
