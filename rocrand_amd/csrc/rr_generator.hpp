// rr_generator.hpp -- generator objects behind the opaque rocrand_generator handle.
//
// One object per rocrand_create_generator* call.  The object holds the engine
// state resident on the device (a 128-bit word position for the counter-based
// engines, S structure-of-arrays engine states for the interleaved ones), the
// stream, ordering and offset, and a per-generator Poisson alias-table cache.
// Behaviour of every setter and status code follows library/src/rocrand.cpp and
// the generator templates of library/src/rng/*.hpp.
#pragma once

#include "rr_launch.hpp"

#include <rocrand.h>

#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

namespace rr
{

// Shape of a transform: engine words consumed and outputs produced per call.
struct dist_shape
{
    int in_w, out_w, esz;
};

// Arguments of one generate call beyond the output buffer.
struct gen_args
{
    double a = 0.0, b = 0.0; // mean/stddev or lambda
    __half ha, hb;           // half mean/stddev
};

// Device-resident Poisson alias table for the most recent lambda
// (library/src/rng/distribution/poisson.hpp:171-412 keeps one table per generator
// and rebuilds it only when lambda changes).
class poisson_cache
{
public:
    ~poisson_cache();
    // Device tables and pinned staging (allocated by rocrand_initialize_generator, so
    // that generate calls under stream capture allocate nothing).
    rocrand_status reserve();
    // Fills p for lambda (<= 4000: table method, else the normal approximation).  A new
    // lambda's table is staged and copied on `stream`; under stream capture the staging
    // is filled by a host-function node, so every replay of the graph uploads that
    // graph's table (poisson.hpp:296-340 does the same with hipLaunchHostFunc).
    rocrand_status get(double lambda, hipStream_t stream, dist_params& p);
    // Quasi-random generators sample by binary search in the cumulative table
    // (sobol.hpp:500-503 selects DISCRETE_METHOD_CDF) instead of the alias table.
    bool cdf_method = false;

private:
    double        m_lambda = -1.0;
    unsigned int  m_size = 0, m_offset = 0;
    double*       m_dprob  = nullptr; // alias probabilities, or the CDF
    unsigned int* m_dalias = nullptr;
    double*       m_hprob  = nullptr; // pinned staging
    unsigned int* m_halias = nullptr;
    uint64_t*     m_dpacked = nullptr; // packed alias table (alias method)
    uint64_t*     m_hpacked = nullptr;
    hipEvent_t    m_copied = nullptr;
    // Tables of captured lambda changes, read by their graphs' host-function nodes; kept
    // for the generator's lifetime (a graph must not outlive its generator).
    struct captured_table
    {
        poisson_cache*            cache;
        unsigned int              size;
        std::vector<double>       prob;
        std::vector<unsigned int> alias;
        std::vector<uint64_t>     packed;
    };
    std::vector<std::unique_ptr<captured_table>> m_captured;
    static void fill_staging(void* table);
};

// Poisson pmf of poisson.hpp:110-161 (unnormalised, truncated where p < 1e-12).
std::vector<double> poisson_pmf(double lambda, unsigned int& size, unsigned int& offset);
// Host computation of the Poisson alias table, same arithmetic as
// poisson.hpp:110-161 (pmf) and discrete.hpp:189-265 (normalize, Vose alias).
unsigned int poisson_table(double lambda, unsigned int& offset, std::vector<double>& prob,
                           std::vector<unsigned int>& alias);
// Packed alias entries for integer sampling: ceil(prob[i] 2^32) | alias[i] << 33.
std::vector<uint64_t> pack_alias(const std::vector<double>& prob, const std::vector<unsigned int>& alias);
// Alias table of an arbitrary probability vector (discrete.hpp:97-140).
void alias_table_from(const std::vector<double>& p, unsigned int size, std::vector<double>& prob,
                      std::vector<unsigned int>& alias, std::vector<double>* cdf);

} // namespace rr

// The opaque handle type of the C ABI (rocrand.h:41).
struct rocrand_generator_base_type
{
    explicit rocrand_generator_base_type(rocrand_rng_type t) : type(t) {}
    virtual ~rocrand_generator_base_type();

    rocrand_rng_type   type;
    int                host_mode = 0; // 0 device, 1 host (stream-ordered), 2 host blocking
    hipStream_t        stream    = 0;
    rocrand_ordering   order     = ROCRAND_ORDERING_PSEUDO_DEFAULT;
    unsigned long long offset    = 0;

    virtual void           set_seed(unsigned long long seed) = 0;
    virtual rocrand_status set_seed_uint4(uint4 seed);
    virtual rocrand_status set_offset(unsigned long long off);
    virtual rocrand_status set_order(rocrand_ordering o);
    virtual rocrand_status set_dimensions(unsigned int d);
    rocrand_status         set_stream(hipStream_t s);
    virtual rocrand_status init() = 0;

    // Entry for every rocrand_generate_*: host generators stage through device memory.
    rocrand_status generate(rr::out_kind k, void* data, size_t n, const rr::gen_args& args);

    // Shape of kind k for this engine, or TYPE_ERROR.
    virtual rocrand_status shape(rr::out_kind k, rr::dist_shape& s) const = 0;

    // rocrand_initialize_generator: engine set-up plus the allocations a generate call
    // could otherwise make (none may happen under stream capture).
    rocrand_status initialize()
    {
        rocrand_status st = init();
        if(st == ROCRAND_STATUS_SUCCESS && host_mode == 0)
            st = m_poisson.reserve();
        return st;
    }

protected:
    virtual rocrand_status generate_device(rr::out_kind k, void* data, size_t n, const rr::gen_args& args) = 0;
    virtual void           reset() = 0;
    rocrand_status         make_params(rr::out_kind k, const rr::gen_args& args, rr::dist_params& p);
    bool                   pseudo_order_ok(rocrand_ordering o) const;

    rr::poisson_cache m_poisson;
    // Host generators: device staging for the kernel's output, and (stream-ordered host
    // generators) pinned staging that a host-function node copies into the caller's buffer.
    // Both are reused by later calls, whose work is ordered after the earlier calls' on
    // m_staged_stream; a regrow or a change of stream waits for that stream first.
    void*             m_staging       = nullptr;
    size_t            m_staging_bytes = 0;
    void*             m_pinned        = nullptr;
    size_t            m_pinned_bytes  = 0;
    hipStream_t       m_staged_stream = nullptr;
    bool              m_staged        = false;
    rocrand_status    stage_host(void* data, size_t bytes, void*& dev);
};

namespace rr
{
// One generate call of a fresh counter-based generator (seed, offset), resolved on the
// host without launching (c.params is left to the caller).
rocrand_status prepare_counter_call(rocrand_rng_type t, unsigned long long seed, unsigned long long offset,
                                    out_kind k, void* data, size_t n, counter_call& c, dist_shape& s,
                                    ctr_kind& kind);
// Seeded state (before any jump) of an interleaved generator's engines, and its shape for k.
rocrand_status interleaved_seed_state(rocrand_rng_type t, unsigned long long seed, out_kind k, uint32_t w[6],
                                      inter_kind& kind, dist_shape& s);

// Factories (rr_api.cpp).
rocrand_generator make_counter_generator(rocrand_rng_type t);
rocrand_generator make_interleaved_generator(rocrand_rng_type t);
rocrand_generator make_mtgp32_generator();
rocrand_generator make_mt19937_generator();
rocrand_generator make_sobol_generator(rocrand_rng_type t);
} // namespace rr
