// rr_batched.hip -- many independent Philox4x32-10 / MRG32k3a generate calls in one launch.
//
// Each stream is the output of one fresh generator's single generate call (see
// include/rocrand_batched.h).  Streams are grouped by (engine, output kind); a group is
// one persistent kernel whose blocks walk a list of tiles:
//   * Philox: a tile is up to 1024 consecutive vecs of one stream, computed exactly as
//     the per-call kernels do (rr_counter_dev.hpp), so outputs match the plain path bit
//     for bit; the stream's head and tail go to its first tile;
//   * MRG32k3a: a tile is up to 256 of a stream's engines.  A lane initialises its engine
//     (the stream's seeded state advanced by offset / S + [engine < offset mod S] steps,
//     jumped by the engine index in subsequences, as the per-call init kernel does) and
//     then writes that engine's vecs v = id, id + S, ... (rr_interleaved_dev.hpp).  The
//     tile's first engine state is computed on the host, so a lane applies one radix-table
//     entry (its offset within the tile) rather than two.  Only the engines a stream uses
//     are set up: min(S, vec_n + 1) instead of S per call.
// The plan records the group launches in a HIP graph once; a launch replays it.  The
// grid size of each group is chosen by timing a few candidates on the first plan of a
// given shape (kept in a process-wide cache).
#include "rr_counter_dev.hpp"
#include "rr_generator.hpp"
#include "rr_interleaved_dev.hpp"

#include <rocrand_batched.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

namespace rr
{
namespace
{

constexpr uint32_t CTR_TILE_VECS = 4096; // 4 wavefronts x 16 chunks of 64
constexpr uint32_t MRG_TILE      = 256;

struct batch_tile
{
    uint32_t stream;
    uint32_t count; // vecs (counter) or engines (interleaved)
    uint64_t first;
};

// One MRG32k3a stream: a fresh generator's single call.  Engine e's state is
// A^(e * 2^76) A^(offset/S + [e < offset mod S]) (seeded state); the step part is the
// same for every engine, so it is applied once on the host: base[0] = A^(offset/S) seeded,
// base[1] = A base[0], and a lane only jumps by its subsequence.
struct inter_call
{
    uint32_t    base[2][6];
    uint32_t    start;    // offset mod S
    uint32_t    shift;    // vecs between the last 128-byte line boundary and vec 0
    void*       data;
    size_t      n, mis, vec_n;
    uint32_t    head, tail;
    dist_params params;
};

// Poisson: the block samples from its copy of the tile's alias table in dynamic LDS
// (sized for the group's largest table), copied again only when a tile refers to another
// table than the one staged last (tables are shared per lambda within a plan).  The
// tile, and so `p`, is block-uniform; every thread of the block calls this.
template<class D>
__device__ __forceinline__ void stage_tile_table(dist_params& p, const uint64_t*& staged)
{
    if constexpr(dist_kind<D>::value == out_kind::poisson)
    {
        extern __shared__ uint64_t lds_packed[];
        if(p.huge)
            return;
        if(p.pt.packed != staged)
        {
            __syncthreads(); // no wavefront still samples the previous table
            for(uint32_t i = threadIdx.x; i < p.pt.size; i += blockDim.x)
                lds_packed[i] = p.pt.packed[i];
            __syncthreads();
            staged = p.pt.packed;
        }
        p.pt.packed = lds_packed;
    }
}

// The plan's descriptors (tiles, calls, tile states) are read-only for the kernels.  Read
// through the constant address space, a tile's descriptors -- block-uniform -- come in
// with scalar loads into SGPRs; as vector loads they landed in VGPRs and every tile's
// first load waited for all the stores of the block's previous tile (vector loads and
// stores retire in order on one counter).
template<class T>
__device__ __forceinline__ T read_const(const T* p, size_t i)
{
    static_assert(sizeof(T) % 4 == 0 && alignof(T) >= 4, "dword-copyable descriptor");
    using cword = const __attribute__((address_space(4))) uint32_t*;
    const cword q = (cword)(p + i);
    T           r;
    uint32_t*   d = reinterpret_cast<uint32_t*>(&r);
#pragma unroll
    for(size_t k = 0; k < sizeof(T) / 4; k++)
        d[k] = q[k];
    return r;
}

template<ctr_kind K, class D>
__global__ __launch_bounds__(256) void batch_counter_kernel(const counter_call* calls,
                                                            const batch_tile*   tiles,
                                                            uint32_t            ntiles)
{
    using namespace ctr_dev;
    const uint64_t* staged = nullptr;
    for(uint32_t ti = blockIdx.x; ti < ntiles; ti += gridDim.x)
    {
        const batch_tile tl = read_const(tiles, ti);
        // (the Poisson kernel keeps a vector copy: with the descriptor in SGPRs its
        // positions went to scratch)
        counter_call c = dist_kind<D>::value == out_kind::poisson ? calls[tl.stream] : read_const(calls, tl.stream);
        const uint64_t rmask = (uint64_t)c.ref_stride - 1;
        stage_tile_table<D>(c.params, staged);
        const size_t end = tl.first + tl.count;
        if(fast_ok<K>(c))
        {
            // the tile's 64-vec chunks on the stream's line-aligned grid (a tile other than
            // the stream's first starts on it; the first starts lane_shift vecs before 0)
            const int64_t  lo   = (int64_t)tl.first, hi = (int64_t)end;
            const int64_t  vb0  = lo - ((lo + c.lane_shift) & 63);
            const uint32_t nch  = (uint32_t)((hi - vb0 + 63) / 64);
            const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
            if(c.vec_base.sub)
            {
                // consecutive chunks per wavefront, so that each run's look-ahead block
                // serves its next chunk
                const uint32_t q  = (nch + 3) / 4;
                const uint32_t m0 = wave * q;
                if(m0 < nch)
                    fast_run<K, D, true>(c, vb0 + 64 * (int64_t)m0, std::min(q, nch - m0), lo, hi);
            }
            else
            {
                for(uint32_t m = wave; m < nch; m += 4)
                    fast_run<K, D, false>(c, vb0 + 64 * (int64_t)m, 1, lo, hi);
            }
        }
        else
        {
            for(size_t v = tl.first + threadIdx.x; v < end; v += 256)
                generic_vec<K, D>(c, v, rmask);
        }
        if(tl.first == 0 && threadIdx.x == 0)
            head_tail<K, D>(c, rmask);
    }
}

template<inter_kind K, class D>
__global__ __launch_bounds__(256) void batch_inter_kernel(const inter_call* calls,
                                                          const batch_tile* tiles,
                                                          const uint32_t*   tile_states,
                                                          uint32_t          ntiles,
                                                          const device_tables t)
{
    using namespace inter_dev;
    using out_t         = typename D::out_t;
    constexpr int IN    = D::in_w;
    constexpr int OUT   = D::out_w;
    constexpr int BYTES = OUT * (int)sizeof(out_t);
    const uint64_t* staged = nullptr;
    for(uint32_t ti = blockIdx.x; ti < ntiles; ti += gridDim.x)
    {
        const batch_tile tl = read_const(tiles, ti);
        inter_call       c  = read_const(calls, tl.stream);
        stage_tile_table<D>(c.params, staged);
        // The tile's engines start `lead` lanes into the block (the stream's first tile
        // only), so that a wavefront's stores start on a 128-byte line of the output.
        const uint32_t lead = (uint32_t)((tl.first + c.shift) & 255u);
        if(threadIdx.x < lead || threadIdx.x - lead >= tl.count)
            continue;
        const uint32_t id = (uint32_t)tl.first + threadIdx.x - lead;
        // engine (id + start) mod S; the host gave the state of the tile's first engine
        // (first + start, or first + start - S when the whole tile has wrapped), so a lane
        // is X^j from it, j = its offset from that engine; in the tile that wraps, the
        // wrapped lanes are X^(j - w) from base[1] (engine 0 with the extra step)
        const uint32_t e0 = (uint32_t)tl.first + c.start;
        const uint32_t jl = threadIdx.x - lead; // offset from the tile's first engine
        const uint32_t w  = e0 < S ? S - e0 : 0u;
        const bool     wr = e0 < S && jl >= w;
        uint32_t       w6[6];
#pragma unroll
        for(int q = 0; q < 6; q++)
            w6[q] = wr ? c.base[1][q] : read_const(tile_states, (size_t)ti * 6 + q);
        ieng<K> g;
        g.set(w6);
        const uint32_t j = wr ? jl - w : jl;
        if(j != 0u)
            g.sub_entry(t, j);
        out_t* const data     = static_cast<out_t*>(c.data);
        out_t* const vec_data = data + c.mis;
        alignas(16) out_t o[OUT];
        uint32_t     in[IN];
        // The tile's last engine takes the fewest vec-steps: that many run for every lane
        // under a scalar counter, storing from a tile-uniform base (buffer stores), and the
        // rest (at most one) per lane.
        const uint32_t     id_last = (uint32_t)tl.first + tl.count - 1;
        const uint32_t     ucnt    = id_last < c.vec_n ? (uint32_t)((c.vec_n - id_last + S - 1) / S) : 0u;
        const uint32_t     off     = id * (uint32_t)BYTES;
        constexpr uint32_t lim     = S * (uint32_t)BYTES;
        char*              base    = reinterpret_cast<char*>(vec_data);
        size_t             index   = id;
        for(uint32_t k = 0; k < ucnt; k++, index += S, base += (size_t)S * BYTES)
        {
#pragma unroll
            for(int i = 0; i < IN; i++)
                in[i] = g.next();
            D::apply(in, o, c.params);
            store_uniform_base<BYTES>(base, off, lim, o);
        }
        for(; index < c.vec_n; index += S, base += (size_t)S * BYTES)
        {
#pragma unroll
            for(int i = 0; i < IN; i++)
                in[i] = g.next();
            D::apply(in, o, c.params);
            store_uniform_base<BYTES>(base, off, lim, o);
        }
        // the engine that would produce vec index vec_n writes the head, then the tail
        // (mrg.hpp:114-153)
        if(OUT > 1 && index == c.vec_n)
        {
            if(c.head > 0)
            {
#pragma unroll
                for(int i = 0; i < IN; i++)
                    in[i] = g.next();
                D::apply(in, o, c.params);
                for(uint32_t j = 0; j < (uint32_t)OUT; j++)
                    if(j < c.head)
                        data[j] = o[j];
            }
            if(c.tail > 0)
            {
#pragma unroll
                for(int i = 0; i < IN; i++)
                    in[i] = g.next();
                D::apply(in, o, c.params);
                for(uint32_t j = 0; j < (uint32_t)OUT; j++)
                    if(j < c.tail)
                        data[c.n - c.tail + j] = o[j];
            }
        }
    }
}

// The device's radix jump (rr_interleaved_dev.hpp mrg_sub_jump) on the host, MRG32k3a.
void host_entry_apply(const uint32_t* entry, uint32_t x[6])
{
    mat3_vec_fold<32, 209>(entry, x);
    mat3_vec_fold<32, 22853>(entry + 9, x + 3);
}
void host_sub_jump(const uint32_t* radix, uint32_t e, uint32_t x[6])
{
    const uint32_t lo = e & (uint32_t)(MRG_RADIX_LO - 1);
    const uint32_t hi = (e >> MRG_RADIX_LO_BITS) & (uint32_t)(MRG_RADIX_HI - 1);
    if(lo != 0u)
        host_entry_apply(radix + lo * MRG_RADIX_WORDS, x);
    if(hi != 0u)
        host_entry_apply(radix + (MRG_RADIX_LO + hi) * MRG_RADIX_WORDS, x);
}

out_kind kind_of(rocrand_batch_kind k)
{
    switch(k)
    {
        case ROCRAND_BATCH_UINT: return out_kind::u32;
        case ROCRAND_BATCH_UNIFORM_FLOAT: return out_kind::uniform_f32;
        case ROCRAND_BATCH_UNIFORM_DOUBLE: return out_kind::uniform_f64;
        case ROCRAND_BATCH_NORMAL_FLOAT: return out_kind::normal_f32;
        case ROCRAND_BATCH_NORMAL_DOUBLE: return out_kind::normal_f64;
        case ROCRAND_BATCH_LOG_NORMAL_FLOAT: return out_kind::lognormal_f32;
        case ROCRAND_BATCH_LOG_NORMAL_DOUBLE: return out_kind::lognormal_f64;
        default: return out_kind::poisson;
    }
}

// `lds` = bytes of dynamic LDS for the group's largest Poisson table (0 otherwise);
// `tile_states` = MRG32k3a groups' per-tile engine states (6 words a tile), else null.
using launcher = hipError_t (*)(const void* calls, const batch_tile* tiles, const uint32_t* tile_states,
                                uint32_t ntiles, uint32_t blocks, uint32_t lds, const device_tables& t,
                                hipStream_t s);

template<ctr_kind K, class D>
hipError_t launch_counter_group(const void* calls, const batch_tile* tiles, const uint32_t*, uint32_t ntiles,
                                uint32_t blocks, uint32_t lds, const device_tables&, hipStream_t s)
{
    hipLaunchKernelGGL((batch_counter_kernel<K, D>), dim3(blocks), dim3(256), lds, s,
                       static_cast<const counter_call*>(calls), tiles, ntiles);
    return hipGetLastError();
}

template<inter_kind K, class D>
hipError_t launch_inter_group(const void* calls, const batch_tile* tiles, const uint32_t* tile_states,
                              uint32_t ntiles, uint32_t blocks, uint32_t lds, const device_tables& t, hipStream_t s)
{
    hipLaunchKernelGGL((batch_inter_kernel<K, D>), dim3(blocks), dim3(256), lds, s,
                       static_cast<const inter_call*>(calls), tiles, tile_states, ntiles, t);
    return hipGetLastError();
}

launcher pick_launcher(bool philox, out_kind k)
{
    constexpr word_family W = word_family::w32, M = word_family::mrg32;
    constexpr ctr_kind    P = ctr_kind::philox4x32;
    constexpr inter_kind  R = inter_kind::mrg32k3a;
#define RR_PICK(KIND)                                                                                  \
    case KIND:                                                                                         \
        return philox ? &launch_counter_group<P, dist<W, KIND>> : &launch_inter_group<R, dist<M, KIND>>;
    switch(k)
    {
        RR_PICK(out_kind::u32)
        RR_PICK(out_kind::uniform_f32)
        RR_PICK(out_kind::uniform_f64)
        RR_PICK(out_kind::normal_f32)
        RR_PICK(out_kind::normal_f64)
        RR_PICK(out_kind::lognormal_f32)
        RR_PICK(out_kind::lognormal_f64)
        RR_PICK(out_kind::poisson)
        default: return nullptr;
    }
#undef RR_PICK
}

struct group
{
    bool                      philox;
    out_kind                  kind;
    launcher                  fn;
    std::vector<counter_call> ccalls;
    std::vector<inter_call>   icalls;
    std::vector<batch_tile>   tiles;
    std::vector<uint32_t>     tile_states; // MRG32k3a: first engine state of each tile
    size_t                    bytes_out = 0;
    size_t                    calls_at  = 0; // offsets in the descriptor image
    size_t                    tiles_at  = 0;
    size_t                    states_at = 0;
    void*                     d_calls   = nullptr;
    batch_tile*               d_tiles   = nullptr;
    const uint32_t*           d_states  = nullptr;
    uint32_t                  blocks    = 2048;
    uint32_t                  tile_vecs = CTR_TILE_VECS; // Philox: vecs per tile
    uint32_t                  lds_bytes = 0; // largest Poisson table of the group, packed
};

// Philox tiles of a group: each stream's vecs in runs of `tile` whose boundaries fall on
// the stream's line-aligned chunk grid (multiples of 64 less its lane_shift), so no chunk
// straddles two tiles; at least one tile per stream (its first tile also writes head and
// tail).
std::vector<batch_tile> counter_tiles(const group& g, uint32_t tile)
{
    std::vector<batch_tile> t;
    for(uint32_t si = 0; si < (uint32_t)g.ccalls.size(); si++)
    {
        const size_t vec_n = g.ccalls[si].vec_n;
        size_t       v     = 0;
        size_t       bound = tile - g.ccalls[si].lane_shift; // lane_shift < 64 <= tile
        do
        {
            const uint32_t cnt = (uint32_t)(std::min(bound, vec_n) - v);
            t.push_back(batch_tile{si, cnt, v});
            v += cnt;
            bound += tile;
        } while(v < vec_n);
    }
    return t;
}

// Schedules chosen per (engine, kind, log2 of the group's output bytes): grid size and,
// for Philox, vecs per tile (the items each thread takes from a tile: tile / 256).
struct schedule
{
    uint32_t blocks, tile_vecs;
};
std::mutex                                    g_tune_mtx;
std::map<std::tuple<int, int, int>, schedule> g_tuned;

std::tuple<int, int, int> tune_key(const group& g)
{
    const int lg = g.bytes_out ? (int)std::log2((double)g.bytes_out) : 0;
    return std::make_tuple(g.philox ? 1 : 0, (int)g.kind, lg);
}

// A batch on the host: groups, the distinct Poisson tables, and the layout of the one
// descriptor image (tables, then each group's calls and tiles) that goes to the device.
struct batch_layout
{
    std::vector<group>                     groups;
    std::vector<std::vector<double>>       prob;
    std::vector<std::vector<unsigned int>> alias;
    std::vector<std::vector<uint64_t>>     packed;
    std::vector<size_t>                    prob_at, alias_at, packed_at;
    size_t                                 bytes = 0;

    static size_t align(size_t x) { return (x + 255) / 256 * 256; }

    void place()
    {
        size_t at = 0;
        for(size_t i = 0; i < prob.size(); i++)
        {
            prob_at.push_back(at);
            at = align(at + prob[i].size() * sizeof(double));
            alias_at.push_back(at);
            at = align(at + alias[i].size() * sizeof(unsigned int));
            packed_at.push_back(at);
            at = align(at + packed[i].size() * sizeof(uint64_t));
        }
        for(auto& g : groups)
        {
            g.calls_at = at;
            at = align(at + (g.philox ? g.ccalls.size() * sizeof(counter_call) : g.icalls.size() * sizeof(inter_call)));
            g.tiles_at = at;
            at         = align(at + g.tiles.size() * sizeof(batch_tile));
            g.states_at = at;
            at          = align(at + g.tile_states.size() * sizeof(uint32_t));
        }
        bytes = std::max<size_t>(at, 256);
    }

    // Writes the image for device address `dev` into `host` (bytes long): Poisson table
    // references (held as table indices until now) become device pointers.
    void pack(char* host, char* dev)
    {
        for(size_t i = 0; i < prob.size(); i++)
        {
            memcpy(host + prob_at[i], prob[i].data(), prob[i].size() * sizeof(double));
            memcpy(host + alias_at[i], alias[i].data(), alias[i].size() * sizeof(unsigned int));
            memcpy(host + packed_at[i], packed[i].data(), packed[i].size() * sizeof(uint64_t));
        }
        auto fix = [&](dist_params& prm) {
            if(prm.pt.size == 0)
                return;
            const size_t t = reinterpret_cast<uintptr_t>(prm.pt.prob);
            prm.pt.prob    = reinterpret_cast<const double*>(dev + prob_at[t]);
            prm.pt.alias   = reinterpret_cast<const unsigned int*>(dev + alias_at[t]);
            prm.pt.packed  = reinterpret_cast<const uint64_t*>(dev + packed_at[t]);
        };
        for(auto& g : groups)
        {
            if(g.philox)
            {
                for(auto& c : g.ccalls)
                    fix(c.params);
                memcpy(host + g.calls_at, g.ccalls.data(), g.ccalls.size() * sizeof(counter_call));
            }
            else
            {
                for(auto& c : g.icalls)
                    fix(c.params);
                memcpy(host + g.calls_at, g.icalls.data(), g.icalls.size() * sizeof(inter_call));
            }
            memcpy(host + g.tiles_at, g.tiles.data(), g.tiles.size() * sizeof(batch_tile));
            memcpy(host + g.states_at, g.tile_states.data(), g.tile_states.size() * sizeof(uint32_t));
            g.d_calls  = dev + g.calls_at;
            g.d_tiles  = reinterpret_cast<batch_tile*>(dev + g.tiles_at);
            g.d_states = g.tile_states.empty() ? nullptr : reinterpret_cast<const uint32_t*>(dev + g.states_at);
        }
    }
};

hipError_t run_groups(const std::vector<group>& groups, const device_tables& tables, hipStream_t s)
{
    for(const auto& g : groups)
    {
        if(g.tiles.empty())
            continue;
        const uint32_t blocks = std::min<uint32_t>(g.blocks, (uint32_t)g.tiles.size());
        hipError_t     e      = g.fn(g.d_calls, g.d_tiles, g.d_states, (uint32_t)g.tiles.size(), blocks, g.lds_bytes, tables, s);
        if(e != hipSuccess)
            return e;
    }
    return hipSuccess;
}

// Streams and events of a fork/join capture: one side stream per group.
struct fork_join
{
    std::vector<hipStream_t> side;
    std::vector<hipEvent_t>  done;
    hipEvent_t               fork = nullptr;

    bool create(size_t n)
    {
        if(hipEventCreateWithFlags(&fork, hipEventDisableTiming) != hipSuccess)
            return false;
        side.assign(n, nullptr);
        done.assign(n, nullptr);
        for(size_t i = 0; i < n; i++)
            if(hipStreamCreateWithFlags(&side[i], hipStreamNonBlocking) != hipSuccess
               || hipEventCreateWithFlags(&done[i], hipEventDisableTiming) != hipSuccess)
                return false;
        return true;
    }
    ~fork_join()
    {
        for(hipStream_t q : side)
            if(q)
                (void)hipStreamDestroy(q);
        for(hipEvent_t e : done)
            if(e)
                (void)hipEventDestroy(e);
        if(fork)
            (void)hipEventDestroy(fork);
    }
};

// run_groups under capture, each group on its side stream between fork and join.
hipError_t run_groups_forked(const std::vector<group>& groups, const device_tables& tables, hipStream_t cap,
                             const fork_join& fj)
{
    hipError_t e = hipEventRecord(fj.fork, cap);
    for(size_t i = 0; i < groups.size() && e == hipSuccess; i++)
    {
        const group& g = groups[i];
        if(g.tiles.empty())
            continue;
        const uint32_t blocks = std::min<uint32_t>(g.blocks, (uint32_t)g.tiles.size());
        if((e = hipStreamWaitEvent(fj.side[i], fj.fork, 0)) != hipSuccess
           || (e = g.fn(g.d_calls, g.d_tiles, g.d_states, (uint32_t)g.tiles.size(), blocks, g.lds_bytes, tables, fj.side[i])) != hipSuccess
           || (e = hipEventRecord(fj.done[i], fj.side[i])) != hipSuccess
           || (e = hipStreamWaitEvent(cap, fj.done[i], 0)) != hipSuccess)
            break;
    }
    return e;
}

// Device blocks with pinned host twins for the one-shot call's descriptor images,
// recycled once the call that used them has completed (its event has fired).  Blocks
// come in power-of-two sizes from 64 KiB and stay allocated for the process lifetime,
// so a steady stream of calls allocates nothing and never waits for the device.
class staging_pool
{
public:
    struct block
    {
        char*  dev  = nullptr;
        char*  host = nullptr;
        size_t size = 0;
        int    device = 0;
    };

    hipError_t acquire(int device, size_t bytes, block& b, hipEvent_t& ev)
    {
        size_t size = 64 << 10;
        while(size < bytes)
            size *= 2;
        std::lock_guard<std::mutex> lock(mu_);
        reclaim();
        auto& fl = free_[std::make_pair(device, size)];
        if(!fl.empty())
        {
            b = fl.back();
            fl.pop_back();
        }
        else
        {
            b        = block{};
            b.size   = size;
            b.device = device;
            hipError_t e = hipMalloc(reinterpret_cast<void**>(&b.dev), size);
            if(e != hipSuccess)
                return e;
            e = hipHostMalloc(reinterpret_cast<void**>(&b.host), size, hipHostMallocDefault);
            if(e != hipSuccess)
            {
                (void)hipFree(b.dev);
                return e;
            }
        }
        auto& evs = events_[device];
        if(!evs.empty())
        {
            ev = evs.back();
            evs.pop_back();
            return hipSuccess;
        }
        return hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    }

    // Hands the block back once `ev` (recorded after the block's last use) has fired.
    void release_after(const block& b, hipEvent_t ev)
    {
        std::lock_guard<std::mutex> lock(mu_);
        pending_.push_back(std::make_pair(b, ev));
    }

private:
    void reclaim()
    {
        size_t keep = 0;
        for(size_t i = 0; i < pending_.size(); i++)
        {
            if(hipEventQuery(pending_[i].second) == hipSuccess)
            {
                const block& b = pending_[i].first;
                free_[std::make_pair(b.device, b.size)].push_back(b);
                events_[b.device].push_back(pending_[i].second);
            }
            else
                pending_[keep++] = pending_[i];
        }
        pending_.resize(keep);
    }

    std::mutex                                        mu_;
    std::map<std::pair<int, size_t>, std::vector<block>> free_;
    std::map<int, std::vector<hipEvent_t>>            events_;
    std::vector<std::pair<block, hipEvent_t>>         pending_;
};

staging_pool& pool()
{
    static staging_pool* p = new staging_pool(); // lives to process exit
    return *p;
}

// Descriptor image of rocrand_generate_batched_noalloc: a fixed array in device memory
// (each GPU the library runs on holds its own copy until the process ends) instead of a
// pooled block.  Every call on a GPU reuses that GPU's copy, so calls that overlap on one
// GPU (different streams, or the next call before the previous one's kernels have read
// their descriptors on another stream) race; calls in order on one stream do not.
constexpr size_t NOALLOC_SCRATCH_BYTES = 16u << 20;
static __device__ __align__(256) char scratch[NOALLOC_SCRATCH_BYTES];

// Device address of `scratch` on the current device (in host code the array's name is a
// host-side stand-in, not a device address), looked up once per device by the first
// no-allocation call there.
char* scratch_address()
{
    static std::mutex m;
    static char*      addr[64] = {};
    int               dev      = 0;
    if(hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64)
        return nullptr;
    std::lock_guard<std::mutex> lock(m);
    if(addr[dev] == nullptr)
    {
        void* p = nullptr;
        if(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess)
            return nullptr;
        addr[dev] = static_cast<char*>(p);
    }
    return addr[dev];
}

} // namespace
} // namespace rr

using namespace rr;

struct rocrand_batch_plan_st
{
    hipStream_t        stream = 0;
    std::vector<group> groups;
    char*              d_image = nullptr;
    device_tables      tables{};
    hipGraph_t         graph    = nullptr;
    hipGraphExec_t     exec     = nullptr;
    hipEvent_t         done     = nullptr; // recorded after each launch
    bool               launched = false;

    ~rocrand_batch_plan_st()
    {
        if(done)
            (void)hipEventDestroy(done);
        if(exec)
            (void)hipGraphExecDestroy(exec);
        if(graph)
            (void)hipGraphDestroy(graph);
        if(d_image)
            (void)hipFree(d_image);
    }
};

namespace
{

// Validates the streams and lays out the batch on the host (no device calls).
rocrand_status build(batch_layout& L, const rocrand_batch_stream* streams, size_t count)
{
    // Poisson alias tables, one per distinct lambda (poisson.hpp:110-161, discrete.hpp:189-265)
    std::map<double, alias_table>         tables;
    std::map<std::pair<int, int>, size_t> gid; // (philox, kind) -> group index
    for(size_t i = 0; i < count; i++)
    {
        const rocrand_batch_stream& b = streams[i];
        const bool philox = b.rng_type == ROCRAND_RNG_PSEUDO_PHILOX4_32_10;
        if(!philox && b.rng_type != ROCRAND_RNG_PSEUDO_MRG32K3A)
            return ROCRAND_STATUS_TYPE_ERROR;
        if((int)b.kind < 0 || (int)b.kind > 7)
            return ROCRAND_STATUS_TYPE_ERROR;
        if(b.n > 0 && b.output == nullptr)
            return ROCRAND_STATUS_OUT_OF_RANGE;
        const out_kind k = kind_of(b.kind);
        dist_params    prm{};
        prm.mf = (float)b.a;
        prm.sf = (float)b.b;
        prm.md = b.a;
        prm.sd = b.b;
        if(k == out_kind::poisson)
        {
            if(!(b.a > 0.0))
                return ROCRAND_STATUS_OUT_OF_RANGE; // rocrand.cpp:371-374
            prm.lambda      = b.a;
            prm.sqrt_lambda = std::sqrt(b.a);
            prm.huge        = b.a > 4000.0 ? 1 : 0;
            if(!prm.huge)
            {
                auto it = tables.find(b.a);
                if(it == tables.end())
                {
                    std::vector<double>       prob;
                    std::vector<unsigned int> alias;
                    unsigned int              off  = 0;
                    const unsigned int        size = poisson_table(b.a, off, prob, alias);
                    alias_table               t{};
                    t.size   = size;
                    t.offset = off;
                    // the table's index until the image is packed
                    t.prob = reinterpret_cast<const double*>(static_cast<uintptr_t>(L.prob.size()));
                    L.packed.push_back(pack_alias(prob, alias));
                    L.prob.push_back(std::move(prob));
                    L.alias.push_back(std::move(alias));
                    it = tables.emplace(b.a, t).first;
                }
                prm.pt = it->second;
            }
        }
        auto key = std::make_pair(philox ? 1 : 0, (int)k);
        auto git = gid.find(key);
        if(git == gid.end())
        {
            group g;
            g.philox = philox;
            g.kind   = k;
            g.fn     = pick_launcher(philox, k);
            if(!g.fn)
                return ROCRAND_STATUS_TYPE_ERROR;
            L.groups.push_back(std::move(g));
            git = gid.emplace(key, L.groups.size() - 1).first;
        }
        group&         g  = L.groups[git->second];
        if(k == out_kind::poisson && !prm.huge)
            g.lds_bytes = std::max<uint32_t>(g.lds_bytes, prm.pt.size * (uint32_t)sizeof(uint64_t));
        const uint32_t si = (uint32_t)(philox ? g.ccalls.size() : g.icalls.size());
        dist_shape     s{};
        if(philox)
        {
            counter_call   c{};
            ctr_kind       ck;
            rocrand_status st = prepare_counter_call(b.rng_type, b.seed, b.offset, k, b.output, b.n, c, s, ck);
            if(st != ROCRAND_STATUS_SUCCESS)
                return st;
            c.params = prm;
            g.ccalls.push_back(c);
        }
        else
        {
            inter_call     c{};
            inter_kind     ik;
            uint32_t       seeded[6];
            rocrand_status st = interleaved_seed_state(b.rng_type, b.seed, k, seeded, ik, s);
            if(st != ROCRAND_STATUS_SUCCESS)
                return st;
            // base[0] = A^(offset / S) seeded, base[1] = A base[0] (host step tables A^(2^i))
            const uint32_t* steps = host_mrg_table(32, 0);
            uint64_t        v     = b.offset / INTERLEAVED_ENGINES;
            for(int i = 0; v > 0; v >>= 1, i++)
                if(v & 1)
                {
                    mat3_vec_fold<32, 209>(steps + (i * 2 + 0) * 9, seeded);
                    mat3_vec_fold<32, 22853>(steps + (i * 2 + 1) * 9, seeded + 3);
                }
            std::copy(seeded, seeded + 6, c.base[0]);
            mat3_vec_fold<32, 209>(steps + 0 * 9, seeded);
            mat3_vec_fold<32, 22853>(steps + 1 * 9, seeded + 3);
            std::copy(seeded, seeded + 6, c.base[1]);
            const uintptr_t addr = reinterpret_cast<uintptr_t>(b.output);
            c.start    = (uint32_t)(b.offset % INTERLEAVED_ENGINES);
            c.data     = b.output;
            c.n        = b.n;
            c.mis      = (s.out_w - (addr / s.esz) % s.out_w) % s.out_w;
            c.head     = (uint32_t)std::min<size_t>(b.n, c.mis);
            c.tail     = (uint32_t)((b.n - c.head) % s.out_w);
            c.vec_n    = (b.n - c.head) / s.out_w;
            c.params   = prm;
            {
                const uintptr_t vaddr  = addr + (uintptr_t)c.mis * s.esz;
                const unsigned  vbytes = s.out_w * s.esz;
                c.shift = 128u % vbytes == 0 ? (uint32_t)((vaddr / vbytes) % (128u / vbytes)) : 0u;
            }
            g.icalls.push_back(c);
            const bool   extra = s.out_w > 1 && (c.head > 0 || c.tail > 0);
            const size_t lanes = std::min<size_t>(INTERLEAVED_ENGINES, c.vec_n + (extra ? 1 : 0));
            // Tile states: X^(first + start) base[0] while the tile's first engine has not
            // wrapped, X^(first + start - S) base[1] after; consecutive tiles of a region
            // differ by X^256 (radix entry 256).
            const uint32_t* radix = host_mrg_radix(32);
            uint32_t        x[6];
            bool            have = false, wrapped = false;
            // Tiles on the grid shifted by c.shift: tile t covers engines [256 t - shift,
            // 256 (t + 1) - shift) of the used ones, its first (active) engine `e`.
            size_t e = 0, prev = 0;
            for(size_t t = 0; e < lanes; t++)
            {
                const size_t end = std::min<size_t>((t + 1) * MRG_TILE - c.shift, lanes); // > e
                g.tiles.push_back(batch_tile{si, (uint32_t)(end - e), e});
                const uint64_t e0 = e + c.start;
                if(!have || (!wrapped && e0 >= INTERLEAVED_ENGINES))
                {
                    wrapped = e0 >= INTERLEAVED_ENGINES;
                    std::copy(c.base[wrapped ? 1 : 0], c.base[wrapped ? 1 : 0] + 6, x);
                    host_sub_jump(radix, (uint32_t)(wrapped ? e0 - INTERLEAVED_ENGINES : e0), x);
                    have = true;
                }
                else // e - prev engines further: 256, or 256 - shift after the first tile
                    host_entry_apply(radix + (e - prev) * MRG_RADIX_WORDS, x);
                g.tile_states.insert(g.tile_states.end(), x, x + 6);
                prev = e;
                e    = end;
            }
        }
        g.bytes_out += b.n * (size_t)s.esz;
    }
    // Philox tiles once the group's size (and so its tuned tile size, if any) is known
    for(auto& g : L.groups)
        if(g.philox)
        {
            {
                std::lock_guard<std::mutex> lock(g_tune_mtx);
                auto                        it = g_tuned.find(tune_key(g));
                if(it != g_tuned.end())
                {
                    g.blocks    = it->second.blocks;
                    g.tile_vecs = it->second.tile_vecs;
                }
            }
            g.tiles = counter_tiles(g, g.tile_vecs);
        }
        else
        {
            std::lock_guard<std::mutex> lock(g_tune_mtx);
            auto                        it = g_tuned.find(tune_key(g));
            if(it != g_tuned.end())
                g.blocks = it->second.blocks;
        }
    L.place();
    return ROCRAND_STATUS_SUCCESS;
}

// Two timing events, destroyed on every return path.
struct event_pair
{
    hipEvent_t a = nullptr, b = nullptr;
    bool       create() { return hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess; }
    ~event_pair()
    {
        if(a)
            (void)hipEventDestroy(a);
        if(b)
            (void)hipEventDestroy(b);
    }
};

// Device memory released on every return path.
struct device_block
{
    void* p = nullptr;
    ~device_block()
    {
        if(p)
            (void)hipFree(p);
    }
};

// Chooses each group's schedule: the cached choice for its shape, else the candidate with
// the lowest median of TUNE_RUNS timed launches (after one untimed launch) on the plan's
// own outputs (the values written are the same for every schedule).  Candidates: grid
// sizes, and for Philox groups tiles of 1024, 4096 and 16384 vecs (4, 16 or 64 vecs per
// thread), each tiling uploaded to a scratch block for its timings.  Sets `relayout` when
// a group's chosen tile size differs from the one the plan was laid out with.
constexpr int TUNE_RUNS = 3;

rocrand_status tune(rocrand_batch_plan_st& p, bool& relayout)
{
    static const uint32_t grids[]       = {1024, 2048, 4096, 8192, 16384};
    static const uint32_t ctr_tiles[]   = {1024, CTR_TILE_VECS, 16384};
    relayout                            = false;
    for(auto& g : p.groups)
    {
        if(g.tiles.empty())
            continue;
        const auto key = tune_key(g);
        {
            std::lock_guard<std::mutex> lock(g_tune_mtx);
            auto                        it = g_tuned.find(key);
            if(it != g_tuned.end())
            {
                g.blocks = it->second.blocks;
                relayout = relayout || (g.philox && it->second.tile_vecs != g.tile_vecs);
                continue;
            }
        }
        event_pair ev;
        if(!ev.create())
            return ROCRAND_STATUS_INTERNAL_ERROR;
        float    best_ms = 1e30f;
        schedule best{g.blocks, g.tile_vecs};
        const size_t ntile_cands = g.philox ? sizeof(ctr_tiles) / sizeof(ctr_tiles[0]) : 1;
        for(size_t tc = 0; tc < ntile_cands; tc++)
        {
            const uint32_t    tile   = g.philox ? ctr_tiles[tc] : g.tile_vecs;
            const batch_tile* tiles  = g.d_tiles;
            uint32_t          ntiles = (uint32_t)g.tiles.size();
            device_block      tmp;
            if(g.philox && tile != g.tile_vecs)
            {
                const std::vector<batch_tile> tt = counter_tiles(g, tile);
                if(hipMalloc(&tmp.p, tt.size() * sizeof(batch_tile)) != hipSuccess)
                    return ROCRAND_STATUS_ALLOCATION_FAILED;
                if(hipMemcpy(tmp.p, tt.data(), tt.size() * sizeof(batch_tile), hipMemcpyHostToDevice) != hipSuccess)
                    return ROCRAND_STATUS_INTERNAL_ERROR;
                tiles  = static_cast<const batch_tile*>(tmp.p);
                ntiles = (uint32_t)tt.size();
            }
            for(uint32_t c : grids)
            {
                const uint32_t blocks = std::min<uint32_t>(c, ntiles);
                float          ms[TUNE_RUNS];
                for(int rep = -1; rep < TUNE_RUNS; rep++)
                {
                    if(rep >= 0 && hipEventRecord(ev.a, p.stream) != hipSuccess)
                        return ROCRAND_STATUS_INTERNAL_ERROR;
                    if(g.fn(g.d_calls, tiles, g.d_states, ntiles, blocks, g.lds_bytes, p.tables, p.stream) != hipSuccess)
                        return ROCRAND_STATUS_LAUNCH_FAILURE;
                    if(rep >= 0
                       && (hipEventRecord(ev.b, p.stream) != hipSuccess || hipEventSynchronize(ev.b) != hipSuccess
                           || hipEventElapsedTime(&ms[rep], ev.a, ev.b) != hipSuccess))
                        return ROCRAND_STATUS_INTERNAL_ERROR;
                }
                std::sort(ms, ms + TUNE_RUNS);
                if(ms[TUNE_RUNS / 2] < best_ms)
                {
                    best_ms = ms[TUNE_RUNS / 2];
                    best    = schedule{c, tile};
                }
                if(blocks == ntiles)
                    break; // larger grids change nothing
            }
            // the scratch tiling's kernels must finish before its block is freed
            if(tmp.p && hipStreamSynchronize(p.stream) != hipSuccess)
                return ROCRAND_STATUS_INTERNAL_ERROR;
        }
        g.blocks = best.blocks;
        relayout = relayout || best.tile_vecs != g.tile_vecs;
        std::lock_guard<std::mutex> lock(g_tune_mtx);
        g_tuned[key] = best;
    }
    return ROCRAND_STATUS_SUCCESS;
}

} // namespace

extern "C" {

rocrand_status ROCRANDAPI rocrand_batch_plan_create(rocrand_batch_plan*         plan,
                                                    const rocrand_batch_stream* streams,
                                                    size_t                      count,
                                                    hipStream_t                 stream)
{
    if(plan == nullptr || (streams == nullptr && count > 0))
        return ROCRAND_STATUS_OUT_OF_RANGE;
    *plan = nullptr;
    rocrand_batch_plan_st* p = nullptr;
    try
    {
        p         = new rocrand_batch_plan_st();
        p->stream = stream;
        // lay the batch out (with the cached schedules) and upload its descriptor image
        auto lay_out = [&]() {
            batch_layout   L;
            rocrand_status r = build(L, streams, count);
            if(r != ROCRAND_STATUS_SUCCESS)
                return r;
            std::vector<char> image(L.bytes);
            if(hipMalloc(reinterpret_cast<void**>(&p->d_image), L.bytes) != hipSuccess)
            {
                p->d_image = nullptr;
                return ROCRAND_STATUS_ALLOCATION_FAILED;
            }
            L.pack(image.data(), p->d_image);
            p->groups = std::move(L.groups);
            return hipMemcpy(p->d_image, image.data(), L.bytes, hipMemcpyHostToDevice) == hipSuccess
                       ? ROCRAND_STATUS_SUCCESS
                       : ROCRAND_STATUS_INTERNAL_ERROR;
        };
        rocrand_status st = lay_out();
        if(st == ROCRAND_STATUS_SUCCESS && get_device_tables(p->tables) != hipSuccess)
            st = ROCRAND_STATUS_INTERNAL_ERROR;
        bool relayout = false;
        if(st == ROCRAND_STATUS_SUCCESS)
            st = tune(*p, relayout);
        if(st == ROCRAND_STATUS_SUCCESS && relayout)
        {
            // a tuned tile size differs from the layout's: lay out again (every timing run
            // has completed: each ended with an event synchronisation)
            (void)hipFree(p->d_image);
            p->d_image = nullptr;
            st         = lay_out();
        }
        if(st == ROCRAND_STATUS_SUCCESS && hipEventCreateWithFlags(&p->done, hipEventDisableTiming) != hipSuccess)
            st = ROCRAND_STATUS_INTERNAL_ERROR;
        if(st == ROCRAND_STATUS_SUCCESS)
        {
            // Record the group launches once.  The null stream cannot be captured: record
            // on a private stream (ordered after the plan's stream) and launch on the
            // plan's stream.
            hipStream_t cap = stream;
            if(cap == 0 && hipStreamCreateWithFlags(&cap, hipStreamNonBlocking) != hipSuccess)
                st = ROCRAND_STATUS_INTERNAL_ERROR;
            if(st == ROCRAND_STATUS_SUCCESS && cap != stream && hipStreamSynchronize(stream) != hipSuccess)
                st = ROCRAND_STATUS_INTERNAL_ERROR;
            // Each group is captured on its own stream forked from and joined back to the
            // capture stream, so the graph's kernel nodes are independent: groups bound by
            // arithmetic (normals, Poisson) and groups bound by stores (uniforms) may run
            // side by side and one group's tail overlaps the next.
            fork_join fj;
            if(st == ROCRAND_STATUS_SUCCESS && !fj.create(p->groups.size()))
                st = ROCRAND_STATUS_INTERNAL_ERROR;
            if(st == ROCRAND_STATUS_SUCCESS && hipStreamBeginCapture(cap, hipStreamCaptureModeRelaxed) != hipSuccess)
                st = ROCRAND_STATUS_INTERNAL_ERROR;
            else if(st == ROCRAND_STATUS_SUCCESS)
            {
                const hipError_t le = run_groups_forked(p->groups, p->tables, cap, fj);
                const hipError_t ce = hipStreamEndCapture(cap, &p->graph);
                if(le != hipSuccess || ce != hipSuccess
                   || hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0) != hipSuccess)
                    st = ROCRAND_STATUS_INTERNAL_ERROR;
            }
            if(cap != stream && cap != 0)
                (void)hipStreamDestroy(cap);
        }
        if(st != ROCRAND_STATUS_SUCCESS)
        {
            (void)hipStreamSynchronize(stream);
            delete p;
            return st;
        }
    }
    catch(const std::bad_alloc&)
    {
        delete p;
        return ROCRAND_STATUS_ALLOCATION_FAILED;
    }
    catch(rocrand_status st)
    {
        delete p;
        return st;
    }
    *plan = p;
    return ROCRAND_STATUS_SUCCESS;
}

rocrand_status ROCRANDAPI rocrand_batch_plan_launch(rocrand_batch_plan plan)
{
    if(plan == nullptr)
        return ROCRAND_STATUS_NOT_CREATED;
    if(plan->groups.empty())
        return ROCRAND_STATUS_SUCCESS;
    if(hipGraphLaunch(plan->exec, plan->stream) != hipSuccess || hipEventRecord(plan->done, plan->stream) != hipSuccess)
        return ROCRAND_STATUS_LAUNCH_FAILURE;
    plan->launched = true;
    return ROCRAND_STATUS_SUCCESS;
}

rocrand_status ROCRANDAPI rocrand_batch_plan_destroy(rocrand_batch_plan plan)
{
    if(plan == nullptr)
        return ROCRAND_STATUS_NOT_CREATED;
    // wait for this plan's last launch only (later work on the stream may still run)
    if(plan->launched)
        (void)hipEventSynchronize(plan->done);
    delete plan;
    return ROCRAND_STATUS_SUCCESS;
}

rocrand_status ROCRANDAPI rocrand_generate_batched(const rocrand_batch_stream* streams, size_t count,
                                                   hipStream_t stream)
{
    if(streams == nullptr && count > 0)
        return ROCRAND_STATUS_OUT_OF_RANGE;
    try
    {
        batch_layout   L;
        rocrand_status st = build(L, streams, count);
        if(st != ROCRAND_STATUS_SUCCESS)
            return st;
        if(L.groups.empty())
            return ROCRAND_STATUS_SUCCESS;
        device_tables tables{};
        int           device = 0;
        if(get_device_tables(tables) != hipSuccess || hipGetDevice(&device) != hipSuccess)
            return ROCRAND_STATUS_INTERNAL_ERROR;
        // Descriptors go through a pooled pinned buffer and a pooled device block, both
        // released to the pool when the event recorded after the kernels has fired: the
        // call returns without waiting for the device.
        staging_pool::block blk;
        hipEvent_t          ev = nullptr;
        if(pool().acquire(device, L.bytes, blk, ev) != hipSuccess)
            return ROCRAND_STATUS_ALLOCATION_FAILED;
        L.pack(blk.host, blk.dev);
        const bool ok = hipMemcpyAsync(blk.dev, blk.host, L.bytes, hipMemcpyHostToDevice, stream) == hipSuccess
                        && run_groups(L.groups, tables, stream) == hipSuccess;
        if(hipEventRecord(ev, stream) != hipSuccess)
        {
            // the block cannot be tracked: wait for the stream, then recycle it
            (void)hipStreamSynchronize(stream);
            (void)hipEventRecord(ev, 0);
            (void)hipEventSynchronize(ev);
        }
        pool().release_after(blk, ev);
        if(!ok)
            return ROCRAND_STATUS_LAUNCH_FAILURE;
    }
    catch(const std::bad_alloc&)
    {
        return ROCRAND_STATUS_ALLOCATION_FAILED;
    }
    catch(rocrand_status st)
    {
        return st;
    }
    return ROCRAND_STATUS_SUCCESS;
}

rocrand_status ROCRANDAPI rocrand_generate_batched_noalloc(const rocrand_batch_stream* streams, size_t count,
                                                           hipStream_t stream)
{
    if(streams == nullptr && count > 0)
        return ROCRAND_STATUS_OUT_OF_RANGE;
    try
    {
        batch_layout   L;
        rocrand_status st = build(L, streams, count);
        if(st != ROCRAND_STATUS_SUCCESS)
            return st;
        if(L.groups.empty())
            return ROCRAND_STATUS_SUCCESS;
        if(L.bytes > NOALLOC_SCRATCH_BYTES)
            return ROCRAND_STATUS_OUT_OF_RANGE; // descriptors beyond the fixed scratch array
        device_tables tables{};
        char*         dev = nullptr;
        if(get_device_tables(tables) != hipSuccess || (dev = scratch_address()) == nullptr)
            return ROCRAND_STATUS_INTERNAL_ERROR;
        // Descriptors are packed in host memory and copied into the scratch array on the
        // stream; a copy from pageable memory has read the host buffer when it returns.
        std::vector<char> image(L.bytes);
        L.pack(image.data(), dev);
        if(hipMemcpyAsync(dev, image.data(), L.bytes, hipMemcpyHostToDevice, stream) != hipSuccess
           || run_groups(L.groups, tables, stream) != hipSuccess)
            return ROCRAND_STATUS_LAUNCH_FAILURE;
    }
    catch(const std::bad_alloc&)
    {
        return ROCRAND_STATUS_ALLOCATION_FAILED;
    }
    catch(rocrand_status st)
    {
        return st;
    }
    return ROCRAND_STATUS_SUCCESS;
}

} // extern "C"
