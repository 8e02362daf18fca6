// rr_serving.cpp -- the continuous-batching request server of include/rocrand_serving.h.
//
// The loop's state lives here: the HBM arena (first fit over one allocation, freed ranges
// coalesced), the FIFO queue, the in-flight batches with their completion events, and the
// sessions' generators.  Each stream holds at most one batch in flight, and a poll
// issues everything admissible as one batch on a free stream, so that requests arriving
// while the streams are busy are coalesced into the next batch.  A batch's
// Philox4x32-10 requests (fresh or session) and fresh MRG32k3a requests are one
// rocrand_generate_batched call, every MRG32k3a session or Sobol64 request one plain call of
// its generator.  Ordering: a request's output depends only on its own call sequence, so
// batches may run concurrently on different streams; an MRG32k3a session, whose engine
// states are read and rewritten on the device by each of its requests, orders its next
// request after its previous one (an event when the stream changes).
#include <rocrand.h>
#include <rocrand_batched.h>
#include <rocrand_serving.h>

#include <hip/hip_runtime.h>

#include "rr_arena.hpp"

#include <deque>
#include <map>
#include <new>
#include <unordered_map>
#include <vector>

namespace
{

constexpr size_t ALIGN = 256;

// Arena bytes of an output: whole 256-byte units, at least one (so an empty request
// still holds a distinct address until it is reported).
size_t slot(size_t bytes) { return bytes == 0 ? ALIGN : (bytes + ALIGN - 1) / ALIGN * ALIGN; }

size_t elem_size(rocrand_serve_dtype d) { return d == ROCRAND_SERVE_DOUBLE ? 8 : 4; }

struct queued
{
    unsigned long long    ticket;
    rocrand_serve_request r;
    size_t                bytes;
};

struct issued
{
    unsigned long long ticket;
    size_t             start, bytes;
};

struct batch
{
    hipEvent_t          done   = nullptr;
    unsigned            stream = 0;
    std::vector<issued> reqs;
};

struct session
{
    rocrand_rng_type   type = ROCRAND_RNG_PSEUDO_PHILOX4_32_10;
    unsigned long long seed = 0;
    unsigned int       dims = 1;
    rocrand_generator  gen  = nullptr; // MRG32k3a, Sobol64
    unsigned long long pos  = 0;       // Philox4x32-10: words consumed
    hipEvent_t         last = nullptr; // MRG32k3a, Sobol64: after its latest request
    hipStream_t        last_stream = nullptr;
};

rocrand_status generate_kind(rocrand_generator g, rocrand_serve_dtype d, void* out, size_t n)
{
    switch(d)
    {
        case ROCRAND_SERVE_UINT: return rocrand_generate(g, static_cast<unsigned int*>(out), n);
        case ROCRAND_SERVE_FLOAT: return rocrand_generate_uniform(g, static_cast<float*>(out), n);
        default: return rocrand_generate_uniform_double(g, static_cast<double*>(out), n);
    }
}

rocrand_batch_kind batch_kind(rocrand_serve_dtype d)
{
    return d == ROCRAND_SERVE_UINT    ? ROCRAND_BATCH_UINT
           : d == ROCRAND_SERVE_FLOAT ? ROCRAND_BATCH_UNIFORM_FLOAT
                                      : ROCRAND_BATCH_UNIFORM_DOUBLE;
}

} // namespace

struct rocrand_server_st
{
    char*                                         mem    = nullptr;
    size_t                                        budget = 0;
    std::vector<hipStream_t>                      streams;
    std::vector<char>                             busy; // a batch of this stream is in flight
    int                                           keep = 0;
    rr::arena                                     free_mem{0};
    std::deque<queued>                            queue;
    std::deque<batch>                             inflight;
    std::vector<hipEvent_t>                       spare;
    std::unordered_map<unsigned long long, session> sessions;
    std::map<unsigned int, rocrand_generator>     sobol_fresh; // per dimension count
    std::unordered_map<unsigned long long, issued> kept;
    std::deque<issued>                            done;     // retired, not yet reported
    std::vector<issued>                           reported; // reported by the latest poll
    unsigned long long                            next_ticket = 1;

    hipEvent_t event()
    {
        if(!spare.empty())
        {
            hipEvent_t e = spare.back();
            spare.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        if(hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
            return nullptr;
        return e;
    }

    void retire()
    {
        for(auto it = inflight.begin(); it != inflight.end();)
        {
            if(hipEventQuery(it->done) != hipSuccess)
            {
                ++it;
                continue;
            }
            for(const issued& q : it->reqs)
                done.push_back(q);
            busy[it->stream] = 0;
            spare.push_back(it->done);
            it = inflight.erase(it);
        }
    }

    // A stream with no batch in flight, or streams.size().
    unsigned free_stream() const
    {
        unsigned i = 0;
        while(i < streams.size() && busy[i])
            i++;
        return i;
    }

    rocrand_status issue(const std::vector<queued>& reqs, const std::vector<size_t>& starts, unsigned si)
    {
        hipStream_t                       s = streams[si];
        std::vector<rocrand_batch_stream> bs;
        for(size_t i = 0; i < reqs.size(); i++)
        {
            const rocrand_serve_request& r   = reqs[i].r;
            void*                        out = mem + starts[i];
            if(r.rng_type == ROCRAND_RNG_PSEUDO_PHILOX4_32_10 || (r.rng_type == ROCRAND_RNG_PSEUDO_MRG32K3A && r.session == 0))
            {
                unsigned long long offset = 0;
                if(r.session != 0)
                {
                    // a counter engine's state is its word position: the session's next call
                    // is a fresh generator's call at that offset (uniform doubles take two words)
                    session& ss = sessions[r.session];
                    offset      = ss.pos;
                    ss.pos += (unsigned long long)r.n * (r.dtype == ROCRAND_SERVE_DOUBLE ? 2 : 1);
                }
                bs.push_back(rocrand_batch_stream{r.rng_type, batch_kind(r.dtype), r.seed, offset, r.n, 0.0, 0.0, out});
            }
        }
        if(!bs.empty() && rocrand_generate_batched(bs.data(), bs.size(), s) != ROCRAND_STATUS_SUCCESS)
            return ROCRAND_STATUS_LAUNCH_FAILURE;
        for(size_t i = 0; i < reqs.size(); i++)
        {
            const rocrand_serve_request& r   = reqs[i].r;
            void*                        out = mem + starts[i];
            rocrand_status               st  = ROCRAND_STATUS_SUCCESS;
            if(r.rng_type == ROCRAND_RNG_QUASI_SOBOL64 && r.session == 0)
            {
                rocrand_generator& g = sobol_fresh[r.dims];
                if(g == nullptr)
                {
                    if((st = rocrand_create_generator(&g, ROCRAND_RNG_QUASI_SOBOL64)) != ROCRAND_STATUS_SUCCESS
                       || (st = rocrand_set_quasi_random_generator_dimensions(g, r.dims)) != ROCRAND_STATUS_SUCCESS)
                    {
                        g = nullptr;
                        return st;
                    }
                }
                if((st = rocrand_set_offset(g, r.seed % (1ull << 20))) != ROCRAND_STATUS_SUCCESS
                   || (st = rocrand_set_stream(g, s)) != ROCRAND_STATUS_SUCCESS
                   || (st = generate_kind(g, r.dtype, out, r.n)) != ROCRAND_STATUS_SUCCESS)
                    return st;
            }
            else if(r.session != 0 && r.rng_type != ROCRAND_RNG_PSEUDO_PHILOX4_32_10)
            {
                session& ss = sessions[r.session];
                if(r.rng_type == ROCRAND_RNG_PSEUDO_MRG32K3A && ss.last != nullptr && ss.last_stream != s
                   && hipStreamWaitEvent(s, ss.last, 0) != hipSuccess)
                    return ROCRAND_STATUS_LAUNCH_FAILURE;
                if((st = rocrand_set_stream(ss.gen, s)) != ROCRAND_STATUS_SUCCESS
                   || (st = generate_kind(ss.gen, r.dtype, out, r.n)) != ROCRAND_STATUS_SUCCESS)
                    return st;
                // after the session's latest request: the next MRG32k3a request waits for
                // it, and ending the session waits for it before destroying the generator
                if(ss.last == nullptr && (ss.last = event()) == nullptr)
                    return ROCRAND_STATUS_INTERNAL_ERROR;
                if(hipEventRecord(ss.last, s) != hipSuccess)
                    return ROCRAND_STATUS_LAUNCH_FAILURE;
                ss.last_stream = s;
            }
        }
        batch b;
        if((b.done = event()) == nullptr)
            return ROCRAND_STATUS_INTERNAL_ERROR;
        if(hipEventRecord(b.done, s) != hipSuccess)
        {
            spare.push_back(b.done);
            return ROCRAND_STATUS_LAUNCH_FAILURE;
        }
        for(size_t i = 0; i < reqs.size(); i++)
            b.reqs.push_back(issued{reqs[i].ticket, starts[i], reqs[i].bytes});
        b.stream  = si;
        busy[si]  = 1;
        inflight.push_back(std::move(b));
        return ROCRAND_STATUS_SUCCESS;
    }

    ~rocrand_server_st()
    {
        for(hipStream_t s : streams)
            (void)hipStreamSynchronize(s);
        for(auto& kv : sessions)
        {
            if(kv.second.gen)
                (void)rocrand_destroy_generator(kv.second.gen);
            if(kv.second.last)
                (void)hipEventDestroy(kv.second.last);
        }
        for(auto& kv : sobol_fresh)
            if(kv.second)
                (void)rocrand_destroy_generator(kv.second);
        for(auto& b : inflight)
            (void)hipEventDestroy(b.done);
        for(hipEvent_t e : spare)
            (void)hipEventDestroy(e);
        for(hipStream_t s : streams)
            (void)hipStreamDestroy(s);
        if(mem)
            (void)hipFree(mem);
    }
};

extern "C" {

rocrand_status ROCRANDAPI rocrand_server_create(rocrand_server* server, size_t budget_bytes, unsigned int streams,
                                                int keep_outputs)
{
    if(server == nullptr || budget_bytes < ALIGN || streams == 0 || streams > 8)
        return ROCRAND_STATUS_OUT_OF_RANGE;
    *server = nullptr;
    rocrand_server_st* s = nullptr;
    try
    {
        s           = new rocrand_server_st();
        s->budget   = budget_bytes / ALIGN * ALIGN;
        s->keep     = keep_outputs;
        s->free_mem = rr::arena(s->budget);
        if(hipMalloc(reinterpret_cast<void**>(&s->mem), s->budget) != hipSuccess)
        {
            s->mem = nullptr;
            delete s;
            return ROCRAND_STATUS_ALLOCATION_FAILED;
        }
        for(unsigned i = 0; i < streams; i++)
        {
            hipStream_t q = nullptr;
            if(hipStreamCreateWithFlags(&q, hipStreamNonBlocking) != hipSuccess)
            {
                delete s;
                return ROCRAND_STATUS_INTERNAL_ERROR;
            }
            s->streams.push_back(q);
        }
        s->busy.assign(streams, 0);
    }
    catch(const std::bad_alloc&)
    {
        delete s;
        return ROCRAND_STATUS_ALLOCATION_FAILED;
    }
    *server = s;
    return ROCRAND_STATUS_SUCCESS;
}

rocrand_status ROCRANDAPI rocrand_server_destroy(rocrand_server server)
{
    if(server == nullptr)
        return ROCRAND_STATUS_NOT_CREATED;
    delete server;
    return ROCRAND_STATUS_SUCCESS;
}

rocrand_status ROCRANDAPI rocrand_server_submit(rocrand_server server, const rocrand_serve_request* requests,
                                                size_t count, unsigned long long* tickets)
{
    if(server == nullptr)
        return ROCRAND_STATUS_NOT_CREATED;
    if(count > 0 && (requests == nullptr || tickets == nullptr))
        return ROCRAND_STATUS_OUT_OF_RANGE;
    try
    {
        // validate everything first: a rejected call queues nothing (sessions a request of
        // this call opens are checked against its later requests too)
        std::unordered_map<unsigned long long, session> opened;
        for(size_t i = 0; i < count; i++)
        {
            const rocrand_serve_request& r = requests[i];
            if(r.rng_type != ROCRAND_RNG_PSEUDO_PHILOX4_32_10 && r.rng_type != ROCRAND_RNG_PSEUDO_MRG32K3A
               && r.rng_type != ROCRAND_RNG_QUASI_SOBOL64)
                return ROCRAND_STATUS_TYPE_ERROR;
            if((int)r.dtype < 0 || (int)r.dtype > 2)
                return ROCRAND_STATUS_TYPE_ERROR;
            const unsigned dims = r.rng_type == ROCRAND_RNG_QUASI_SOBOL64 ? r.dims : 1u;
            if(r.rng_type == ROCRAND_RNG_QUASI_SOBOL64 && (dims < 1 || dims > 20000))
                return ROCRAND_STATUS_OUT_OF_RANGE;
            if(r.n % dims != 0)
                return ROCRAND_STATUS_LENGTH_NOT_MULTIPLE;
            if(r.session != 0)
            {
                auto it = server->sessions.find(r.session);
                if(it == server->sessions.end())
                {
                    it = opened.find(r.session);
                    if(it == opened.end())
                    {
                        session ss;
                        ss.type = r.rng_type;
                        ss.seed = r.seed;
                        ss.dims = dims;
                        it      = opened.emplace(r.session, ss).first;
                    }
                }
                if(it->second.type != r.rng_type || it->second.seed != r.seed || it->second.dims != dims)
                    return ROCRAND_STATUS_TYPE_ERROR;
            }
        }
        // then the generators of new sessions (a failure here queues nothing) ...
        for(size_t i = 0; i < count; i++)
        {
            const rocrand_serve_request& r = requests[i];
            if(r.session == 0 || slot(r.n * elem_size(r.dtype)) > server->budget
               || server->sessions.find(r.session) != server->sessions.end())
                continue;
            session ss;
            ss.type = r.rng_type;
            ss.seed = r.seed;
            ss.dims = r.rng_type == ROCRAND_RNG_QUASI_SOBOL64 ? r.dims : 1u;
            if(r.rng_type != ROCRAND_RNG_PSEUDO_PHILOX4_32_10)
            {
                rocrand_status st = rocrand_create_generator(&ss.gen, r.rng_type);
                if(st != ROCRAND_STATUS_SUCCESS)
                    return st;
                if(r.rng_type == ROCRAND_RNG_QUASI_SOBOL64)
                {
                    st = rocrand_set_quasi_random_generator_dimensions(ss.gen, ss.dims);
                    if(st == ROCRAND_STATUS_SUCCESS)
                        st = rocrand_set_offset(ss.gen, r.seed % (1ull << 20));
                }
                else
                    st = rocrand_set_seed(ss.gen, r.seed);
                if(st != ROCRAND_STATUS_SUCCESS)
                {
                    (void)rocrand_destroy_generator(ss.gen);
                    return st;
                }
            }
            server->sessions[r.session] = ss;
        }
        // ... then the requests
        for(size_t i = 0; i < count; i++)
        {
            const rocrand_serve_request& r     = requests[i];
            const size_t                 bytes = r.n * elem_size(r.dtype);
            if(slot(bytes) > server->budget)
            {
                tickets[i] = 0; // can never fit
                continue;
            }
            tickets[i] = server->next_ticket++;
            server->queue.push_back(queued{tickets[i], r, bytes});
        }
    }
    catch(const std::bad_alloc&)
    {
        return ROCRAND_STATUS_ALLOCATION_FAILED;
    }
    return ROCRAND_STATUS_SUCCESS;
}

rocrand_status ROCRANDAPI rocrand_server_poll(rocrand_server server, rocrand_serve_completion* completions,
                                              size_t capacity, size_t* completed, size_t* pending)
{
    if(server == nullptr)
        return ROCRAND_STATUS_NOT_CREATED;
    if(completed == nullptr || (capacity > 0 && completions == nullptr))
        return ROCRAND_STATUS_OUT_OF_RANGE;
    try
    {
        // the previous poll's outputs were valid until now
        for(const issued& q : server->reported)
            server->free_mem.release(q.start, slot(q.bytes));
        server->reported.clear();
        server->retire();
        size_t k = 0;
        while(k < capacity && !server->done.empty())
        {
            const issued& q  = server->done.front();
            completions[k++] = rocrand_serve_completion{q.ticket, server->mem + q.start, q.bytes};
            if(server->keep)
                server->kept[q.ticket] = q;
            else
                server->reported.push_back(q);
            server->done.pop_front();
        }
        *completed = k;
        // admission, when a stream is free: first come first served while the outputs fit;
        // while every stream is busy, arrivals wait and then go out together
        std::vector<queued> reqs;
        std::vector<size_t> starts;
        const unsigned      si = server->free_stream();
        while(si < server->streams.size() && !server->queue.empty())
        {
            size_t start = 0;
            if(!server->free_mem.alloc(slot(server->queue.front().bytes), start))
                break;
            reqs.push_back(server->queue.front());
            starts.push_back(start);
            server->queue.pop_front();
        }
        if(!reqs.empty())
        {
            const rocrand_status st = server->issue(reqs, starts, si);
            if(st != ROCRAND_STATUS_SUCCESS)
            {
                // the batch is dropped: let whatever it launched finish, then free its memory
                for(hipStream_t q : server->streams)
                    (void)hipStreamSynchronize(q);
                for(size_t i = 0; i < reqs.size(); i++)
                    server->free_mem.release(starts[i], slot(reqs[i].bytes));
                return st;
            }
        }
        if(pending != nullptr)
        {
            size_t n = server->queue.size() + server->done.size();
            for(const batch& b : server->inflight)
                n += b.reqs.size();
            *pending = n;
        }
    }
    catch(const std::bad_alloc&)
    {
        return ROCRAND_STATUS_ALLOCATION_FAILED;
    }
    return ROCRAND_STATUS_SUCCESS;
}

rocrand_status ROCRANDAPI rocrand_server_release(rocrand_server server, unsigned long long ticket)
{
    if(server == nullptr)
        return ROCRAND_STATUS_NOT_CREATED;
    auto it = server->kept.find(ticket);
    if(it == server->kept.end())
        return ROCRAND_STATUS_OUT_OF_RANGE;
    server->free_mem.release(it->second.start, slot(it->second.bytes));
    server->kept.erase(it);
    return ROCRAND_STATUS_SUCCESS;
}

rocrand_status ROCRANDAPI rocrand_server_end_session(rocrand_server server, unsigned long long session)
{
    if(server == nullptr)
        return ROCRAND_STATUS_NOT_CREATED;
    auto it = server->sessions.find(session);
    if(it == server->sessions.end())
        return ROCRAND_STATUS_OUT_OF_RANGE;
    // its queued requests still need it
    for(const queued& q : server->queue)
        if(q.r.session == session)
            return ROCRAND_STATUS_OUT_OF_RANGE;
    if(it->second.last != nullptr)
    {
        (void)hipEventSynchronize(it->second.last);
        server->spare.push_back(it->second.last);
    }
    if(it->second.gen != nullptr)
        (void)rocrand_destroy_generator(it->second.gen);
    server->sessions.erase(it);
    return ROCRAND_STATUS_SUCCESS;
}

} // extern "C"
