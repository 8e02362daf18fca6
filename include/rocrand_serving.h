/*
 * rocrand_serving.h -- a continuous-batching request server (MI355X extension).
 *
 * No counterpart in the reference.  It serves open-loop streams of generate requests of
 * mixed engines and sizes under a fixed HBM budget: requests are queued first come first
 * served, admitted while their outputs fit the server's memory, and issued in batches on
 * the server's streams in turn: one rocrand_generate_batched call for the Philox4x32-10
 * requests and the fresh MRG32k3a requests, plus one generator call per Sobol64 request
 * and per MRG32k3a session request.  A poll retires the batches whose completion event
 * has fired.  The host never waits for the device.
 *
 * Every request's output is defined by a plain call sequence:
 *  - session 0 (a fresh stream): rocrand_create_generator(rng_type), rocrand_set_seed(seed)
 *    (Sobol64: rocrand_set_quasi_random_generator_dimensions(dims) and
 *    rocrand_set_offset(seed mod 2^20) instead), then one rocrand_generate / _uniform /
 *    _uniform_double of n values;
 *  - session s != 0: the generator of session s, created by its first request as above and
 *    kept on the device between requests, makes its next such call: a session continues
 *    where its previous request ended (MRG32k3a sessions keep their 131072 engine states
 *    resident in HBM, Philox4x32-10 sessions their 128-bit word position, Sobol64
 *    sessions their point offset).  Requests of one session must keep its engine,
 *    dimensions and seed.
 */
#ifndef ROCRAND_SERVING_H_
#define ROCRAND_SERVING_H_

#include "rocrand.h"

#if defined(__cplusplus)
extern "C" {
#endif

typedef struct rocrand_server_st* rocrand_server;

typedef enum rocrand_serve_dtype
{
    ROCRAND_SERVE_UINT   = 0, /* rocrand_generate                (unsigned int) */
    ROCRAND_SERVE_FLOAT  = 1, /* rocrand_generate_uniform        (float)        */
    ROCRAND_SERVE_DOUBLE = 2  /* rocrand_generate_uniform_double (double)       */
} rocrand_serve_dtype;

typedef struct rocrand_serve_request
{
    rocrand_rng_type    rng_type; /* PSEUDO_PHILOX4_32_10, PSEUDO_MRG32K3A or QUASI_SOBOL64 */
    rocrand_serve_dtype dtype;
    unsigned long long  seed;
    unsigned long long  session; /* 0: a fresh stream; else the session it continues */
    size_t              n;       /* values; Sobol64: a multiple of dims */
    unsigned int        dims;    /* Sobol64 dimensions (1 otherwise) */
} rocrand_serve_request;

typedef struct rocrand_serve_completion
{
    unsigned long long ticket;
    void*              output; /* device memory of the server, n values of the dtype */
    size_t             bytes;
} rocrand_serve_completion;

/* A server owning budget_bytes (rounded down to whole 256-byte units; OUT_OF_RANGE below
 * 256 or for streams outside 1..8) of device memory and `streams` HIP streams.  With
 * keep_outputs = 0 a reported output stays valid until the next rocrand_server_poll call,
 * which returns its memory to the server; otherwise it stays until rocrand_server_release.
 * Calls on one server are not thread-safe. */
ROCRANDAPI rocrand_status rocrand_server_create(rocrand_server* server, size_t budget_bytes,
                                                unsigned int streams, int keep_outputs);
/* Waits for the server's work, then releases its memory, streams and sessions. */
ROCRANDAPI rocrand_status rocrand_server_destroy(rocrand_server server);
/* Queues `count` requests; tickets[i] receives request i's ticket (increasing from 1), or 0
 * when its output cannot fit the budget (rejected).  TYPE_ERROR for an unsupported engine
 * or a session request that changes its session's engine, seed or dimensions (also
 * against an earlier request of the same call); LENGTH_NOT_MULTIPLE for a Sobol64 length
 * that is not a multiple of dims.  A call that returns an error queues nothing. */
ROCRANDAPI rocrand_status rocrand_server_submit(rocrand_server               server,
                                                const rocrand_serve_request* requests,
                                                size_t                       count,
                                                unsigned long long*          tickets);
/* One turn of the loop: reports up to `capacity` completed requests (oldest batches first;
 * the rest are reported by later polls), then admits queued requests into free memory
 * and issues them as one batch.  *pending receives the number of requests queued, in
 * flight or completed but not yet reported.  When issuing fails the status is returned
 * and that batch's requests are dropped (neither queued nor reported). */
ROCRANDAPI rocrand_status rocrand_server_poll(rocrand_server            server,
                                              rocrand_serve_completion* completions,
                                              size_t                    capacity,
                                              size_t*                   completed,
                                              size_t*                   pending);
/* Returns a completed request's memory (keep_outputs servers). */
ROCRANDAPI rocrand_status rocrand_server_release(rocrand_server server, unsigned long long ticket);
/* Drops a session's generator (after its issued requests complete). */
ROCRANDAPI rocrand_status rocrand_server_end_session(rocrand_server server, unsigned long long session);

#if defined(__cplusplus)
}
#endif

#endif /* ROCRAND_SERVING_H_ */
