/*
 * rocrand_batched.h -- many independent generate calls in one launch (MI355X extension).
 *
 * No counterpart in the reference: it serves callers that issue thousands of small,
 * independent generate calls (per-request serving, per-sample streams), for which the
 * reference pays one generator set-up, one engine-initialisation kernel and one launch
 * per call.  Every stream here is defined by a plain reference call sequence and
 * produces exactly its output:
 *
 *     rocrand_create_generator(&g, rng_type);
 *     rocrand_set_seed(g, seed);
 *     rocrand_set_offset(g, offset);
 *     rocrand_generate_<kind>(g, output, n, a, b);   // a, b: mean/stddev or lambda
 *
 * (including the head/tail split that the output pointer's alignment selects).
 * A plan validates and lays out a batch once and records its kernels in a HIP graph;
 * each rocrand_batch_plan_launch replays that graph on the plan's stream and writes
 * every stream's output again (the plan is stateless).
 */
#ifndef ROCRAND_BATCHED_H_
#define ROCRAND_BATCHED_H_

#include "rocrand.h"

#if defined(__cplusplus)
extern "C" {
#endif

/* Output kind of a stream: which rocrand_generate_* call it stands for. */
typedef enum rocrand_batch_kind
{
    ROCRAND_BATCH_UINT              = 0, /* rocrand_generate                 (unsigned int) */
    ROCRAND_BATCH_UNIFORM_FLOAT     = 1, /* rocrand_generate_uniform         (float)        */
    ROCRAND_BATCH_UNIFORM_DOUBLE    = 2, /* rocrand_generate_uniform_double  (double)       */
    ROCRAND_BATCH_NORMAL_FLOAT      = 3, /* rocrand_generate_normal          (a=mean, b=sd) */
    ROCRAND_BATCH_NORMAL_DOUBLE     = 4, /* rocrand_generate_normal_double                  */
    ROCRAND_BATCH_LOG_NORMAL_FLOAT  = 5, /* rocrand_generate_log_normal                     */
    ROCRAND_BATCH_LOG_NORMAL_DOUBLE = 6, /* rocrand_generate_log_normal_double              */
    ROCRAND_BATCH_POISSON           = 7  /* rocrand_generate_poisson         (a=lambda)     */
} rocrand_batch_kind;

typedef struct rocrand_batch_stream
{
    rocrand_rng_type   rng_type; /* ROCRAND_RNG_PSEUDO_PHILOX4_32_10 or ROCRAND_RNG_PSEUDO_MRG32K3A */
    rocrand_batch_kind kind;
    unsigned long long seed;
    unsigned long long offset;
    size_t             n;
    double             a, b;
    void*              output; /* device memory for n values of the kind's type */
} rocrand_batch_stream;

typedef struct rocrand_batch_plan_st* rocrand_batch_plan;

/*
 * Lays out `count` streams for launches on `stream`.  Status: OUT_OF_RANGE for a null
 * argument, lambda <= 0 or a null output with n > 0; TYPE_ERROR for another engine or
 * kind; ALLOCATION_FAILED when device memory for the plan cannot be had.  The first
 * plan of a given shape times a few launch configurations (autotuning), which writes
 * the outputs; later plans of that shape reuse the choice.
 */
ROCRANDAPI rocrand_status rocrand_batch_plan_create(rocrand_batch_plan*         plan,
                                                    const rocrand_batch_stream* streams,
                                                    size_t                      count,
                                                    hipStream_t                 stream);
/* Writes every stream's output (asynchronously on the plan's stream). */
ROCRANDAPI rocrand_status rocrand_batch_plan_launch(rocrand_batch_plan plan);
/* Releases the plan once its last launch has finished (work queued later on the stream
 * is not waited for). */
ROCRANDAPI rocrand_status rocrand_batch_plan_destroy(rocrand_batch_plan plan);
/* One-shot: validates and lays out the batch and launches it without a graph,
 * asynchronously on `stream` (like any rocrand_generate_* call).  The descriptors go
 * through pooled pinned and device memory that is recycled once the launch has
 * completed, so repeated calls neither allocate in steady state nor wait for the device;
 * `streams` may be reused as soon as the call returns. */
ROCRANDAPI rocrand_status rocrand_generate_batched(const rocrand_batch_stream* streams,
                                                   size_t                      count,
                                                   hipStream_t                 stream);
/* No counterpart in the reference: the one-shot call for callers that cannot allocate
 * device or pinned memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Same streams, same outputs as rocrand_generate_batched; the
 * descriptors go into a fixed 16 MiB array in device memory, one per GPU, held until the
 * process ends, and no allocation routine is called.  OUT_OF_RANGE when a batch's
 * descriptors exceed that array (about 10^6 tiles of 16 Ki values, or fewer streams
 * with large Poisson tables).  Every call on a GPU shares that GPU's array: calls that
 * overlap on one GPU (issued on different streams) race; calls in order on one stream
 * do not. */
ROCRANDAPI rocrand_status rocrand_generate_batched_noalloc(const rocrand_batch_stream* streams,
                                                           size_t                      count,
                                                           hipStream_t                 stream);

#if defined(__cplusplus)
}
#endif

#endif /* ROCRAND_BATCHED_H_ */
