/*
 * bs_mi355.h -- C ABI of libbs_mi355.so, the MI355X (gfx950) Black-Scholes map
 * (and, beside it, the suite's swaptions map).
 *
 * Reference: ParaGroup/p3arsec, PARSEC blackscholes
 *   (parsec-ff/pkgs/apps/blackscholes/src/blackscholes.c, "BS" below) and
 *   swaptions (parsec-ff/pkgs/apps/swaptions/src, "SW" below).
 * The reference compiles its pricing path into one C++ translation unit with
 * no exported header; each entry point below names the interface it replaces.
 *
 * Conventions
 *   - Every function returns int status: BS_OK (0) or a negative BS_ERR_*.
 *     bs_last_error() gives a message for the calling thread.  No C++
 *     exception crosses this boundary.
 *   - Book arrays are structure-of-arrays: spot s, strike k, rate r,
 *     volatility v, time t (years), option type (0 = call, nonzero = put;
 *     the reference maps 'P' -> 1, everything else -> 0, BS:761).
 *   - Functions taking `bs_stream_t` expect DEVICE pointers (hipMalloc'd or
 *     pinned host memory) and run asynchronously on that HIP stream (NULL =
 *     the null stream).  Functions named *_host take host pointers and are
 *     synchronous.  Pointers are validated (hipPointerGetAttributes) unless
 *     bs_config_set(BS_CFG_CHECK_POINTERS, 0).
 *   - Arrays aligned for two elements per load (8 bytes for float arrays,
 *     16 for double, 2 for uint8 types, 8 for int32 types) take the vector
 *     path; anything else the one-option-per-work-item path.  Results are
 *     identical either way.
 *   - Arithmetic: FP32 entry points follow the reference's float statements
 *     and its order of operations: correctly rounded division, reciprocal
 *     and square root, the double literals applied with one rounding, exp
 *     and log from the hardware's FP32 exp2/log2.  Every price is within
 *     1e-4 of the reference (largest difference seen 8.4e-5 over 50M
 *     options).  *_fast_* uses hardware transcendentals throughout.  FP64
 *     entry points are the fptype=double build of the same algorithm.
 */
#ifndef BS_MI355_H
#define BS_MI355_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void *bs_stream_t; /* a hipStream_t */

enum {
    BS_OK = 0,
    BS_ERR_INVALID = -1,   /* bad argument (null pointer, n < 0, ...)        */
    BS_ERR_HIP = -2,       /* a HIP runtime call failed                      */
    BS_ERR_NOMEM = -3,     /* device allocation failed                       */
    BS_ERR_BUDGET = -4,    /* request larger than the context's HBM budget   */
    BS_ERR_NOT_FOUND = -5, /* unknown book or ticket                         */
    BS_ERR_CANCELLED = -6, /* ticket was cancelled                           */
    BS_ERR_TIMEOUT = -7,   /* bs_wait timed out                              */
    BS_ERR_BUSY = -8,      /* book still referenced by queued requests       */
    BS_ERR_NODEVICE = -9,  /* no HIP device / device index out of range      */
    BS_ERR_FAILED = -10,   /* a request's launch failed                      */
    BS_ERR_SHUTDOWN = -11  /* context is being destroyed                     */
};

const char *bs_status_string(int status);
const char *bs_last_error(void);
int bs_version(void); /* 10000*major + 100*minor + patch */

/* ---- configuration (process-wide) -------------------------------------- */
enum {
    BS_CFG_CHECK_POINTERS = 1, /* 1 (default): validate device pointers      */
    BS_CFG_MAX_BLOCKS = 2,     /* grid cap for streaming kernels, 0 = auto    */
    BS_CFG_BLOCK_THREADS = 3,  /* streaming block size: 0 (default), 64, 128
                                  or 256; bs_ctx_create reads it for the
                                  serving context's fused kernels          */
    BS_CFG_VECTOR_WIDTH = 4,   /* options per work-item: 0 (default), 2, 4   */
    BS_CFG_NONTEMPORAL = 5,    /* 1 (default): non-temporal loads and stores  */
    BS_CFG_PIPELINE = 6,       /* -1 (default), 0, 1: software pipelining     */
    BS_CFG_LDS_PER_WAVE = 7    /* LDS bytes reserved per wave of a block of the
                                  batch maps and portfolio sums (at most 64 KB
                                  per block), capping the waves resident on a
                                  CU: -1 (default: each kernel's measured
                                  best), 0 (none) or 1..65536               */
};
int bs_config_set(int key, int64_t value);
int bs_config_get(int key, int64_t *value);

/* ---- per-option functions ------------------------------------------------
 * fptype CNDF(fptype InputX)                                   BS:126-184
 * fptype BlkSchlsEqEuroNoDiv(fptype sptprice, fptype strike, fptype rate,
 *        fptype volatility, fptype time, int otype, float timet) BS:190-258
 * Batched over device arrays (the reference's per-option calls become one
 * map).  timet is unused by the reference and is not taken here.
 */
int bs_cndf_f32(const float *x, float *out, int64_t n, bs_stream_t stream);
int bs_cndf_f64(const double *x, double *out, int64_t n, bs_stream_t stream);

/* Scalar convenience forms with the reference's signatures: one device round
 * trip per call (for API parity and spot checks, not for throughput). */
int bs_cndf_scalar_f32(float x, float *out);
int bs_price_scalar_f32(float sptprice, float strike, float rate, float volatility,
                        float time, int otype, float timet, float *out);

/* ---- batch pricing -------------------------------------------------------
 * Replaces one pass of the FastFlow ParallelFor over the book
 * (BS:315-348; ff_Map::parallel_for_thid, ff/map.hpp:120-124), the SkePU2
 * Map<7> call (blackscholes_skepu_omp.cpp:725-728) and the SSE pointer batch
 * BlkSchlsEqEuroNoDiv(fptype*, int, fptype*, ..., int*, float)
 * (blackscholes.simd.c:221-223): price[i] = BlkSchlsEqEuroNoDiv(s[i], k[i],
 * r[i], v[i], t[i], otype[i], 0) for 0 <= i < n.
 */
int bs_price_f32(const float *s, const float *k, const float *r, const float *v,
                 const float *t, const uint8_t *otype, float *price, int64_t n,
                 bs_stream_t stream);
/* Same with the reference's int32 option-type array (BS:757-761). */
int bs_price_i32type_f32(const float *s, const float *k, const float *r, const float *v,
                         const float *t, const int32_t *otype, float *price, int64_t n,
                         bs_stream_t stream);
/* The reference's NUM_RUNS loop (BS:315-348, the FastFlow map's passes;
 * BS:607-658, bs_thread's): prices the book `passes` times on `stream`, each
 * pass starting after the previous one ends, with the same output as calling
 * bs_price_f32 (otype_bytes = 1, uint8 types) or bs_price_i32type_f32
 * (otype_bytes = 4, int32 types) `passes` times.  For books of up to 2^25
 * options the passes are one hipGraph of `passes` kernel nodes, captured on
 * the first call for these arrays, n, passes and device, and replayed after
 * that (the library keeps the 16 most recently used): small books are bound
 * by launch cost, and larger ones lose less to the cache maintenance between
 * separately launched passes.  Larger books launch pass by pass. */
int bs_price_passes_f32(const float *s, const float *k, const float *r, const float *v,
                        const float *t, const void *otype, int otype_bytes, float *price,
                        int64_t n, int passes, bs_stream_t stream);
/* Fast variant: hardware transcendentals (v_exp_f32, v_log_f32, v_rcp_f32,
 * v_sqrt_f32) and a Horner polynomial; about a third of the VALU work of
 * bs_price_f32, but prices may differ from the reference by slightly more
 * than 1e-4 (mean 6e-6). */
int bs_price_fast_f32(const float *s, const float *k, const float *r, const float *v,
                      const float *t, const uint8_t *otype, float *price, int64_t n,
                      bs_stream_t stream);
/* fptype = double build of the same map (BS:85 set to double). */
int bs_price_f64(const double *s, const double *k, const double *r, const double *v,
                 const double *t, const uint8_t *otype, double *price, int64_t n,
                 bs_stream_t stream);

/* Fused adjoint: price plus delta = dP/dS, vega = dP/dsigma, rho = dP/dr and
 * theta = -dP/dT (per year) of the same FP32 price, in one pass.  No
 * reference counterpart (the reference prices only); it extends BS:190-258. */
int bs_price_greeks_f32(const float *s, const float *k, const float *r, const float *v,
                        const float *t, const uint8_t *otype, float *price, float *delta,
                        float *vega, float *rho, float *theta, int64_t n,
                        bs_stream_t stream);

/* Portfolio reduction over a book: sums[0] = sum of w_i * price_i and, with
 * greeks != 0, sums[1..4] = the sums of w_i * delta_i, vega_i, rho_i, theta_i,
 * each per-option value exactly the one bs_price_f32 / bs_price_greeks_f32
 * computes.  weights may be NULL (every w_i = 1).  sums is a device array of 1
 * or 5 doubles.  Summed in FP64 in a fixed order: bit-reproducible for a given
 * n.  n == 0 writes zeros.  No reference counterpart: the MapReduce step that
 * follows the map (SkePU2 MapReduce, skepu2/backend/impl/mapreduce/). */
int bs_portfolio_f32(const float *s, const float *k, const float *r, const float *v,
                     const float *t, const uint8_t *otype, const float *weights, int64_t n,
                     int greeks, double *sums, bs_stream_t stream);

/* bs_portfolio_f32 that allocates nothing.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (work under HIP graph
 * capture, serving loops and harnesses that forbid hipMalloc).  The same
 * arguments, checks and sums, bit for bit, for n <= 2^25 options
 * (BS_ERR_INVALID past that); the partial sums live in a fixed device array
 * of the library (2.6 MB on each device it runs on, held until the process
 * ends) instead of stream-ordered memory.  Calls that overlap on one device
 * share that array and race: issue them on one stream, or let one finish
 * before the next starts. */
int bs_portfolio_f32_noalloc(const float *s, const float *k, const float *r, const float *v,
                             const float *t, const uint8_t *otype, const float *weights,
                             int64_t n, int greeks, double *sums, bs_stream_t stream);

/* AoS -> SoA repack on the device (BS:747-767).  `aos` points at n records
 * laid out as the reference's OptionData (BS:89-100): float s, strike, r,
 * divq, v, t; char OptionType; float divs, DGrefval -- 36 bytes each.  The
 * type becomes 1 for 'P' and 0 otherwise. */
int bs_repack_aos_f32(const void *aos, int64_t n, float *s, float *k, float *r, float *v,
                      float *t, uint8_t *otype, bs_stream_t stream);

/* Host-pointer forms (the SkePU2 CUDA back-end's updateDevice / kernel /
 * updateHost sequence, skepu2/backend/impl/map/map_cu.inl:15-58): copy in,
 * run `passes` passes (>= 1), copy the price out.  Synchronous. */
int bs_price_host_f32(const float *s, const float *k, const float *r, const float *v,
                      const float *t, const int32_t *otype, float *price, int64_t n,
                      int passes, int device);

/* bs_thread (BS:607-658) drop-in.  bs_bind_globals() records the caller's
 * globals (BS:102-113); bs_thread_compat(&tid) then prices thread tid's block
 * [tid*(N/nthreads), +N/nthreads) NUM_RUNS = 100 times on the GPU, or the
 * whole book when nthreads == 1 (serial and OpenMP drivers).  Returns 0. */
int bs_bind_globals(int num_options, int nthreads, float *sptprice, float *strike,
                    float *rate, float *volatility, float *otime, int *otype,
                    float *prices);
int bs_thread_compat(void *tid_ptr);

/* ---- swaptions: HJM Monte-Carlo -------------------------------------------
 * The suite's other computational-finance map (SW = parsec-ff/pkgs/apps/
 * swaptions/src).  One record per swaption, the fields of the reference's
 * `parm` (HJM_Securities.cpp:275-296) that HJM_Swaption_Blocking reads
 * (HJM_Swaption_Blocking.cpp:20-40); yield and factors are HOST pointers. */
typedef struct {
    double strike;           /* dStrike                                        */
    double compounding;      /* dCompounding: 0 = continuous, else periods/yr  */
    double maturity;         /* dMaturity: years to the swap's start           */
    double tenor;            /* dTenor: years of the swap                      */
    double payment_interval; /* dPaymentInterval: years between payments       */
    double years;            /* dYears: HJM horizon                            */
    int32_t n_steps;         /* iN: HJM time points, 2..64                     */
    int32_t n_factors;       /* iFactors: 1..8, the same for every record      */
    const double *yield;     /* pdYield[n_steps]                               */
    const double *factors;   /* ppdFactors: n_factors rows of n_steps - 1      */
} bs_swaption_t;

/* Replaces the FastFlow ParallelFor over swaptions (SW HJM_Securities.cpp:
 * 311-323) and each call HJM_Swaption_Blocking(pdSwaptionPrice, ...,
 * iRndSeed = seed + i, lTrials = trials, BLOCKSIZE = block_size)
 * (HJM_Swaption_Blocking.cpp:20-222): price[i] and std_error[i] are swaption
 * i's Monte-Carlo mean and standard error, FP64, from the same random
 * stream as the reference (ceil(trials / block_size) * block_size paths,
 * sums divided by trials).  Every path of every swaption runs on the GPU.
 * Host arrays; synchronous on `stream`.  Limits: trials in [1, 2^40],
 * block_size >= 1, seed in [0, 2^42) and seed + n + paths * (n_steps - 1) *
 * n_factors below 2^42, at most 2^24 blocks of 1024 paths in one call
 * (n * ceil(paths / 1024) <= 2^24, about 2^34 paths in all: one launch of at
 * most 2^32 work-items), the swap's start step x n_factors <= 64; a record
 * the reference would index out of its arrays with (maturity past the
 * horizon, tenor past the curve, payment interval under one step) is
 * BS_ERR_INVALID.  A call's device buffers (about 1.5 KB per swaption plus
 * 16 bytes per 1024 paths) come from a stream-ordered pool the library keeps
 * per device, which holds up to 1 GiB of freed memory for later calls until
 * the process ends.  Calls on different threads and streams are independent. */
int bs_swaptions_host_f64(const bs_swaption_t *sw, int64_t n, int64_t trials, int64_t seed,
                          int32_t block_size, double *price, double *std_error,
                          bs_stream_t stream);

/* ---- device plumbing for hosts without a GPU runtime of their own ------- */
int bs_device_count(int *count);
int bs_set_device(int device);
int bs_malloc(void **ptr, size_t bytes);
int bs_free(void *ptr);
int bs_memcpy_h2d(void *dst, const void *src, size_t bytes, bs_stream_t stream);
int bs_memcpy_d2h(void *dst, const void *src, size_t bytes, bs_stream_t stream);
int bs_memcpy_d2d(void *dst, const void *src, size_t bytes, bs_stream_t stream);
int bs_memset(void *dst, int value, size_t bytes, bs_stream_t stream);
int bs_stream_create(bs_stream_t *stream);
int bs_stream_destroy(bs_stream_t stream);
int bs_stream_sync(bs_stream_t stream);
int bs_device_sync(void);
typedef void *bs_event_t;
int bs_event_create(bs_event_t *ev);
/* An event for timing only: its record performs no system-scope fence (no
 * cache writeback and invalidation at the marker, so the time between two
 * such events holds the work between them and not the fences).  bs_event_sync on it waits for the GPU to pass the marker but does
 * not make the work's results visible to the host: synchronize the stream or
 * device before reading them. */
int bs_event_create_timing(bs_event_t *ev);
int bs_event_destroy(bs_event_t ev);
int bs_event_record(bs_event_t ev, bs_stream_t stream);
int bs_event_sync(bs_event_t ev);
/* Work queued on `stream` after this call waits for `ev`'s last record. */
int bs_stream_wait_event(bs_stream_t stream, bs_event_t ev);
int bs_event_elapsed_ms(bs_event_t start, bs_event_t stop, float *ms);

/* ---- serving: resident books, continuous batching -------------------------
 * No reference counterpart: the reference is a single-shot batch benchmark
 * (BS:664-960).  A context owns one HBM arena of `hbm_budget_bytes`; books
 * registered into it stay resident; requests against resident books are
 * queued, admitted when their output buffers fit in the arena, fused with the
 * other queued requests of their lane into one batch (a fused kernel per
 * request kind, launched directly or replayed from a captured hipGraph), and
 * completed in the order their batches finish.  Requests of at most 64K
 * options form a lane of their own (bs_serve_opts.small_request_options).
 * Every call on a context may come from any thread at once, except
 * bs_ctx_destroy, which must not overlap any other call on that context: it
 * finishes the batches in flight, fails the queued requests with
 * BS_ERR_SHUTDOWN and frees the arena (result pointers from
 * bs_result_device die with it). */
typedef struct bs_ctx_s *bs_ctx_t;
typedef int64_t bs_book_t;
typedef int64_t bs_ticket_t;

enum { BS_REQ_PRICE = 1, BS_REQ_GREEKS = 2 };
enum {
    BS_TICKET_QUEUED = 0,
    BS_TICKET_RUNNING = 1,
    BS_TICKET_DONE = 2,
    BS_TICKET_FAILED = 3,
    BS_TICKET_CANCELLED = 4
};

typedef struct {
    int64_t max_batch_requests; /* requests fused per launch (<= 4096)       */
    int64_t max_batch_options;  /* options fused per launch                  */
    int32_t use_graph;          /* 1: replay a captured hipGraph per batch;
                                   0: direct launches (default, faster)     */
    int32_t grid_blocks;        /* cap on each section's grid (one block per
                                   512-option tile; blocks loop past the
                                   cap), 0 = 2^20                            */
    int32_t completion_queue;   /* 1: finished requests queue for bs_reap    */
    int32_t small_request_options; /* requests of at most this many options
                                   form batches of their own, on a stream of
                                   their own, so they do not wait for large
                                   requests' batches; 0 = 65536, -1 = one
                                   lane for all requests                     */
} bs_serve_opts;

/* One finished request from the completion queue.  Times are steady-clock
 * (CLOCK_MONOTONIC) nanoseconds. */
typedef struct {
    bs_ticket_t ticket;
    int64_t user;   /* the caller's tag from bs_submit_many (0 otherwise)      */
    int64_t n;      /* options of the request's book                         */
    int32_t state;  /* BS_TICKET_DONE, BS_TICKET_FAILED or BS_TICKET_CANCELLED */
    int32_t status; /* BS_OK or the request's error code                     */
    int64_t submit_ns, done_ns;
} bs_completion;

typedef struct {
    int64_t arena_bytes, arena_used, arena_peak;
    int64_t books_resident, requests_queued, requests_running;
    int64_t requests_done, requests_failed, requests_cancelled;
    int64_t launches, options_priced, admission_waits;
    int64_t queue_ns_total; /* sum over admitted requests of submit -> launch  */
    int64_t batch_ns_total; /* sum over batches of launch -> completion seen   */
    /* dispatcher thread: time admitting (lock wait + admission + table),
     * launching, retiring; time with no batch in flight and the launches
     * that ended such a stretch */
    int64_t dispatch_admit_ns, dispatch_launch_ns, dispatch_retire_ns;
    int64_t gpu_idle_ns, idle_launches;
} bs_serve_stats;

int bs_ctx_create(int device, size_t hbm_budget_bytes, const bs_serve_opts *opts,
                  bs_ctx_t *ctx);
int bs_ctx_destroy(bs_ctx_t ctx);
/* Uploads a book (host arrays) into the arena; BS_ERR_BUDGET when it cannot
 * fit next to the resident books and admitted outputs. */
int bs_book_register(bs_ctx_t ctx, const float *s, const float *k, const float *r,
                     const float *v, const float *t, const uint8_t *otype, int64_t n,
                     bs_book_t *book);
/* Books still referenced by queued or running requests: BS_ERR_BUSY. */
int bs_book_release(bs_ctx_t ctx, bs_book_t book);
int bs_submit(bs_ctx_t ctx, bs_book_t book, int kind, bs_ticket_t *ticket);
/* Submits count requests under one lock, all or none: an unknown book, a bad
 * kind or outputs that can never fit fail the call before any is queued.
 * Requests of one call join the queue in order.  users (may be NULL) gives
 * each request a tag that its bs_completion carries back.  A context holds
 * at most 2^24 requests between submit and release (queued, running or
 * finished and not yet released); past that, bs_submit and bs_submit_many
 * return BS_ERR_NOMEM. */
int bs_submit_many(bs_ctx_t ctx, const bs_book_t *books, const int *kinds,
                   const int64_t *users, int64_t count, bs_ticket_t *tickets);
/* With opts.completion_queue set: pops up to max finished requests in the
 * order they finished, waiting up to timeout_us for the first one (0 polls,
 * < 0 waits forever).  *count = 0 after a timeout.  Outputs stay resident
 * until released.  BS_ERR_INVALID when the context has no completion queue. */
int bs_reap(bs_ctx_t ctx, bs_completion *out, int64_t max, int64_t timeout_us, int64_t *count);
int bs_poll(bs_ctx_t ctx, bs_ticket_t ticket, int *state);
/* timeout_us < 0 waits forever. */
int bs_wait(bs_ctx_t ctx, bs_ticket_t ticket, int64_t timeout_us);
/* Copies a finished request's outputs to host arrays (NULL skips one; the
 * Greeks arrays only for BS_REQ_GREEKS). */
int bs_result_copy(bs_ctx_t ctx, bs_ticket_t ticket, float *price, float *delta,
                   float *vega, float *rho, float *theta);
/* Device pointers of a finished request's outputs (valid until release). */
int bs_result_device(bs_ctx_t ctx, bs_ticket_t ticket, const float **price,
                     const float **delta, const float **vega, const float **rho,
                     const float **theta);
/* Frees the request's outputs in the arena and forgets the ticket. */
int bs_result_release(bs_ctx_t ctx, bs_ticket_t ticket);
/* Cancels a queued request (running ones finish). */
int bs_cancel(bs_ctx_t ctx, bs_ticket_t ticket);
/* Nanosecond timestamps (CLOCK_MONOTONIC) of submit and completion. */
/* Releases count tickets under one lock.  Every ticket is tried; the return
 * value is the first failure (BS_ERR_NOT_FOUND, BS_ERR_BUSY) or BS_OK. */
int bs_result_release_many(bs_ctx_t ctx, const bs_ticket_t *tickets, int64_t count);
int bs_ticket_times(bs_ctx_t ctx, bs_ticket_t ticket, int64_t *submit_ns,
                    int64_t *done_ns);
int bs_ctx_stats(bs_ctx_t ctx, bs_serve_stats *stats);

#ifdef __cplusplus
}
#endif
#endif /* BS_MI355_H */
