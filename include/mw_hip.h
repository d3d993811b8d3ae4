/*
 * mw_hip.h -- C ABI of the MI355X (gfx950) separation likelihood library
 * (libmw_separation_hip.so).
 *
 * Two layers:
 *
 * 1. Drop-in entry points with the reference evaluator's signatures and side
 *    effects.  A client selects them where it selects cpu_evaluate today
 *    (src/milkyway_client_app.c:263-274) and calls them through the same
 *    EvalFunc pointer (milkyway_separation/cpu/simple_evaluator.h:26).
 *
 * 2. A reentrant context API for batched evaluation and serving: a context
 *    keeps one wedge's tables and star list resident on the current GPU and
 *    evaluates many parameter vectors per call.  It returns status codes and
 *    never exits.
 *
 * Limits of this build: the drop-in entry points take any number of streams
 * (as the reference's calculate_integrals and calculate_likelihood do);
 * contexts take at most MW_MAX_STREAMS streams per model (the size of
 * mw_result's arrays), batches and the request server at most 4, and the
 * stream separation 1 to 4 (the reference's prob_ok, probability.c:310-403,
 * handles 1 to 4).  Models of more than 4 streams are evaluated in passes of
 * 4 streams on the exact quadrature.  At most MW_MAX_CONVOLVE convolution
 * points.
 */
#ifndef MW_HIP_H
#define MW_HIP_H

#include <stddef.h>

#include "mw_separation_types.h"

#ifdef __cplusplus
extern "C" {
#endif

#define MW_MAX_STREAMS 16
#define MW_MAX_CONVOLVE 1024

/* Status codes */
#define MW_OK 0
#define MW_E_INVALID 1      /* inputs the evaluator cannot handle */
#define MW_E_HIP 2          /* a HIP runtime call failed */
#define MW_E_NOMEM 3        /* device or host allocation failed */
#define MW_E_NODEVICE 4     /* no GPU visible */

/* Precision modes */
#define MW_PREC_FP64 0      /* FP64 throughout */
#define MW_PREC_MIXED 1     /* FP32 exp for the streams (their terms summed over 8 points in FP32 in rows
                               of >= 100 points); FP64 positions, distances, background, other sums, logs */
/* MW_PREC_MIXED on a resampled quadrature (context, batch and server calls;
 * not the drop-in entry points).  The reference's integral sums, per sight,
 * r_steps x convolve weighted densities at distances that are the same for
 * every sight.  Here those weights are folded, by 16-node Lagrange
 * interpolation, onto a uniform grid in distance or in magnitude (whichever
 * needs fewer nodes) whose spacing follows the narrowest feature the model
 * can have: the thinnest stream, and the sights' closest approach to the
 * Galactic centre.  The sum is taken over the grid: the same quadrature to
 * about 1e-10 per sight, from up to hundreds of times fewer density
 * evaluations on large wedges.  Regions where it would not pay, and
 * background modes other than the Hernquist profile, use the exact tables.
 * In the star pass the streams keep the exact convolution; each star's
 * background term is folded over the convolution's magnitude offsets where
 * that takes fewer nodes. */
#define MW_PREC_RESAMPLED 2

/* The weight folding behind MW_PREC_RESAMPLED, on the host, for any point
 * set: points g[n] (magnitudes) with weights w[n] become weights K[m] on the
 * grid *g_lo + m delta, so that sum_m K[m] f(grid_m) equals
 * sum_p w[p] f(g[p]) for every polynomial f of the given degree, and to the
 * interpolation error otherwise.  Returns the grid size M (K is written
 * only when capacity >= M), or -1 for bad input. */
long mw_resample_weights(const double* g, const double* w, long n, double delta, int degree,
                         double* g_lo, double* K, long capacity);
/* The same with the grid starting pad nodes below the lowest point and
 * ending pad nodes above the highest (mw_resample_weights: pad = degree / 2
 * + 1, every stencil centred; a smaller pad shifts the stencils at the ends
 * inwards, as the library's background and star-pass tables do with 4). */
long mw_resample_weights_padded(const double* g, const double* w, long n, double delta, int degree,
                                int pad, double* g_lo, double* K, long capacity);

/* ------------------------------------------------------------------------
 * Drop-in entry points
 * ------------------------------------------------------------------------ */

/* Replaces cpu_evaluate (milkyway_separation/cpu/evaluation_optimized.h:59-62,
 * evaluation_optimized.c:756-784).  Maps the parameter vector into ap
 * (set_astronomy_parameters, parameters.c:332-363), resets es, computes the
 * integrals and the star pass with the same side effects on es, and returns
 * es->prob_sum / (number_stars - bad_jacobians); the client subtracts 3.0.
 * On failure it prints "APP: error calculating ..." and exits with the error
 * code, as the reference does through mw_finish. */
double mw_hip_evaluate(double* parameters, ASTRONOMY_PARAMETERS* ap, EVALUATION_STATE* es,
                       STAR_POINTS* sp);

/* Replaces calculate_integrals (evaluation_optimized.h:49, .c:564-611).
 * Honours es->current_integral and each area's current_calculation /
 * max_calculation (partial and resumed ranges), adds to the areas' existing
 * partial integrals and writes es->background_integral and
 * es->stream_integrals as region 0 minus the cut regions.  Returns 0, or a
 * status code (the reference always returns 0 and exits on failure). */
int mw_hip_calculate_integrals(const ASTRONOMY_PARAMETERS* ap, EVALUATION_STATE* es,
                               const STAR_POINTS* sp);

/* Replaces calculate_likelihood (evaluation_optimized.h:50, .c:613-753).
 * Stars from es->current_star_point on; writes es->prob_sum, es->num_zero and
 * es->current_star_point.  Returns 0 or a status code. */
int mw_hip_calculate_likelihood(const ASTRONOMY_PARAMETERS* ap, EVALUATION_STATE* es,
                                const STAR_POINTS* sp);

/* Precision of the drop-in entry points (default MW_PREC_FP64, or the
 * environment variable MW_HIP_PRECISION=mixed).  No reference counterpart:
 * the reference picks precision at build time (DOUBLE_PRECISION,
 * include/config.h.in:30). */
int mw_hip_set_precision(int precision);

/* ------------------------------------------------------------------------
 * Context API (no reference counterpart; the reference's nearest equivalent
 * is the per-rank integral_f / likelihood_f plugin pair of
 * milkyway_separation/cpu/astronomy_worker.c:35-124)
 *
 * Threads: the calls on one context are not reentrant (each evaluation
 * reuses the context's staging and result buffers), so a caller that shares
 * a context between threads serialises its calls; separate contexts may be
 * used from separate threads.  The drop-in entry points serialise
 * themselves on one process-wide lock around their cached context.
 * ------------------------------------------------------------------------ */

typedef struct mw_ctx mw_ctx;

typedef struct {
    double likelihood;                 /* prob_sum / number_stars, without the -3.0 */
    double prob_sum;                   /* sum of log10 P, -238 per zero star */
    int num_zero;                      /* stars with P == 0 */
    int number_streams;
    double background_integral;        /* region 0 minus cut regions */
    double stream_integrals[MW_MAX_STREAMS];
    double bg_only_likelihood;         /* mean background-only log10 (no -3.0) */
    double stream_only_likelihood[MW_MAX_STREAMS];
} mw_result;

typedef struct {
    double integral_ms;                /* integral kernel(s) of the last mw_ctx_eval */
    double star_ms;                    /* star kernel(s) */
    double total_ms;                   /* first to last kernel, including reductions */
    long integral_blocks, star_blocks; /* workgroups launched */
    long integral_points;              /* integrand evaluations of the integral pass: cells x
                                          convolve points, or sights x grid nodes where the
                                          quadrature is resampled */
    long integral_points_recurrence;   /* of those, the ones on grids uniform in distance whose
                                          stream terms came from the FP64 recurrence instead of
                                          an exponential per point (resampled only) */
    long integral_points_background;   /* background evaluations: integral_points, except on
                                          distance grids whose background is sampled on every
                                          2nd or 4th node or on a grid of its own */
    long integral_stream_terms;        /* stream terms the recurrence ran over: per sight, the
                                          grid's nodes times the streams, or, where each stream
                                          has a grid of its own, the sum of those grids' nodes
                                          (rows a stream does not reach still counted) */
} mw_timing;

/* Flags for mw_ctx_create */
#define MW_CTX_TIMING 1                /* record HIP events around every kernel */

/* Builds the wedge described by ap (regions, convolve, wedge, sgr_coordinates,
 * aux_bg_profile, number of streams, optimize masks) and the star list on the
 * current device.  ap's parameter values are the defaults that parameter
 * vectors override.  sp may be NULL (integrals only). */
int mw_ctx_create(const ASTRONOMY_PARAMETERS* ap, const STAR_POINTS* sp, int flags,
                  mw_ctx** out);

/* Number of doubles in one parameter vector (get_optimized_parameter_count,
 * parameters.c:307-330). */
int mw_ctx_parameter_count(const mw_ctx* ctx);

/* Evaluates n_vec parameter vectors (params: n_vec x mw_ctx_parameter_count
 * doubles, laid out as set_astronomy_parameters reads them) in one batch.
 * results: n_vec entries.  log10p: NULL or n_vec x number_stars doubles that
 * receive every star's log10 P (-238 where P == 0). */
int mw_ctx_eval(mw_ctx* ctx, const double* params, int n_vec, int precision,
                mw_result* results, double* log10p);

/* Region-level integrals of the last mw_ctx_eval: out[v][g][1 + MW_MAX_STREAMS]
 * for vector v and region g (background first).  n is the capacity in doubles. */
int mw_ctx_region_integrals(const mw_ctx* ctx, double* out, size_t n);

int mw_ctx_last_timing(const mw_ctx* ctx, mw_timing* out);

/* One evaluation that allocates nothing.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The same
 * kernels as mw_ctx_eval for one parameter vector (params: the context's
 * parameter count; its result is bitwise mw_ctx_eval's), MW_PREC_FP64 or
 * MW_PREC_MIXED, models of at most 4 streams and 16 regions; no per-star
 * output.  Its working memory (constants, work items, block partials,
 * reduced rows) is a fixed array of MW_NOALLOC_SCRATCH_BYTES that every GPU
 * the library runs on holds until the process ends; an evaluation that needs
 * more (about 80 bytes per logical block -- a 256-column tile, or one of its
 * r-windows on a region of few tiles -- and 112 per 256 stars) fails
 * with MW_E_INVALID.  Calls on one GPU share that array: calls that overlap
 * on one GPU race, so the caller serialises them.  stream: a hipStream_t, or
 * NULL for the null stream; the call waits for it before returning. */
#define MW_NOALLOC_SCRATCH_BYTES ((size_t)4 << 20)
int mw_ctx_eval_noalloc(mw_ctx* ctx, const double* params, int precision, void* stream,
                        mw_result* result);

/* Device bytes the context holds resident. */
size_t mw_ctx_resident_bytes(const mw_ctx* ctx);

void mw_ctx_destroy(mw_ctx* ctx);

/* ------------------------------------------------------------------------
 * Batched evaluation (no reference counterpart): many parameter vectors over
 * several resident wedges of uneven size in one persistent launch.  Work
 * items (one 256-column tile of one region of one vector) are ordered by
 * cost, largest first, and taken from a device counter by a grid sized to
 * the GPU; every vector's result is bitwise the same as its single
 * evaluation.  With MW_BATCH_GRAPH the launch sequence is captured once as a
 * HIP graph and replayed by later runs with new parameter values.
 * ------------------------------------------------------------------------ */

typedef struct mw_batch mw_batch;

#define MW_BATCH_GRAPH 1

/* n_vectors[i] parameter vectors for context ctxs[i]; all contexts on one
 * device.  The contexts must outlive the batch. */
int mw_batch_create(mw_ctx* const* ctxs, const int* n_vectors, int n_ctxs, int precision,
                    int flags, mw_batch** out);

/* params: the vectors of ctxs[0], then of ctxs[1], ..., each
 * mw_ctx_parameter_count(ctx) doubles; results: one per vector, same order. */
int mw_batch_run(mw_batch* batch, const double* params, mw_result* results);

/* Kernel variant (inner-loop unroll and occupancy target, 0 ..
 * mw_batch_variant_count() - 1) and persistent blocks per compute unit
 * (1-16), as chosen by the autotuner. */
int mw_batch_set_variant(mw_batch* batch, int variant, int blocks_per_cu);
int mw_batch_variant_count(void);
/* Scratch (register spill) bytes per work-item of the integral kernels this
 * batch launches at the given variant (after its first mw_batch_run, which
 * plans the launch); the autotuner prefers variants that spill nothing. */
int mw_batch_variant_scratch(const mw_batch* batch, int variant, size_t* bytes);
int mw_batch_vector_count(const mw_batch* batch);
/* Timing of the last run: total_ms from the constants' upload to the results'
 * download; integral_ms the persistent integral launch(es) alone (equal to
 * total_ms for a replayed graph, where only the whole sequence is timed). */
int mw_batch_last_timing(const mw_batch* batch, mw_timing* out);
void mw_batch_destroy(mw_batch* batch);

/* ------------------------------------------------------------------------
 * Request server (no reference counterpart; the reference evaluates one work
 * unit per process).  Wedges are made resident once; requests are queued as
 * they arrive and a dispatcher thread admits them, in arrival order, into
 * batches whenever a runner (a stream with its own buffers) is free and the
 * batch's scratch fits the HBM budget next to the resident wedges.  Up to
 * max_inflight batches run concurrently on the device.  A request's result is
 * bitwise the same alone and in any mix.  Each request is collected once
 * (poll or wait).
 * ------------------------------------------------------------------------ */

typedef struct mw_server mw_server;

typedef struct {
    size_t hbm_budget_bytes;     /* resident wedges + in-flight scratch; 0 = no limit */
    int max_inflight;            /* batches on the device at once, 1-16 */
    int max_batch_vectors;       /* requests admitted per batch */
    int variant;                 /* kernel variant (see mw_batch_set_variant) */
    int blocks_per_cu;           /* persistent workgroups per compute unit (0 = 2) */
} mw_server_config;

typedef struct {
    double submit_s, dispatch_s, done_s;   /* steady-clock seconds */
} mw_request_times;

typedef struct {
    size_t resident_bytes, inflight_scratch_bytes;
    long pending, batches, completed, largest_batch;
    int inflight_batches;
} mw_server_stats;

int mw_server_create(const mw_server_config* cfg, mw_server** out);
/* Makes a wedge resident (as mw_ctx_create); fails with MW_E_NOMEM beyond the budget. */
int mw_server_add_wedge(mw_server* server, const ASTRONOMY_PARAMETERS* ap, const STAR_POINTS* sp,
                        int* wedge_id);
/* Queues one evaluation (params: mw_ctx_parameter_count doubles); never blocks. */
int mw_server_submit(mw_server* server, int wedge_id, const double* params, int precision,
                     long* request_id);
/* *finished = 1 and the result when the request is done, else 0. */
int mw_server_poll(mw_server* server, long request_id, mw_result* out, mw_request_times* times,
                   int* finished);
int mw_server_wait(mw_server* server, long request_id, mw_result* out, mw_request_times* times);
int mw_server_get_stats(mw_server* server, mw_server_stats* out);
/* Finishes queued and in-flight requests, then frees everything. */
void mw_server_destroy(mw_server* server);

/* ------------------------------------------------------------------------
 * Stream separation (milkyway_separation/cpu/separation.c:37-254)
 *
 * After an evaluation, each star gets each stream's membership probability
 *   p_s = (eps_s S_s / I_s) / (sum_t eps_t S_t / I_t + eps_b B / I_b),
 *   eps_s = exp(w_s) / (1 + sum_t exp(w_t)),  eps_b = 1 / (1 + sum_t exp(w_t)),
 * with B and S_s the star's convolved densities and I the integrals.  It is
 * then assigned to stream s (1-based) or to the background (0) by the i-th
 * drand48() number after srand48(seed) (prob_ok, probability.c:310-403).
 * The reference seeds with time(NULL); a caller that wants the reference's
 * behaviour passes time(NULL), and any fixed seed gives the assignment that
 * srand48(seed) would give the reference.  The output position is
 * transform_point of the star (stVector.c:155-169): the star's direction at
 * distance g = 5 log10(100 r) - 6.8, rotated so that the wedge's normal is +z.
 * ------------------------------------------------------------------------ */

typedef struct {
    int number_streams;
    int number_stars;
    double epsilon_b;
    double epsilon_s[MW_MAX_STREAMS];
    double expected[MW_MAX_STREAMS];   /* nstars: sum over stars of p_s, in star order */
    int separated[MW_MAX_STREAMS];     /* q: stars assigned to stream s */
    double background_integral;        /* the integrals used */
    double stream_integrals[MW_MAX_STREAMS];
} mw_separation_summary;

/* Separation of the context's stars for one parameter vector (params: the
 * context's parameter count, or NULL for the parameter file's own values).
 * integrals: [1 + S] background then stream integrals, or NULL to compute
 * them first (one evaluation).  Outputs, each optional (NULL skips it):
 * assignment [N], xyz [N][3], probability [N][S].  Needs 1 to 4 streams. */
int mw_ctx_separation(mw_ctx* ctx, const double* params, int precision, const double* integrals,
                      unsigned long seed, int* assignment, double* xyz, double* probability,
                      mw_separation_summary* summary);

/* Replaces separation (separation.c:37-254), which reads ap and sp from
 * globals: writes "s x y z" per star to filename ("%d %lf %lf %lf", as
 * :229) and prints the reference's summary lines to stdout.  Uses the
 * parameter values held in ap and the drop-in precision. */
int mw_hip_separation(const char* filename, double background_integral,
                      const double* stream_integrals, const ASTRONOMY_PARAMETERS* ap,
                      const STAR_POINTS* sp, unsigned long seed);

/* Message of the last failure on this thread ("" if none). */
const char* mw_last_error(void);

/* Number of visible GPUs (0 when none). */
int mw_device_count(void);

/* Makes GPU `device` (0 .. mw_device_count() - 1) the calling thread's
 * current device: contexts, batches and the drop-in calls made on this thread
 * afterwards run there (one process per GPU picks its GPU with it). */
int mw_set_device(int device);

/* Waits for all work of the calling thread's current device. */
int mw_device_synchronize(void);

#ifdef __cplusplus
}
#endif

#endif /* MW_HIP_H */
