/*
 * paddi.h -- C ABI of the MI355X (gfx950) discord engine.
 *
 * One GPU's share of PADDi's discord search: for every subsequence length m in
 * minL..maxL (step `step`), the top-k subsequences whose nearest
 * non-overlapping neighbour is farthest away under z-normalised Euclidean
 * distance.  All distances crossing this boundary are SQUARED z-normalised
 * distances, as in the reference (struct Discord, src/PALMAD_types.hpp:7-11;
 * r is passed squared, src/main.cu:144).
 *
 * Conventions (shared by every entry point):
 *   - plain pointers and sizes only; host pointers unless the name says
 *     "device"; the library owns its device scratch;
 *   - calls block until the result is on the host and return a status code
 *     (never exit(); the reference's CHECK() does, src/common.h:6-16);
 *   - a context may be used by one host thread at a time; separate contexts
 *     are independent;
 *   - row i and column j form a valid pair iff |i - 1 - j| >= m
 *     (src/cpu_kernels/cpu_kernels.cpp:57; paddi_set_exclusion selects the
 *     reference GPU kernels' |i - j| >= m - 1 instead); a row is pruned by a valid pair
 *     with ED^2 <= r2 (cpu_kernels.cpp:59-72); a window with zero standard
 *     deviation has ED^2 = 2m to everything (src/PD3_kernels.cuh:63-64).
 */
#ifndef PADDI_H
#define PADDI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t paddi_status;
enum {
    PADDI_OK = 0,
    PADDI_ERR_INVALID = 1,      /* bad argument (sizes, ranges, null pointers) */
    PADDI_ERR_DEVICE = 2,       /* HIP runtime error */
    PADDI_ERR_NOMEM = 3,        /* device allocation failed */
    PADDI_ERR_CAPACITY = 4,     /* output buffer too small; *count says how many */
    PADDI_ERR_UNSUPPORTED = 5   /* request outside the engine's limits: m > 2048, a call of
                                   2^31 rows or more (rows = windows at the shortest length,
                                   padded to 64 per series, plus 128 per series) */
};

/* Output record: struct Discord {int ind; DATA_TYPE dist;} (src/PALMAD_types.hpp:7-11)
 * plus the length it was found at and, for batches, the series it belongs to. */
typedef struct {
    int64_t index;      /* 0-based start of the subsequence in its series */
    int32_t length;     /* m */
    int32_t series;     /* series number in a batch call, 0 otherwise */
    float nn_dist2;     /* squared z-normalised distance to the nearest neighbour */
    float range2;       /* squared range r of the attempt that returned it */
} paddi_discord;

enum {
    PADDI_MODE_PLAIN = 0,    /* each series alone, per length the r-descent of main.cu:136-233 */
    PADDI_MODE_BATCHED = 1,  /* every series of the call in each launch; per length one pass at
                                an r taken from the previous length's discords (exact values at
                                the new length), retried lower on failure; a pass's kernels run
                                without a host round trip and may replay as a hipGraph
                                (paddi_tunables.use_graph); with paddi_tunables.device_lengths
                                the lengths after the first run on the device, one captured
                                graph per length (its range, pass and exact top-k), the host
                                synchronising every device_lengths lengths */
    PADDI_MODE_LEARNED = 2   /* batched, with the first length's r from the learned proposal;
                                the descent remains the fallback */
};

typedef struct {
    int32_t mode;            /* PADDI_MODE_* */
    int32_t max_attempts;    /* r-descent attempts per length before r = 0 (default 64) */
    int32_t flags;           /* reserved, 0 */
    int32_t reserved;
    const char *model_path;  /* learned mode: weights file, NULL = built-in default */
} paddi_opts;

/* Per-call counters (optional output). */
typedef struct {
    int64_t attempts;        /* (m, r) passes run, summed over lengths */
    int64_t survivors;       /* rows that survived selection, summed */
    int64_t exact_refines;   /* candidates refined in double precision */
    int64_t pairs;           /* pair evaluations counted on the device (estimate) */
    double seconds;          /* wall time of the call */
    double kernel_ms[4];     /* derive, select, refine, exact: summed kernel time (ms, hipEvent) */
    int64_t launches[4];     /* launches of the same four kernels */
    int64_t live_pairs[2];   /* of the select and refine scans' pair evaluations, those on rows
                                still alive when their 64-column tile started */
    int64_t scan_pairs[2];   /* the select and refine scans' pair evaluations */
} paddi_stats;

typedef struct paddi_ctx paddi_ctx;
typedef struct paddi_series paddi_series;

/* Schedule tunables of a context (chosen by paddi_amd/autotune.py; no counterpart in the
 * reference, whose tile shapes are compile-time macros, src/parameters.hpp:4-24). */
typedef struct {
    int64_t select_window;   /* columns each 64-row block scans in k_scan_select */
    int64_t refine_chunk;    /* longest k_scan_refine work item (columns) */
    int64_t min_chunk;       /* shortest k_scan_refine work item (columns) */
    int64_t refine_items;    /* target work items per refine round */
    int32_t witness_tiles;   /* witness ranges scanned first (0 = off) */
    int32_t speculate;       /* candidates refined per exact round */
    int32_t use_graph;       /* capture the per-length device pipeline in a hipGraph */
    int32_t prev_shift;      /* the next length's r: previous discords evaluated at i and at
                                i-1 .. i-prev_shift (0..4) */
    int32_t bootstrap;       /* learned form, first length: r also bounded from below by the
                                exact NN of the best-scored candidate rows (0/1) */
    int32_t bound_candidates;/* those candidate rows per series (1..256; 0 = the default 32) */
    int32_t direct_budget;   /* a refine work item whose block has L live rows computes them on
                                the matrix cores alone when L (m + 15) <= direct_budget (0 = never) */
    int32_t direct_rows;     /* and L <= direct_rows (0..4) */
    int32_t device_lengths;  /* batched and learned forms: run the lengths after the first on the
                                device, each one captured graph (range from the previous length's
                                discords, pass, exact top-k), the host synchronising every
                                device_lengths lengths (0 = every length on the host) */
    int32_t device_rounds;   /* exact rounds a length's graph holds (1..4); a length needing more
                                is finished by the host */
    int32_t refine_rounds;   /* refine rounds per pass (3..4): cyclic offsets past the select
                                window up to 2, 8 (and 32) x refine_chunk, then the rest */
    int32_t seed_witnesses;  /* batched and learned forms, a call's first length: scan first the
                                tile of a row with the same SAX word (0/1) */
    int32_t refine_order;    /* refine work items block-major (0), chunk-major (1: every
                                block's nearest chunk first), or chunk-major at a call's first
                                length only (2) */
} paddi_tunables;

const char *paddi_status_str(paddi_status s);
const char *paddi_version(void);
/* 1 if this build supports PADDI_MODE_* `mode` (the learned mode needs its model).
 * No counterpart in the reference, which has a single form. */
int32_t paddi_has_mode(int32_t mode);

/* Context on HIP device `device`.  Replaces the per-GPU allocation block of
 * src/main.cu:54-125 (the library owns all device buffers). */
paddi_status paddi_create(int32_t device, paddi_ctx **out);
void paddi_destroy(paddi_ctx *ctx);
paddi_status paddi_get_tunables(paddi_ctx *ctx, paddi_tunables *out);
paddi_status paddi_set_tunables(paddi_ctx *ctx, const paddi_tunables *t);

/* Exclusion rule of a context's later calls.  The reference has two: its CPU path pairs row
 * i with column j iff |i - 1 - j| >= m, i.e. j - i >= m - 1 or i - j >= m + 1
 * (src/cpu_kernels/cpu_kernels.cpp:57; the default here), its GPU kernels iff
 * |i - j| >= m - 1 (src/PD3_kernels.cuh:70), which also pairs i - j = m - 1 and i - j = m.
 * The top-k's non-overlap rule between accepted discords (|i - a| >= m) is the same under
 * both.  Invalid rule: PADDI_ERR_INVALID. */
enum {
    PADDI_EXCL_CPU = 0,
    PADDI_EXCL_SYMMETRIC = 1
};
paddi_status paddi_set_exclusion(paddi_ctx *ctx, int32_t rule);
paddi_status paddi_get_exclusion(paddi_ctx *ctx, int32_t *rule);

/* Sliding mean and standard deviation of the n-m+1 windows of T.
 * Replaces compute_statistics / compute_statistics_cpu
 * (src/preprocessing.cuh:68-76, src/cpu_kernels/cpu_kernels.cpp:7-31);
 * computed from double-precision prefix sums, returned as float. */
paddi_status paddi_compute_statistics(paddi_ctx *ctx, const float *T, int64_t n, int32_t m,
                                      float *mean, float *std);

/* Every entry point that takes series returns PADDI_ERR_INVALID when a series holds a NaN
 * or an infinite value. */

/* Range discords at squared range r2 over the n-m+1 windows of T.
 * Replaces find_discords_cpu (src/cpu_kernels/cpu_kernels.cpp:33-78) and the
 * single-GPU do_PD3 + refine pipeline (src/PD3.cuh:15-61):
 * is_discord[i] = 1 iff every valid pair of row i has ED^2 > r2, and then
 * nn_dist2[i] is row i's nearest-neighbour ED^2 (FLT_MAX when row i has no
 * valid pair): the double-precision value for rows within the single-precision scans'
 * error bound of r2 and for distances below m/5 (up to 65536 such rows per call), the
 * scans' single-precision value elsewhere (its error is a few 1e-6 of 2m, below 1e-4
 * relative above m/5).  nn_dist2 of rows with is_discord[i] = 0 is unspecified. */
paddi_status paddi_range_discords(paddi_ctx *ctx, const float *T, int64_t n, int32_t m, float r2,
                                  int32_t *is_discord, float *nn_dist2);

/* paddi_range_discords without any allocation: no hipMalloc, hipFree or pinned-host allocation
 * during the call.  No counterpart in the reference; for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid allocation).  Same
 * arguments, results and errors as paddi_range_discords; every device buffer of the call is a
 * slice of a file-scope device array of PADDI_NOALLOC_SCRATCH_BYTES that each GPU holds from
 * the library's first use on it until the process ends, and the pinned staging the call
 * uses is allocated by paddi_create.  Series longer than PADDI_NOALLOC_MAX_POINTS return
 * PADDI_ERR_UNSUPPORTED.  Every call on one GPU shares that GPU's array (whatever context it
 * comes from) and leaves its contents there: calls that overlap on one GPU, from two
 * contexts or two threads, race. */
#define PADDI_NOALLOC_SCRATCH_BYTES (32u << 20)
#define PADDI_NOALLOC_MAX_POINTS 300000
paddi_status paddi_range_discords_noalloc(paddi_ctx *ctx, const float *T, int64_t n, int32_t m,
                                          float r2, int32_t *is_discord, float *nn_dist2);

/* One (m, r) attempt of a search: a row of the reference's times table (src/IOdata.hpp:140-177),
 * which src/main.cu:144,176,219 fills once per do_PALMAD_one_length call.  The phases of one
 * GPU (statistics, scans, exact refinement) and the counts.  In the batched forms an attempt is
 * one pass over the call's series. */
typedef struct {
    int32_t m;
    int32_t attempt;         /* 1, 2, .. within the length */
    float r2;                /* squared range (batched forms: the smallest over the pass's series) */
    int32_t discords;        /* discords accepted by the attempt (summed over the series that
                                finished the length with it) */
    int64_t survivors;       /* rows left by the scans (the table's "# Candidates") */
    double ms_derive;        /* k_derive: window statistics (on the length's first attempt) */
    double ms_scan;          /* k_scan_select + k_scan_refine */
    double ms_exact;         /* k_exact: double-precision refinement (the first attempt also
                                carries the length's range estimate) */
    double ms_total;         /* host wall time since the previous attempt ended */
} paddi_attempt_stats;

/* Top-k discords for every length m = minL, minL+step, .., <= maxL.
 * Replaces the PALMAD length loop of src/main.cu:136-233 (top-1 there; for
 * k > 1 discords are ranked by (nn_dist2 desc, index asc) and accepted when
 * |i - a| >= m for every accepted a).  out[] receives up to top_k records per
 * length, lengths ascending, each length's records in rank order; *count is
 * the number written.  opts and stats may be NULL.  A wait for device work
 * longer than PADDI_WAIT_LIMIT_S seconds (environment, read when the context
 * is created; default 600) returns PADDI_ERR_DEVICE. */
paddi_status paddi_discords(paddi_ctx *ctx, const float *T, int64_t n, int32_t minL, int32_t maxL,
                            int32_t step, int32_t top_k, const paddi_opts *opts,
                            paddi_discord *out, int64_t cap, int64_t *count, paddi_stats *stats);

/* paddi_discords with a record per (m, r) attempt, the rows of the reference's times table
 * (src/IOdata.hpp:140-177; no counterpart beyond the reference's timers): per_attempt receives
 * up to cap_attempts records in the order the attempts ran, *n_attempts their number (when it
 * exceeds cap_attempts, the later records are dropped and the call still succeeds). */
paddi_status paddi_discords_timed(paddi_ctx *ctx, const float *T, int64_t n, int32_t minL,
                                  int32_t maxL, int32_t step, int32_t top_k,
                                  const paddi_opts *opts, paddi_discord *out, int64_t cap,
                                  int64_t *count, paddi_attempt_stats *per_attempt,
                                  int64_t cap_attempts, int64_t *n_attempts);

/* Many independent series (uneven lengths) in one call; records carry
 * `series`.  Same per-series semantics as paddi_discords. */
paddi_status paddi_discords_batch(paddi_ctx *ctx, const float *const *T, const int64_t *n,
                                  int32_t nseries, int32_t minL, int32_t maxL, int32_t step,
                                  int32_t top_k, const paddi_opts *opts, paddi_discord *out,
                                  int64_t cap, int64_t *count, paddi_stats *stats);

/* Device-resident series: upload once, search many times (the bench times
 * paddi_discords_series with the series already in HBM).  These have no
 * counterpart in the reference, whose driver copies the series once per run
 * (src/main.cu:129-133). */
paddi_status paddi_series_upload(paddi_ctx *ctx, const float *T, int64_t n, paddi_series **out);
void paddi_series_free(paddi_series *s);
paddi_status paddi_discords_series(paddi_ctx *ctx, paddi_series *const *series, int32_t nseries,
                                   int32_t minL, int32_t maxL, int32_t step, int32_t top_k,
                                   const paddi_opts *opts, paddi_discord *out, int64_t cap,
                                   int64_t *count, paddi_stats *stats);

/* ---- learned range proposal (no counterpart in the reference, whose range schedule is
 * the blind r-descent of src/main.cu:136-233) ----------------------------------------
 * Features of (series, m): m/512, log10(global std)/4 and the 32-bin histogram of
 * ED^2/(2m) over 1024 seeded random non-overlapping window pairs.  Network:
 * nfeat -> hidden (ReLU) -> hidden (ReLU) -> 1, single precision, parameters as one flat
 * array W1[nfeat][hidden] b1 W2[hidden][hidden] b2 W3[hidden] b3; inputs standardised by
 * (X - fmu) / fsd.  Loss: mean squared error. */
paddi_status paddi_mlp_sizes(int32_t *nfeat, int32_t *hidden, int32_t *nparam);
paddi_status paddi_features(paddi_ctx *ctx, paddi_series *const *series, int32_t nseries,
                            int32_t m, int32_t top_k, float *out /* [nseries][nfeat] */);
paddi_status paddi_mlp_forward(paddi_ctx *ctx, const float *params, const float *fmu,
                               const float *fsd, const float *X, int32_t B, float *Y);
paddi_status paddi_mlp_gradient(paddi_ctx *ctx, const float *params, const float *fmu,
                                const float *fsd, const float *X, const float *T, int32_t B,
                                float *grad /* [nparam] */, float *loss);
/* Full-batch Adam on the device for `steps` steps; params updated in place; losses[steps]
 * (optional) receives the loss before each step. */
paddi_status paddi_mlp_train(paddi_ctx *ctx, float *params, const float *fmu, const float *fsd,
                             const float *X, const float *T, int32_t B, int32_t steps, float lr,
                             float weight_decay, float *losses);

#ifdef __cplusplus
}
#endif
#endif /* PADDI_H */
