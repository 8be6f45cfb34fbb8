/*
 * lpe.h - C ABI of the MI355X batched LP engine (liblpe.so).
 *
 * One handle owns one batch of B linear programs that share a sparsity
 * pattern (and, except for a declared set of per-LP "variable" entries, the
 * coefficient values) of an m x n constraint matrix:
 *
 *     minimise   c_b . x
 *     subject to A_b[:neq] x == rhs_b[:neq]
 *                A_b[neq:] x <= rhs_b[neq:]
 *                lower_b <= x <= upper_b              (b = 0..B-1)
 *
 * Row duals follow HiGHS: y_ub <= 0, reduced costs r = c - A^T y.
 *
 * Conventions for every entry point:
 *  - return 0 (LPE_OK) or a negative code; nothing is mutated on error and
 *    lpe_last_error() describes the failure;
 *  - "device" pointers are HIP device addresses (torch tensors' data_ptr());
 *    per-LP arrays are batch-major [B][len] FP64 (row b is LP b);
 *  - the stream argument is the caller's hipStream_t (0 = null stream); the
 *    engine enqueues on it and synchronises it only where stated;
 *  - a handle is single-owner and not thread-safe.
 *
 * The reference's engine is Python over CuPy; each entry point names the
 * reference interface it replaces.
 */
#ifndef LPE_H
#define LPE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LPE_OK 0
#define LPE_EINVAL (-1)   /* invalid argument / input values              */
#define LPE_EDEVICE (-2)  /* HIP runtime failure                          */
#define LPE_ESTATE (-3)   /* call not valid in the handle's current state */

/* iterate precision of the PDHG bulk work */
#define LPE_PREC_FP64 0
#define LPE_PREC_FP32 1

/* per-LP status codes written by lpe_state() */
#define LPE_STATUS_RUNNING 0
#define LPE_STATUS_OPTIMAL 1        /* FP64 certificate passed             */
#define LPE_STATUS_LIMIT 2          /* budget exhausted, not certified     */
#define LPE_STATUS_PRIMAL_INFEASIBLE 3
#define LPE_STATUS_DUAL_INFEASIBLE 4
#define LPE_STATUS_NUMERICAL 5

/* algorithm families */
#define LPE_METHOD_FIXED 0      /* fixed-metric PDHG, reference semantics  */
#define LPE_METHOD_RESTARTED 1  /* restarted-averaged PDHG (Ruiz scaling)  */
#define LPE_METHOD_HALPERN 2    /* restarted reflected Halpern PDHG (Ruiz
                                 * scaling, same checks and restarts)       */

typedef struct lpe_handle lpe_handle;

/* Problem structure.  Host pointers; copied during lpe_create. */
typedef struct lpe_desc {
    int32_t m, n, neq, batch;
    int64_t nnz;
    const int32_t *row_ptr;   /* [m+1] CSR, sorted unique columns per row */
    const int32_t *col_idx;   /* [nnz]                                     */
    const double *values;     /* [nnz] shared coefficient values            */
    int32_t n_var;            /* number of per-LP coefficient positions     */
    const int64_t *var_pos;   /* [n_var] positions in values[] that vary    */
    int32_t precision;        /* LPE_PREC_*                                 */
    int32_t method;           /* LPE_METHOD_*                               */
    double step_safety;       /* eta in the diagonal metric, (0,1)          */
    double theta;             /* extrapolation, [0,1]                       */
    double primal_weight;     /* initial omega > 0; 0 = ||c^||/||b^|| per LP  */
    int32_t ruiz_iterations;  /* 0 disables equilibration                   */
    int32_t shared_c;         /* 1: c is one [n] vector for every LP        */
} lpe_desc;

/* Per-LP vectors, device pointers, batch-major FP64. */
typedef struct lpe_vectors {
    const double *c;          /* [B][n] or [n] when shared_c                */
    const double *rhs;        /* [B][m]                                     */
    const double *lower;      /* [B][n]  (-inf allowed)                     */
    const double *upper;      /* [B][n]  (+inf allowed)                     */
    const double *var_values; /* [B][n_var] or NULL (keep previous)         */
} lpe_vectors;

/* Solve controls (lpe_solve). */
typedef struct lpe_solve_opts {
    int64_t max_iterations;   /* PDHG iteration budget per LP                */
    int32_t check_interval;   /* iterations between device checks (chunk)    */
    int32_t polish;           /* 1: FP64 active-set polish                   */
    double polish_kkt;        /* relative KKT at which polish is first tried */
    int32_t polish_cg_iterations; /* CG steps per polish attempt             */
    int32_t use_graph;        /* 1: replay chunks as hipGraphs               */
    int32_t rows_per_wave;    /* autotuned launch shape (0 = default)        */
    int32_t lps_per_lane;     /* 1, 2 or 4 LPs per lane (0 = default)        */
    double eps_certify_primal;/* reference gates (0 = 1e-5, 1e-7, 1e-7)      */
    double eps_certify_dual;
    double eps_certify_gap;
    int32_t stop_after_failed_polishes; /* > 0: end the solve once at most
                                 * 1/16 of the batch is still running and each
                                 * of those LPs has failed this many polish
                                 * attempts (their status is then
                                 * LPE_STATUS_LIMIT; 0 = run to the budget)   */
    double polish_kkt_start;  /* relative KKT for a polish of the starting
                                 point itself (0 = 1e-3 * polish_kkt)         */
    int32_t polish_stable_checks; /* > 0: first polish only after the active
                                 * set held for this many checks (or 4 chunks
                                 * below polish_kkt)                          */
    int32_t first_check_interval; /* > 0: iterations before the first check
                                 * (a warm start may be polished sooner); the
                                 * graph chunk is then this length when it
                                 * divides check_interval (0 = check_interval) */
    int32_t infeasibility_interval; /* restarted / Halpern methods: test the
                                 * iterate differences since the last restart
                                 * for a primal or dual infeasibility
                                 * certificate every this many checks from the
                                 * 8th check on (0 = 4, < 0 = never); a
                                 * certified LP ends LPE_STATUS_PRIMAL_INFEASIBLE
                                 * or LPE_STATUS_DUAL_INFEASIBLE               */
    double eps_infeasible;    /* certificate tolerance: largest violation of
                                 * the ray <= eps * |ray objective| (0 = 1e-8) */
} lpe_solve_opts;

/* Summary of the last lpe_solve. */
typedef struct lpe_solve_info {
    int64_t iterations;       /* PDHG iterations enqueued (max over LPs)     */
    int32_t certified;        /* LPs with LPE_STATUS_OPTIMAL                 */
    int32_t polish_attempts;
    int32_t restarts;         /* total restarts over all LPs                 */
    int32_t chunks;
    double seconds;           /* wall time of the solve (host clock)         */
    double pdhg_ms;           /* device time of PDHG chunks (hipEvents)      */
    double check_ms;          /* device time of checks                       */
    double polish_ms;         /* device time of polish + certificates        */
    int64_t pdhg_kernel_launches;
    int64_t cg_steps;         /* preconditioned-CG steps of the polish rounds */
    int64_t cg_lp_steps;      /* sum over CG blocks of the LPs with an active
                                 system times the block's steps (work measure) */
} lpe_solve_info;

/* Replaces GpuPdhgCorrector.__init__ / GpuPdhgWorkspace.__init__
 * (src/gpu_pdhg_corrector.py:99-218, src/gpu_pdhg_workspace.py:65-155):
 * validates the pattern, builds the transposed pattern, equilibration and
 * diagonal metric, and allocates every device buffer. */
int lpe_create(const lpe_desc *desc, lpe_handle **out);

/* Replaces GpuPdhgWorkspace.update_problem (src/gpu_pdhg_workspace.py:209-267):
 * validated, all-or-nothing copy of c, rhs, bounds and per-LP coefficients.
 * Invalidates the iterate (lpe_reset is required afterwards).
 * Synchronises the stream (validation result). */
int lpe_update(lpe_handle *h, const lpe_vectors *v, void *stream);

/* Replaces GpuPdhgWorkspace.reset (src/gpu_pdhg_workspace.py:269-293):
 * cold (NULL) or warm start, x projected into the box and inequality duals
 * to y <= 0; clears restart state and accepted flags. */
int lpe_reset(lpe_handle *h, const double *x0, const double *y0, void *stream);

/* Replaces GpuPdhgWorkspace.run (src/gpu_pdhg_workspace.py:340-367): enqueue
 * exactly `iterations` PDHG iterations (no checks, no restarts); graph chunks
 * of `chunk` iterations when use_graph, the tail as plain launches. */
int lpe_run(lpe_handle *h, int64_t iterations, int32_t chunk, int32_t use_graph, void *stream);

/* Replaces GpuPdhgWorkspace.set_accepted (src/gpu_pdhg_workspace.py:295-308):
 * host mask [B] of LPs to freeze; unfreezing requires lpe_reset. */
int lpe_set_accepted(lpe_handle *h, const uint8_t *mask, void *stream);

/* Replaces the GpuPdhgCorrector.solve loop (src/gpu_pdhg_corrector.py:273-438)
 * and the resident service loop (src/gpu_resident_reduced_pdhg.py:104-177):
 * checks, restarts, polish and certification on the device, host control
 * between chunks.  Synchronises the stream. */
int lpe_solve(lpe_handle *h, const lpe_solve_opts *opts, lpe_solve_info *info, void *stream);

/* Replaces certify_blocks_device (src/gpu_block_lp.py:52-132): FP64
 * certificate of the given iterates (or of the engine's iterate when x and y
 * are NULL) in original units.  metrics: device [B][LPE_NMETRIC]. */
#define LPE_NMETRIC 8
#define LPE_M_PRIMAL 0
#define LPE_M_DUAL 1
#define LPE_M_GAP 2
#define LPE_M_OBJECTIVE 3
#define LPE_M_PASSED 4
#define LPE_M_FINITE 5
#define LPE_M_BOUND_GAP 6
#define LPE_M_ROW_GAP 7
int lpe_certify(lpe_handle *h, const double *x, const double *y, double *metrics,
                double eps_primal, double eps_dual, double eps_gap, void *stream);

/* A caller-owned batch of B LPs sharing one sparsity pattern, all device
 * pointers (batch-major per-LP arrays, FP64 values, int32 indices). */
typedef struct lpe_batch_problem {
    int32_t m, n, neq, batch;
    int64_t nnz;
    const int32_t *row_ptr;   /* [m+1] CSR of A, sorted columns per row       */
    const int32_t *col_idx;   /* [nnz]                                        */
    const int32_t *col_ptr;   /* [n+1] transposed pattern (CSR of A^T) ...    */
    const int32_t *row_idx;   /* [nnz] ... rows ascending inside each column  */
    const int32_t *t_pos;     /* [nnz] position in the CSR values of each A^T entry */
    const double *values;     /* [nnz], or [B][nnz] when values_per_lp        */
    int32_t values_per_lp;
    int32_t c_per_lp;         /* c is [n] (0) or [B][n] (1)                   */
    const double *c;
    const double *rhs;        /* [B][m]                                       */
    const double *lower;      /* [B][n]  (-inf allowed)                       */
    const double *upper;      /* [B][n]  (+inf allowed)                       */
} lpe_batch_problem;

/* No counterpart in the reference: the variant of lpe_certify for callers
 * that cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The same FP64 certificate and gates as
 * lpe_certify, of (x [B][n], y [B][m]) for a caller-owned problem, with no
 * handle; metrics [B][LPE_NMETRIC] on the device.  Its block partials live
 * in a static device array of lpe_certify_noalloc_capacity() doubles (8 MiB)
 * per GPU, so it calls no hipMalloc/hipFree and touches no memory but its
 * arguments and that array; a batch needing more than the array holds
 * (batch * lpe_certify_noalloc_need(m, n) doubles) is rejected with
 * LPE_EINVAL before any device work.  Enqueued on `stream`, not
 * synchronised.  Calls that overlap on one GPU race on the shared array. */
int lpe_certify_noalloc(const lpe_batch_problem *p, const double *x, const double *y, double *metrics,
                        double eps_primal, double eps_dual, double eps_gap, void *stream);
int64_t lpe_certify_noalloc_capacity(void);
int64_t lpe_certify_noalloc_need(int32_t m, int32_t n);

/* Replaces GpuPdhgWorkspace.state (src/gpu_pdhg_workspace.py:369-380) with
 * owned copies: x [B][n], y [B][m] in original units (FP64), status [B]. */
int lpe_state(lpe_handle *h, double *x, double *y, int32_t *status, void *stream);

/* FP64 active-set polish of the current iterate for the LPs selected by the
 * host mask (NULL = all running LPs); no reference counterpart - the
 * reference has no polish (src/gpu_pdhg_corrector.py:12-14 fails closed). */
int lpe_polish(lpe_handle *h, const uint8_t *mask, int32_t cg_iterations, void *stream);

/* Batched sparse products with the original coefficients, FP64:
 * transpose = 0: out[B][m] = A_b x[B][n];  1: out[B][n] = A_b^T x[B][m].
 * precision selects the kernel arithmetic (FP32 rounds inputs and values).
 * Replaces the csr_warp_spmv launches (src/gpu_pdhg_workspace.py:23-37,313,316). */
int lpe_spmv(lpe_handle *h, int32_t transpose, int32_t precision, const double *in,
             double *out, void *stream);

/* Scaled problem data used by the iteration (for inspection and tests):
 * row_scale [m], col_scale [n], sigma [m], tau [n] on the host.  Replaces
 * the metric attributes of GpuPdhgCorrector (src/gpu_pdhg_corrector.py:185-196). */
int lpe_get_scaling(lpe_handle *h, double *row_scale, double *col_scale, double *sigma,
                    double *tau);

/* Replaces the tau / sigma / theta arguments of GpuPdhgWorkspace.__init__ and
 * update_problem (src/gpu_pdhg_workspace.py:65-68, :209-267): a caller-supplied
 * diagonal metric for the fixed-metric method, tau [B, n] and sigma [B, m]
 * (device, row-major per LP, finite and strictly positive; validated before
 * anything changes) and theta in [0, 1].  Kept until the next lpe_update (which
 * restores the default metric); lpe_reset is required afterwards. */
int lpe_set_metric(lpe_handle *h, const double *tau, const double *sigma, double theta, void *stream);

/* Per-LP primal weights [B] (host). */
int lpe_get_primal_weight(lpe_handle *h, double *omega);

/* Set per-LP primal weights [B] (host, > 0) after lpe_reset, e.g. the
 * weights a previous solve of the same environments converged to (warm
 * start of the restarted method).  A desc.primal_weight of 0 selects the
 * PDLP default ||c^||/||b^|| per LP at every reset.  No reference
 * counterpart (src/gpu_pdhg_corrector.py:99-106 takes one fixed weight). */
int lpe_set_primal_weight(lpe_handle *h, const double *omega, void *stream);

/* Launch-shape controls used by the autotuner. */
int lpe_set_launch(lpe_handle *h, int32_t rows_per_wave, int32_t lps_per_lane);

/* Time `iterations` PDHG iterations on the stream with hipEvents; returns ms. */
int lpe_time_iterations(lpe_handle *h, int64_t iterations, int32_t chunk, int32_t use_graph,
                        double *ms, void *stream);

/* Average duration (ms) of `reps` back-to-back launches of one kernel on
 * the engine's stream, with hipEvents: kind 0 = dual step (the fused A x
 * SpMV + dual update), 1 = primal step (fused A^T y SpMV + primal update +
 * extrapolation + averaging), 3 = the polish's CG columns pass q = D A^T p
 * with the |q|^2 partials, 4 = its CG rows pass s = R A q with the CG update
 * (w += alpha p, r -= alpha s) and the residual partials, 5 = the CG direction
 * p = D^-1 r + beta p (3-5: both CG systems, every LP of the batch active; the
 * polish scratch vectors are overwritten); kind 2 (a separate CG update
 * pass) no longer exists and is rejected.  The iterate is advanced; reset
 * before solving again.  No
 * reference counterpart (the reference times whole graph chunks,
 * scripts/benchmark_pdhg_graph_chunk.py:217-234). */
int lpe_time_kernel(lpe_handle *h, int32_t kind, int32_t reps, double *ms, void *stream);

int lpe_destroy(lpe_handle *h);

/* Last error message of the calling thread (NUL-terminated, truncated). */
int lpe_last_error(char *buf, int64_t size);

/* Library version string, e.g. "lpe 0.1 gfx950". */
const char *lpe_version(void);

#ifdef __cplusplus
}
#endif
#endif
