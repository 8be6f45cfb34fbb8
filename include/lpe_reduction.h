/*
 * lpe_reduction.h - exact homogeneous equality reduction for the batched LP
 * engine (part of liblpe.so).
 *
 * A reduction plan is built once per sparsity pattern on the host: the
 * equality rows with exactly two entries, zero right-hand side and a bounded
 * coefficient ratio form a spanning forest over the columns; every tree is
 * replaced by one reduced column z_g, with x_j = w_j z_g for its members, so
 * the eliminated rows hold identically.  Per-LP data (bounds, objective,
 * kept rows' right-hand sides, per-LP matrix values) are reduced, and reduced
 * solutions expanded and their duals lifted to the original rows, by batched
 * kernels on the device - or, with on_device = 0, by the same arithmetic on
 * the host over host arrays.
 *
 * Rows that hold a per-LP coefficient (a var_pos entry) are never
 * eliminated, so one plan serves a whole batch whose LPs differ there.
 * Conventions as in lpe.h: return LPE_OK or a negative code, lpe_last_error()
 * describes a failure, per-LP arrays are batch-major [B][len] FP64.
 */
#ifndef LPE_REDUCTION_H
#define LPE_REDUCTION_H

#include <stdint.h>

#include "lpe.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lpe_reduction lpe_reduction;

/* Original pattern (host pointers, copied). */
typedef struct lpe_reduction_desc {
    int32_t m, n, neq;
    int64_t nnz;
    const int32_t *row_ptr;   /* [m+1] CSR, sorted unique columns per row   */
    const int32_t *col_idx;   /* [nnz]                                      */
    const double *values;     /* [nnz] template values                      */
    const double *rhs;        /* [m] template rhs (nonzero rows are kept)   */
    int32_t n_var;            /* per-LP coefficient positions               */
    const int64_t *var_pos;   /* [n_var] entries of values[] that vary      */
    double max_scale_ratio;   /* >= 1: largest |a|max/|a|min of a tree row and
                                 largest spread of a tree's weights          */
} lpe_reduction_desc;

typedef struct lpe_reduction_dims {
    int32_t m, n, neq;        /* reduced LP                                 */
    int64_t nnz;              /* reduced pattern                            */
    int32_t n_var;            /* per-LP positions of the reduced pattern    */
    int32_t n_eliminated;     /* eliminated equality rows                   */
    int64_t lift_nnz;         /* entries of the dual-lift map               */
} lpe_reduction_dims;

/* Per-LP status bits written by lpe_reduction_reduce. */
#define LPE_RED_INFEASIBLE 1      /* a reduced lower bound exceeds its upper  */
#define LPE_RED_OVERFLOW 2        /* a finite bound became non-finite         */
#define LPE_RED_INHOMOGENEOUS 4   /* an eliminated row has nonzero rhs        */

/* Reduced per-LP data (outputs of lpe_reduction_reduce). */
typedef struct lpe_reduced {
    double *c;                /* [B][rn], or [rn] when shared_c             */
    double *rhs;              /* [B][rm]                                    */
    double *lower, *upper;    /* [B][rn]                                    */
    double *var_values;       /* [B][r_n_var] (may be NULL when r_n_var = 0) */
    int32_t *lower_witness;   /* [B][rn] original column inducing the lower
                                 endpoint (-1: none finite)                 */
    int32_t *upper_witness;   /* [B][rn]                                    */
    int32_t *status;          /* [B] LPE_RED_* bits                         */
} lpe_reduced;

/* Replaces HomogeneousEqualityReduction.__init__
 * (src/lp_equality_reduction.py:295-461): union-find forest over the
 * two-entry homogeneous equality rows, null-space weights, groups,
 * compression, dual-lift map and the identity checks E T = 0, L E^T = I.
 * Host only (no device call). */
int lpe_reduction_create(const lpe_reduction_desc *desc, lpe_reduction **out);

int lpe_reduction_get_dims(const lpe_reduction *r, lpe_reduction_dims *dims);

/* Plan arrays (host outputs; any pointer may be NULL): group [n] (reduced
 * column of each original column), weight [n], representative [rn],
 * kept_rows [rm], eliminated_rows [n_eliminated].  Replaces the attributes
 * original_to_reduced / weights / representatives / kept_rows /
 * eliminated_rows (src/lp_equality_reduction.py:411-450). */
int lpe_reduction_get_plan(const lpe_reduction *r, int32_t *group, double *weight, int32_t *representative,
                           int32_t *kept_rows, int32_t *eliminated_rows);

/* Reduced pattern A[kept] T with template values, and the positions of its
 * per-LP entries (host outputs).  Replaces the reduced matrix of
 * HomogeneousEqualityReduction.reduce (src/lp_equality_reduction.py:533-538). */
int lpe_reduction_get_pattern(const lpe_reduction *r, int32_t *row_ptr, int32_t *col_idx, double *values,
                              int64_t *var_pos);

/* Dual-lift map L [n_eliminated x n] in CSR (host outputs).  Replaces
 * dual_lift_map (src/lp_equality_reduction.py:463-517). */
int lpe_reduction_get_lift_map(const lpe_reduction *r, int32_t *row_ptr, int32_t *col_idx, double *values);

/* Reduce B LPs: bounds intersected per group with their witnesses, c' = T^T c,
 * rhs' = rhs[kept], per-LP reduced matrix values.  in->var_values NULL = the
 * template values.  Replaces HomogeneousEqualityReduction.reduce and
 * _SegmentedBounds.reduce (src/lp_equality_reduction.py:127-160,519-560);
 * instead of raising, per-LP failures are status bits. */
int lpe_reduction_reduce(lpe_reduction *r, int32_t batch, int32_t shared_c, const lpe_vectors *in,
                         const lpe_reduced *out, int32_t on_device, void *stream);

/* x [B][n] = T z [B][rn].  Replaces expand_primal (src/lp_equality_reduction.py:578-586). */
int lpe_reduction_expand(lpe_reduction *r, int32_t batch, const double *z, double *x, int32_t on_device,
                         void *stream);

/* z [B][rn] = (T^T T)^-1 T^T x [B][n].  Replaces compress_primal
 * (src/lp_equality_reduction.py:562-576) without its optional check. */
int lpe_reduction_compress(lpe_reduction *r, int32_t batch, const double *x, double *z, int32_t on_device,
                           void *stream);

/* Original (x, y) from a reduced (z, y'): x = T z, y[kept] = y', q = c - A^T y,
 * the group's reduced cost T^T q assigned to its witness bound (lower when
 * >= 0, else upper, the representative when missing), eliminated-row duals
 * L (q - normal).  On the device, A^T y runs on `original` (an engine handle
 * of the original pattern with the same batch, holding the LPs' current
 * values; var_values is ignored); on the host (original = NULL) the plan's
 * pattern with var_values (NULL = template) is used.  Replaces
 * GpuReducedPdhgCorrector.expand_and_lift (src/gpu_reduced_pdhg.py:81-103). */
int lpe_reduction_lift(lpe_reduction *r, lpe_handle *original, int32_t batch, int32_t shared_c, const double *c,
                       const double *var_values, const double *z, const double *y_reduced,
                       const int32_t *lower_witness, const int32_t *upper_witness, double *x, double *y,
                       int32_t on_device, void *stream);

int lpe_reduction_destroy(lpe_reduction *r);

#ifdef __cplusplus
}
#endif

#endif /* LPE_REDUCTION_H */
