/*
 * warmstart.h - C ABI of the warm-start network kernels (libwarmstart.so).
 *
 * A multilayer perceptron maps each LP's parameters (medium concentrations
 * and exchange bounds) to an initial primal, an initial dual and active-set
 * logits.  GEMMs run in BF16 on the MFMA units with FP32 accumulation; the
 * master weights, biases, Adam moments, activations kept for the backward
 * pass and the loss are FP32.  All pointers are device pointers, matrices are
 * row-major, `stream` is a hipStream_t.  No reference counterpart: the
 * reference's learned warm starts are torch modules
 * (src/temporal_lp_model.py:58-129, src/supervised_graph_lp.py:13-91).
 */
#ifndef WARMSTART_H
#define WARMSTART_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* C[M][N] = op(A) op(B) (+ C if accumulate).  A, B: BF16 (uint16_t).
 * op(A) is A[M][K] (trans_a = 0, leading dimension lda) or the transpose of
 * A[K][M] (trans_a = 1); op(B) is B[K][N] (trans_b = 0) or the transpose of
 * B[N][K] (trans_b = 1).  Returns 0, or -1 on invalid arguments. */
int ws_gemm_bf16(int M, int N, int K, const uint16_t *A, int lda, int trans_a, const uint16_t *B, int ldb,
                 int trans_b, float *C, int ldc, int accumulate, void *stream);

/* C[M][N] = A B^T (+ C if accumulate) for k-contiguous A[M][K] and B[N][K]
 * (leading dimensions multiples of 8, 16-byte aligned) with K split into
 * `splits` ranges multiplied by separate blocks into work[s][M][N] (FP32,
 * splits x M x N floats; unused when splits = 1) and summed into C in range
 * order (deterministic).  For products whose output has few tiles but a long
 * K (the backward product through the output layer).  Returns 0, or -1 on
 * invalid arguments. */
int ws_gemm_bf16_splitk(int M, int N, int K, const uint16_t *A, int lda, const uint16_t *B, int ldb, float *C,
                        int ldc, int accumulate, float *work, int splits, void *stream);

/* Inference products with fused epilogues (op(A) = A[M][K], op(B) = the
 * transpose of B[N][K], both k-contiguous BF16):
 * ws_gemm_bf16_act: H[M][N] (leading dimension ldh) = bf16(gelu(A B^T + bias)).
 * ws_gemm_bf16_start: the warm-start output layer.  B's N = 2 nx + ny rows
 * interleave (x_j, logit_j) for j < nx, then y_i; with v = (A B^T + bias)
 * per column, x0[r][j] = clamp(v * scale + shift, lower[r][j], upper[r][j]),
 * moved to the nearer finite bound when snap and logit_j > 0, and y0[r][i] =
 * v * scale + shift (clipped to <= 0 for i >= neq), stored FP64 batch-major
 * (lower, upper: FP64 [M][nx]).  Returns 0, -1 on invalid arguments. */
int ws_gemm_bf16_act(int M, int N, int K, const uint16_t *A, int lda, const uint16_t *B, int ldb,
                     const float *bias, uint16_t *H, int ldh, void *stream);
int ws_gemm_bf16_start(int M, int nx, int ny, int K, const uint16_t *A, int lda, const uint16_t *B, int ldb,
                       const float *bias, const float *scale, const float *shift, const double *lower,
                       const double *upper, int neq, int snap, double *x0, double *y0, void *stream);

/* X[M][ldx] = bf16((F[M][n_in] - mean) / std), columns n_in..ldx-1 zero. */
int ws_standardize(const float *F, const float *mean, const float *std, uint16_t *X, int M, int n_in, int ldx,
                   void *stream);

/* Z[M][N] += bias[N]; H = gelu(Z) as BF16 (tanh form) into H[M][N]; Z keeps
 * the pre-activation for the backward pass.  gelu == 0: H = bf16(Z).  H may be
 * NULL (output layer: only Z is needed). */
int ws_bias_act(float *Z, const float *bias, uint16_t *H, int M, int N, int gelu, void *stream);

/* dZ = dH * gelu'(Z) as BF16 (gelu == 0: dZ = bf16(dH)); db[N] += column sums of dH * gelu'(Z)
 * (atomic adds over 64-row slabs: the last bits of db depend on their order). */
int ws_act_backward(const float *dH, const float *Z, uint16_t *dZ, float *db, int M, int N, int gelu,
                    void *stream);

/* Loss of the network output O[M][N] (N = nx + ny + na): mean squared error on
 * the first nx and next ny columns against T, binary cross-entropy on the last
 * na columns (logits) against labels in T; FP32 gradient dO of the mean loss
 * (weights wx, wy, wa per part); per-row loss into row_loss[M] (may be NULL). */
int ws_loss(const float *O, const float *T, float *dO, float *row_loss, int M, int nx, int ny, int na,
            float wx, float wy, float wa, void *stream);

/* Adam (Kingma & Ba) on FP32 master weights with bias correction at step t;
 * writes the BF16 copy used by the GEMMs. */
int ws_adam(float *w, const float *g, float *m, float *v, uint16_t *w_bf16, int64_t n, float lr, float beta1,
            float beta2, float eps, int t, void *stream);

/* Y[C][R] = transpose of X[R][C] (BF16). */
int ws_transpose_bf16(const uint16_t *X, uint16_t *Y, int R, int C, void *stream);

/* Round FP32 to BF16 (round to nearest even). */
int ws_to_bf16(const float *x, uint16_t *y, int64_t n, void *stream);

const char *ws_version(void);

#ifdef __cplusplus
}
#endif
#endif
