// Warm-start network kernels for gfx950 (see include/warmstart.h).
//
// GEMM: 128x128 output tile per 256-thread block, four waves in a 2x2
// arrangement, each wave 64x64 = 2x2 tiles of the 32x32x8 BF16 MFMA with FP32
// accumulation.  K advances by 32: the block stages op(A) as [m][k] and op(B)
// as [n][k] in LDS (the transposes of the backward products are folded into
// the staging), so each MFMA operand is one 8-byte LDS read per lane.
//
// MFMA 32x32x8 bf16 lane layout (64 lanes): A operand lane l holds
// A[l % 32][4*(l / 32) + 0..3]; B operand lane l holds B[4*(l / 32) + 0..3][l % 32];
// accumulator element r of lane l is D[8*(r / 4) + 4*(l / 32) + (r % 4)][l % 32].
#include "warmstart.h"

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace {

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float16_t __attribute__((ext_vector_type(16)));

constexpr int BM = 128, BN = 128, BK = 32, PAD = 8;

__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);   // quiet NaN
    u += 0x7fffu + ((u >> 16) & 1u);                                               // round to nearest even
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

template <int TA, int TB>
__global__ void __launch_bounds__(256) k_gemm(int M, int N, int K, const uint16_t* __restrict__ A, int lda,
                                              const uint16_t* __restrict__ B, int ldb, float* __restrict__ C,
                                              int ldc, int accumulate) {
    __shared__ __align__(16) uint16_t As[BM][BK + PAD];
    __shared__ __align__(16) uint16_t Bs[BN][BK + PAD];
    const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int wm = (wave / 2) * 64, wn = (wave % 2) * 64;
    float16_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

    for (int k0 = 0; k0 < K; k0 += BK) {
        // stage op(A) -> As[m][k]
        if (!TA) {
            const int row = tid / 2, kk = (tid % 2) * 16;
            const int gm = m0 + row;
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                const int gk = k0 + kk + e;
                As[row][kk + e] = (gm < M && gk < K) ? A[(long)gm * lda + gk] : (uint16_t)0;
            }
        } else {
            const int kr = tid / 8, mm = (tid % 8) * 16;
            const int gk = k0 + kr;
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                const int gm = m0 + mm + e;
                As[mm + e][kr] = (gm < M && gk < K) ? A[(long)gk * lda + gm] : (uint16_t)0;
            }
        }
        // stage op(B) -> Bs[n][k]
        if (!TB) {
            const int kr = tid / 8, nn = (tid % 8) * 16;
            const int gk = k0 + kr;
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                const int gn = n0 + nn + e;
                Bs[nn + e][kr] = (gn < N && gk < K) ? B[(long)gk * ldb + gn] : (uint16_t)0;
            }
        } else {
            const int col = tid / 2, kk = (tid % 2) * 16;
            const int gn = n0 + col;
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                const int gk = k0 + kk + e;
                Bs[col][kk + e] = (gn < N && gk < K) ? B[(long)gn * ldb + gk] : (uint16_t)0;
            }
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < BK / 8; ++ks) {
            short4_t a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i)
                a[i] = *reinterpret_cast<const short4_t*>(&As[wm + i * 32 + lane % 32][ks * 8 + (lane / 32) * 4]);
#pragma unroll
            for (int j = 0; j < 2; ++j)
                b[j] = *reinterpret_cast<const short4_t*>(&Bs[wn + j * 32 + lane % 32][ks * 8 + (lane / 32) * 4]);
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = m0 + wm + i * 32 + (r / 4) * 8 + (lane / 32) * 4 + (r % 4);
                const int col = n0 + wn + j * 32 + lane % 32;
                if (row < M && col < N) {
                    const long o = (long)row * ldc + col;
                    C[o] = accumulate ? C[o] + acc[i][j][r] : acc[i][j][r];
                }
            }
}

constexpr float SQRT_2_OVER_PI = 0.7978845608028654f;
constexpr float GELU_C = 0.044715f;

__device__ __forceinline__ float gelu(float z) {
    const float u = SQRT_2_OVER_PI * (z + GELU_C * z * z * z);
    return 0.5f * z * (1.0f + tanhf(u));
}

__device__ __forceinline__ float gelu_grad(float z) {
    const float u = SQRT_2_OVER_PI * (z + GELU_C * z * z * z);
    const float t = tanhf(u);
    return 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * SQRT_2_OVER_PI * (1.0f + 3.0f * GELU_C * z * z);
}

__global__ void k_bias_act(float* Z, const float* bias, uint16_t* H, long total, int N, int act) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const float z = Z[i] + bias[i % N];
    Z[i] = z;
    if (H) H[i] = f2bf(act ? gelu(z) : z);
}

// grid (columns / 256, rows / ACT_ROWS): each thread walks ACT_ROWS rows of one
// column; the column sums for db are accumulated with one atomic add per thread.
constexpr int ACT_ROWS = 64;

__global__ void k_act_backward(const float* dH, const float* Z, uint16_t* dZ, float* db, int M, int N, int act) {
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= N) return;
    const int r0 = blockIdx.y * ACT_ROWS, r1 = min(M, r0 + ACT_ROWS);
    float s = 0.0f;
    for (int r = r0; r < r1; ++r) {
        const long i = (long)r * N + col;
        const float d = act ? dH[i] * gelu_grad(Z[i]) : dH[i];
        dZ[i] = f2bf(d);
        s += d;
    }
    if (db) atomicAdd(db + col, s);
}

// one block per row: squared error / BCE-with-logits and their gradients
__global__ void __launch_bounds__(256) k_loss(const float* O, const float* T, float* dO, float* row_loss, int M,
                                              int nx, int ny, int na, float wx, float wy, float wa) {
    __shared__ float part[256];
    const int row = blockIdx.x;
    const int N = nx + ny + na;
    float acc = 0.0f;
    const float sx = nx > 0 ? wx / ((float)M * nx) : 0.0f;
    const float sy = ny > 0 ? wy / ((float)M * ny) : 0.0f;
    const float sa = na > 0 ? wa / ((float)M * na) : 0.0f;
    for (int j = threadIdx.x; j < N; j += blockDim.x) {
        const long i = (long)row * N + j;
        const float o = O[i], t = T[i];
        float g, l;
        if (j < nx + ny) {
            const float s = j < nx ? sx : sy;
            const float d = o - t;
            l = s * d * d;
            g = 2.0f * s * d;
        } else {
            // softplus(o) - t*o, gradient sigmoid(o) - t
            const float sp = o > 0.0f ? o + log1pf(expf(-o)) : log1pf(expf(o));
            l = sa * (sp - t * o);
            g = sa * (1.0f / (1.0f + expf(-o)) - t);
        }
        acc += l;
        dO[i] = g;
    }
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0 && row_loss) row_loss[row] = part[0];
}

__global__ void k_adam(float* w, const float* g, float* m, float* v, uint16_t* wb, long n, float lr, float b1,
                       float b2, float eps, float c1, float c2) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float upd = lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    const float wi = w[i] - upd;
    w[i] = wi;
    wb[i] = f2bf(wi);
}

__global__ void k_to_bf16(const float* x, uint16_t* y, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = f2bf(x[i]);
}

inline int blocks(long n, int t) { return (int)((n + t - 1) / t); }

}  // namespace

extern "C" {

const char* ws_version(void) { return "warmstart 0.1 gfx950 mfma_f32_32x32x8bf16_1k"; }

int ws_gemm_bf16(int M, int N, int K, const uint16_t* A, int lda, int trans_a, const uint16_t* B, int ldb,
                 int trans_b, float* C, int ldc, int accumulate, void* stream) {
    if (M < 0 || N < 0 || K < 0 || !A || !B || !C) return -1;
    if (M == 0 || N == 0) return 0;
    dim3 grid(blocks(N, BN), blocks(M, BM));
    hipStream_t s = (hipStream_t)stream;
    if (!trans_a && !trans_b) hipLaunchKernelGGL((k_gemm<0, 0>), grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, C, ldc, accumulate);
    else if (!trans_a && trans_b) hipLaunchKernelGGL((k_gemm<0, 1>), grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, C, ldc, accumulate);
    else if (trans_a && !trans_b) hipLaunchKernelGGL((k_gemm<1, 0>), grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, C, ldc, accumulate);
    else hipLaunchKernelGGL((k_gemm<1, 1>), grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, C, ldc, accumulate);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int ws_bias_act(float* Z, const float* bias, uint16_t* H, int M, int N, int act, void* stream) {
    if (!Z || !bias || M < 0 || N < 0) return -1;
    const long total = (long)M * N;
    if (total == 0) return 0;
    hipLaunchKernelGGL(k_bias_act, dim3(blocks(total, 256)), dim3(256), 0, (hipStream_t)stream, Z, bias, H, total,
                       N, act);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int ws_act_backward(const float* dH, const float* Z, uint16_t* dZ, float* db, int M, int N, int act, void* stream) {
    if (!dH || !Z || !dZ || M < 0 || N < 0) return -1;
    if (N == 0) return 0;
    if (M == 0) return 0;
    hipLaunchKernelGGL(k_act_backward, dim3(blocks(N, 256), blocks(M, ACT_ROWS)), dim3(256), 0, (hipStream_t)stream,
                       dH, Z, dZ, db, M, N, act);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int ws_loss(const float* O, const float* T, float* dO, float* row_loss, int M, int nx, int ny, int na, float wx,
            float wy, float wa, void* stream) {
    if (!O || !T || !dO || M < 0 || nx < 0 || ny < 0 || na < 0) return -1;
    if (M == 0) return 0;
    hipLaunchKernelGGL(k_loss, dim3(M), dim3(256), 0, (hipStream_t)stream, O, T, dO, row_loss, M, nx, ny, na, wx,
                       wy, wa);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int ws_adam(float* w, const float* g, float* m, float* v, uint16_t* wb, int64_t n, float lr, float b1, float b2,
            float eps, int t, void* stream) {
    if (!w || !g || !m || !v || !wb || n < 0 || t < 1) return -1;
    if (n == 0) return 0;
    const float c1 = 1.0f - powf(b1, (float)t), c2 = 1.0f - powf(b2, (float)t);
    hipLaunchKernelGGL(k_adam, dim3(blocks(n, 256)), dim3(256), 0, (hipStream_t)stream, w, g, m, v, wb, (long)n, lr,
                       b1, b2, eps, c1, c2);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int ws_to_bf16(const float* x, uint16_t* y, int64_t n, void* stream) {
    if (!x || !y || n < 0) return -1;
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_to_bf16, dim3(blocks(n, 256)), dim3(256), 0, (hipStream_t)stream, x, y, (long)n);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // extern "C"
