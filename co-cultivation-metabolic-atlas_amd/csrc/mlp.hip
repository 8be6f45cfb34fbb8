// Warm-start network kernels for gfx950 (see include/warmstart.h).
//
// GEMM: 128x128 output tile per 256-thread block, four waves in a 2x2
// arrangement, each wave 64x64 = 2x2 tiles of the 32x32x16 BF16 MFMA with
// FP32 accumulation.  K advances by 32: the block stages op(A) as [m][k] and
// op(B) as [n][k] in LDS (the transposes of the backward products are folded
// into the staging), so each MFMA operand is one 16-byte LDS read per lane.
// With 16-byte aligned operands the next K block is loaded into registers
// while the MFMAs of the current one run.
//
// MFMA 32x32x16 bf16 lane layout (64 lanes): A operand lane l holds
// A[l % 32][8*(l / 32) + 0..7]; B operand lane l holds B[8*(l / 32) + 0..7][l % 32];
// accumulator element r of lane l is D[8*(r / 4) + 4*(l / 32) + (r % 4)][l % 32].
#include "warmstart.h"

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float float16_t __attribute__((ext_vector_type(16)));

constexpr int BM = 128, BN = 128, BK = 32, PAD = 8;

__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);   // quiet NaN
    u += 0x7fffu + ((u >> 16) & 1u);                                               // round to nearest even
    return (uint16_t)(u >> 16);
}

// Operand tiles are 128 (tile rows) x 32 (k).  KMAJOR: the operand is X[row][k]
// (k contiguous); otherwise X[k][row].  Entries past K are zeroed (they meet
// the other operand's k); entries past R only reach output rows/columns that
// are never stored.

// 16-byte path (ldx % 8 == 0, X 16-byte aligned): thread tid owns chunks
// tid and tid + 256 of the tile, 8 elements each.
template <bool KMAJOR>
__device__ __forceinline__ void load_vec(uint4 (&v)[2], const uint16_t* __restrict__ X, int ldx, int r0, int R,
                                         int k0, int K, int tid) {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int c = tid + q * 256;
        v[q] = make_uint4(0, 0, 0, 0);
        if (KMAJOR) {
            const int gr = r0 + c / 4, gk = k0 + (c % 4) * 8;
            if (gr < R && gk < K) {
                v[q] = *reinterpret_cast<const uint4*>(X + (long)gr * ldx + gk);
                if (gk + 8 > K) {
                    uint16_t* e = reinterpret_cast<uint16_t*>(&v[q]);
#pragma unroll
                    for (int t = 0; t < 8; ++t)
                        if (gk + t >= K) e[t] = 0;
                }
            }
        } else {
            const int gk = k0 + c / 16, gr = r0 + (c % 16) * 8;
            if (gk < K && gr < R) v[q] = *reinterpret_cast<const uint4*>(X + (long)gk * ldx + gr);
        }
    }
}

template <bool KMAJOR>
__device__ __forceinline__ void store_vec(uint16_t (*S)[BK + PAD], const uint4 (&v)[2], int tid) {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int c = tid + q * 256;
        if (KMAJOR) {
            *reinterpret_cast<uint4*>(&S[c / 4][(c % 4) * 8]) = v[q];
        } else {
            const int kr = c / 16, rc = (c % 16) * 8;
            const uint16_t* e = reinterpret_cast<const uint16_t*>(&v[q]);
#pragma unroll
            for (int t = 0; t < 8; ++t) S[rc + t][kr] = e[t];
        }
    }
}

// Element path for unaligned operands.
template <bool KMAJOR>
__device__ __forceinline__ void stage_scalar(uint16_t (*S)[BK + PAD], const uint16_t* __restrict__ X, int ldx,
                                             int r0, int R, int k0, int K, int tid) {
    if (KMAJOR) {
        const int row = tid / 2, kk = (tid % 2) * 16;
        const int gr = r0 + row;
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const int gk = k0 + kk + e;
            S[row][kk + e] = (gr < R && gk < K) ? X[(long)gr * ldx + gk] : (uint16_t)0;
        }
    } else {
        const int kr = tid / 8, rr = (tid % 8) * 16;
        const int gk = k0 + kr;
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const int gr = r0 + rr + e;
            S[rr + e][kr] = (gr < R && gk < K) ? X[(long)gk * ldx + gr] : (uint16_t)0;
        }
    }
}

constexpr float SQRT_2_OVER_PI = 0.7978845608028654f;
constexpr float GELU_C = 0.044715f;

__device__ __forceinline__ float gelu(float z) {
    const float u = SQRT_2_OVER_PI * (z + GELU_C * z * z * z);
    return 0.5f * z * (1.0f + tanhf(u));
}

// Fused epilogues of the inference products (EPI 1 and 2 of k_gemm).
struct Epilogue {
    const float* bias;      // [N]
    // EPI 1: H[row][col] = bf16(gelu(acc + bias))
    uint16_t* H; int ldh;
    // EPI 2 (warm start): columns interleave (x_j, logit_j) for j < nx, then
    // y_i; x_j = clamp((acc + bias) * scale + shift, lower, upper), snapped to
    // the nearer finite bound when snap and logit_j > 0; y_i = (acc + bias) *
    // scale + shift, clipped to <= 0 on inequality rows (i >= neq); stored FP64
    // batch-major x0[row][j], y0[row][i] (row = LP)
    const float* scale; const float* shift;
    const double* lower; const double* upper;
    double* x0; double* y0;
    int nx, ny, neq, snap;
};

// Stores a wave's 64 x 64 accumulator tile (2 x 2 MFMA tiles of 32 x 32)
// whose first row / column are r0 / c0, through the epilogue EPI.
template <int EPI>
__device__ __forceinline__ void store_tile(const float16_t (&acc)[2][2], int r0, int c0, int lane, int M, int N,
                                           float* __restrict__ C, int ldc, int accumulate, const Epilogue& ep) {
    // per-column parameters of the lane's two columns, loaded once
    float bias[2] = {0.0f, 0.0f}, scale[2] = {0.0f, 0.0f}, shift[2] = {0.0f, 0.0f};
    if constexpr (EPI != 0) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int col = c0 + j * 32 + lane % 32;
            if (col < N) {
                bias[j] = ep.bias[col];
                if constexpr (EPI == 2) { scale[j] = ep.scale[col]; shift[j] = ep.shift[col]; }
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = r0 + i * 32 + (r / 4) * 8 + (lane / 32) * 4 + (r % 4);
                const int col = c0 + j * 32 + lane % 32;
                if constexpr (EPI == 0) {
                    if (row < M && col < N) {
                        const long o = (long)row * ldc + col;
                        C[o] = accumulate ? C[o] + acc[i][j][r] : acc[i][j][r];
                    }
                } else if constexpr (EPI == 1) {
                    if (row < M && col < N) ep.H[(long)row * ep.ldh + col] = f2bf(gelu(acc[i][j][r] + bias[j]));
                } else {
                    // lanes l and l ^ 1 hold columns col and col ^ 1 of the same row:
                    // the even lane of a (x_j, logit_j) pair reads the logit from its
                    // neighbour (every lane takes part in the exchange)
                    const float v = col < N ? acc[i][j][r] + bias[j] : 0.0f;
                    const float other = __shfl_xor(v, 1);
                    if (row < M && col < N) {
                        if (col < 2 * ep.nx) {
                            if ((col & 1) == 0) {
                                const int xj = col >> 1;
                                const long o = (long)row * ep.nx + xj;
                                const double lo = ep.lower[o], hi = ep.upper[o];
                                double xv = (double)(v * scale[j] + shift[j]);
                                xv = fmin(fmax(xv, lo), hi);
                                if (ep.snap && other > 0.0f) {
                                    const bool flo = isfinite(lo), fhi = isfinite(hi);
                                    if (flo && (!fhi || xv - lo <= hi - xv)) xv = lo;
                                    else if (fhi) xv = hi;
                                }
                                ep.x0[o] = xv;
                            }
                        } else if (col < 2 * ep.nx + ep.ny) {
                            const int yi = col - 2 * ep.nx;
                            double yv = (double)(v * scale[j] + shift[j]);
                            if (yi >= ep.neq) yv = fmin(yv, 0.0);
                            ep.y0[(long)row * ep.ny + yi] = yv;
                        }
                    }
                }
            }
}

template <int TA, int TB, int EPI = 0>
__global__ void __launch_bounds__(256) k_gemm(int M, int N, int K, const uint16_t* __restrict__ A, int lda,
                                              const uint16_t* __restrict__ B, int ldb, float* __restrict__ C,
                                              int ldc, int accumulate, int va, int vb, Epilogue ep) {
    __shared__ __align__(16) uint16_t As[BM][BK + PAD];
    __shared__ __align__(16) uint16_t Bs[BN][BK + PAD];
    constexpr bool AK = !TA, BKM = TB != 0;   // k-contiguous operands
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

    uint4 ra[2], rb[2];
    if (va) load_vec<AK>(ra, A, lda, m0, M, 0, K, tid);
    if (vb) load_vec<BKM>(rb, B, ldb, n0, N, 0, K, tid);
    for (int k0 = 0; k0 < K; k0 += BK) {
        if (va) store_vec<AK>(As, ra, tid);
        else stage_scalar<AK>(As, A, lda, m0, M, k0, K, tid);
        if (vb) store_vec<BKM>(Bs, rb, tid);
        else stage_scalar<BKM>(Bs, B, ldb, n0, N, k0, K, tid);
        __syncthreads();
        if (k0 + BK < K) {                      // next block in flight during the MFMAs
            if (va) load_vec<AK>(ra, A, lda, m0, M, k0 + BK, K, tid);
            if (vb) load_vec<BKM>(rb, B, ldb, n0, N, k0 + BK, K, tid);
        }
#pragma unroll
        for (int ks = 0; ks < BK / 16; ++ks) {
            bf16x8_t a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i)
                a[i] = *reinterpret_cast<const bf16x8_t*>(&As[wm + i * 32 + lane % 32][ks * 16 + (lane / 32) * 8]);
#pragma unroll
            for (int j = 0; j < 2; ++j)
                b[j] = *reinterpret_cast<const bf16x8_t*>(&Bs[wn + j * 32 + lane % 32][ks * 16 + (lane / 32) * 8]);
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
    store_tile<EPI>(acc, m0 + wm, n0 + wn, lane, M, N, C, ldc, accumulate, ep);
}

// ---------------------------------------------------------------------------
// GEMM with both operands k-contiguous (A [M][K], B [N][K], 16-byte aligned
// rows): the form of every product the network runs (inference, the forward
// pass and both backward products, whose operands are transposed first).
// 128 x 128 block tile, four waves of 64 x 64, K advances by 64 through two
// LDS buffers: the next K block is loaded from global memory into registers
// while the MFMAs run on the current buffer, then stored into the other one,
// so each K block costs one barrier (the 32-wide kernel above costs two per
// 32).  16 MFMAs per wave per barrier.
constexpr int KB = 32, KPAD = 8, KCH = KB / 8, KQ = BM * KCH / 256;

// thread tid owns 16-byte chunks tid + 256 q (q < 4) of a 128 x 64 tile:
// row c / 8, k offset (c % 8) * 8; entries past K are zeroed
__device__ __forceinline__ void kk_load(uint4 (&v)[KQ], const uint16_t* __restrict__ X, int ldx, int r0, int R,
                                        int k0, int K, int tid) {
#pragma unroll
    for (int q = 0; q < KQ; ++q) {
        const int c = tid + q * 256;
        const int gr = r0 + c / KCH, gk = k0 + (c % KCH) * 8;
        v[q] = make_uint4(0, 0, 0, 0);
        if (gr < R && gk < K) {
            v[q] = *reinterpret_cast<const uint4*>(X + (long)gr * ldx + gk);
            if (gk + 8 > K) {
                uint16_t* e = reinterpret_cast<uint16_t*>(&v[q]);
#pragma unroll
                for (int t = 0; t < 8; ++t)
                    if (gk + t >= K) e[t] = 0;
            }
        }
    }
}

__device__ __forceinline__ void kk_store(uint16_t (*S)[KB + KPAD], const uint4 (&v)[KQ], int tid) {
#pragma unroll
    for (int q = 0; q < KQ; ++q) {
        const int c = tid + q * 256;
        *reinterpret_cast<uint4*>(&S[c / KCH][(c % KCH) * 8]) = v[q];
    }
}

// Split K (gridDim.z > 1, EPI 0 only): block z multiplies the k range
// [z kchunk, (z + 1) kchunk) into its own slice C + z zstride.
template <int EPI = 0>
__global__ void __launch_bounds__(256) k_gemm_kk(int M, int N, int K, const uint16_t* __restrict__ A, int lda,
                                                 const uint16_t* __restrict__ B, int ldb, float* __restrict__ C,
                                                 int ldc, int accumulate, Epilogue ep, int kchunk = 0,
                                                 long zstride = 0) {
    __shared__ __align__(16) uint16_t As[2][BM][KB + KPAD];
    __shared__ __align__(16) uint16_t Bs[2][BN][KB + KPAD];
    if (gridDim.z > 1) {
        const int kb = blockIdx.z * kchunk;
        A += kb;
        B += kb;
        C += blockIdx.z * zstride;
        K = min(K - kb, kchunk);
    }
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
    const int nk = (K + KB - 1) / KB;
    uint4 ra[KQ], rb[KQ];
    kk_load(ra, A, lda, m0, M, 0, K, tid);
    kk_load(rb, B, ldb, n0, N, 0, K, tid);
    kk_store(As[0], ra, tid);
    kk_store(Bs[0], rb, tid);
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
        const int cur = kt & 1;
        const bool more = kt + 1 < nk;
        if (more) {
            kk_load(ra, A, lda, m0, M, (kt + 1) * KB, K, tid);
            kk_load(rb, B, ldb, n0, N, (kt + 1) * KB, K, tid);
        }
#pragma unroll
        for (int ks = 0; ks < KB / 16; ++ks) {
            bf16x8_t a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i)
                a[i] = *reinterpret_cast<const bf16x8_t*>(&As[cur][wm + i * 32 + lane % 32][ks * 16 + (lane / 32) * 8]);
#pragma unroll
            for (int j = 0; j < 2; ++j)
                b[j] = *reinterpret_cast<const bf16x8_t*>(&Bs[cur][wn + j * 32 + lane % 32][ks * 16 + (lane / 32) * 8]);
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        if (more) {
            // the other buffer was last read in iteration kt - 1, which every
            // wave finished before the barrier that ended it
            kk_store(As[cur ^ 1], ra, tid);
            kk_store(Bs[cur ^ 1], rb, tid);
        }
        __syncthreads();
    }
    store_tile<EPI>(acc, m0 + wm, n0 + wn, lane, M, N, C, ldc, accumulate, ep);
}

// ---------------------------------------------------------------------------
// k-contiguous GEMM with direct global-to-LDS loads (gfx950
// global_load_lds_dwordx4: each lane's 16 bytes land at the wave's LDS base +
// 16 x lane, no registers involved), four LDS stages, K in blocks of 32 (K a
// multiple of 32).  A 128 x 32 operand tile is 8 KB = eight 1 KB wave loads;
// row r of the tile holds its four 16-byte k parts at slots kp ^ ((r >> 1) & 3)
// (the load picks each lane's k part accordingly), so the MFMA operand reads
// of 32 consecutive rows hit distinct banks.  Each stage is its own LDS array,
// so the compiler's wait before reading a stage covers only that stage's loads.
// Rows past M (N) read row M - 1 (N - 1): their products reach only outputs
// that are never stored.
constexpr int DL_KB = 32, DL_NS = 4;

struct DlStage {
    uint16_t* a;    // [128][32] swizzled
    uint16_t* b;
};

__device__ __forceinline__ void dl_issue(uint16_t* lds, const uint16_t* __restrict__ X, int ldx, int r0, int R,
                                         int k0, int wave, int lane) {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int c = wave * 2 + q;                     // 1 KB chunk of the tile: rows 16c .. 16c + 15
        const int row = c * 16 + (lane >> 2);
        const int kp = (lane & 3) ^ ((row >> 1) & 3);
        const int gr = min(r0 + row, R - 1);
        const uint16_t* src = X + (long)gr * ldx + k0 + kp * 8;
        __builtin_amdgcn_global_load_lds((const void*)src,
                                         (__attribute__((address_space(3))) void*)(lds + c * 512), 16, 0, 0);
    }
}

__device__ __forceinline__ bf16x8_t dl_read(const uint16_t* lds, int row, int kp) {
    return *reinterpret_cast<const bf16x8_t*>(lds + row * 32 + ((kp ^ ((row >> 1) & 3)) << 3));
}

// Split K as in k_gemm_kk (gridDim.z > 1, EPI 0): block z multiplies the k
// range [z kchunk, (z + 1) kchunk) into C + z zstride (kchunk a multiple of 32).
template <int EPI = 0>
__global__ void __launch_bounds__(256) k_gemm_dl(int M, int N, int K, const uint16_t* __restrict__ A, int lda,
                                                 const uint16_t* __restrict__ B, int ldb, float* __restrict__ C,
                                                 int ldc, int accumulate, Epilogue ep, int kchunk = 0,
                                                 long zstride = 0) {
    __shared__ __align__(16) uint16_t A0[128 * 32], A1[128 * 32], A2[128 * 32], A3[128 * 32];
    __shared__ __align__(16) uint16_t B0[128 * 32], B1[128 * 32], B2[128 * 32], B3[128 * 32];
    if (gridDim.z > 1) {
        const int kb = blockIdx.z * kchunk;
        A += kb;
        B += kb;
        C += blockIdx.z * zstride;
        K = min(K - kb, kchunk);
    }
    uint16_t* const As[DL_NS] = {A0, A1, A2, A3};
    uint16_t* const Bs[DL_NS] = {B0, B1, B2, B3};
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
    const int nk = K / DL_KB;
#pragma unroll
    for (int s = 0; s < DL_NS - 1; ++s)
        if (s < nk) {
            dl_issue(As[s], A, lda, m0, M, s * DL_KB, wave, lane);
            dl_issue(Bs[s], B, ldb, n0, N, s * DL_KB, wave, lane);
        }
    for (int kt0 = 0; kt0 < nk; kt0 += DL_NS) {
#pragma unroll
        for (int j = 0; j < DL_NS; ++j) {
            const int kt = kt0 + j;
            if (kt < nk) {
                // stage kt's loads (this wave's) are complete; the barrier makes
                // every wave's visible and ends the reads of stage kt - 1
                const int later = min(DL_NS - 2, nk - 1 - kt);
                if (later >= 2) __builtin_amdgcn_s_waitcnt(0xF78);        // vmcnt(8)
                else if (later == 1) __builtin_amdgcn_s_waitcnt(0xF74);   // vmcnt(4)
                else __builtin_amdgcn_s_waitcnt(0xF70);                   // vmcnt(0)
                // a bare barrier: __syncthreads' fence would wait for every
                // load in flight, the later stages' included
                __builtin_amdgcn_s_barrier();
                if (kt + DL_NS - 1 < nk) {
                    const int ns = (j + DL_NS - 1) % DL_NS;
                    dl_issue(As[ns], A, lda, m0, M, (kt + DL_NS - 1) * DL_KB, wave, lane);
                    dl_issue(Bs[ns], B, ldb, n0, N, (kt + DL_NS - 1) * DL_KB, wave, lane);
                }
                const uint16_t* a_s = As[j];
                const uint16_t* b_s = Bs[j];
#pragma unroll
                for (int ks = 0; ks < DL_KB / 16; ++ks) {
                    bf16x8_t a[2], b[2];
#pragma unroll
                    for (int i = 0; i < 2; ++i) a[i] = dl_read(a_s, wm + i * 32 + lane % 32, ks * 2 + lane / 32);
#pragma unroll
                    for (int jj = 0; jj < 2; ++jj) b[jj] = dl_read(b_s, wn + jj * 32 + lane % 32, ks * 2 + lane / 32);
#pragma unroll
                    for (int i = 0; i < 2; ++i)
#pragma unroll
                        for (int jj = 0; jj < 2; ++jj)
                            acc[i][jj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[jj], acc[i][jj], 0, 0, 0);
                }
            }
        }
    }
    store_tile<EPI>(acc, m0 + wm, n0 + wn, lane, M, N, C, ldc, accumulate, ep);
}

// C[r][c] = (accumulate ? C[r][c] : 0) + sum over s = 0 .. S-1 of W[s][r][c], in that order.
__global__ void k_splitk_reduce(const float* __restrict__ W, float* __restrict__ C, int M, int N, int ldc, int S,
                                int accumulate) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long total = (long)M * N;
    if (i >= total) return;
    const long r = i / N, c = i - r * N;
    float v = accumulate ? C[r * ldc + c] : 0.0f;
    for (int s = 0; s < S; ++s) v += W[s * total + i];
    C[r * ldc + c] = v;
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

// X[r][c] = bf16((F[r][c] - mean[c]) / std[c]) for c < n_in, 0 for n_in <= c < ldx.
__global__ void k_standardize(const float* F, const float* mean, const float* sd, uint16_t* X, long M, int n_in,
                              int ldx) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M * ldx) return;
    const long r = i / ldx;
    const int c = (int)(i - r * ldx);
    X[i] = c < n_in ? f2bf((F[r * n_in + c] - mean[c]) / sd[c]) : (uint16_t)0;
}

// Y[C][R] = X[R][C]^T through a 64 x 65 LDS tile (coalesced reads and writes).
__global__ void __launch_bounds__(256) k_transpose(const uint16_t* __restrict__ X, uint16_t* __restrict__ Y, int R,
                                                   int C) {
    __shared__ uint16_t tile[64][65];
    const int c0 = blockIdx.x * 64, r0 = blockIdx.y * 64;
    const int tx = threadIdx.x % 64, ty = threadIdx.x / 64;
    for (int i = ty; i < 64; i += 4) {
        const int r = r0 + i, c = c0 + tx;
        if (r < R && c < C) tile[i][tx] = X[(long)r * C + c];
    }
    __syncthreads();
    for (int i = ty; i < 64; i += 4) {
        const int c = c0 + i, r = r0 + tx;
        if (c < C && r < R) Y[(long)c * R + r] = tile[tx][i];
    }
}

inline int blocks(long n, int t) { return (int)((n + t - 1) / t); }

}  // namespace

extern "C" {

const char* ws_version(void) { return "warmstart 0.2 gfx950 mfma_f32_32x32x16_bf16"; }

int ws_gemm_bf16(int M, int N, int K, const uint16_t* A, int lda, int trans_a, const uint16_t* B, int ldb,
                 int trans_b, float* C, int ldc, int accumulate, void* stream) {
    if (M < 0 || N < 0 || K < 0 || !A || !B || !C) return -1;
    if (M == 0 || N == 0) return 0;
    dim3 grid(blocks(N, BN), blocks(M, BM));
    hipStream_t s = (hipStream_t)stream;
    const int va = (lda % 8 == 0 && ((uintptr_t)A & 15) == 0) ? 1 : 0;
    const int vb = (ldb % 8 == 0 && ((uintptr_t)B & 15) == 0) ? 1 : 0;
    if (!trans_a && !trans_b)
        hipLaunchKernelGGL((k_gemm<0, 0>), grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, C, ldc, accumulate, va, vb, Epilogue{});
    else if (!trans_a && trans_b)
        if (va && vb && K % DL_KB == 0)
            hipLaunchKernelGGL((k_gemm_dl<0>), grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, C, ldc, accumulate, Epilogue{}, 0, 0L);
        else if (va && vb)
            hipLaunchKernelGGL((k_gemm_kk<0>), grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, C, ldc, accumulate, Epilogue{}, 0, 0L);
        else
            hipLaunchKernelGGL((k_gemm<0, 1>), grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, C, ldc, accumulate, va, vb, Epilogue{});
    else if (trans_a && !trans_b)
        hipLaunchKernelGGL((k_gemm<1, 0>), grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, C, ldc, accumulate, va, vb, Epilogue{});
    else
        hipLaunchKernelGGL((k_gemm<1, 1>), grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, C, ldc, accumulate, va, vb, Epilogue{});
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int ws_gemm_bf16_splitk(int M, int N, int K, const uint16_t* A, int lda, const uint16_t* B, int ldb, float* C,
                        int ldc, int accumulate, float* work, int splits, void* stream) {
    if (M < 0 || N < 0 || K < 0 || !A || !B || !C || splits < 1 || (splits > 1 && !work) || ldc < N) return -1;
    if (lda % 8 || ldb % 8 || ((uintptr_t)A & 15) || ((uintptr_t)B & 15)) return -1;
    if (M == 0 || N == 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    // k chunks are whole K blocks, so every slice's operand rows stay 16-byte aligned
    const int kchunk = ((K + splits - 1) / splits + KB - 1) / KB * KB;
    const int S = K == 0 ? 1 : (K + kchunk - 1) / kchunk;
    const bool dl = K % DL_KB == 0;      // direct-to-LDS kernel (kchunk is a multiple of 32 too)
    if (S <= 1) {
        if (dl)
            hipLaunchKernelGGL((k_gemm_dl<0>), dim3(blocks(N, BN), blocks(M, BM)), dim3(256), 0, s, M, N, K, A, lda,
                               B, ldb, C, ldc, accumulate, Epilogue{}, 0, 0L);
        else
            hipLaunchKernelGGL((k_gemm_kk<0>), dim3(blocks(N, BN), blocks(M, BM)), dim3(256), 0, s, M, N, K, A, lda,
                               B, ldb, C, ldc, accumulate, Epilogue{}, 0, 0L);
        return hipGetLastError() == hipSuccess ? 0 : -2;
    }
    const long zstride = (long)M * N;
    if (dl)
        hipLaunchKernelGGL((k_gemm_dl<0>), dim3(blocks(N, BN), blocks(M, BM), S), dim3(256), 0, s, M, N, K, A, lda,
                           B, ldb, work, N, 0, Epilogue{}, kchunk, zstride);
    else
        hipLaunchKernelGGL((k_gemm_kk<0>), dim3(blocks(N, BN), blocks(M, BM), S), dim3(256), 0, s, M, N, K, A, lda,
                           B, ldb, work, N, 0, Epilogue{}, kchunk, zstride);
    hipLaunchKernelGGL(k_splitk_reduce, dim3(blocks(zstride, 256)), dim3(256), 0, s, (const float*)work, C, M, N,
                       ldc, S, accumulate);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int ws_gemm_bf16_act(int M, int N, int K, const uint16_t* A, int lda, const uint16_t* B, int ldb, const float* bias,
                     uint16_t* H, int ldh, void* stream) {
    if (M < 0 || N < 0 || K < 0 || !A || !B || !bias || !H || ldh < N) return -1;
    if (M == 0 || N == 0) return 0;
    dim3 grid(blocks(N, BN), blocks(M, BM));
    const int va = (lda % 8 == 0 && ((uintptr_t)A & 15) == 0) ? 1 : 0;
    const int vb = (ldb % 8 == 0 && ((uintptr_t)B & 15) == 0) ? 1 : 0;
    Epilogue ep{};
    ep.bias = bias; ep.H = H; ep.ldh = ldh;
    if (va && vb && K % DL_KB == 0)
        hipLaunchKernelGGL((k_gemm_dl<1>), grid, dim3(256), 0, (hipStream_t)stream, M, N, K, A, lda, B, ldb,
                           (float*)nullptr, 0, 0, ep, 0, 0L);
    else if (va && vb)
        hipLaunchKernelGGL((k_gemm_kk<1>), grid, dim3(256), 0, (hipStream_t)stream, M, N, K, A, lda, B, ldb,
                           (float*)nullptr, 0, 0, ep, 0, 0L);
    else
        hipLaunchKernelGGL((k_gemm<0, 1, 1>), grid, dim3(256), 0, (hipStream_t)stream, M, N, K, A, lda, B, ldb,
                           (float*)nullptr, 0, 0, va, vb, ep);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int ws_gemm_bf16_start(int M, int nx, int ny, int K, const uint16_t* A, int lda, const uint16_t* B, int ldb,
                       const float* bias, const float* scale, const float* shift, const double* lower,
                       const double* upper, int neq, int snap, double* x0, double* y0, void* stream) {
    if (M < 0 || nx < 0 || ny < 0 || K < 0 || neq < 0 || neq > ny || !A || !B || !bias || !scale || !shift ||
        (nx > 0 && (!lower || !upper || !x0)) || (ny > 0 && !y0))
        return -1;
    const int N = 2 * nx + ny;
    if (M == 0 || N == 0) return 0;
    dim3 grid(blocks(N, BN), blocks(M, BM));
    const int va = (lda % 8 == 0 && ((uintptr_t)A & 15) == 0) ? 1 : 0;
    const int vb = (ldb % 8 == 0 && ((uintptr_t)B & 15) == 0) ? 1 : 0;
    Epilogue ep{};
    ep.bias = bias; ep.scale = scale; ep.shift = shift; ep.lower = lower; ep.upper = upper;
    ep.x0 = x0; ep.y0 = y0; ep.nx = nx; ep.ny = ny; ep.neq = neq; ep.snap = snap;
    if (va && vb && K % DL_KB == 0)
        hipLaunchKernelGGL((k_gemm_dl<2>), grid, dim3(256), 0, (hipStream_t)stream, M, N, K, A, lda, B, ldb,
                           (float*)nullptr, 0, 0, ep, 0, 0L);
    else if (va && vb)
        hipLaunchKernelGGL((k_gemm_kk<2>), grid, dim3(256), 0, (hipStream_t)stream, M, N, K, A, lda, B, ldb,
                           (float*)nullptr, 0, 0, ep, 0, 0L);
    else
        hipLaunchKernelGGL((k_gemm<0, 1, 2>), grid, dim3(256), 0, (hipStream_t)stream, M, N, K, A, lda, B, ldb,
                           (float*)nullptr, 0, 0, va, vb, ep);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int ws_standardize(const float* F, const float* mean, const float* sd, uint16_t* X, int M, int n_in, int ldx,
                   void* stream) {
    if (!F || !mean || !sd || !X || M < 0 || n_in < 0 || ldx < n_in) return -1;
    const long total = (long)M * ldx;
    if (total == 0) return 0;
    hipLaunchKernelGGL(k_standardize, dim3(blocks(total, 256)), dim3(256), 0, (hipStream_t)stream, F, mean, sd, X,
                       (long)M, n_in, ldx);
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

int ws_transpose_bf16(const uint16_t* X, uint16_t* Y, int R, int C, void* stream) {
    if (!X || !Y || R < 0 || C < 0) return -1;
    if (R == 0 || C == 0) return 0;
    hipLaunchKernelGGL(k_transpose, dim3(blocks(C, 64), blocks(R, 64)), dim3(256), 0, (hipStream_t)stream, X, Y, R, C);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int ws_to_bf16(const float* x, uint16_t* y, int64_t n, void* stream) {
    if (!x || !y || n < 0) return -1;
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_to_bf16, dim3(blocks(n, 256)), dim3(256), 0, (hipStream_t)stream, x, y, (long)n);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // extern "C"
