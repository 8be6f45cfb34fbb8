// Device code of the MI355X batched LP engine.
//
// Data layout: every per-LP vector is stored batch-minor, element (j, b) of a
// length-L vector at [j * ld + b] with ld = padded batch.  A 64-lane wavefront
// therefore owns one matrix row (or column) for 64 * LPL consecutive LPs: the
// row's structure and shared coefficients are wave-uniform (scalar loads),
// and every vector access - including the gathers x[col][b] - is a contiguous
// 256-2048 byte segment.  The shared matrix is read once per wave instead of
// once per LP.
//
// Work decomposition: a block of 4 waves covers `rpw` rows per wave for one
// batch tile of W = 64 * LPL LPs.  Blocks are remapped so that each XCD walks
// a contiguous range of (tile, row-block) pairs; the gathered vectors of a
// tile then stay in that XCD's L2.
//
// Floating point: compiled with -ffp-contract=off so that every product and
// sum is rounded as written (the order of the reference's separate FP64
// operations, src/gpu_pdhg_workspace.py:122-134 "--fmad=false").
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <float.h>

namespace lpe {

constexpr int WAVE = 64;
constexpr int WAVES_PER_BLOCK = 4;
constexpr int BLOCK = WAVE * WAVES_PER_BLOCK;

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
template <typename T, int N>
__device__ __forceinline__ void load_vec(const T* __restrict__ p, T (&v)[N]) {
    if constexpr (N == 1) {
        v[0] = p[0];
    } else {
        struct alignas(sizeof(T) * N) Pack { T e[N]; };
        Pack q = *reinterpret_cast<const Pack*>(p);
#pragma unroll
        for (int t = 0; t < N; ++t) v[t] = q.e[t];
    }
}

template <typename T, int N>
__device__ __forceinline__ void store_vec(T* __restrict__ p, const T (&v)[N]) {
    if constexpr (N == 1) {
        p[0] = v[0];
    } else {
        struct alignas(sizeof(T) * N) Pack { T e[N]; };
        Pack q;
#pragma unroll
        for (int t = 0; t < N; ++t) q.e[t] = v[t];
        *reinterpret_cast<Pack*>(p) = q;
    }
}

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Load of the shared sparse structure (row pointers, indices, slots, shared
// values) through the constant address space.  With a wave-uniform index the
// compiler emits scalar loads - batched per chunk and served by the scalar
// cache - instead of a vector load, a wait and a readfirstlane per entry,
// which serialised the structure reads ahead of every gather.  The structure
// is never written while a kernel reads it.
template <typename T>
__device__ __forceinline__ T sload(const T* p, long i) {
    return ((const __attribute__((address_space(4))) T*)p)[i];
}

// XCD-aware block remap.  The grid is padded to a multiple of 8; block g is
// dispatched to XCD g % 8, which receives the contiguous virtual range
// [x * G/8, (x+1) * G/8).  Returns false for padding blocks.
__device__ __forceinline__ bool map_block(int n_real, int n_rb, int& tile, int& rb) {
    const int g = blockIdx.x;
    const int G = gridDim.x;
    const int v = (G % 8 == 0) ? (g % 8) * (G / 8) + g / 8 : g;
    if (v >= n_real) return false;
    tile = v / n_rb;
    rb = v - tile * n_rb;
    return true;
}

struct Shape {
    int rows;      // rows of the operator this launch walks (m or n)
    int ld;        // leading dimension (padded batch)
    int batch;     // real batch
    int rpw;       // rows per wave
    int n_rb;      // row blocks
    int n_tiles;
    int n_real;    // n_tiles * n_rb
};

// Gather-accumulate one sparse row for LPL consecutive LPs.
// acc[t] = sum_k a_k * x[idx_k][b0 + t], sequential in storage order.
template <typename TV, typename TX, typename TA, int LPL, int D = 8>
__device__ __forceinline__ void row_dot(const int* __restrict__ ptr, const int* __restrict__ idx,
                                        const int* __restrict__ slot, const TV* __restrict__ val,
                                        const TV* __restrict__ var, const TX* __restrict__ x,
                                        long ld, int b0, int row, TA (&acc)[LPL]) {
#pragma unroll
    for (int t = 0; t < LPL; ++t) acc[t] = TA(0);
    const int s = sload(ptr, row), e = sload(ptr, row + 1);
    // Chunks of D entries: all (index, slot, value) scalars of the chunk are
    // loaded first, then all gathers are issued, then the products are
    // accumulated in storage order - the gathers of one chunk are in flight
    // together instead of one dependent load at a time.  The depth trades
    // loads in flight per wave against registers (waves per SIMD); it does
    // not change the result.
    for (int k0 = s; k0 < e; k0 += D) {
        const int cnt = e - k0 < D ? e - k0 : D;
        int col[D], sl[D];
        TV a[D];
#pragma unroll
        for (int u = 0; u < D; ++u) {
            if (u < cnt) { col[u] = sload(idx, k0 + u); sl[u] = sload(slot, k0 + u); a[u] = sload(val, k0 + u); }
            else { col[u] = 0; sl[u] = -1; a[u] = TV(0); }
        }
        TX xv[D][LPL];
#pragma unroll
        for (int u = 0; u < D; ++u)
            if (u < cnt) load_vec<TX, LPL>(x + (long)col[u] * ld + b0, xv[u]);
        TV vv[D][LPL];
#pragma unroll
        for (int u = 0; u < D; ++u)
            if (u < cnt && sl[u] >= 0) load_vec<TV, LPL>(var + (long)sl[u] * ld + b0, vv[u]);
#pragma unroll
        for (int u = 0; u < D; ++u) {
            if (u < cnt) {
                if (sl[u] < 0) {
                    const TA av = (TA)a[u];
#pragma unroll
                    for (int t = 0; t < LPL; ++t) acc[t] = acc[t] + av * (TA)xv[u][t];
                } else {
#pragma unroll
                    for (int t = 0; t < LPL; ++t) acc[t] = acc[t] + (TA)vv[u][t] * (TA)xv[u][t];
                }
            }
        }
    }
}

// Gather-accumulate one sparse row for the two polish CG systems of LPL
// consecutive LPs whose vectors are interleaved, x[k][b] = (x1, x2): one
// 16-byte load per LP and entry serves both systems.  acc1 and acc2 are
// summed in storage order exactly as row_dot sums each system alone.
template <int LPL, int D>
__device__ __forceinline__ void row_dot2(const int* __restrict__ ptr, const int* __restrict__ idx,
                                         const int* __restrict__ slot, const double* __restrict__ val,
                                         const double* __restrict__ var, const double2* __restrict__ x, long ld,
                                         int b0, int row, double (&acc1)[LPL], double (&acc2)[LPL]) {
#pragma unroll
    for (int t = 0; t < LPL; ++t) { acc1[t] = 0.0; acc2[t] = 0.0; }
    const int s = sload(ptr, row), e = sload(ptr, row + 1);
    for (int k0 = s; k0 < e; k0 += D) {
        const int cnt = e - k0 < D ? e - k0 : D;
        int col[D], sl[D];
        double a[D];
#pragma unroll
        for (int u = 0; u < D; ++u) {
            if (u < cnt) { col[u] = sload(idx, k0 + u); sl[u] = sload(slot, k0 + u); a[u] = sload(val, k0 + u); }
            else { col[u] = 0; sl[u] = -1; a[u] = 0.0; }
        }
        double2 xv[D][LPL];
#pragma unroll
        for (int u = 0; u < D; ++u)
            if (u < cnt) {
#pragma unroll
                for (int t = 0; t < LPL; ++t) xv[u][t] = x[(long)col[u] * ld + b0 + t];
            }
        double vv[D][LPL];
#pragma unroll
        for (int u = 0; u < D; ++u)
            if (u < cnt && sl[u] >= 0) load_vec<double, LPL>(var + (long)sl[u] * ld + b0, vv[u]);
#pragma unroll
        for (int u = 0; u < D; ++u) {
            if (u < cnt) {
#pragma unroll
                for (int t = 0; t < LPL; ++t) {
                    const double av = sl[u] < 0 ? a[u] : vv[u][t];
                    acc1[t] = acc1[t] + av * xv[u][t].x;
                    acc2[t] = acc2[t] + av * xv[u][t].y;
                }
            }
        }
    }
}

// Sparse operator in one orientation (CSR of A, or CSR of A^T).
template <typename TV>
struct Op {
    const int* ptr;
    const int* idx;
    const int* slot;
    const TV* val;
    const TV* var;   // [n_var][ld]
};

// ---------------------------------------------------------------------------
// PDHG half steps
// ---------------------------------------------------------------------------
template <typename T>
struct DualArgs {
    Shape sh;
    Op<T> A;
    const T* xbar;       // [n][ld]
    T* y;                // [m][ld]
    const T* rhs;        // [m][ld]
    const T* sigma;      // [m] shared, or [m][ld] when REF
    const T* omega;      // [ld] primal weight (RESTARTED); unused when REF
    const int* flags;    // [ld] nonzero = frozen
    T* ysum;             // [m][ld] running sum (RESTARTED); T(z)_y when HAL and store_t
    int neq;
    // Halpern (HAL): y <- lam ((1 + refl) y+ - refl y) + (1 - lam) y0,
    // lam = (k + 1) / (k + 2), k = hk[b] + j
    const T* anchor;     // [m][ld] y0 (the last restart point)
    const int* hk;       // [ld] Halpern iterations since the restart, before this launch group
    int j;               // iteration index within the launch group
    int store_t;         // 1: also store y+ = T(z)_y into ysum; 2: only that (probe)
    int reverse = 0;     // walk the LP tiles last to first
    T refl;
};

// y+ = y + sigma*(rhs - A xbar), inequality rows clipped to <= 0, NaN passes,
// frozen LPs keep y (src/gpu_pdhg_workspace.py:122-127; corrector :340-344).
// HAL: the reflected Halpern step of restarted Halpern PDHG (Lu & Yang,
// "Restarted Halpern PDHG for linear programming", 2024) on the dual block.
template <typename T, int LPL, bool REF, bool HAL = false>
__global__ void __launch_bounds__(BLOCK) k_dual_step(DualArgs<T> a) {
    int tile, rb;
    if (!map_block(a.sh.n_real, a.sh.n_rb, tile, rb)) return;
    if (a.reverse) tile = a.sh.n_tiles - 1 - tile;
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int b0 = tile * WAVE * LPL + lane * LPL;
    if (b0 >= a.sh.ld) return;
    const long ld = a.sh.ld;
    int fr[LPL];
    T om[LPL], lam[LPL];
#pragma unroll
    for (int t = 0; t < LPL; ++t) {
        fr[t] = a.flags[b0 + t];
        om[t] = REF ? T(1) : a.omega[b0 + t];
        if (HAL) {
            const int k = a.hk[b0 + t] + a.j;
            lam[t] = T(k + 1) / T(k + 2);
        }
    }
    bool all_frozen = true;
#pragma unroll
    for (int t = 0; t < LPL; ++t) all_frozen = all_frozen && fr[t];
    if (__all(all_frozen)) return;
    const int r0 = (rb * WAVES_PER_BLOCK + wave) * a.sh.rpw;
    for (int r = 0; r < a.sh.rpw; ++r) {
        const int row = uniform(r0 + r);
        if (row >= a.sh.rows) break;
        T ax[LPL];
        row_dot<T, T, T, LPL>(a.A.ptr, a.A.idx, a.A.slot, a.A.val, a.A.var, a.xbar, ld, b0, row, ax);
        T yv[LPL], bv[LPL], sg[LPL];
        load_vec<T, LPL>(a.y + row * ld + b0, yv);
        load_vec<T, LPL>(a.rhs + row * ld + b0, bv);
        if (REF) {
            load_vec<T, LPL>(a.sigma + row * ld + b0, sg);
        } else {
            const T s = a.sigma[row];
#pragma unroll
            for (int t = 0; t < LPL; ++t) sg[t] = s * om[t];
        }
        const bool ineq = row >= a.neq;
        T out[LPL];
#pragma unroll
        for (int t = 0; t < LPL; ++t) {
            const T delta = sg[t] * (bv[t] - ax[t]);
            const T cand = yv[t] + delta;
            T nv = cand;
            if (ineq && !__builtin_isnan(cand)) nv = __builtin_fmin(cand, T(0));
            out[t] = fr[t] ? yv[t] : nv;
        }
        if (HAL && a.store_t == 2) {        // probe: T(z)_y only
            store_vec<T, LPL>(a.ysum + row * ld + b0, out);
            continue;
        }
        if (HAL) {
            T y0[LPL], hv[LPL];
            load_vec<T, LPL>(a.anchor + row * ld + b0, y0);
#pragma unroll
            for (int t = 0; t < LPL; ++t) {
                const T refl_pt = (T(1) + a.refl) * out[t] - a.refl * yv[t];
                hv[t] = fr[t] ? yv[t] : lam[t] * refl_pt + (T(1) - lam[t]) * y0[t];
            }
            store_vec<T, LPL>(a.y + row * ld + b0, hv);
            if (a.store_t) store_vec<T, LPL>(a.ysum + row * ld + b0, out);
            continue;
        }
        store_vec<T, LPL>(a.y + row * ld + b0, out);
        if (!REF) {
            T s[LPL];
            load_vec<T, LPL>(a.ysum + row * ld + b0, s);
#pragma unroll
            for (int t = 0; t < LPL; ++t) s[t] = fr[t] ? s[t] : s[t] + out[t];
            store_vec<T, LPL>(a.ysum + row * ld + b0, s);
        }
    }
}

template <typename T>
struct PrimalArgs {
    Shape sh;
    Op<T> AT;
    const T* y;          // [m][ld]
    T* x;                // [n][ld]
    T* xbar;             // [n][ld]
    const T* c;          // [n] (c_shared) or [n][ld]
    const T* lower;      // [n][ld]
    const T* upper;      // [n][ld]
    const T* tau;        // [n] shared, or [n][ld] when REF
    const T* inv_omega;  // [ld]
    const int* flags;
    T* xsum;             // [n][ld]; T(z)_x when HAL and store_t
    T theta;
    int c_shared;
    // Halpern (HAL): x <- lam ((1 + refl) x+ - refl x) + (1 - lam) x0; xbar
    // still extrapolates x+ for the dual step of the same iteration
    const T* anchor;     // [n][ld] x0
    const int* hk;       // [ld]
    int j;
    int store_t;         // 1: also store x+ into xsum; 2: only x+ and xbar (probe)
    T refl;
};

// x+ = clip(x - tau*(c - A^T y+), l, u); xbar = x+ + theta*(x+ - x); frozen
// LPs keep x (src/gpu_pdhg_workspace.py:128-134; corrector :345-357).
template <typename T, int LPL, bool REF, bool HAL = false>
__global__ void __launch_bounds__(BLOCK) k_primal_step(PrimalArgs<T> a) {
    int tile, rb;
    if (!map_block(a.sh.n_real, a.sh.n_rb, tile, rb)) return;
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int b0 = tile * WAVE * LPL + lane * LPL;
    if (b0 >= a.sh.ld) return;
    const long ld = a.sh.ld;
    int fr[LPL];
    T iw[LPL], lam[LPL];
#pragma unroll
    for (int t = 0; t < LPL; ++t) {
        fr[t] = a.flags[b0 + t];
        iw[t] = REF ? T(1) : a.inv_omega[b0 + t];
        if (HAL) {
            const int k = a.hk[b0 + t] + a.j;
            lam[t] = T(k + 1) / T(k + 2);
        }
    }
    bool all_frozen = true;
#pragma unroll
    for (int t = 0; t < LPL; ++t) all_frozen = all_frozen && fr[t];
    if (__all(all_frozen)) return;
    const int r0 = (rb * WAVES_PER_BLOCK + wave) * a.sh.rpw;
    for (int r = 0; r < a.sh.rpw; ++r) {
        const int col = uniform(r0 + r);
        if (col >= a.sh.rows) break;
        T aty[LPL];
        row_dot<T, T, T, LPL>(a.AT.ptr, a.AT.idx, a.AT.slot, a.AT.val, a.AT.var, a.y, ld, b0, col, aty);
        T xv[LPL], lo[LPL], hi[LPL], cv[LPL], tv[LPL];
        load_vec<T, LPL>(a.x + col * ld + b0, xv);
        load_vec<T, LPL>(a.lower + col * ld + b0, lo);
        load_vec<T, LPL>(a.upper + col * ld + b0, hi);
        if (a.c_shared) {
            const T cc = a.c[col];
#pragma unroll
            for (int t = 0; t < LPL; ++t) cv[t] = cc;
        } else {
            load_vec<T, LPL>(a.c + col * ld + b0, cv);
        }
        if (REF) {
            load_vec<T, LPL>(a.tau + col * ld + b0, tv);
        } else {
            const T tt = a.tau[col];
#pragma unroll
            for (int t = 0; t < LPL; ++t) tv[t] = tt * iw[t];
        }
        T nx[LPL], nb[LPL];
#pragma unroll
        for (int t = 0; t < LPL; ++t) {
            const T step = tv[t] * (cv[t] - aty[t]);
            const T cand = xv[t] - step;
            T nxt;
            if (fr[t]) {
                nxt = xv[t];
            } else if (__builtin_isnan(cand)) {
                nxt = cand;
            } else {
                nxt = __builtin_fmin(__builtin_fmax(cand, lo[t]), hi[t]);
            }
            const T ext = a.theta * (nxt - xv[t]);
            nx[t] = nxt;
            nb[t] = nxt + ext;
        }
        if (HAL && a.store_t == 2) {        // probe: T(z)_x and its extrapolation only
            store_vec<T, LPL>(a.xsum + col * ld + b0, nx);
            store_vec<T, LPL>(a.xbar + col * ld + b0, nb);
            continue;
        }
        if (HAL) {
            T x0[LPL], hv[LPL];
            load_vec<T, LPL>(a.anchor + col * ld + b0, x0);
#pragma unroll
            for (int t = 0; t < LPL; ++t) {
                const T refl_pt = (T(1) + a.refl) * nx[t] - a.refl * xv[t];
                hv[t] = fr[t] ? xv[t] : lam[t] * refl_pt + (T(1) - lam[t]) * x0[t];
            }
            store_vec<T, LPL>(a.x + col * ld + b0, hv);
            store_vec<T, LPL>(a.xbar + col * ld + b0, nb);
            if (a.store_t) store_vec<T, LPL>(a.xsum + col * ld + b0, nx);
            continue;
        }
        store_vec<T, LPL>(a.x + col * ld + b0, nx);
        store_vec<T, LPL>(a.xbar + col * ld + b0, nb);
        if (!REF) {
            T s[LPL];
            load_vec<T, LPL>(a.xsum + col * ld + b0, s);
#pragma unroll
            for (int t = 0; t < LPL; ++t) s[t] = fr[t] ? s[t] : s[t] + nx[t];
            store_vec<T, LPL>(a.xsum + col * ld + b0, s);
        }
    }
}

// ---------------------------------------------------------------------------
// Block-level reduction of per-LP partial values.  Every wave holds NV values
// for its lane's LPL LPs; the block combines the 4 waves in fixed order and
// writes partial[(rb * NV + v) * ld + b].
// ---------------------------------------------------------------------------
enum RedOp { RSUM = 0, RMAX = 1 };

template <int NV, int LPL, int WPB = WAVES_PER_BLOCK>
__device__ __forceinline__ void block_partials(double (&val)[NV][LPL], const int (&ops)[NV],
                                               double* __restrict__ partial, int rb, long ld, int b0,
                                               double* lds, bool valid, int only = -1) {
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    // lds: [WAVES][NV][64*LPL]
    constexpr int W = WAVE * LPL;
#pragma unroll
    for (int v = 0; v < NV; ++v)
#pragma unroll
        for (int t = 0; t < LPL; ++t) lds[(wave * NV + v) * W + lane * LPL + t] = val[v][t];
    __syncthreads();
    if (wave == 0 && valid) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
#pragma unroll
            for (int t = 0; t < LPL; ++t) {
                double acc = lds[(0 * NV + v) * W + lane * LPL + t];
                for (int w = 1; w < WPB; ++w) {
                    const double o = lds[(w * NV + v) * W + lane * LPL + t];
                    acc = ops[v] == RSUM ? acc + o : (o > acc || __builtin_isnan(o) ? o : acc);
                }
                if (only < 0 || v == only) partial[((long)rb * NV + v) * ld + b0 + t] = acc;
            }
        }
    }
}

// out[v * ld + b] = reduce over rb of partial[(rb * NV + v) * ld + b].  A
// 1024-thread block covers 64 LPs of one value v.  Wave w walks the row blocks
// rb = w, w + 16, ... with four independent accumulators (four loads in flight
// per lane), combines them as ((a0, a1), (a2, a3)), and wave 0 combines the 16
// wave results in wave order: the summation order is fixed (deterministic).
constexpr int RED_WAVES = 16;

__device__ __forceinline__ double red_combine(int op, double acc, double o) {
    return op == RSUM ? acc + o : (o > acc || __builtin_isnan(o) ? o : acc);
}

__global__ void __launch_bounds__(RED_WAVES * 64) k_reduce_partials(const double* __restrict__ partial,
                                                                     double* __restrict__ out,
                                                                     const int* __restrict__ ops, int nv,
                                                                     int n_rb, int ld) {
    __shared__ double part[RED_WAVES][64];
    const int lane = threadIdx.x % 64, w = threadIdx.x / 64;
    const int b = blockIdx.x * 64 + lane;
    const int v = blockIdx.y;
    const int op = ops[v];
    const double init = op == RSUM ? 0.0 : -INFINITY;
    double acc = init;
    if (b < ld) {
        const long stride = (long)nv * ld;
        const double* p = partial + (long)v * ld + b;
        double a0 = init, a1 = init, a2 = init, a3 = init;
        int r = w;
        for (; r + 3 * RED_WAVES < n_rb; r += 4 * RED_WAVES) {
            const double o0 = p[r * stride];
            const double o1 = p[(r + RED_WAVES) * stride];
            const double o2 = p[(r + 2 * RED_WAVES) * stride];
            const double o3 = p[(r + 3 * RED_WAVES) * stride];
            a0 = red_combine(op, a0, o0);
            a1 = red_combine(op, a1, o1);
            a2 = red_combine(op, a2, o2);
            a3 = red_combine(op, a3, o3);
        }
        for (; r < n_rb; r += RED_WAVES) a0 = red_combine(op, a0, p[r * stride]);
        acc = red_combine(op, red_combine(op, a0, a1), red_combine(op, a2, a3));
    }
    part[w][lane] = acc;
    __syncthreads();
    if (w == 0 && b < ld) {
        double t = part[0][lane];
        for (int q = 1; q < RED_WAVES; ++q) t = red_combine(op, t, part[q][lane]);
        out[(long)v * ld + b] = t;
    }
}


// ---------------------------------------------------------------------------
// Device checks of restarted PDHG (FP64 arithmetic on the scaled problem).
// Current iterate (x, y) and average (xsum/cnt, ysum/cnt) are evaluated in
// one pass; each pass writes NV block partials per LP.
// ---------------------------------------------------------------------------
// Integer weight in [1, 2^20] of index i for the active-set signatures: sums of
// at most 2 * 2^20 * rows are exact in FP64, so the signature is independent of
// the summation order.
__device__ __forceinline__ double asig_weight(unsigned i) {
    return (double)(((i * 2654435761u) >> 12) & 0xFFFFFu) + 1.0;
}

enum CheckRowVal {
    CR_RP2 = 0,    // sum of squared scaled primal residual
    CR_PERR,       // max original-unit row error (eq |Ax-b|, ineq (Ax-b)+)
    CR_BY,         // sum b^ . y^
    CR_ROWGAP,     // sum |y (b - Ax)| over inequality rows
    CR_YPOS,       // max original y over inequality rows
    CR_DY2,        // sum (y - y_last)^2
    CR_NONFIN,     // count of non-finite activity / y
    CR_ASIG,       // active-set signature: hashed sum over inequality rows with y < 0
    CR_FPY,        // Halpern: sum (y - T(z)_y)^2 (fixed-point residual, current iterate)
    CR_N
};
enum CheckColVal {
    CC_RD2 = 0,    // sum of squared scaled dual residual
    CC_DVIOL,      // max original dual violation on infinite bounds
    CC_BOUNDTERM,  // sum of l^ r^+ - u^ r^- (finite bounds)
    CC_POBJ,       // sum c^ x^
    CC_BOUNDGAP,   // sum |r^ (x^ - target)|
    CC_BVIOL,      // max original bound violation
    CC_DX2,        // sum (x - x_last)^2
    CC_NONFIN,
    CC_ASIG,       // active-set signature: hashed sum over columns at a bound (1 lower, 2 upper)
    CC_FPX,        // Halpern: sum (x - T(z)_x)^2
    CC_N
};

template <typename T>
struct CheckRowsArgs {
    Shape sh;
    Op<double> A;            // scaled FP64 values
    const T* x;              // [n][ld]
    const T* xsum;
    const T* y;              // [m][ld]
    const T* ysum;
    const T* ylast;
    const double* rhs;       // scaled FP64 [m][ld]
    const double* dr;        // [m]
    const int* flags;
    const int* cnt;          // iterations summed in xsum/ysum before this chunk
    int K;                   // iterations in this chunk
    int neq;
    int with_avg;
    int hal;                 // 1: the second candidate (xsum, ysum) is T(z), not a sum
    double* partial;         // [n_rb][2*CR_N][ld]
};

template <typename T>
__global__ void __launch_bounds__(BLOCK) k_check_rows(CheckRowsArgs<T> a) {
    __shared__ double lds[WAVES_PER_BLOCK * 2 * CR_N * WAVE];
    int tile, rb;
    if (!map_block(a.sh.n_real, a.sh.n_rb, tile, rb)) return;
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int b0 = tile * WAVE + lane;
    const bool valid = b0 < a.sh.ld;
    const long ld = a.sh.ld;
    double v[2 * CR_N][1];
    for (int i = 0; i < 2 * CR_N; ++i) v[i][0] = 0.0;
    const bool active = valid && !a.flags[valid ? b0 : 0];
    const double inv = active ? (a.hal ? 1.0 : 1.0 / (double)(a.cnt[b0] + a.K)) : 0.0;
    if (active) {
        const int r0 = (rb * WAVES_PER_BLOCK + wave) * a.sh.rpw;
        for (int r = 0; r < a.sh.rpw; ++r) {
            const int row = uniform(r0 + r);
            if (row >= a.sh.rows) break;
            const bool ineq = row >= a.neq;
            const double b = a.rhs[row * ld + b0];
            const double d = a.dr[row];
            // unrolled so that the candidate index is a constant: the partials
            // stay in registers (a dynamic index put them in scratch memory)
#pragma unroll
            for (int which = 0; which < 2; ++which) {
                if (which == 1 && !a.with_avg) break;
                double ax[1];
                double yv;
                if (which == 0) {
                    row_dot<double, T, double, 1>(a.A.ptr, a.A.idx, a.A.slot, a.A.val, a.A.var, a.x, ld, b0, row, ax);
                    yv = (double)a.y[row * ld + b0];
                } else {
                    row_dot<double, T, double, 1>(a.A.ptr, a.A.idx, a.A.slot, a.A.val, a.A.var, a.xsum, ld, b0, row, ax);
                    ax[0] = ax[0] * inv;
                    yv = (double)a.ysum[row * ld + b0] * inv;
                }
                double (*w)[1] = v + which * CR_N;
                const double res = b - ax[0];
                const double rp = ineq ? (res < 0.0 ? res : 0.0) : res;
                w[CR_RP2][0] += rp * rp;
                const double err = ineq ? -res : __builtin_fabs(res);
                const double perr = err / d;
                if (perr > w[CR_PERR][0] || __builtin_isnan(perr)) w[CR_PERR][0] = perr;
                w[CR_BY][0] += b * yv;
                if (ineq) {
                    w[CR_ROWGAP][0] += __builtin_fabs(yv * res);
                    const double yo = yv * d;
                    if (yo > w[CR_YPOS][0] || __builtin_isnan(yo)) w[CR_YPOS][0] = yo;
                }
                const double dy = yv - (double)a.ylast[row * ld + b0];
                w[CR_DY2][0] += dy * dy;
                if (!__builtin_isfinite(ax[0]) || !__builtin_isfinite(yv)) w[CR_NONFIN][0] += 1.0;
                if (which == 0 && ineq && yv < 0.0) w[CR_ASIG][0] += asig_weight((unsigned)row + 7919u);
                if (which == 0 && a.hal) {
                    const double dt = yv - (double)a.ysum[row * ld + b0];
                    w[CR_FPY][0] += dt * dt;
                }
            }
        }
    }
    const int ops[2 * CR_N] = {RSUM, RMAX, RSUM, RSUM, RMAX, RSUM, RSUM, RSUM, RSUM,
                               RSUM, RMAX, RSUM, RSUM, RMAX, RSUM, RSUM, RSUM, RSUM};
    block_partials<2 * CR_N, 1>(v, ops, a.partial, rb, ld, b0, lds, valid);
}

template <typename T>
struct CheckColsArgs {
    Shape sh;
    Op<double> AT;           // scaled FP64
    const T* x;
    const T* xsum;
    const T* xlast;
    const T* y;
    const T* ysum;
    const double* c;         // scaled FP64, [n] or [n][ld]
    const double* lower;     // scaled FP64 [n][ld]
    const double* upper;
    const double* dc;        // [n]
    const int* flags;
    const int* cnt;
    int K;
    int c_shared;
    int with_avg;
    int hal;
    double* partial;         // [n_rb][2*CC_N][ld]
};

template <typename T>
__global__ void __launch_bounds__(BLOCK) k_check_cols(CheckColsArgs<T> a) {
    __shared__ double lds[WAVES_PER_BLOCK * 2 * CC_N * WAVE];
    int tile, rb;
    if (!map_block(a.sh.n_real, a.sh.n_rb, tile, rb)) return;
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int b0 = tile * WAVE + lane;
    const bool valid = b0 < a.sh.ld;
    const long ld = a.sh.ld;
    double v[2 * CC_N][1];
    for (int i = 0; i < 2 * CC_N; ++i) v[i][0] = 0.0;
    const bool active = valid && !a.flags[valid ? b0 : 0];
    const double inv = active ? (a.hal ? 1.0 : 1.0 / (double)(a.cnt[b0] + a.K)) : 0.0;
    if (active) {
        const int r0 = (rb * WAVES_PER_BLOCK + wave) * a.sh.rpw;
        for (int r = 0; r < a.sh.rpw; ++r) {
            const int col = uniform(r0 + r);
            if (col >= a.sh.rows) break;
            const double cc = a.c_shared ? a.c[col] : a.c[col * ld + b0];
            const double lo = a.lower[col * ld + b0];
            const double hi = a.upper[col * ld + b0];
            const double d = a.dc[col];
            // unrolled: the candidate index is a constant (partials in registers)
#pragma unroll
            for (int which = 0; which < 2; ++which) {
                if (which == 1 && !a.with_avg) break;
                double aty[1];
                double xv;
                if (which == 0) {
                    row_dot<double, T, double, 1>(a.AT.ptr, a.AT.idx, a.AT.slot, a.AT.val, a.AT.var, a.y, ld, b0, col, aty);
                    xv = (double)a.x[col * ld + b0];
                } else {
                    row_dot<double, T, double, 1>(a.AT.ptr, a.AT.idx, a.AT.slot, a.AT.val, a.AT.var, a.ysum, ld, b0, col, aty);
                    aty[0] = aty[0] * inv;
                    xv = (double)a.xsum[col * ld + b0] * inv;
                }
                double (*w)[1] = v + which * CC_N;
                const double rc = cc - aty[0];
                const bool lfin = __builtin_isfinite(lo), ufin = __builtin_isfinite(hi);
                const double rdl = (!lfin && rc > 0.0) ? rc : 0.0;
                const double rdu = (!ufin && rc < 0.0) ? -rc : 0.0;
                w[CC_RD2][0] += rdl * rdl + rdu * rdu;
                const double dv = (rdl > rdu ? rdl : rdu) / d;
                if (dv > w[CC_DVIOL][0] || __builtin_isnan(dv)) w[CC_DVIOL][0] = dv;
                double bt = 0.0;
                if (rc > 0.0 && lfin) bt = lo * rc;
                else if (rc < 0.0 && ufin) bt = hi * rc;
                w[CC_BOUNDTERM][0] += bt;
                w[CC_POBJ][0] += cc * xv;
                const double tgt = rc >= 0.0 ? lo : hi;
                if (__builtin_isfinite(tgt)) w[CC_BOUNDGAP][0] += __builtin_fabs(rc * (xv - tgt));
                const double bv = ((lo - xv) > (xv - hi) ? (lo - xv) : (xv - hi)) * d;
                if (bv > w[CC_BVIOL][0] || __builtin_isnan(bv)) w[CC_BVIOL][0] = bv;
                const double dx = xv - (double)a.xlast[col * ld + b0];
                w[CC_DX2][0] += dx * dx;
                if (!__builtin_isfinite(rc) || !__builtin_isfinite(xv)) w[CC_NONFIN][0] += 1.0;
                if (which == 0 && a.hal) {
                    const double dt = xv - (double)a.xsum[col * ld + b0];
                    w[CC_FPX][0] += dt * dt;
                }
                if (which == 0) {
                    const double width = (lfin && ufin && hi - lo > 1.0) ? hi - lo : 1.0;
                    const bool at_lo = lfin && xv - lo <= 1e-6 * width;
                    const bool at_hi = !at_lo && ufin && hi - xv <= 1e-6 * width;
                    if (at_lo || at_hi) w[CC_ASIG][0] += (at_lo ? 1.0 : 2.0) * asig_weight((unsigned)col);
                }
            }
        }
    }
    const int ops[2 * CC_N] = {RSUM, RMAX, RSUM, RSUM, RSUM, RMAX, RSUM, RSUM, RSUM, RSUM,
                               RSUM, RMAX, RSUM, RSUM, RSUM, RMAX, RSUM, RSUM, RSUM, RSUM};
    block_partials<2 * CC_N, 1>(v, ops, a.partial, rb, ld, b0, lds, valid);
}


// ---------------------------------------------------------------------------
// Infeasibility certificates from the iterate differences since the last
// restart (PDLP: Applegate et al., "Practical large-scale linear programming
// using primal-dual hybrid gradient", 2021, sec. 5; cuPDLP: Lu & Yang, 2023),
// on the scaled LP, in FP64.  With dx = x - x_last, dy = y - y_last (the
// launcher passes the last step's end points as x and x_last for dx):
//  * dual ray (primal infeasible): r = -A^T dy, objective
//      F = b.dy + sum_j (r_j > 0 ? r_j l_j : r_j u_j)   (finite bounds only)
//    violations: dy_i > 0 on inequality rows, r_j > 0 with l_j = -inf,
//    r_j < 0 with u_j = +inf;
//  * primal ray (dual infeasible, unbounded): objective c.dx < 0, violations:
//    |A dx| on equality rows, (A dx)+ on inequality rows, dx_j < 0 with l_j
//    finite, dx_j > 0 with u_j finite.
// A ray is a certificate when its objective has the right sign and its
// largest violation is at most eps times |objective| (both scale with the ray).
// ---------------------------------------------------------------------------
enum InfRowVal { IR_BDY = 0, IR_DYPOS, IR_EQV, IR_UBV, IR_N };
enum InfColVal { IC_BT = 0, IC_RV, IC_CDX, IC_DXV, IC_N };

template <typename T>
struct InfRowsArgs {
    Shape sh;
    Op<double> A;            // scaled FP64
    const T* x; const T* xlast; const T* y; const T* ylast;
    const double* rhs;       // scaled [m][ld]
    const int* flags;
    int neq;
    double* partial;         // [n_rb][IR_N][ld]
};

template <typename T>
__global__ void __launch_bounds__(BLOCK) k_infeas_rows(InfRowsArgs<T> a) {
    __shared__ double lds[WAVES_PER_BLOCK * IR_N * WAVE];
    int tile, rb;
    if (!map_block(a.sh.n_real, a.sh.n_rb, tile, rb)) return;
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int b0 = tile * WAVE + lane;
    const bool valid = b0 < a.sh.ld;
    const long ld = a.sh.ld;
    double v[IR_N][1] = {{0.0}, {0.0}, {0.0}, {0.0}};
    if (valid && !a.flags[b0]) {
        const int r0 = (rb * WAVES_PER_BLOCK + wave) * a.sh.rpw;
        for (int r = 0; r < a.sh.rpw; ++r) {
            const int row = uniform(r0 + r);
            if (row >= a.sh.rows) break;
            const long i = row * ld + b0;
            double ax[1], axl[1];
            row_dot<double, T, double, 1>(a.A.ptr, a.A.idx, a.A.slot, a.A.val, a.A.var, a.x, ld, b0, row, ax);
            row_dot<double, T, double, 1>(a.A.ptr, a.A.idx, a.A.slot, a.A.val, a.A.var, a.xlast, ld, b0, row, axl);
            const double adx = ax[0] - axl[0];
            const double dy = (double)a.y[i] - (double)a.ylast[i];
            v[IR_BDY][0] += a.rhs[i] * dy;
            if (row < a.neq) {
                v[IR_EQV][0] = fmax(v[IR_EQV][0], __builtin_fabs(adx));
            } else {
                v[IR_DYPOS][0] = fmax(v[IR_DYPOS][0], dy);
                v[IR_UBV][0] = fmax(v[IR_UBV][0], adx);
            }
        }
    }
    const int ops[IR_N] = {RSUM, RMAX, RMAX, RMAX};
    block_partials<IR_N, 1>(v, ops, a.partial, rb, ld, b0, lds, valid);
}

template <typename T>
struct InfColsArgs {
    Shape sh;
    Op<double> AT;           // scaled FP64
    const T* x; const T* xlast; const T* y; const T* ylast;
    const double* c;         // scaled, [n] or [n][ld]
    const double* lower; const double* upper;   // scaled [n][ld]
    const int* flags;
    int c_shared;
    double* partial;         // [n_rb][IC_N][ld]
};

template <typename T>
__global__ void __launch_bounds__(BLOCK) k_infeas_cols(InfColsArgs<T> a) {
    __shared__ double lds[WAVES_PER_BLOCK * IC_N * WAVE];
    int tile, rb;
    if (!map_block(a.sh.n_real, a.sh.n_rb, tile, rb)) return;
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int b0 = tile * WAVE + lane;
    const bool valid = b0 < a.sh.ld;
    const long ld = a.sh.ld;
    double v[IC_N][1] = {{0.0}, {0.0}, {0.0}, {0.0}};
    if (valid && !a.flags[b0]) {
        const int r0 = (rb * WAVES_PER_BLOCK + wave) * a.sh.rpw;
        for (int r = 0; r < a.sh.rpw; ++r) {
            const int col = uniform(r0 + r);
            if (col >= a.sh.rows) break;
            const long i = col * ld + b0;
            double aty[1] = {0.0}, atyl[1] = {0.0};
            if (a.y) {
                row_dot<double, T, double, 1>(a.AT.ptr, a.AT.idx, a.AT.slot, a.AT.val, a.AT.var, a.y, ld, b0, col, aty);
                row_dot<double, T, double, 1>(a.AT.ptr, a.AT.idx, a.AT.slot, a.AT.val, a.AT.var, a.ylast, ld, b0, col, atyl);
            }
            const double rr = -(aty[0] - atyl[0]);
            const double lo = a.lower[i], hi = a.upper[i];
            const bool lfin = __builtin_isfinite(lo), ufin = __builtin_isfinite(hi);
            if (rr > 0.0) {
                if (lfin) v[IC_BT][0] += rr * lo;
                else v[IC_RV][0] = fmax(v[IC_RV][0], rr);
            } else if (rr < 0.0) {
                if (ufin) v[IC_BT][0] += rr * hi;
                else v[IC_RV][0] = fmax(v[IC_RV][0], -rr);
            }
            const double dx = (double)a.x[i] - (double)a.xlast[i];
            const double cc = a.c_shared ? a.c[col] : a.c[i];
            v[IC_CDX][0] += cc * dx;
            if (lfin && dx < 0.0) v[IC_DXV][0] = fmax(v[IC_DXV][0], -dx);
            if (ufin && dx > 0.0) v[IC_DXV][0] = fmax(v[IC_DXV][0], dx);
        }
    }
    const int ops[IC_N] = {RSUM, RMAX, RSUM, RMAX};
    block_partials<IC_N, 1>(v, ops, a.partial, rb, ld, b0, lds, valid);
}

// ---------------------------------------------------------------------------
// Per-LP restart / acceptance decisions (one thread per LP).
//
// Restart rule (PDLP: Applegate et al., NeurIPS 2021, sec. 3; cuPDLP.jl, Lu &
// Yang 2023): the candidate is the current iterate or the average, whichever
// has the smaller weighted KKT error
//     K_w(z) = sqrt(w^2 |r_p|^2 + |r_d|^2 / w^2 + gap^2);
// restart when K(cand) <= 0.2 K(last restart), or K(cand) <= 0.8 K(last) and
// K(cand) > K(previous candidate), or the restart length reaches 0.36 of the
// iterations run.  The primal weight is smoothed at restarts:
//     w <- exp(0.5 log(|dy|/|dx|) + 0.5 log w).
// ---------------------------------------------------------------------------
struct LpState {
    double omega;        // primal weight
    double kkt_last;     // weighted KKT at the last restart point
    double kkt_prev;     // weighted KKT of the previous candidate
    double rel_kkt;      // relative KKT of the better of (current, average)
    double avg_div;      // divisor of the averages at the last check
    long long iters;     // PDHG iterations since reset
    long long accepted_iter;
    int status;          // LPE_STATUS_*
    int action;          // Action for k_check_apply
    int restarts;
    int source;          // 0: result from iterate, 1: from polish
    double asig;         // active-set signature of the current iterate at the last check
    int asig_same;       // consecutive checks with an unchanged signature
    int pad_;
};

enum Action { ACT_NONE = 0, ACT_RESTART_CUR = 1, ACT_RESTART_AVG = 2, ACT_ACCEPT_CUR = 3, ACT_ACCEPT_AVG = 4 };

// Per-LP solver state at a reset: LPs b < batch running with primal weight
// omega0, padding LPs frozen; iteration counters and result sources zeroed.
__global__ void k_init_state(LpState* st, int* flags, int* cnt, int* source, double* om, float* om_f, double* iom,
                             float* iom_f, double omega0, int batch, int ld) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= ld) return;
    LpState s;
    s.omega = omega0;
    s.kkt_last = INFINITY;
    s.kkt_prev = INFINITY;
    s.rel_kkt = INFINITY;
    s.avg_div = 1.0;
    s.iters = 0;
    s.accepted_iter = -1;
    s.status = b < batch ? 0 : 1;
    s.action = 0;
    s.restarts = 0;
    s.source = 0;
    s.asig = -1.0;
    s.asig_same = 0;
    s.pad_ = 0;
    st[b] = s;
    flags[b] = b < batch ? 0 : 1;
    cnt[b] = 0;
    source[b] = 0;
    om[b] = omega0; iom[b] = 1.0 / omega0; om_f[b] = (float)omega0; iom_f[b] = (float)(1.0 / omega0);
}

// One thread per LP: a certificate marks a running LP primal (3) or dual (4)
// infeasible and freezes it.
__global__ void k_infeas_decide(LpState* st, int* flags, const double* rows, const double* cols, int batch, int ld,
                                int have_rows, double eps) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= batch || flags[b] || st[b].status != 0) return;
    const double bdy = have_rows ? rows[IR_BDY * (long)ld + b] : 0.0;
    const double dypos = have_rows ? rows[IR_DYPOS * (long)ld + b] : 0.0;
    const double eqv = have_rows ? rows[IR_EQV * (long)ld + b] : 0.0;
    const double ubv = have_rows ? rows[IR_UBV * (long)ld + b] : 0.0;
    const double fobj = bdy + cols[IC_BT * (long)ld + b];
    const double dviol = fmax(dypos, cols[IC_RV * (long)ld + b]);
    const double cdx = cols[IC_CDX * (long)ld + b];
    const double pviol = fmax(fmax(eqv, ubv), cols[IC_DXV * (long)ld + b]);
    int status = 0;
    if (isfinite(fobj) && isfinite(dviol) && fobj > 0.0 && dviol <= eps * fobj) status = 3;
    else if (isfinite(cdx) && isfinite(pviol) && cdx < 0.0 && pviol <= eps * (-cdx)) status = 4;
    if (status) {
        st[b].status = status;
        flags[b] = 1;
    }
}


struct DecideArgs {
    const double* rows;     // [2*CR_N][ld] reduced row values
    const double* cols;     // [2*CC_N][ld]
    LpState* st;
    int* flags;
    int* cnt;
    double* omega64;
    float* omega32;
    double* inv_omega64;
    float* inv_omega32;
    const double* bnorm;    // [ld]
    const double* cnorm;    // [ld]
    int batch;
    int ld;
    int K;
    int restarted;
    int hal;                // Halpern: cnt already counts this chunk, candidate 2 is T(z)
    int certify;
    double eps_p, eps_d, eps_g;
    // primal-weight update at a restart: log w <- theta log(|dy|/|dx|) +
    // (1 - theta) log w, the change limited to +-step when step > 0
    double pw_theta, pw_step;
};

__device__ __forceinline__ double next_primal_weight(double om, double dx, double dy, double theta, double step) {
    if (!(dx > 1e-10 && dy > 1e-10 && isfinite(dx) && isfinite(dy))) return om;
    const double lo = log(om);
    double nl = theta * log(dy / dx) + (1.0 - theta) * lo;
    if (step > 0.0) nl = fmin(fmax(nl, lo - step), lo + step);
    return fmin(fmax(exp(nl), 1e-8), 1e8);
}

__global__ void k_check_decide(DecideArgs a) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= a.batch) return;
    LpState s = a.st[b];
    s.action = ACT_NONE;
    if (s.status != 0) { a.st[b] = s; return; }
    s.iters += a.K;
    const long ld = a.ld;
    double om = s.omega;
    double kkt[2] = {0.0, 0.0}, rel[2] = {0.0, 0.0};
    bool pass[2] = {false, false};
    const int nw = a.restarted ? 2 : 1;
    for (int w = 0; w < nw; ++w) {
        const double* R = a.rows + (long)w * CR_N * ld + b;
        const double* C = a.cols + (long)w * CC_N * ld + b;
        const double rp = sqrt(R[CR_RP2 * ld]);
        const double rd = sqrt(C[CC_RD2 * ld]);
        const double pobj = C[CC_POBJ * ld];
        const double dobj = R[CR_BY * ld] + C[CC_BOUNDTERM * ld];
        const double gap = fabs(pobj - dobj);
        kkt[w] = sqrt(om * om * rp * rp + rd * rd / (om * om) + gap * gap);
        double r1 = rp / (1.0 + a.bnorm[b]);
        const double r2 = rd / (1.0 + a.cnorm[b]);
        const double r3 = gap / (1.0 + fabs(pobj) + fabs(dobj));
        r1 = r2 > r1 ? r2 : r1;
        rel[w] = r3 > r1 ? r3 : r1;
        const bool finite = R[CR_NONFIN * ld] == 0.0 && C[CC_NONFIN * ld] == 0.0;
        double primal = fmax(R[CR_PERR * ld], C[CC_BVIOL * ld]);
        primal = fmax(primal, 0.0);
        double dual = fmax(R[CR_YPOS * ld], C[CC_DVIOL * ld]);
        dual = fmax(dual, 0.0);
        const double cg = (C[CC_BOUNDGAP * ld] + R[CR_ROWGAP * ld]) / fmax(1.0, fabs(pobj));
        pass[w] = finite && isfinite(cg) && primal <= a.eps_p && dual <= a.eps_d && cg <= a.eps_g;
    }
    const int best = (nw == 2 && kkt[1] < kkt[0]) ? 1 : 0;
    s.rel_kkt = rel[best];
    const double sig = a.cols[CC_ASIG * ld + b] + a.rows[CR_ASIG * ld + b];
    s.asig_same = (a.K > 0 && sig == s.asig) ? s.asig_same + 1 : 0;
    s.asig = sig;
    const int since = a.hal ? a.cnt[b] : a.cnt[b] + a.K;   // iterations since the last restart
    s.avg_div = a.hal ? 1.0 : (double)since;
    if (a.certify && (pass[0] || pass[1])) {
        s.status = 1;
        s.action = pass[0] ? ACT_ACCEPT_CUR : ACT_ACCEPT_AVG;
        s.accepted_iter = s.iters;
        a.flags[b] = 1;
    } else if (a.restarted && a.hal) {
        // restarted Halpern (Lu & Yang 2024): restart to T(z) on the weighted
        // fixed-point residual |z - T(z)|_w = sqrt(w |dx|^2 + |dy|^2 / w)
        s.rel_kkt = fmin(rel[0], rel[1]);
        const double fp = sqrt(om * a.cols[CC_FPX * ld + b] + a.rows[CR_FPY * ld + b] / om);
        bool restart = false;
        if (!(s.kkt_last < INFINITY)) {
            s.kkt_last = fp;
        } else {
            restart = fp <= 0.2 * s.kkt_last || (fp <= 0.8 * s.kkt_last && fp > s.kkt_prev) ||
                      (double)since >= 0.36 * (double)s.iters;
        }
        s.kkt_prev = fp;
        if (restart) {
            s.action = ACT_RESTART_AVG;
            const double dx = sqrt(a.cols[(long)(CC_N + CC_DX2) * ld + b]);
            const double dy = sqrt(a.rows[(long)(CR_N + CR_DY2) * ld + b]);
            om = next_primal_weight(om, dx, dy, a.pw_theta, a.pw_step);
            s.omega = om;
            s.kkt_last = fp;
            s.kkt_prev = INFINITY;
            s.restarts += 1;
            a.cnt[b] = 0;
        }
        a.omega64[b] = om;
        a.omega32[b] = (float)om;
        a.inv_omega64[b] = 1.0 / om;
        a.inv_omega32[b] = (float)(1.0 / om);
    } else if (a.restarted) {
        const double kc = kkt[best];
        bool restart = false;
        if (!(s.kkt_last < INFINITY)) {
            s.kkt_last = kc;
        } else {
            restart = kc <= 0.2 * s.kkt_last || (kc <= 0.8 * s.kkt_last && kc > s.kkt_prev) ||
                      (double)since >= 0.36 * (double)s.iters;
        }
        s.kkt_prev = kc;
        if (restart) {
            s.action = best ? ACT_RESTART_AVG : ACT_RESTART_CUR;
            const double dx = sqrt(a.cols[(long)(best * CC_N + CC_DX2) * ld + b]);
            const double dy = sqrt(a.rows[(long)(best * CR_N + CR_DY2) * ld + b]);
            om = next_primal_weight(om, dx, dy, a.pw_theta, a.pw_step);
            s.omega = om;
            s.kkt_last = kc;
            s.kkt_prev = INFINITY;
            s.restarts += 1;
            a.cnt[b] = 0;
        } else if (!a.hal) {
            a.cnt[b] += a.K;
        }
        a.omega64[b] = om;
        a.omega32[b] = (float)om;
        a.inv_omega64[b] = 1.0 / om;
        a.inv_omega32[b] = (float)(1.0 / om);
    }
    a.st[b] = s;
}

// Apply restart / acceptance actions to one family of vectors.
template <typename T>
struct ApplyArgs {
    int rows;
    int ld;
    int rows_per_block;
    const LpState* st;
    T* v;        // x or y
    T* vsum;
    T* vlast;
    T* vbar;     // xbar, or nullptr for duals
    int keep_sum;   // Halpern: vsum holds T(z) (the polish input), not a running sum
};

template <typename T>
__global__ void k_check_apply(ApplyArgs<T> a) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= a.ld) return;
    const int act = a.st[b].action;
    if (act == ACT_NONE) return;
    const double inv = 1.0 / a.st[b].avg_div;
    const long ld = a.ld;
    const int r0 = blockIdx.y * a.rows_per_block;
    const int r1 = min(a.rows, r0 + a.rows_per_block);
    for (int r = r0; r < r1; ++r) {
        const long i = r * ld + b;
        T v = a.v[i];
        if (act == ACT_RESTART_AVG || act == ACT_ACCEPT_AVG) {
            v = (T)((double)a.vsum[i] * inv);
            a.v[i] = v;
        }
        if (a.vbar) a.vbar[i] = v;
        if (act == ACT_RESTART_AVG || act == ACT_RESTART_CUR) {
            a.vlast[i] = v;
            if (!a.keep_sum) a.vsum[i] = T(0);
        }
    }
}

// ---------------------------------------------------------------------------
// FP64 certificate in original units: restates certify_blocks_device
// (src/gpu_block_lp.py:52-132) and the host _certificate
// (src/cpu_repeated_lp.py:150-178).
// ---------------------------------------------------------------------------
enum CertRowVal { XR_EQ = 0, XR_UB, XR_ROWGAP, XR_YUB, XR_NONFIN, XR_N };
enum CertColVal { XC_LOW = 0, XC_UP, XC_RLOW, XC_RUP, XC_BOUNDGAP, XC_OBJ, XC_NONFIN_X, XC_NONFIN_R, XC_N };

__device__ __forceinline__ double nmax(double acc, double o) {
    return (o > acc || __builtin_isnan(o)) ? o : acc;
}

struct CertRowsArgs {
    Shape sh;
    Op<double> A;           // original FP64
    const double* x;        // [n][ld]
    const double* y;        // [m][ld]
    const double* rhs;      // [m][ld]
    int neq;
    double* partial;        // [n_rb][XR_N][ld]
};

__global__ void __launch_bounds__(BLOCK) k_cert_rows(CertRowsArgs a) {
    __shared__ double lds[WAVES_PER_BLOCK * XR_N * WAVE];
    int tile, rb;
    if (!map_block(a.sh.n_real, a.sh.n_rb, tile, rb)) return;
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int b0 = tile * WAVE + lane;
    const bool valid = b0 < a.sh.ld;
    const long ld = a.sh.ld;
    double v[XR_N][1] = {{0.0}, {-INFINITY}, {0.0}, {-INFINITY}, {0.0}};
    if (valid) {
        const int r0 = (rb * WAVES_PER_BLOCK + wave) * a.sh.rpw;
        for (int r = 0; r < a.sh.rpw; ++r) {
            const int row = uniform(r0 + r);
            if (row >= a.sh.rows) break;
            double act[1];
            row_dot<double, double, double, 1>(a.A.ptr, a.A.idx, a.A.slot, a.A.val, a.A.var, a.x, ld, b0, row, act);
            const double err = act[0] - a.rhs[row * ld + b0];
            const double yv = a.y[row * ld + b0];
            if (row < a.neq) {
                v[XR_EQ][0] = nmax(v[XR_EQ][0], __builtin_fabs(err));
            } else {
                v[XR_UB][0] = nmax(v[XR_UB][0], err);
                v[XR_ROWGAP][0] += __builtin_fabs(yv * (a.rhs[row * ld + b0] - act[0]));
                v[XR_YUB][0] = nmax(v[XR_YUB][0], yv);
            }
            if (!__builtin_isfinite(act[0]) || !__builtin_isfinite(yv)) v[XR_NONFIN][0] += 1.0;
        }
    }
    const int ops[XR_N] = {RMAX, RMAX, RSUM, RMAX, RSUM};
    block_partials<XR_N, 1>(v, ops, a.partial, rb, ld, b0, lds, valid);
}

struct CertColsArgs {
    Shape sh;
    Op<double> AT;          // original FP64
    const double* x;
    const double* y;
    const double* c;        // [n] or [n][ld]
    const double* lower;
    const double* upper;
    int c_shared;
    double* partial;        // [n_rb][XC_N][ld]
};

__global__ void __launch_bounds__(BLOCK) k_cert_cols(CertColsArgs a) {
    __shared__ double lds[WAVES_PER_BLOCK * XC_N * WAVE];
    int tile, rb;
    if (!map_block(a.sh.n_real, a.sh.n_rb, tile, rb)) return;
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int b0 = tile * WAVE + lane;
    const bool valid = b0 < a.sh.ld;
    const long ld = a.sh.ld;
    double v[XC_N][1] = {{-INFINITY}, {-INFINITY}, {-INFINITY}, {-INFINITY}, {0.0}, {0.0}, {0.0}, {0.0}};
    if (valid) {
        const int r0 = (rb * WAVES_PER_BLOCK + wave) * a.sh.rpw;
        for (int r = 0; r < a.sh.rpw; ++r) {
            const int col = uniform(r0 + r);
            if (col >= a.sh.rows) break;
            double aty[1];
            row_dot<double, double, double, 1>(a.AT.ptr, a.AT.idx, a.AT.slot, a.AT.val, a.AT.var, a.y, ld, b0, col, aty);
            const double cc = a.c_shared ? a.c[col] : a.c[col * ld + b0];
            const double xv = a.x[col * ld + b0];
            const double lo = a.lower[col * ld + b0], hi = a.upper[col * ld + b0];
            const double rc = cc - aty[0];
            v[XC_LOW][0] = nmax(v[XC_LOW][0], lo - xv);
            v[XC_UP][0] = nmax(v[XC_UP][0], xv - hi);
            v[XC_RLOW][0] = nmax(v[XC_RLOW][0], __builtin_isfinite(lo) ? 0.0 : rc);
            v[XC_RUP][0] = nmax(v[XC_RUP][0], __builtin_isfinite(hi) ? 0.0 : -rc);
            const double tgt = rc >= 0.0 ? lo : hi;
            const double t2 = __builtin_isfinite(tgt) ? tgt : xv;
            v[XC_BOUNDGAP][0] += __builtin_fabs(rc * (xv - t2));
            v[XC_OBJ][0] += cc * xv;
            if (!__builtin_isfinite(xv)) v[XC_NONFIN_X][0] += 1.0;
            if (!__builtin_isfinite(rc)) v[XC_NONFIN_R][0] += 1.0;
        }
    }
    const int ops[XC_N] = {RMAX, RMAX, RMAX, RMAX, RSUM, RSUM, RSUM, RSUM};
    block_partials<XC_N, 1>(v, ops, a.partial, rb, ld, b0, lds, valid);
}

// metrics[b][LPE_NMETRIC] (batch-major, device) and pass flags.
struct CertFinalArgs {
    const double* rows;     // [XR_N][ld] reduced (nullptr when m == 0)
    const double* cols;     // [XC_N][ld]
    double* metrics;        // [B][8]
    int* passed;            // [ld]
    int batch, ld;
    double eps_p, eps_d, eps_g;
};

__global__ void k_cert_final(CertFinalArgs a) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= a.batch) return;
    const long ld = a.ld;
    double row_eq = 0.0, row_ub = 0.0, row_gap = 0.0, dual_rows = 0.0, nonfin_rows = 0.0;
    if (a.rows) {
        row_eq = a.rows[XR_EQ * ld + b];
        row_ub = a.rows[XR_UB * ld + b];
        row_gap = a.rows[XR_ROWGAP * ld + b];
        dual_rows = a.rows[XR_YUB * ld + b];
        nonfin_rows = a.rows[XR_NONFIN * ld + b];
        if (row_ub == -INFINITY) row_ub = 0.0;          // empty inequality group
        if (dual_rows == -INFINITY) dual_rows = 0.0;
    }
    const double low = a.cols[XC_LOW * ld + b], up = a.cols[XC_UP * ld + b];
    const double rlow = a.cols[XC_RLOW * ld + b], rup = a.cols[XC_RUP * ld + b];
    const double bound_gap = a.cols[XC_BOUNDGAP * ld + b];
    const double obj = a.cols[XC_OBJ * ld + b];
    // primal = max(0, max(max(row_eq, row_ub), max(max(l-x), max(x-u))))
    const double primal = nmax(0.0, nmax(nmax(row_eq, row_ub), nmax(low, up)));
    const double dual = nmax(0.0, nmax(dual_rows, nmax(rlow, rup)));
    const double gap = (bound_gap + row_gap) / fmax(1.0, fabs(obj));
    const bool finite_primal = a.cols[XC_NONFIN_X * ld + b] == 0.0 && isfinite(obj) && nonfin_rows == 0.0;
    const bool finite = finite_primal && a.cols[XC_NONFIN_R * ld + b] == 0.0;
    const bool ok = finite && isfinite(gap) && primal <= a.eps_p && dual <= a.eps_d && gap <= a.eps_g;
    double* m = a.metrics + (long)b * 8;
    m[0] = primal; m[1] = dual; m[2] = gap; m[3] = obj; m[4] = ok ? 1.0 : 0.0;
    m[5] = finite ? 1.0 : 0.0; m[6] = bound_gap; m[7] = row_gap;
    if (a.passed) a.passed[b] = ok ? 1 : 0;
}

// ---------------------------------------------------------------------------
// FP64 active-set polish.
//
// With x from the iteration, variables within tol*max(1, u-l) of a bound are
// fixed there (set N); the rest are free (F).  Inequality rows are active when
// their dual is negative or their slack is small.  In the equilibrated space
// (A^ = Dr A Dc) the polish solves two least-squares problems with the same
// normal-equations operator M = R A^ D A^T R (R, D: 0/1 masks of active rows
// and free columns) by conjugate gradients:
//   primal: minimum-norm correction  A^_F dz = Dr (b - A x_N)  -> x_F += Dc dz
//   dual:   min |D (c^ - A^T u)|, warm-started at u = y^        -> y = Dr u
// A solution that identifies the optimal active set satisfies every KKT
// condition to rounding; the FP64 certificate decides acceptance.
// ---------------------------------------------------------------------------
template <typename T>
struct PolishColsInitArgs {
    Shape sh;
    Op<double> AT;           // scaled FP64 (orig == 0) or original FP64 (orig == 1)
    const T* x;              // iterate: scaled (orig == 0) or original units
    const T* y;
    const double* c;         // c^ scaled (orig == 0) or original c
    int c_shared;
    int orig;
    const double* dc;
    const double* lower;     // original [n][ld]
    const double* upper;
    const int* sel;          // [ld] selected LPs
    double* xs;              // [n][ld] snapped original
    uint8_t* fmask;          // [n][ld]
    double tol, tol_far, tol_r;
};

// A variable is fixed at a bound when it lies within tol*max(1, u-l) of it, or
// within tol_far*max(1, u-l) while its reduced cost r = c - A^T y points at
// that bound by more than tol_r*(1 + |c|).
template <typename T>
__global__ void __launch_bounds__(BLOCK) k_polish_cols_init(PolishColsInitArgs<T> a) {
    int tile, rb;
    if (!map_block(a.sh.n_real, a.sh.n_rb, tile, rb)) return;
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int b0 = tile * WAVE + lane;
    if (b0 >= a.sh.ld || !a.sel[b0]) return;
    const long ld = a.sh.ld;
    const int r0 = (rb * WAVES_PER_BLOCK + wave) * a.sh.rpw;
    for (int r = 0; r < a.sh.rpw; ++r) {
        const int j = uniform(r0 + r);
        if (j >= a.sh.rows) break;
        const long i = j * ld + b0;
        // the column's own operands are loaded before the gathers
        const double cs = a.c_shared ? a.c[j] : a.c[i];
        const double xraw = (double)a.x[i];
        const double lo = a.lower[i], hi = a.upper[i];
        double aty[1];
        row_dot<double, T, double, 1>(a.AT.ptr, a.AT.idx, a.AT.slot, a.AT.val, a.AT.var, a.y, ld, b0, j, aty);
        const double d = a.dc[j];
        const double rc = a.orig ? cs - aty[0] : (cs - aty[0]) / d;
        const double co = a.orig ? cs : cs / d;
        const double thr = a.tol_r * (1.0 + fabs(co));
        double xv = a.orig ? xraw : xraw * d;
        double w = hi - lo;
        w = (isfinite(w) && w > 1.0) ? w : 1.0;
        uint8_t free_ = 1;
        if (xv - lo <= a.tol * w) { xv = lo; free_ = 0; }
        else if (hi - xv <= a.tol * w) { xv = hi; free_ = 0; }
        else if (rc > thr && xv - lo <= a.tol_far * w) { xv = lo; free_ = 0; }
        else if (rc < -thr && hi - xv <= a.tol_far * w) { xv = hi; free_ = 0; }
        a.xs[i] = xv;
        a.fmask[i] = free_;
    }
}

template <typename T>
struct PolishRowsInitArgs {
    Shape sh;
    Op<double> A;            // original FP64 values
    Op<double> As;           // scaled FP64 values (Jacobi diagonal)
    const double* xs;
    const T* y;              // dual iterate: scaled (orig == 0) or original units
    const double* rhs;       // original [m][ld]
    const double* dr;
    int orig;
    const int* sel;
    const uint8_t* fmask;
    uint8_t* rmask;
    double* rp; double2* p; double* wp;   // p: the CG directions, interleaved (p1, p2); writes p1
    double* wd;
    float* invdiag;          // [m][ld] 1 / (R A^ D A^T R)_ii in FP32, 0 on inactive rows and on
                             // active rows without a free column
    int neq;
    double tol_y, tol_s;
    int plain_orig;          // 1: |r|^2 in original units (r / dr); 2: max_i r_i^2 in original
                             // units (the infinity norm of the active rows' residual, squared); 0: |r|^2 scaled
    double* partial;         // [n_rb][2][ld]: r.D^-1 r and the primal stop measure
};

template <typename T>
__global__ void __launch_bounds__(BLOCK) k_polish_rows_init(PolishRowsInitArgs<T> a) {
    __shared__ double lds[WAVES_PER_BLOCK * 2 * WAVE];
    int tile, rb;
    if (!map_block(a.sh.n_real, a.sh.n_rb, tile, rb)) return;
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int b0 = tile * WAVE + lane;
    const bool valid = b0 < a.sh.ld;
    const long ld = a.sh.ld;
    double v[2][1] = {{0.0}, {0.0}};
    if (valid && a.sel[b0]) {
        const int r0 = (rb * WAVES_PER_BLOCK + wave) * a.sh.rpw;
        for (int r = 0; r < a.sh.rpw; ++r) {
            const int row = uniform(r0 + r);
            if (row >= a.sh.rows) break;
            const long i = row * ld + b0;
            // the row's own operands are loaded before the gathers (their
            // memory round trips overlap)
            const double bb = a.rhs[i];
            const double yraw = (double)a.y[i];
            double act[1];
            row_dot<double, double, double, 1>(a.A.ptr, a.A.idx, a.A.slot, a.A.val, a.A.var, a.xs, ld, b0, row, act);
            const double res = bb - act[0];
            const double yv = a.orig ? yraw / a.dr[row] : yraw;
            uint8_t on = 1;
            if (row >= a.neq) {
                const double yo = a.orig ? yraw : yv * a.dr[row];
                on = (yo < -a.tol_y || res <= a.tol_s * fmax(1.0, fabs(bb))) ? 1 : 0;
            }
            a.rmask[i] = on;
            // Jacobi diagonal sum_k a^_ik^2 over the free columns, in storage
            // order; the free-column masks (and per-LP values) of 8 entries
            // are loaded together instead of one dependent load per entry
            double dg = 0.0;
            const int k1 = sload(a.As.ptr, row + 1);
            for (int k0 = sload(a.As.ptr, row); k0 < k1; k0 += 8) {
                const int cnt = k1 - k0 < 8 ? k1 - k0 : 8;
                int cl[8], sl[8];
                double av[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    if (u < cnt) { cl[u] = sload(a.As.idx, k0 + u); sl[u] = sload(a.As.slot, k0 + u); av[u] = sload(a.As.val, k0 + u); }
                    else { cl[u] = 0; sl[u] = -1; av[u] = 0.0; }
                }
                uint8_t fm[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) fm[u] = u < cnt ? a.fmask[(long)cl[u] * ld + b0] : 0;
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    if (u < cnt && sl[u] >= 0) av[u] = a.As.var[(long)sl[u] * ld + b0];
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    if (fm[u]) dg = dg + av[u] * av[u];
            }
            // The Jacobi preconditioner is stored in FP32 (any positive diagonal
            // is a valid preconditioner; every use reads the same rounded value):
            // half the preconditioner's bytes in the CG rows and direction passes.
            // A positive entry stays positive and finite.
            float ivf = (on && dg > 0.0) ? (float)(1.0 / dg) : 0.0f;
            if (on && dg > 0.0) ivf = fminf(fmaxf(ivf, FLT_MIN), FLT_MAX);
            a.invdiag[i] = ivf;
            const double inv = (double)ivf;
            const double rv = on ? a.dr[row] * res : 0.0;
            a.rp[i] = rv; a.p[i].x = rv * inv; a.wp[i] = 0.0;
            a.wd[i] = on ? yv : 0.0;
            v[0][0] += rv * rv * inv;
            const double ro = a.plain_orig ? (on ? res : 0.0) : rv;
            if (a.plain_orig == 2) v[1][0] = fmax(v[1][0], ro * ro);
            else v[1][0] += ro * ro;
        }
    }
    const int ops[2] = {RSUM, a.plain_orig == 2 ? RMAX : RSUM};
    block_partials<2, 1>(v, ops, a.partial, rb, ld, b0, lds, valid);
}

// q_j = fmask * (c^_j - sum_i a^_ij p_i): the dual system's set-up product
// (q = D (c^ - A^T u)), one system per launch.
struct PolishColsArgs {
    Shape sh;
    Op<double> AT;           // scaled FP64
    const double* p1;        // [m][ld]
    double* q1;              // [n][ld]
    const double* c;         // scaled c^
    int c_shared;
    const uint8_t* fmask;
    const int* sel;
};

template <int LPL>
__global__ void __launch_bounds__(BLOCK) k_polish_cols(PolishColsArgs a) {
    int tile, rb;
    if (!map_block(a.sh.n_real, a.sh.n_rb, tile, rb)) return;
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int b0 = tile * WAVE * LPL + lane * LPL;
    if (b0 >= a.sh.ld) return;
    const long ld = a.sh.ld;
    bool sel[LPL];
    bool anys = false;
#pragma unroll
    for (int t = 0; t < LPL; ++t) {
        sel[t] = a.sel[b0 + t] != 0;
        anys = anys || sel[t];
    }
    if (!anys) return;
    const int r0 = (rb * WAVES_PER_BLOCK + wave) * a.sh.rpw;
    for (int r = 0; r < a.sh.rpw; ++r) {
        const int col = uniform(r0 + r);
        if (col >= a.sh.rows) break;
        const long i = col * ld + b0;
        double s1[LPL];
        row_dot<double, double, double, LPL>(a.AT.ptr, a.AT.idx, a.AT.slot, a.AT.val, a.AT.var, a.p1, ld, b0, col, s1);
#pragma unroll
        for (int t = 0; t < LPL; ++t) {
            const double cc = a.c_shared ? a.c[col] : a.c[i + t];
            if (sel[t]) a.q1[i + t] = (a.fmask[i + t] ? 1.0 : 0.0) * (cc - s1[t]);
        }
    }
}

// Dual set-up, rows: r_d = R A^ q, p_d = r_d * invdiag and the partial of
// r_d.p_d (the second partial of each row block is 0, so the two-value
// reduction of the primal set-up applies).
struct PolishRowsArgs {
    Shape sh;
    Op<double> A;            // scaled FP64
    const double* q1;
    double* setup_r;
    double2* setup_p;        // interleaved CG directions: writes p2
    const float* invdiag;
    const uint8_t* rmask;
    const int* sel;
    double* partial;         // [n_rb][2][ld]
};

template <int LPL>
__global__ void __launch_bounds__(BLOCK) k_polish_rows(PolishRowsArgs a) {
    __shared__ double lds[WAVES_PER_BLOCK * 2 * WAVE * LPL];
    int tile, rb;
    if (!map_block(a.sh.n_real, a.sh.n_rb, tile, rb)) return;
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int b0 = tile * WAVE * LPL + lane * LPL;
    const bool valid = b0 < a.sh.ld;
    const long ld = a.sh.ld;
    double v[2][LPL];
    bool sel[LPL];
    bool anys = false;
#pragma unroll
    for (int t = 0; t < LPL; ++t) {
        v[0][t] = 0.0; v[1][t] = 0.0;
        sel[t] = valid && a.sel[b0 + t] != 0;
        anys = anys || sel[t];
    }
    if (anys) {
        const int r0 = (rb * WAVES_PER_BLOCK + wave) * a.sh.rpw;
        for (int r = 0; r < a.sh.rpw; ++r) {
            const int row = uniform(r0 + r);
            if (row >= a.sh.rows) break;
            const long i = row * ld + b0;
            double t1[LPL];
            row_dot<double, double, double, LPL>(a.A.ptr, a.A.idx, a.A.slot, a.A.val, a.A.var, a.q1, ld, b0, row, t1);
#pragma unroll
            for (int t = 0; t < LPL; ++t) {
                if (sel[t]) {
                    const double sv1 = (a.rmask[i + t] ? 1.0 : 0.0) * t1[t];
                    const double z = sv1 * (double)a.invdiag[i + t];
                    a.setup_r[i + t] = sv1; a.setup_p[i + t].y = z;
                    v[0][t] += sv1 * z;
                }
            }
        }
    }
    const int ops[2] = {RSUM, RSUM};
    block_partials<2, LPL>(v, ops, a.partial, rb, ld, b0, lds, valid);
}

// ---------------------------------------------------------------------------
// One Jacobi-preconditioned CG step for both systems (0: primal correction,
// 1: dual least squares) is three passes over the batch:
//   columns  q = D A^T p, with the partials of |q|^2 - p is zero off the
//            active rows and D is a 0/1 mask, so p.Mp = p.R A^ D A^T R p =
//            |q|^2 and alpha = rho / |q|^2 is known before the rows pass;
//   rows     s = R A^ q, applied at once: w += alpha p, r -= alpha s (s is
//            never stored), with the partials of r.D^-1 r and of the primal
//            stop measure;
//   direction p = D^-1 r + beta p.
// ---------------------------------------------------------------------------
struct CgColsArgs {
    Shape sh;
    Op<double> AT;           // scaled FP64
    const double2* p;        // [m][ld] interleaved (p1, p2)
    double2* q;              // [n][ld] interleaved (q1, q2)
    const uint8_t* fmask;
    const int* sel;
    const int* act;          // [2][ld] CG activity
    double* partial;         // [n_rb][2][ld]: |q1|^2, |q2|^2
};

template <int LPL, int D, int WPB>
__global__ void __launch_bounds__(WPB * WAVE) k_cg_cols(CgColsArgs a) {
    __shared__ double lds[WPB * 2 * WAVE * LPL];
    int tile, rb;
    if (!map_block(a.sh.n_real, a.sh.n_rb, tile, rb)) return;
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int b0 = tile * WAVE * LPL + lane * LPL;
    const bool valid = b0 < a.sh.ld;
    const long ld = a.sh.ld;
    double v[2][LPL];
    bool on1[LPL], on2[LPL];
    bool any = false;
#pragma unroll
    for (int t = 0; t < LPL; ++t) {
        v[0][t] = 0.0; v[1][t] = 0.0;
        const bool sel = valid && a.sel[b0 + t] != 0;
        on1[t] = sel && a.act[b0 + t];
        on2[t] = sel && a.act[ld + b0 + t];
        any = any || on1[t] || on2[t];
    }
    if (any) {
        const int r0 = (rb * WPB + wave) * a.sh.rpw;
        for (int r = 0; r < a.sh.rpw; ++r) {
            const int col = uniform(r0 + r);
            if (col >= a.sh.rows) break;
            const long i = col * ld + b0;
            // the column's mask is loaded before the gathers so that the two
            // memory round trips overlap
            double f[LPL];
#pragma unroll
            for (int t = 0; t < LPL; ++t) f[t] = (on1[t] || on2[t]) && a.fmask[i + t] ? 1.0 : 0.0;
            double s1[LPL], s2[LPL];
            row_dot2<LPL, D>(a.AT.ptr, a.AT.idx, a.AT.slot, a.AT.val, a.AT.var, a.p, ld, b0, col, s1, s2);
#pragma unroll
            for (int t = 0; t < LPL; ++t) {
                if (!(on1[t] || on2[t])) continue;
                // both components are written; the rows pass reads only those
                // of active systems
                const double2 q = make_double2(f[t] * s1[t], f[t] * s2[t]);
                if (on1[t]) v[0][t] += q.x * q.x;
                if (on2[t]) v[1][t] += q.y * q.y;
                a.q[i + t] = q;
            }
        }
    }
    const int ops[2] = {RSUM, RSUM};
    block_partials<2, LPL, WPB>(v, ops, a.partial, rb, ld, b0, lds, valid);
}

struct CgRowsArgs {
    Shape sh;
    Op<double> A;            // scaled FP64
    const double2* q;        // [n][ld] interleaved (q1, q2)
    const double2* p;        // [m][ld] interleaved (p1, p2)
    double* w1; double* r1;  // [m][ld], primal system
    double* w2; double* r2;  // dual system
    const double* alpha;     // [2][ld]
    const float* invdiag;    // Jacobi preconditioner, 0 off the active rows (and on
                             // active rows without a free column, whose A^ q is 0)
    const double* dr;        // row scale: the primal stop measure in original units (r / dr), or nullptr (scaled)
    const int* sel;
    const int* act;          // [2][ld]
    double* partial;         // [n_rb][3][ld]: r_p.D^-1 r_p, r_d.D^-1 r_d, primal stop measure
    int inf_norm;            // 1: stop measure max_i (r_i / dr_i)^2, else sum_i (r_i [/ dr_i])^2
    int reverse;             // walk the LP tiles last to first
};

template <int LPL, int D, int WPB>
__global__ void __launch_bounds__(WPB * WAVE) k_cg_rows(CgRowsArgs a) {
    __shared__ double lds[WPB * 3 * WAVE * LPL];
    int tile, rb;
    if (!map_block(a.sh.n_real, a.sh.n_rb, tile, rb)) return;
    if (a.reverse) tile = a.sh.n_tiles - 1 - tile;
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int b0 = tile * WAVE * LPL + lane * LPL;
    const bool valid = b0 < a.sh.ld;
    const long ld = a.sh.ld;
    double v[3][LPL];
    double al1[LPL], al2[LPL];
    bool on1[LPL], on2[LPL];
    bool any = false;
#pragma unroll
    for (int t = 0; t < LPL; ++t) {
        v[0][t] = 0.0; v[1][t] = 0.0; v[2][t] = 0.0;
        const bool sel = valid && a.sel[b0 + t] != 0;
        on1[t] = sel && a.act[b0 + t];
        on2[t] = sel && a.act[ld + b0 + t];
        any = any || on1[t] || on2[t];
        al1[t] = on1[t] ? a.alpha[b0 + t] : 0.0;
        al2[t] = on2[t] ? a.alpha[ld + b0 + t] : 0.0;
    }
    if (any) {
        const int r0 = (rb * WPB + wave) * a.sh.rpw;
        for (int r = 0; r < a.sh.rpw; ++r) {
            const int row = uniform(r0 + r);
            if (row >= a.sh.rows) break;
            const long i = row * ld + b0;
            const double drow = a.dr ? a.dr[row] : 1.0;
            // the row's streamed operands are loaded before the gathers so
            // that their memory round trip overlaps the gathers'
            double on[LPL], iv[LPL], w1v[LPL], r1v[LPL], w2v[LPL], r2v[LPL];
            double2 pv[LPL];
#pragma unroll
            for (int t = 0; t < LPL; ++t) {
                on[t] = 0.0; iv[t] = 0.0; w1v[t] = 0.0; r1v[t] = 0.0; w2v[t] = 0.0; r2v[t] = 0.0;
                pv[t] = make_double2(0.0, 0.0);
                if (!(on1[t] || on2[t])) continue;
                // the row mask is the preconditioner's support: a row off it is
                // inactive or has no free column (its R A^ q is 0 either way)
                iv[t] = (double)a.invdiag[i + t];
                on[t] = iv[t] != 0.0 ? 1.0 : 0.0;
                pv[t] = a.p[i + t];
                if (on1[t]) { w1v[t] = a.w1[i + t]; r1v[t] = a.r1[i + t]; }
                if (on2[t]) { w2v[t] = a.w2[i + t]; r2v[t] = a.r2[i + t]; }
            }
            double t1[LPL], t2[LPL];
            row_dot2<LPL, D>(a.A.ptr, a.A.idx, a.A.slot, a.A.val, a.A.var, a.q, ld, b0, row, t1, t2);
#pragma unroll
            for (int t = 0; t < LPL; ++t) {
                if (!(on1[t] || on2[t])) continue;
                if (on1[t]) {
                    a.w1[i + t] = w1v[t] + al1[t] * pv[t].x;
                    const double n1 = r1v[t] - al1[t] * (on[t] * t1[t]);
                    a.r1[i + t] = n1;
                    v[0][t] += n1 * n1 * iv[t];
                    const double u = a.dr ? n1 / drow : n1;
                    if (a.inf_norm) v[2][t] = fmax(v[2][t], u * u);
                    else v[2][t] += u * u;
                }
                if (on2[t]) {
                    a.w2[i + t] = w2v[t] + al2[t] * pv[t].y;
                    const double n2 = r2v[t] - al2[t] * (on[t] * t2[t]);
                    a.r2[i + t] = n2;
                    v[1][t] += n2 * n2 * iv[t];
                }
            }
        }
    }
    const int ops[3] = {RSUM, RSUM, a.inf_norm ? RMAX : RSUM};
    block_partials<3, LPL, WPB>(v, ops, a.partial, rb, ld, b0, lds, valid);
}

// Per-LP CG scalars (Jacobi-preconditioned CG, rho = r . D^-1 r), two systems
// (0: primal correction, 1: dual least squares).  mode 0: initialise rho and
// rho0 from the reduced setup dots and mark active systems; mode 1: alpha =
// rho / (p . M p); mode 2: beta = rho_new / rho, and a system stops (stays
// inactive) once rho <= rel * rho0.  act[s*ld + b] != 0 while system s of LP b
// still iterates; the kernels skip LPs whose two systems have stopped.
struct CgScalarArgs {
    const double* d0;       // [ld] primal-system dot (rho or p.Mp)
    const double* d1;       // [ld] dual-system dot
    const double* plain;    // [ld] |r_p|^2 (modes 0 and 2)
    const double* ptol2;    // [ld] primal stop: |r_p|^2 <= ptol2[b]
    double* rho;            // [2][ld]
    double* rho0;           // [2][ld]
    double* alpha;          // [2][ld]
    double* beta;           // [2][ld]
    int* act;               // [2][ld]
    const double* floor2;   // [2][ld] absolute floor of rho (rounding level)
    const int* sel;
    double rel;
    int batch, ld, mode;
};

__global__ void k_cg_scalars(CgScalarArgs a) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= a.batch) return;
    if (!a.sel[b]) { a.act[b] = 0; a.act[a.ld + b] = 0; return; }
    for (int s = 0; s < 2; ++s) {
        const long i = (long)s * a.ld + b;
        const double v = s == 0 ? a.d0[b] : a.d1[b];
        const bool small = s == 0 && a.plain && a.plain[b] <= a.ptol2[b];
        if (a.mode == 0) {
            a.rho[i] = v;
            a.rho0[i] = v;
            a.act[i] = (v > 0.0 && isfinite(v) && !small) ? 1 : 0;
            a.alpha[i] = 0.0;
            a.beta[i] = 0.0;
        } else if (!a.act[i]) {
            a.alpha[i] = 0.0;
            a.beta[i] = 0.0;
        } else if (a.mode == 1) {
            const double r = a.rho[i];
            a.alpha[i] = (v > 0.0 && isfinite(v) && r > 0.0) ? r / v : 0.0;
            if (a.alpha[i] == 0.0) a.act[i] = 0;
        } else {
            const double r = a.rho[i];
            a.beta[i] = (r > 0.0 && isfinite(v)) ? v / r : 0.0;
            a.rho[i] = v;
            if (!(v > a.rel * a.rho0[i]) || !(v > a.floor2[i]) || !isfinite(v) || small) a.act[i] = 0;
        }
    }
}

// p = r + beta p (two systems).
struct CgDirArgs {
    Shape sh;               // rows = m; one row per wave, 64 LPs per wave
    double2* p;             // interleaved (p1, p2)
    const double* r1; const double* r2;
    const double* beta;
    const float* invdiag;
    const int* sel;
    const int* act;
    int reverse;            // walk the LP tiles last to first
};

// One row per wave and 64 LPs per wave, so each wave streams whole row
// segments (1 KB of p, 512 bytes of the others).
__global__ void __launch_bounds__(BLOCK) k_cg_direction(CgDirArgs a) {
    int tile, rb;
    if (!map_block(a.sh.n_real, a.sh.n_rb, tile, rb)) return;
    if (a.reverse) tile = a.sh.n_tiles - 1 - tile;
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int b = tile * WAVE + lane;
    if (b >= a.sh.ld || !a.sel[b]) return;
    const long ld = a.sh.ld;
    const bool on1 = a.act[b], on2 = a.act[ld + b];
    if (!on1 && !on2) return;
    const double be1 = a.beta[b], be2 = a.beta[ld + b];
    const int r0 = (rb * WAVES_PER_BLOCK + wave) * a.sh.rpw;
    for (int r = 0; r < a.sh.rpw; ++r) {
        const int row = r0 + r;
        if (row >= a.sh.rows) break;
        const long i = row * ld + b;
        const double iv = (double)a.invdiag[i];
        double2 pv = a.p[i];
        if (on1) pv.x = a.r1[i] * iv + be1 * pv.x;
        if (on2) pv.y = a.r2[i] * iv + be2 * pv.y;
        a.p[i] = pv;
    }
}

// ---------------------------------------------------------------------------
// Primal-only CG passes.  The dual least-squares system stops several steps
// before the primal correction (in steady rollout steps: after about 25 of 40
// steps, for every LP); from then on the interleaved (p1, p2) / (q1, q2)
// vectors would move twice the bytes the primal system needs.  Once no LP
// has an active dual system, p1 is copied out of the interleaved p
// (k_cg_split) and the remaining steps run these kernels on the contiguous
// p1, q1.  Each computes the primal half of its two-system counterpart with
// the same operations in the same order (row_dot accumulates exactly as
// row_dot2 does per system), so the primal answer does not change; the dual
// partials are written as 0.
// ---------------------------------------------------------------------------
__global__ void k_cg_split(const double2* __restrict__ p, double* __restrict__ p1, long count) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p1[i] = p[i].x;
}

struct CgColsPArgs {
    Shape sh;
    Op<double> AT;           // scaled FP64
    const double* p1;        // [m][ld]
    double* q1;              // [n][ld]
    const uint8_t* fmask;
    const int* sel;
    const int* act;          // [2][ld] (system 0 read)
    double* partial;         // [n_rb][2][ld]: |q1|^2, 0
};

template <int LPL, int D, int WPB>
__global__ void __launch_bounds__(WPB * WAVE) k_cg_cols_p(CgColsPArgs a) {
    __shared__ double lds[WPB * 2 * WAVE * LPL];
    int tile, rb;
    if (!map_block(a.sh.n_real, a.sh.n_rb, tile, rb)) return;
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int b0 = tile * WAVE * LPL + lane * LPL;
    const bool valid = b0 < a.sh.ld;
    const long ld = a.sh.ld;
    double v[2][LPL];
    bool on[LPL];
    bool any = false;
#pragma unroll
    for (int t = 0; t < LPL; ++t) {
        v[0][t] = 0.0; v[1][t] = 0.0;
        on[t] = valid && a.sel[b0 + t] != 0 && a.act[b0 + t];
        any = any || on[t];
    }
    if (any) {
        const int r0 = (rb * WPB + wave) * a.sh.rpw;
        for (int r = 0; r < a.sh.rpw; ++r) {
            const int col = uniform(r0 + r);
            if (col >= a.sh.rows) break;
            const long i = col * ld + b0;
            double f[LPL];
#pragma unroll
            for (int t = 0; t < LPL; ++t) f[t] = on[t] && a.fmask[i + t] ? 1.0 : 0.0;
            double s1[LPL];
            row_dot<double, double, double, LPL, D>(a.AT.ptr, a.AT.idx, a.AT.slot, a.AT.val, a.AT.var, a.p1, ld, b0,
                                                    col, s1);
#pragma unroll
            for (int t = 0; t < LPL; ++t) {
                if (!on[t]) continue;
                const double q = f[t] * s1[t];
                v[0][t] += q * q;
                a.q1[i + t] = q;
            }
        }
    }
    const int ops[2] = {RSUM, RSUM};
    block_partials<2, LPL, WPB>(v, ops, a.partial, rb, ld, b0, lds, valid);
}

struct CgRowsPArgs {
    Shape sh;
    Op<double> A;            // scaled FP64
    const double* q1;        // [n][ld]
    const double* p1;        // [m][ld]
    double* w1; double* r1;  // [m][ld]
    const double* alpha;     // [2][ld] (system 0 read)
    const float* invdiag;
    const double* dr;        // row scale for the primal stop measure, or nullptr
    const int* sel;
    const int* act;
    double* partial;         // [n_rb][3][ld]: r_p.D^-1 r_p, 0, primal stop measure
    int inf_norm;
    int reverse;
};

template <int LPL, int D, int WPB>
__global__ void __launch_bounds__(WPB * WAVE) k_cg_rows_p(CgRowsPArgs a) {
    __shared__ double lds[WPB * 3 * WAVE * LPL];
    int tile, rb;
    if (!map_block(a.sh.n_real, a.sh.n_rb, tile, rb)) return;
    if (a.reverse) tile = a.sh.n_tiles - 1 - tile;
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int b0 = tile * WAVE * LPL + lane * LPL;
    const bool valid = b0 < a.sh.ld;
    const long ld = a.sh.ld;
    double v[3][LPL];
    double al[LPL];
    bool on1[LPL];
    bool any = false;
#pragma unroll
    for (int t = 0; t < LPL; ++t) {
        v[0][t] = 0.0; v[1][t] = 0.0; v[2][t] = 0.0;
        on1[t] = valid && a.sel[b0 + t] != 0 && a.act[b0 + t];
        al[t] = on1[t] ? a.alpha[b0 + t] : 0.0;
        any = any || on1[t];
    }
    if (any) {
        const int r0 = (rb * WPB + wave) * a.sh.rpw;
        for (int r = 0; r < a.sh.rpw; ++r) {
            const int row = uniform(r0 + r);
            if (row >= a.sh.rows) break;
            const long i = row * ld + b0;
            const double drow = a.dr ? a.dr[row] : 1.0;
            // the row's streamed operands are loaded before the gathers
            double iv[LPL], pv[LPL], wv[LPL], rv[LPL];
#pragma unroll
            for (int t = 0; t < LPL; ++t) {
                iv[t] = 0.0; pv[t] = 0.0; wv[t] = 0.0; rv[t] = 0.0;
                if (!on1[t]) continue;
                iv[t] = (double)a.invdiag[i + t];
                pv[t] = a.p1[i + t]; wv[t] = a.w1[i + t]; rv[t] = a.r1[i + t];
            }
            double t1[LPL];
            row_dot<double, double, double, LPL, D>(a.A.ptr, a.A.idx, a.A.slot, a.A.val, a.A.var, a.q1, ld, b0, row,
                                                    t1);
#pragma unroll
            for (int t = 0; t < LPL; ++t) {
                if (!on1[t]) continue;
                const double on = iv[t] != 0.0 ? 1.0 : 0.0;
                a.w1[i + t] = wv[t] + al[t] * pv[t];
                const double n1 = rv[t] - al[t] * (on * t1[t]);
                a.r1[i + t] = n1;
                v[0][t] += n1 * n1 * iv[t];
                const double u = a.dr ? n1 / drow : n1;
                if (a.inf_norm) v[2][t] = fmax(v[2][t], u * u);
                else v[2][t] += u * u;
            }
        }
    }
    const int ops[3] = {RSUM, RSUM, a.inf_norm ? RMAX : RSUM};
    block_partials<3, LPL, WPB>(v, ops, a.partial, rb, ld, b0, lds, valid);
}

struct CgDirPArgs {
    Shape sh;               // rows = m; one row per wave, 64 LPs per wave
    double* p1;
    const double* r1;
    const double* beta;
    const float* invdiag;
    const int* sel;
    const int* act;
    int reverse;
};

__global__ void __launch_bounds__(BLOCK) k_cg_direction_p(CgDirPArgs a) {
    int tile, rb;
    if (!map_block(a.sh.n_real, a.sh.n_rb, tile, rb)) return;
    if (a.reverse) tile = a.sh.n_tiles - 1 - tile;
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int b = tile * WAVE + lane;
    if (b >= a.sh.ld || !a.sel[b] || !a.act[b]) return;
    const long ld = a.sh.ld;
    const double be = a.beta[b];
    const int r0 = (rb * WAVES_PER_BLOCK + wave) * a.sh.rpw;
    for (int r = 0; r < a.sh.rpw; ++r) {
        const int row = r0 + r;
        if (row >= a.sh.rows) break;
        const long i = row * ld + b;
        a.p1[i] = a.r1[i] * (double)a.invdiag[i] + be * a.p1[i];
    }
}

// x = clip(xs + Dc * fmask * A^T w_p, l, u) (original units).
struct PolishFinalColsArgs {
    Shape sh;
    Op<double> AT;          // scaled FP64
    const double* wp;
    const double* xs;
    const double* dc;
    const double* lower; const double* upper;
    const uint8_t* fmask;
    const int* sel;
    double* xout;
};

__global__ void __launch_bounds__(BLOCK) k_polish_final_cols(PolishFinalColsArgs a) {
    int tile, rb;
    if (!map_block(a.sh.n_real, a.sh.n_rb, tile, rb)) return;
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int b0 = tile * WAVE + lane;
    if (b0 >= a.sh.ld || !a.sel[b0]) return;
    const long ld = a.sh.ld;
    const int r0 = (rb * WAVES_PER_BLOCK + wave) * a.sh.rpw;
    for (int r = 0; r < a.sh.rpw; ++r) {
        const int col = uniform(r0 + r);
        if (col >= a.sh.rows) break;
        const long i = col * ld + b0;
        double xv = a.xs[i];
        if (a.fmask[i]) {
            // the bounds are loaded before the gathers (overlapping round trips)
            const double lo = a.lower[i], hi = a.upper[i];
            double s[1];
            row_dot<double, double, double, 1>(a.AT.ptr, a.AT.idx, a.AT.slot, a.AT.val, a.AT.var, a.wp, ld, b0, col, s);
            xv = xv + a.dc[col] * s[0];
            xv = fmin(fmax(xv, lo), hi);
        }
        a.xout[i] = xv;
    }
}

// y = rmask * Dr * w_d, inequality rows clipped to <= 0.
struct PolishFinalRowsArgs {
    int rows, ld, rows_per_block, neq;
    const double* wd;
    const double* dr;
    const uint8_t* rmask;
    const int* sel;
    double* yout;
};

__global__ void k_polish_final_rows(PolishFinalRowsArgs a) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= a.ld || !a.sel[b]) return;
    const long ld = a.ld;
    const int r0 = blockIdx.y * a.rows_per_block;
    const int r1 = min(a.rows, r0 + a.rows_per_block);
    for (int r = r0; r < r1; ++r) {
        const long i = r * ld + b;
        double yv = a.rmask[i] ? a.dr[r] * a.wd[i] : 0.0;
        if (r >= a.neq) yv = fmin(yv, 0.0);
        a.yout[i] = yv;
    }
}

// ---------------------------------------------------------------------------
// Layout conversion between the caller's batch-major [B][len] FP64 arrays and
// the internal batch-minor [len][ld] arrays, with optional diagonal scaling:
// mode 0: copy, 1: multiply by scale[j], 2: divide by scale[j].
// ---------------------------------------------------------------------------
template <typename TO>
__global__ void __launch_bounds__(256) k_to_internal(const double* __restrict__ in, TO* __restrict__ out,
                                                     int len, int batch, int ld, const double* __restrict__ scale,
                                                     int mode, double fill) {
    __shared__ double tile[64][65];
    const int j0 = blockIdx.x * 64, b0 = blockIdx.y * 64;
    const int tx = threadIdx.x % 64, ty = threadIdx.x / 64;
    for (int k = ty; k < 64; k += 4) {
        const int b = b0 + k, j = j0 + tx;
        double v = fill;
        if (b < batch && j < len) v = in[(long)b * len + j];
        tile[k][tx] = v;
    }
    __syncthreads();
    for (int k = ty; k < 64; k += 4) {
        const int j = j0 + k, b = b0 + tx;
        if (j < len && b < ld) {
            double v = tile[tx][k];
            if (b < batch) {
                if (mode == 1) v = v * scale[j];
                else if (mode == 2) v = v / scale[j];
            }
            out[(long)j * ld + b] = (TO)v;
        }
    }
}

template <typename TI>
__global__ void __launch_bounds__(256) k_to_external(const TI* __restrict__ in, double* __restrict__ out,
                                                     int len, int batch, int ld, const double* __restrict__ scale,
                                                     int mode) {
    __shared__ double tile[64][65];
    const int j0 = blockIdx.x * 64, b0 = blockIdx.y * 64;
    const int tx = threadIdx.x % 64, ty = threadIdx.x / 64;
    for (int k = ty; k < 64; k += 4) {
        const int j = j0 + k, b = b0 + tx;
        double v = 0.0;
        if (j < len && b < ld) v = (double)in[(long)j * ld + b];
        tile[k][tx] = v;
    }
    __syncthreads();
    for (int k = ty; k < 64; k += 4) {
        const int b = b0 + k, j = j0 + tx;
        if (b < batch && j < len) {
            double v = tile[tx][k];
            if (mode == 1) v = v * scale[j];
            else if (mode == 2) v = v / scale[j];
            out[(long)b * len + j] = v;
        }
    }
}

// Elementwise over [rows][ld]: dst = src * scale[row] (mode 1) or copy, per
// selected LP (sel == nullptr: all).  TI -> double.
template <typename TI>
__global__ void k_scale_copy(const TI* __restrict__ src, double* __restrict__ dst, int rows, int ld,
                             int rows_per_block, const double* __restrict__ scale, const int* __restrict__ sel,
                             int sel_value) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= ld) return;
    if (sel && sel[b] != sel_value) return;
    const long l = ld;
    const int r0 = blockIdx.y * rows_per_block;
    const int r1 = min(rows, r0 + rows_per_block);
    for (int r = r0; r < r1; ++r) {
        double v = (double)src[r * l + b];
        if (scale) v = v * scale[r];
        dst[r * l + b] = v;
    }
}

// Generic batched SpMV: out[r][b] = sum a_rk in[idx][b].
template <typename TV, typename TA>
struct SpmvArgs {
    Shape sh;
    Op<TV> A;
    const TA* x;
    TA* y;
};

template <typename TV, typename TA>
__global__ void __launch_bounds__(BLOCK) k_spmv(SpmvArgs<TV, TA> a) {
    int tile, rb;
    if (!map_block(a.sh.n_real, a.sh.n_rb, tile, rb)) return;
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int b0 = tile * WAVE + lane;
    if (b0 >= a.sh.ld) return;
    const long ld = a.sh.ld;
    const int r0 = (rb * WAVES_PER_BLOCK + wave) * a.sh.rpw;
    for (int r = 0; r < a.sh.rpw; ++r) {
        const int row = uniform(r0 + r);
        if (row >= a.sh.rows) break;
        TA s[1];
        row_dot<TV, TA, TA, 1>(a.A.ptr, a.A.idx, a.A.slot, a.A.val, a.A.var, a.x, ld, b0, row, s);
        a.y[row * ld + b0] = s[0];
    }
}

// Input validation of the caller's batch-major arrays (src/gpu_pdhg_workspace.py:185-207,
// :230-244): counts violations into err[0..3].
__global__ void k_validate(const double* __restrict__ c, long nc, const double* __restrict__ rhs, long nr,
                           const double* __restrict__ lower, const double* __restrict__ upper, long nb,
                           const double* __restrict__ var, long nv, int* err) {
    const long stride = (long)gridDim.x * blockDim.x;
    const long t0 = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int e_c = 0, e_r = 0, e_b = 0, e_v = 0;
    for (long i = t0; i < nc; i += stride) e_c += !isfinite(c[i]);
    for (long i = t0; i < nr; i += stride) e_r += !isfinite(rhs[i]);
    for (long i = t0; i < nb; i += stride) {
        const double lo = lower[i], hi = upper[i];
        e_b += (isnan(lo) || isnan(hi) || lo > hi || lo == INFINITY || hi == -INFINITY);
    }
    for (long i = t0; i < nv; i += stride) e_v += !isfinite(var[i]);
    if (e_c) atomicAdd(err + 0, e_c);
    if (e_r) atomicAdd(err + 1, e_r);
    if (e_b) atomicAdd(err + 2, e_b);
    if (e_v) atomicAdd(err + 3, e_v);
}

// Caller-supplied diagonal metric: err[0] / err[1] count tau / sigma entries
// that are not finite and strictly positive.
__global__ void k_validate_metric(const double* __restrict__ tau, long nt, const double* __restrict__ sigma, long ns,
                                  int* err) {
    const long stride = (long)gridDim.x * blockDim.x;
    const long t0 = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int e_t = 0, e_s = 0;
    for (long i = t0; i < nt; i += stride) e_t += !(tau[i] > 0.0 && isfinite(tau[i]));
    for (long i = t0; i < ns; i += stride) e_s += !(sigma[i] > 0.0 && isfinite(sigma[i]));
    if (e_t) atomicAdd(err + 0, e_t);
    if (e_s) atomicAdd(err + 1, e_s);
}

// Per-LP values of the variable coefficients, scaled: out[s][b] = in * dr[row_s] * dc[col_s].
template <typename TO>
__global__ void k_var_scale(const double* __restrict__ in, TO* __restrict__ out, int nvar, int ld,
                            const int* __restrict__ vrow, const int* __restrict__ vcol,
                            const double* __restrict__ dr, const double* __restrict__ dc) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    if (b >= ld || s >= nvar) return;
    const double v = in[(long)s * ld + b];
    out[(long)s * ld + b] = (TO)(v * dr[vrow[s]] * dc[vcol[s]]);
}

// Shared PDHG metric from the batch's largest per-LP coefficient magnitudes
// (after k_var_absmax): sigma_i = eta / sum_k |a^_ik| over row i, tau_j =
// eta / sum_k |a^_kj| over column j, an entry with a per-LP value counting
// var_absmax[slot]; 1 where the sum is 0.  Rows sum in CSR order, columns in
// CSC order (ascending row), each term |a_k dr_i dc_j| formed as on the host,
// so the values equal the host restatement's bit for bit.  Thread t < m: row
// t; m <= t < m + n: column t - m.
__global__ void k_shared_metric(const int* __restrict__ rp, const int* __restrict__ ci, const int* __restrict__ rslot,
                                const double* __restrict__ val, const int* __restrict__ cp,
                                const int* __restrict__ ri, const int* __restrict__ cslot,
                                const double* __restrict__ tval, const double* __restrict__ dr,
                                const double* __restrict__ dc, const double* __restrict__ var_absmax, double eta,
                                int m, int n, double* __restrict__ sig, double* __restrict__ tau,
                                float* __restrict__ sig_f, float* __restrict__ tau_f) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < m) {
        double rs = 0.0;
        for (int k = rp[t]; k < rp[t + 1]; ++k) {
            const int sl = rslot[k];
            rs += sl < 0 ? fabs(val[k] * dr[t] * dc[ci[k]]) : var_absmax[sl];
        }
        const double v = rs > 0.0 ? eta / rs : 1.0;
        sig[t] = v;
        sig_f[t] = (float)v;
    } else if (t < m + n) {
        const int j = t - m;
        double cs = 0.0;
        for (int k = cp[j]; k < cp[j + 1]; ++k) {
            const int sl = cslot[k];
            cs += sl < 0 ? fabs(tval[k] * dr[ri[k]] * dc[j]) : var_absmax[sl];
        }
        const double v = cs > 0.0 ? eta / cs : 1.0;
        tau[j] = v;
        tau_f[j] = (float)v;
    }
}

// out[s] = max over the real LPs b < batch of |var[s][b]| (one block per slot;
// the maximum does not depend on the order).
__global__ void __launch_bounds__(256) k_var_absmax(const double* __restrict__ var, double* __restrict__ out,
                                                    int batch, int ld) {
    __shared__ double part[256];
    const int s = blockIdx.x;
    double mx = 0.0;
    for (int b = threadIdx.x; b < batch; b += blockDim.x) mx = fmax(mx, fabs(var[(long)s * ld + b]));
    part[threadIdx.x] = mx;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] = fmax(part[threadIdx.x], part[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[s] = part[0];
}

// Shared scaled objective: cs[j] = c[j] * dc[j] (and its FP32 copy when cf).
__global__ void k_scale_shared(const double* __restrict__ c, const double* __restrict__ dc, double* __restrict__ cs,
                               float* __restrict__ cf, int n) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const double v = c[j] * dc[j];
    cs[j] = v;
    if (cf) cf[j] = (float)v;
}

// Fixed-metric per-LP diagonal metric (src/gpu_pdhg_corrector.py:185-196):
// out[r][b] = (sum_k |a_rk| > 0 ? eta / sum : 1) * mul   (mul = w for sigma,
// 1/w for tau, applied as in the reference: base * w or base / w).
struct RefMetricArgs {
    Shape sh;
    Op<double> A;
    double* out;
    double eta;
    double weight;
    int divide;      // 0: base * weight, 1: base / weight
};

__global__ void __launch_bounds__(BLOCK) k_ref_metric(RefMetricArgs a) {
    int tile, rb;
    if (!map_block(a.sh.n_real, a.sh.n_rb, tile, rb)) return;
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int b0 = tile * WAVE + lane;
    if (b0 >= a.sh.ld) return;
    const long ld = a.sh.ld;
    const int r0 = (rb * WAVES_PER_BLOCK + wave) * a.sh.rpw;
    for (int r = 0; r < a.sh.rpw; ++r) {
        const int row = uniform(r0 + r);
        if (row >= a.sh.rows) break;
        double s = 0.0;
        const int k0 = sload(a.A.ptr, row), k1 = sload(a.A.ptr, row + 1);
        for (int k = k0; k < k1; ++k) {
            const int sl = sload(a.A.slot, k);
            const double v = sl < 0 ? sload(a.A.val, k) : a.A.var[(long)sl * ld + b0];
            s = s + __builtin_fabs(v);
        }
        const double base = s > 0.0 ? a.eta / s : 1.0;
        a.out[row * ld + b0] = a.divide ? base / a.weight : base * a.weight;
    }
}

// Per-LP 2-norm of an internal [rows][ld] FP64 vector (or of a shared [rows]
// vector when shared != 0).
// Per-LP Euclidean norms of a [rows][ld] (or shared [rows]) vector in two
// launches: squared sums of NORM_CHUNK-row chunks (grid: LP tiles x chunks),
// then the chunks summed in order per LP (deterministic).
constexpr int NORM_CHUNK = 64;

// mask/cap: entries whose |mask| is at least cap are left out (the primal
// weight's right-hand-side norm leaves out rows with unlimited supply).
__global__ void k_norms_part(const double* __restrict__ v, int rows, int ld, int shared, double* __restrict__ part,
                             const double* __restrict__ mask = nullptr, double cap = 0.0) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= ld) return;
    const int r0 = blockIdx.y * NORM_CHUNK, r1 = min(rows, r0 + NORM_CHUNK);
    double s = 0.0;
    for (int r = r0; r < r1; ++r) {
        const long i = shared ? r : (long)r * ld + b;
        if (mask && !(fabs(mask[i]) < cap)) continue;
        const double x = v[i];
        s += x * x;
    }
    part[(long)blockIdx.y * ld + b] = s;
}

__global__ void k_norms_final(const double* __restrict__ part, int chunks, int ld, double* __restrict__ out) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= ld) return;
    double s = 0.0;
    for (int k = 0; k < chunks; ++k) s += part[(long)k * ld + b];
    out[b] = sqrt(s);
}

// Project a starting point: x = clip(x, l, u) (scaled), y <= 0 on inequality rows.
template <typename T>
__global__ void k_project(T* __restrict__ x, T* __restrict__ xbar, T* __restrict__ xlast, T* __restrict__ xsum,
                          const T* __restrict__ lower, const T* __restrict__ upper, int rows, int ld,
                          int rows_per_block, int is_dual, int neq) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= ld) return;
    const long l = ld;
    const int r0 = blockIdx.y * rows_per_block;
    const int r1 = min(rows, r0 + rows_per_block);
    for (int r = r0; r < r1; ++r) {
        const long i = r * l + b;
        T v = x[i];
        if (is_dual) {
            if (r >= neq) v = __builtin_fmin(v, T(0));
        } else {
            v = __builtin_fmin(__builtin_fmax(v, lower[i]), upper[i]);
        }
        x[i] = v;
        if (xbar) xbar[i] = v;
        if (xlast) xlast[i] = v;
        if (xsum) xsum[i] = T(0);
    }
}

// Masked copy: dst[r][b] = src[r][b] for LPs with mask[b] == value.
template <typename T>
__global__ void k_masked_copy_t(const T* __restrict__ src, T* __restrict__ dst, int rows, int ld,
                                int rows_per_block, const int* __restrict__ mask, int value) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= ld || mask[b] != value) return;
    const long l = ld;
    const int r0 = blockIdx.y * rows_per_block;
    const int r1 = min(rows, r0 + rows_per_block);
    for (int r = r0; r < r1; ++r) dst[r * l + b] = src[r * l + b];
}

__global__ void k_masked_copy(const double* __restrict__ src, double* __restrict__ dst, int rows, int ld,
                              int rows_per_block, const int* __restrict__ mask, int value) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= ld || mask[b] != value) return;
    const long l = ld;
    const int r0 = blockIdx.y * rows_per_block;
    const int r1 = min(rows, r0 + rows_per_block);
    for (int r = r0; r < r1; ++r) dst[r * l + b] = src[r * l + b];
}

// dst[r][j] = src[r][perm[j]] (perm[j] >= 0), 0 otherwise.
template <typename T>
__global__ void k_gather_cols(const T* __restrict__ src, T* __restrict__ dst, int rows, int ld_src, int ld_dst,
                              const int* __restrict__ perm, int rows_per_block) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ld_dst) return;
    const int p = perm[j];
    const int r0 = blockIdx.y * rows_per_block;
    const int r1 = min(rows, r0 + rows_per_block);
    for (int r = r0; r < r1; ++r) dst[(long)r * ld_dst + j] = p >= 0 ? src[(long)r * ld_src + p] : T(0);
}

// dst[r][perm[j]] = src[r][j] for compacted LPs j with mask[j] != 0.
template <typename T>
__global__ void k_scatter_cols(const T* __restrict__ src, T* __restrict__ dst, int rows, int ld_src, int ld_dst,
                               const int* __restrict__ perm, const int* __restrict__ mask, int rows_per_block) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ld_src || perm[j] < 0 || !mask[j]) return;
    const int p = perm[j];
    const int r0 = blockIdx.y * rows_per_block;
    const int r1 = min(rows, r0 + rows_per_block);
    for (int r = r0; r < r1; ++r) dst[(long)r * ld_dst + p] = src[(long)r * ld_src + j];
}

template <typename TI, typename TO>
__global__ void k_convert(const TI* __restrict__ in, TO* __restrict__ out, long count) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) out[i] = (TO)in[i];
}

// CG stopping floors: rho = r.D^-1 r below (1e-13 (1 + |b^|))^2 (primal,
// rounding level) or (f (1 + |c^|))^2 (dual least squares: reduced costs
// small enough for the 1e-7 relative gap gate on nearly every LP; the few
// that miss it by little are refined by the polish's second round, whose
// dual floor is 100x lower).  f = LPE_CG_FLOOR_D times the caller's dscale.
// perm maps the view's columns to LPs (nullptr: identity).
#ifndef LPE_CG_FLOOR_P
#define LPE_CG_FLOOR_P 1e-13
#endif
#ifndef LPE_CG_FLOOR_D
#define LPE_CG_FLOOR_D 1e-11
#endif
// floor2[0, ld): the primal system's rho floor, [ld, 2 ld): the dual's,
// [2 ld, 3 ld): the primal stop threshold ptol2 * pscale2.  dscale multiplies
// the dual floor.
__global__ void k_cg_floor(const double* bnorm, const double* cnorm, const int* perm, double* floor2, int ld,
                           double dscale, double ptol2, double pscale2) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ld) return;
    const int b = perm ? perm[j] : j;
    const double bn = b >= 0 ? bnorm[b] : 0.0, cn = b >= 0 ? cnorm[b] : 0.0;
    const double fp = LPE_CG_FLOOR_P * (1.0 + bn), fd = dscale * LPE_CG_FLOOR_D * (1.0 + cn);
    floor2[j] = fp * fp;
    floor2[ld + j] = fd * fd;
    floor2[2 * ld + j] = ptol2 * pscale2;
}

// Per-LP primal weight: the given value, or ||c^|| / ||b^|| (the PDLP default
// for a cold start) when `given` is not positive and both norms are nonzero.
__global__ void k_init_omega(LpState* st, double* om, float* om_f, double* iom, float* iom_f,
                             const double* bnorm, const double* cnorm, const double* given, double fallback,
                             int batch) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= batch) return;
    double w = given ? given[b] : fallback;
    if (!(w > 0.0) || !isfinite(w)) {
        const double bn = bnorm[b], cn = cnorm[b];
        w = (bn > 0.0 && cn > 0.0 && isfinite(bn) && isfinite(cn)) ? cn / bn : 1.0;
        w = fmin(fmax(w, 1e-8), 1e8);
    }
    st[b].omega = w;
    om[b] = w; om_f[b] = (float)w; iom[b] = 1.0 / w; iom_f[b] = (float)(1.0 / w);
}

// Halpern: the restart reference of an LP that restarted at this check is the
// fixed-point residual at its new point z0 = T(z) (measured by a second probe
// and check pass): |z0 - T(z0)|_w.
__global__ void k_hal_restart_ref(LpState* st, const double* rows, const double* cols, int batch, int ld) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= batch || st[b].action != ACT_RESTART_AVG) return;
    const double om = st[b].omega;
    st[b].kkt_last = sqrt(om * cols[(long)CC_FPX * ld + b] + rows[(long)CR_FPY * ld + b] / om);
}

// Halpern: the iterations since the last restart advance by the launch
// group's length (the kernels of the group read cnt + j).
__global__ void k_hal_advance(int* cnt, const int* flags, int count, int batch) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < batch && !flags[b]) cnt[b] += count;
}

// Mark polished LPs optimal and frozen.
__global__ void k_accept_polished(LpState* st, int* flags, int* source, const int* sel, int ld) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= ld || !sel[b]) return;
    st[b].status = 1;
    st[b].source = 1;
    st[b].accepted_iter = st[b].iters;
    flags[b] = 1;
    source[b] = 1;
}

}  // namespace lpe
