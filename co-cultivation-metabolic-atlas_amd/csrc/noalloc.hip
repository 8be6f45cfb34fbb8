// FP64 certificate of a caller-owned LP batch with no allocation at all
// (lpe_certify_noalloc in include/lpe.h).
//
// The handle-based lpe_certify keeps its block partials in buffers the
// library allocates at lpe_create.  This entry point takes the problem as
// device pointers and keeps its partials in one file-scope device array,
// `scratch`, whose address is looked up once per device with
// hipGetSymbolAddress and cached; the host never passes the array's name to
// a kernel.  Every call on a device shares that device's array: calls that
// overlap on one GPU (two streams, two threads) race on it.  Each call writes
// every partial it later reads, so nothing has to be cleared between calls.
//
// Layout: batch-major, as every other entry point of the C ABI (x [B][n],
// y [B][m], rhs [B][m], lower/upper [B][n], metrics [B][8]).  A block of 256
// threads covers 1024 rows (or columns) of one LP; the partials of LP b sit
// at scratch[b * per_lp ...], rows first, and are reduced in chunk order, so
// the result is deterministic.  Each row product sums the CSR entries in
// storage order and each reduced cost the transposed entries in row order -
// the order of the engine's certificate kernels - so the activities, reduced
// costs and every max-based metric equal lpe_certify's bit for bit; the sums
// (objective, gaps) differ from it only in summation order.
#include "lpe.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>

namespace lpe {
int set_error(int code, const std::string& msg);
}  // namespace lpe

namespace {

constexpr int NA_THREADS = 256;
constexpr int NA_PER_THREAD = 4;
constexpr int NA_CHUNK = NA_THREADS * NA_PER_THREAD;   // rows (columns) per block
constexpr int NA_RN = 5;   // row values: |eq residual| max, ub residual max, row-gap sum, y_ub max, non-finite count
constexpr int NA_CN = 8;   // column values: l-x max, x-u max, r on free lower max, -r on free upper max,
                           // bound-gap sum, objective sum, non-finite x count, non-finite r count
constexpr long NA_SCRATCH_DOUBLES = 1L << 20;           // 8 MiB per device
constexpr int NA_MAX_DEVICES = 64;

static __device__ __align__(256) double scratch[NA_SCRATCH_DOUBLES];

std::mutex g_mu;
double* g_scratch[NA_MAX_DEVICES] = {};

// The current device's copy of `scratch` (looked up on the first call on
// each device, with that device current).
int device_scratch(double** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return lpe::set_error(LPE_EDEVICE, std::string("hipGetDevice failed: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= NA_MAX_DEVICES) return lpe::set_error(LPE_EDEVICE, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return lpe::set_error(LPE_EDEVICE, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
        g_scratch[dev] = static_cast<double*>(p);
    }
    *out = g_scratch[dev];
    return LPE_OK;
}

__device__ __forceinline__ double nmax(double acc, double o) { return (o > acc || __builtin_isnan(o)) ? o : acc; }

struct NaArgs {
    lpe_batch_problem p;
    const double* x;
    const double* y;
    double* part;      // device scratch
    long per_lp;       // partial doubles per LP
    int chunks_r, chunks_c;
};

// Fixed-order block reduction of NV values per thread (max or sum); thread 0
// writes them to out[0..NV).
template <int NV>
__device__ void block_reduce(double (&v)[NV], const bool (&is_max)[NV], double* out) {
    __shared__ double lds[NV][NA_THREADS];
    const int t = threadIdx.x;
#pragma unroll
    for (int k = 0; k < NV; ++k) lds[k][t] = v[k];
    __syncthreads();
    for (int s = NA_THREADS / 2; s > 0; s >>= 1) {
        if (t < s) {
#pragma unroll
            for (int k = 0; k < NV; ++k) lds[k][t] = is_max[k] ? nmax(lds[k][t], lds[k][t + s]) : lds[k][t] + lds[k][t + s];
        }
        __syncthreads();
    }
    if (t == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) out[k] = lds[k][0];
    }
}

// grid (B, chunks_r): activities A x, residual maxima, row gap, y_ub maximum.
__global__ void __launch_bounds__(NA_THREADS) k_na_rows(NaArgs a) {
    const lpe_batch_problem& p = a.p;
    const long b = blockIdx.x;
    const int chunk = blockIdx.y;
    const double* xb = a.x + b * p.n;
    const double* yb = a.y + b * p.m;
    const double* rb = p.rhs + b * p.m;
    const double* vals = p.values + (p.values_per_lp ? b * p.nnz : 0);
    double v[NA_RN] = {0.0, -INFINITY, 0.0, -INFINITY, 0.0};
    for (int r = 0; r < NA_PER_THREAD; ++r) {
        const int row = chunk * NA_CHUNK + r * NA_THREADS + threadIdx.x;
        if (row >= p.m) break;
        double act = 0.0;
        for (int k = p.row_ptr[row]; k < p.row_ptr[row + 1]; ++k) act = act + vals[k] * xb[p.col_idx[k]];
        const double err = act - rb[row];
        const double yv = yb[row];
        if (row < p.neq) {
            v[0] = nmax(v[0], __builtin_fabs(err));
        } else {
            v[1] = nmax(v[1], err);
            v[2] += __builtin_fabs(yv * (rb[row] - act));
            v[3] = nmax(v[3], yv);
        }
        if (!__builtin_isfinite(act) || !__builtin_isfinite(yv)) v[4] += 1.0;
    }
    const bool is_max[NA_RN] = {true, true, false, true, false};
    block_reduce<NA_RN>(v, is_max, a.part + b * a.per_lp + (long)chunk * NA_RN);
}

// grid (B, chunks_c): reduced costs c - A^T y, bound violations, dual
// violations on free bounds, bound gap, objective.
__global__ void __launch_bounds__(NA_THREADS) k_na_cols(NaArgs a) {
    const lpe_batch_problem& p = a.p;
    const long b = blockIdx.x;
    const int chunk = blockIdx.y;
    const double* xb = a.x + b * p.n;
    const double* yb = p.m > 0 ? a.y + b * p.m : nullptr;
    const double* lob = p.lower + b * p.n;
    const double* upb = p.upper + b * p.n;
    const double* cb = p.c + (p.c_per_lp ? b * p.n : 0);
    const double* vals = p.values + (p.values_per_lp ? b * p.nnz : 0);
    double v[NA_CN] = {-INFINITY, -INFINITY, -INFINITY, -INFINITY, 0.0, 0.0, 0.0, 0.0};
    for (int r = 0; r < NA_PER_THREAD; ++r) {
        const int col = chunk * NA_CHUNK + r * NA_THREADS + threadIdx.x;
        if (col >= p.n) break;
        double aty = 0.0;
        if (yb)
            for (int k = p.col_ptr[col]; k < p.col_ptr[col + 1]; ++k) aty = aty + vals[p.t_pos[k]] * yb[p.row_idx[k]];
        const double cc = cb[col];
        const double xv = xb[col];
        const double lo = lob[col], hi = upb[col];
        const double rc = cc - aty;
        v[0] = nmax(v[0], lo - xv);
        v[1] = nmax(v[1], xv - hi);
        v[2] = nmax(v[2], __builtin_isfinite(lo) ? 0.0 : rc);
        v[3] = nmax(v[3], __builtin_isfinite(hi) ? 0.0 : -rc);
        const double tgt = rc >= 0.0 ? lo : hi;
        const double t2 = __builtin_isfinite(tgt) ? tgt : xv;
        v[4] += __builtin_fabs(rc * (xv - t2));
        v[5] += cc * xv;
        if (!__builtin_isfinite(xv)) v[6] += 1.0;
        if (!__builtin_isfinite(rc)) v[7] += 1.0;
    }
    const bool is_max[NA_CN] = {true, true, true, true, false, false, false, false};
    block_reduce<NA_CN>(v, is_max, a.part + b * a.per_lp + (long)a.chunks_r * NA_RN + (long)chunk * NA_CN);
}

// One thread per LP: chunk partials reduced in order, then the gates of
// lpe_certify (src/gpu_block_lp.py:52-132).
__global__ void k_na_final(NaArgs a, double* metrics, double eps_p, double eps_d, double eps_g) {
    const long b = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= a.p.batch) return;
    const double* pr = a.part + b * a.per_lp;
    const double* pc = pr + (long)a.chunks_r * NA_RN;
    double row_eq = 0.0, row_ub = -INFINITY, row_gap = 0.0, dual_rows = -INFINITY, nonfin_rows = 0.0;
    for (int k = 0; k < a.chunks_r; ++k) {
        const double* q = pr + (long)k * NA_RN;
        row_eq = nmax(row_eq, q[0]);
        row_ub = nmax(row_ub, q[1]);
        row_gap += q[2];
        dual_rows = nmax(dual_rows, q[3]);
        nonfin_rows += q[4];
    }
    if (row_ub == -INFINITY) row_ub = 0.0;          // empty inequality group
    if (dual_rows == -INFINITY) dual_rows = 0.0;
    double low = -INFINITY, up = -INFINITY, rlow = -INFINITY, rup = -INFINITY;
    double bound_gap = 0.0, obj = 0.0, nonfin_x = 0.0, nonfin_r = 0.0;
    for (int k = 0; k < a.chunks_c; ++k) {
        const double* q = pc + (long)k * NA_CN;
        low = nmax(low, q[0]);
        up = nmax(up, q[1]);
        rlow = nmax(rlow, q[2]);
        rup = nmax(rup, q[3]);
        bound_gap += q[4];
        obj += q[5];
        nonfin_x += q[6];
        nonfin_r += q[7];
    }
    const double primal = nmax(0.0, nmax(nmax(row_eq, row_ub), nmax(low, up)));
    const double dual = nmax(0.0, nmax(dual_rows, nmax(rlow, rup)));
    const double gap = (bound_gap + row_gap) / fmax(1.0, fabs(obj));
    const bool finite_primal = nonfin_x == 0.0 && isfinite(obj) && nonfin_rows == 0.0;
    const bool finite = finite_primal && nonfin_r == 0.0;
    const bool ok = finite && isfinite(gap) && primal <= eps_p && dual <= eps_d && gap <= eps_g;
    double* m = metrics + b * LPE_NMETRIC;
    m[LPE_M_PRIMAL] = primal; m[LPE_M_DUAL] = dual; m[LPE_M_GAP] = gap; m[LPE_M_OBJECTIVE] = obj;
    m[LPE_M_PASSED] = ok ? 1.0 : 0.0; m[LPE_M_FINITE] = finite ? 1.0 : 0.0;
    m[LPE_M_BOUND_GAP] = bound_gap; m[LPE_M_ROW_GAP] = row_gap;
}

}  // namespace

extern "C" {

int64_t lpe_certify_noalloc_capacity(void) { return NA_SCRATCH_DOUBLES; }

int64_t lpe_certify_noalloc_need(int32_t m, int32_t n) {
    if (m < 0 || n < 1) return -1;
    const long cr = (m + NA_CHUNK - 1) / NA_CHUNK, cc = (n + NA_CHUNK - 1) / NA_CHUNK;
    return cr * NA_RN + cc * NA_CN;
}

int lpe_certify_noalloc(const lpe_batch_problem* p, const double* x, const double* y, double* metrics,
                        double eps_primal, double eps_dual, double eps_gap, void* stream) {
    // every check before any HIP call: a rejected call touches no device
    if (!p || !x || !metrics) return lpe::set_error(LPE_EINVAL, "problem, x and metrics are required");
    if (p->m < 0 || p->n < 1 || p->batch < 1 || p->neq < 0 || p->neq > p->m || p->nnz < 0)
        return lpe::set_error(LPE_EINVAL, "invalid dimensions (need n >= 1, batch >= 1, 0 <= neq <= m)");
    if (p->m > 0 && (!y || !p->rhs)) return lpe::set_error(LPE_EINVAL, "y and rhs are required when m > 0");
    if (!p->row_ptr || !p->col_ptr || !p->c || !p->lower || !p->upper ||
        (p->nnz > 0 && (!p->col_idx || !p->row_idx || !p->t_pos || !p->values)))
        return lpe::set_error(LPE_EINVAL, "CSR, transposed pattern, c and bounds are required");
    const long per_lp = lpe_certify_noalloc_need(p->m, p->n);
    const long chunks_r = (p->m + NA_CHUNK - 1) / NA_CHUNK, chunks_c = (p->n + NA_CHUNK - 1) / NA_CHUNK;
    if (chunks_r > 65535 || chunks_c > 65535) return lpe::set_error(LPE_EINVAL, "m or n exceeds 65535 * 1024");
    if ((long)p->batch * per_lp > NA_SCRATCH_DOUBLES)
        return lpe::set_error(LPE_EINVAL, "batch * partials per LP (" + std::to_string(per_lp) +
                                              ") exceeds the static scratch of " + std::to_string(NA_SCRATCH_DOUBLES) +
                                              " doubles; at most " + std::to_string(NA_SCRATCH_DOUBLES / per_lp) +
                                              " LPs of this shape per call");
    double* part = nullptr;
    int rc = device_scratch(&part);
    if (rc != LPE_OK) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    NaArgs a;
    a.p = *p; a.x = x; a.y = y; a.part = part; a.per_lp = per_lp;
    a.chunks_r = (int)chunks_r; a.chunks_c = (int)chunks_c;
    if (chunks_r > 0) hipLaunchKernelGGL(k_na_rows, dim3(p->batch, (unsigned)chunks_r), dim3(NA_THREADS), 0, s, a);
    hipLaunchKernelGGL(k_na_cols, dim3(p->batch, (unsigned)chunks_c), dim3(NA_THREADS), 0, s, a);
    hipLaunchKernelGGL(k_na_final, dim3((p->batch + 255) / 256), dim3(256), 0, s, a, metrics,
                       eps_primal > 0 ? eps_primal : 1e-5, eps_dual > 0 ? eps_dual : 1e-7,
                       eps_gap > 0 ? eps_gap : 1e-7);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return lpe::set_error(LPE_EDEVICE, std::string("kernel launch failed: ") + hipGetErrorString(e));
    return LPE_OK;
}

}  // extern "C"
