// Exact homogeneous equality reduction (see include/lpe_reduction.h).
//
// The plan (forest, weights, groups, reduced pattern, dual-lift map) is built
// on the host once per pattern.  Every per-LP operation is one batched
// elementwise kernel over [B][len] FP64 arrays: one thread per (LP, group),
// (LP, row) or (LP, column), the members of a group contiguous in a CSR-like
// list.  The arithmetic of each element is a __host__ __device__ function, so
// the host path (on_device = 0) computes bit-identical results.
#include "lpe.h"
#include "lpe_reduction.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <map>
#include <string>
#include <utility>
#include <vector>

namespace lpe {
int set_error(int code, const std::string& msg);
int handle_dims(const lpe_handle* h, int* m, int* n, int* batch);
}  // namespace lpe

namespace {

using lpe::set_error;

#define RHIPCHK(expr)                                                                                  \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            return set_error(LPE_EDEVICE, std::string(#expr) + " failed: " + hipGetErrorString(e_));   \
    } while (0)

#define RCHK(expr)                       \
    do {                                 \
        int rc_ = (expr);                \
        if (rc_ != LPE_OK) return rc_;   \
    } while (0)

// Raw view of a plan, on the host or the device.
struct PlanView {
    int m, n, rm, rn, n_elim, n_var, r_n_var;
    const int32_t* group;    // [n]
    const double* weight;    // [n]
    const double* cmpw;      // [n] w_j / sum_{k in g} w_k^2
    const int32_t* rep;      // [rn]
    const int32_t* gptr;     // [rn+1]
    const int32_t* gmem;     // [n] members, ascending column
    const int32_t* gord;     // [n] members, descending |w| then ascending column
    const int32_t* rowmap;   // [m] kept index, or -1 - eliminated index
    const int32_t* elim;     // [n_elim]
    const int32_t* cptr;     // [r_n_var+1] contributions of each per-LP reduced entry
    const int32_t* csrc;     // per-LP value index, or -1 for a constant
    const double* cw;        // weight (per-LP) or template product (constant)
    const int32_t* lptr;     // [n_elim+1] dual-lift map
    const int32_t* lcol;
    const double* lval;
};

// finite test usable on host and device (inf - inf and nan - nan are nan)
__host__ __device__ inline bool fin(double v) { return v - v == 0.0; }

// Reduced bounds, witnesses and objective of group g of one LP
// (src/lp_equality_reduction.py:127-160, :540).
__host__ __device__ inline int red_group(const PlanView& p, int g, const double* lo, const double* up,
                                         const double* c, double* rl, double* ru, int32_t* lw, int32_t* uw,
                                         double* rc) {
    int flags = 0;
    double L = -INFINITY, U = INFINITY;
    const int k0 = p.gptr[g], k1 = p.gptr[g + 1];
    for (int k = k0; k < k1; ++k) {
        const int j = p.gmem[k];
        const double w = p.weight[j];
        const double a = w > 0.0 ? lo[j] : up[j], b = w > 0.0 ? up[j] : lo[j];
        const double il = a / w, iu = b / w;
        if ((fin(a) && !fin(il)) || (fin(b) && !fin(iu))) flags |= LPE_RED_OVERFLOW;
        if (il > L) L = il;
        if (iu < U) U = iu;
    }
    if (L > U) flags |= LPE_RED_INFEASIBLE;
    int wl = -1, wu = -1;
    for (int k = k0; k < k1; ++k) {
        const int j = p.gord[k];
        const double w = p.weight[j];
        const double il = (w > 0.0 ? lo[j] : up[j]) / w, iu = (w > 0.0 ? up[j] : lo[j]) / w;
        if (wl < 0 && fin(L) && il == L) wl = j;
        if (wu < 0 && fin(U) && iu == U) wu = j;
    }
    *rl = L; *ru = U; *lw = wl; *uw = wu;
    if (c) {
        double s = 0.0;
        for (int k = k0; k < k1; ++k) {
            const int j = p.gmem[k];
            s += p.weight[j] * c[j];
        }
        *rc = s;
    }
    return flags;
}

// One per-LP reduced matrix value: the row's entries of the group summed in
// ascending column order, as A[kept] @ T accumulates them.
__host__ __device__ inline double red_value(const PlanView& p, int q, const double* var) {
    double s = 0.0;
    for (int k = p.cptr[q]; k < p.cptr[q + 1]; ++k) {
        const int v = p.csrc[k];
        s += v >= 0 ? var[v] * p.cw[k] : p.cw[k];
    }
    return s;
}

__host__ __device__ inline double compress_group(const PlanView& p, int g, const double* x) {
    double s = 0.0;
    for (int k = p.gptr[g]; k < p.gptr[g + 1]; ++k) {
        const int j = p.gmem[k];
        s += p.cmpw[j] * x[j];
    }
    return s;
}

// Stationarity residual d = q - normal of group g, q = c - A^T y
// (src/gpu_reduced_pdhg.py:89-101).
__host__ __device__ inline void lift_group(const PlanView& p, int g, const double* c, const double* aty,
                                           int32_t lw, int32_t uw, double* d) {
    double rc = 0.0;
    const int k0 = p.gptr[g], k1 = p.gptr[g + 1];
    for (int k = k0; k < k1; ++k) {
        const int j = p.gmem[k];
        rc += p.weight[j] * (c[j] - aty[j]);
    }
    int wit = rc >= 0.0 ? lw : uw;
    if (wit < 0) wit = p.rep[g];
    for (int k = k0; k < k1; ++k) {
        const int j = p.gmem[k];
        const double q = c[j] - aty[j];
        d[j] = j == wit ? q - rc / p.weight[j] : q;
    }
}

__host__ __device__ inline double lift_row(const PlanView& p, int e, const double* d) {
    double s = 0.0;
    for (int k = p.lptr[e]; k < p.lptr[e + 1]; ++k) s += p.lval[k] * d[p.lcol[k]];
    return s;
}

// ---------------------------------------------------------------------------
// kernels: one thread per element of a [B][len] array

__global__ void k_zero_status(int32_t* status, int B) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < B) status[b] = 0;
}

__global__ void k_red_groups(PlanView p, int B, const double* c, int c_shared, const double* lo, const double* up,
                             double* rl, double* ru, int32_t* lw, int32_t* uw, double* rc, int32_t* status) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)B * p.rn) return;
    const int b = (int)(t / p.rn), g = (int)(t % p.rn);
    const long o = (long)b * p.n, ro = (long)b * p.rn;
    const bool want_c = c && (!c_shared || b == 0);
    double dummy;
    const int f = red_group(p, g, lo + o, up + o, want_c ? (c_shared ? c : c + o) : nullptr, rl + ro + g, ru + ro + g,
                            lw + ro + g, uw + ro + g, want_c ? (c_shared ? rc + g : rc + ro + g) : &dummy);
    if (f) atomicOr(&status[b], f);
}

__global__ void k_red_rows(PlanView p, int B, const double* rhs, double* rrhs, int32_t* status) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)B * p.m) return;
    const int b = (int)(t / p.m), i = (int)(t % p.m);
    const int r = p.rowmap[i];
    const double v = rhs[(long)b * p.m + i];
    if (r >= 0) rrhs[(long)b * p.rm + r] = v;
    else if (v != 0.0) atomicOr(&status[b], LPE_RED_INHOMOGENEOUS);
}

__global__ void k_red_values(PlanView p, int B, const double* var, long var_ld, double* rvar) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)B * p.r_n_var) return;
    const int b = (int)(t / p.r_n_var), q = (int)(t % p.r_n_var);
    rvar[t] = red_value(p, q, var + b * var_ld);
}

__global__ void k_expand(PlanView p, int B, const double* z, double* x) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)B * p.n) return;
    const int b = (int)(t / p.n), j = (int)(t % p.n);
    x[t] = p.weight[j] * z[(long)b * p.rn + p.group[j]];
}

__global__ void k_compress(PlanView p, int B, const double* x, double* z) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)B * p.rn) return;
    const int b = (int)(t / p.rn), g = (int)(t % p.rn);
    z[t] = compress_group(p, g, x + (long)b * p.n);
}

__global__ void k_lift_rows(PlanView p, int B, const double* yr, double* y) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)B * p.m) return;
    const int b = (int)(t / p.m), i = (int)(t % p.m);
    const int r = p.rowmap[i];
    y[t] = r >= 0 ? yr[(long)b * p.rm + r] : 0.0;
}

__global__ void k_lift_groups(PlanView p, int B, const double* c, int c_shared, const double* aty, const int32_t* lw,
                              const int32_t* uw, double* d) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)B * p.rn) return;
    const int b = (int)(t / p.rn), g = (int)(t % p.rn);
    const long o = (long)b * p.n;
    lift_group(p, g, c_shared ? c : c + o, aty + o, lw[t], uw[t], d + o);
}

__global__ void k_lift_elim(PlanView p, int B, const double* d, double* y) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)B * p.n_elim) return;
    const int b = (int)(t / p.n_elim), e = (int)(t % p.n_elim);
    y[(long)b * p.m + p.elim[e]] = lift_row(p, e, d + (long)b * p.n);
}

inline unsigned blocks(long count) { return (unsigned)std::max<long>(1, (count + 255) / 256); }

}  // namespace

struct lpe_reduction {
    // original pattern
    int m = 0, n = 0, neq = 0;
    std::vector<int32_t> rp, ci;
    std::vector<double> val;
    std::vector<int32_t> var_of;     // [nnz] per-LP index or -1
    std::vector<double> tvar;        // template values of the per-LP entries
    int n_var = 0;
    // plan
    std::vector<int32_t> group, rep, gptr, gmem, gord, rowmap, kept, elim;
    std::vector<double> weight, cmpw;
    // reduced pattern
    int rm = 0, rn = 0, rneq = 0;
    std::vector<int32_t> rrp, rci;
    std::vector<double> rval;
    std::vector<int64_t> rvar_pos;
    std::vector<int32_t> cptr, csrc;
    std::vector<double> cw;
    std::vector<int32_t> lptr, lcol;
    std::vector<double> lval;
    // device mirror (built at the first device call)
    int dev = -1;
    std::vector<void*> allocs;
    PlanView dview{};
    double* d_tvar = nullptr;
    double* d_aty = nullptr;
    double* d_res = nullptr;
    long scratch = 0;

    PlanView host_view() const {
        PlanView p;
        p.m = m; p.n = n; p.rm = rm; p.rn = rn; p.n_elim = (int)elim.size(); p.n_var = n_var;
        p.r_n_var = (int)rvar_pos.size();
        p.group = group.data(); p.weight = weight.data(); p.cmpw = cmpw.data(); p.rep = rep.data();
        p.gptr = gptr.data(); p.gmem = gmem.data(); p.gord = gord.data(); p.rowmap = rowmap.data();
        p.elim = elim.data(); p.cptr = cptr.data(); p.csrc = csrc.data(); p.cw = cw.data();
        p.lptr = lptr.data(); p.lcol = lcol.data(); p.lval = lval.data();
        return p;
    }
    ~lpe_reduction() {
        for (void* q : allocs) (void)hipFree(q);
    }
};

namespace {

// Null-space forest of the two-entry homogeneous equality rows
// (src/lp_equality_reduction.py:320-409) and the tree dual map (:463-517).
struct Edge {
    int row, left, right;
    double al, ar;
};

int build_plan(lpe_reduction* r, const double* rhs, double ratio) {
    const int m = r->m, n = r->n;
    std::vector<char> var_row(m, 0);
    for (int i = 0; i < m; ++i)
        for (int k = r->rp[i]; k < r->rp[i + 1]; ++k)
            if (r->var_of[k] >= 0) var_row[i] = 1;
    // union-find, smaller root wins, rows in order
    std::vector<int> root(n);
    for (int j = 0; j < n; ++j) root[j] = j;
    auto find = [&](int v) {
        while (root[v] != v) {
            root[v] = root[root[v]];
            v = root[v];
        }
        return v;
    };
    std::vector<Edge> edges;
    std::vector<int> edge_of_row(m, -1);
    for (int i = 0; i < r->neq; ++i) {
        const int k = r->rp[i];
        if (r->rp[i + 1] - k != 2 || rhs[i] != 0.0 || var_row[i]) continue;
        const double a0 = std::fabs(r->val[k]), a1 = std::fabs(r->val[k + 1]);
        const double lo = std::min(a0, a1), hi = std::max(a0, a1);
        if (lo == 0.0 || hi / lo > ratio) continue;
        int u = find(r->ci[k]), v = find(r->ci[k + 1]);
        if (u == v) continue;   // a cycle: the row stays explicit
        if (u > v) std::swap(u, v);
        root[v] = u;
        edge_of_row[i] = (int)edges.size();
        edges.push_back({i, r->ci[k], r->ci[k + 1], r->val[k], r->val[k + 1]});
    }
    std::vector<std::vector<std::pair<int, int>>> nbrs(n);   // (neighbour, edge)
    for (int e = 0; e < (int)edges.size(); ++e) {
        nbrs[edges[e].left].push_back({edges[e].right, e});
        nbrs[edges[e].right].push_back({edges[e].left, e});
    }
    // depth-first weights from each tree's smallest column
    std::vector<char> seen(n, 0), has_alpha(n, 0), in_tree(n, 0);
    std::vector<double> alpha(n, 0.0), w(n, 1.0);
    std::vector<int> tree_of(n, -1);
    struct Tree {
        std::vector<int> nodes, edges;
        int rep;
    };
    std::vector<Tree> trees;
    for (int start = 0; start < n; ++start) {
        if (nbrs[start].empty() || seen[start]) continue;
        std::vector<int> order, todo{start}, tedges;
        std::vector<char> edge_seen;
        alpha[start] = 1.0;
        has_alpha[start] = 1;
        bool ok = true;
        while (!todo.empty()) {
            const int v = todo.back();
            todo.pop_back();
            if (seen[v]) continue;
            seen[v] = 1;
            order.push_back(v);
            for (const auto& ue : nbrs[v]) {
                tedges.push_back(ue.second);
                const int u = ue.first;
                if (has_alpha[u]) continue;
                const Edge& E = edges[ue.second];
                const double val = v == E.left ? (-E.al * alpha[v]) / E.ar : (-E.ar * alpha[v]) / E.al;
                if (!std::isfinite(val) || val == 0.0) ok = false;
                alpha[u] = val;
                has_alpha[u] = 1;
                todo.push_back(u);
            }
        }
        if (!ok || order.size() < 2) continue;
        double top = 0.0, low = INFINITY;
        bool finite = true;
        for (int v : order) {
            const double a = std::fabs(alpha[v]);
            finite = finite && std::isfinite(a);
            top = std::max(top, a);
            low = std::min(low, a);
        }
        if (!finite || top / low > ratio) continue;
        int rep = n;
        for (int v : order)
            if (std::fabs(alpha[v]) == top) rep = std::min(rep, v);
        const double s = alpha[rep];
        for (int v : order) w[v] = alpha[v] / s;
        w[rep] = 1.0;
        Tree t;
        t.nodes = order;
        std::sort(t.nodes.begin(), t.nodes.end());
        std::sort(tedges.begin(), tedges.end());
        tedges.erase(std::unique(tedges.begin(), tedges.end()), tedges.end());
        t.edges = tedges;   // edge indices ascend with their rows
        t.rep = rep;
        for (int v : order) {
            in_tree[v] = 1;
            tree_of[v] = (int)trees.size();
        }
        trees.push_back(std::move(t));
    }
    // groups ordered by smallest member: valid trees and singletons
    r->group.assign(n, -1);
    r->weight.assign(n, 1.0);
    r->rep.clear();
    std::vector<int> elim_rows;
    std::vector<int> tree_group(trees.size(), -1);
    for (int v = 0; v < n; ++v) {
        if (in_tree[v]) {
            const Tree& t = trees[tree_of[v]];
            if (t.nodes[0] != v) continue;
            const int g = (int)r->rep.size();
            tree_group[tree_of[v]] = g;
            for (int u : t.nodes) {
                r->group[u] = g;
                r->weight[u] = w[u];
            }
            int first_one = -1;
            for (int u : t.nodes)
                if (w[u] == 1.0) { first_one = u; break; }
            r->rep.push_back(first_one);
            for (int e : t.edges) elim_rows.push_back(edges[e].row);
        } else {
            r->group[v] = (int)r->rep.size();
            r->rep.push_back(v);
        }
    }
    r->rn = (int)r->rep.size();
    r->gptr.assign(r->rn + 1, 0);
    for (int j = 0; j < n; ++j) r->gptr[r->group[j] + 1]++;
    for (int g = 0; g < r->rn; ++g) r->gptr[g + 1] += r->gptr[g];
    r->gmem.assign(n, 0);
    {
        std::vector<int> fill(r->gptr.begin(), r->gptr.end() - 1);
        for (int j = 0; j < n; ++j) r->gmem[fill[r->group[j]]++] = j;
    }
    r->gord = r->gmem;
    for (int g = 0; g < r->rn; ++g) {
        auto b = r->gord.begin() + r->gptr[g], e = r->gord.begin() + r->gptr[g + 1];
        std::stable_sort(b, e, [&](int x, int y) { return std::fabs(r->weight[x]) > std::fabs(r->weight[y]); });
    }
    std::vector<double> denom(r->rn, 0.0);
    for (int j = 0; j < n; ++j) denom[r->group[j]] += r->weight[j] * r->weight[j];
    r->cmpw.resize(n);
    for (int j = 0; j < n; ++j) r->cmpw[j] = r->weight[j] / denom[r->group[j]];
    // kept and eliminated rows
    std::sort(elim_rows.begin(), elim_rows.end());
    r->elim.assign(elim_rows.begin(), elim_rows.end());
    r->rowmap.assign(m, 0);
    std::vector<char> gone(m, 0);
    for (int e = 0; e < (int)r->elim.size(); ++e) {
        gone[r->elim[e]] = 1;
        r->rowmap[r->elim[e]] = -1 - e;
    }
    r->kept.clear();
    for (int i = 0; i < m; ++i)
        if (!gone[i]) {
            r->rowmap[i] = (int)r->kept.size();
            r->kept.push_back(i);
        }
    r->rm = (int)r->kept.size();
    r->rneq = r->neq - (int)r->elim.size();
    // E T = 0 (src/lp_equality_reduction.py:452-454)
    for (int row : r->elim) {
        const Edge& E = edges[edge_of_row[row]];
        const double s = 0.0 + E.al * r->weight[E.left] + E.ar * r->weight[E.right];
        if (!(std::fabs(s) <= 1e-12)) return set_error(LPE_EINVAL, "Constructed primal map is not in the equality null space");
    }
    // dual-lift map: leaf-to-root solve on a breadth-first tree from the
    // representative; row e of L is sparse over the subtree below edge e
    std::vector<std::vector<std::pair<int, double>>> expr(r->elim.size());
    std::vector<int> elim_pos(m, -1);
    for (int e = 0; e < (int)r->elim.size(); ++e) elim_pos[r->elim[e]] = e;
    for (const Tree& t : trees) {
        if (tree_group[&t - trees.data()] < 0) continue;
        std::map<int, std::vector<std::pair<int, int>>> adj;   // node -> (neighbour, edge)
        for (int e : t.edges) {
            adj[edges[e].left].push_back({edges[e].right, e});
            adj[edges[e].right].push_back({edges[e].left, e});
        }
        const int root_v = r->rep[tree_group[&t - trees.data()]];
        std::map<int, int> parent, up;
        std::vector<int> bfs{root_v};
        parent[root_v] = -1;
        for (size_t h = 0; h < bfs.size(); ++h) {
            const int v = bfs[h];
            for (const auto& ue : adj[v]) {
                if (parent.count(ue.first)) continue;
                parent[ue.first] = v;
                up[ue.first] = ue.second;
                bfs.push_back(ue.first);
            }
        }
        if (bfs.size() != t.nodes.size()) return set_error(LPE_EINVAL, "Equality forest is disconnected");
        std::map<int, std::vector<int>> kids;
        for (size_t h = 1; h < bfs.size(); ++h) kids[parent[bfs[h]]].push_back(bfs[h]);
        for (size_t h = bfs.size(); h-- > 1;) {
            const int v = bfs[h];
            const Edge& E = edges[up[v]];
            const double coef = v == E.left ? E.al : E.ar;
            std::vector<std::pair<int, double>> ex{{v, 1.0 / coef}};
            for (int k : kids[v]) {
                const Edge& K = edges[up[k]];
                const double f = -(v == K.left ? K.al : K.ar) / coef;
                for (const auto& cv : expr[elim_pos[K.row]]) ex.push_back({cv.first, f * cv.second});
            }
            std::sort(ex.begin(), ex.end());
            expr[elim_pos[E.row]] = std::move(ex);
        }
    }
    r->lptr.assign(r->elim.size() + 1, 0);
    r->lcol.clear();
    r->lval.clear();
    for (size_t e = 0; e < r->elim.size(); ++e) {
        for (const auto& cv : expr[e])
            if (cv.second != 0.0) {
                r->lcol.push_back(cv.first);
                r->lval.push_back(cv.second);
            }
        r->lptr[e + 1] = (int)r->lcol.size();
    }
    // L E^T = I (src/lp_equality_reduction.py:455-461)
    {
        std::vector<std::vector<int>> rows_of_col(n);
        for (size_t e = 0; e < r->elim.size(); ++e) {
            const Edge& E = edges[edge_of_row[r->elim[e]]];
            rows_of_col[E.left].push_back((int)e);
            rows_of_col[E.right].push_back((int)e);
        }
        std::map<int, double> acc;
        for (size_t e = 0; e < r->elim.size(); ++e) {
            acc.clear();
            for (int k = r->lptr[e]; k < r->lptr[e + 1]; ++k)
                for (int e2 : rows_of_col[r->lcol[k]]) {
                    const Edge& E = edges[edge_of_row[r->elim[e2]]];
                    acc[e2] += r->lval[k] * (r->lcol[k] == E.left ? E.al : E.ar);
                }
            acc[(int)e] -= 1.0;
            for (const auto& kv : acc)
                if (!(std::fabs(kv.second) <= 1e-10))
                    return set_error(LPE_EINVAL, "Constructed tree dual map is not a left inverse of the eliminated transpose");
        }
    }
    // reduced pattern A[kept] T: per kept row, its groups ascending; a
    // constant entry that sums to zero is dropped as the reference's sparse
    // product drops it
    r->rrp.assign(r->rm + 1, 0);
    r->rci.clear();
    r->rval.clear();
    r->rvar_pos.clear();
    r->cptr.assign(1, 0);
    r->csrc.clear();
    r->cw.clear();
    for (int ri = 0; ri < r->rm; ++ri) {
        const int i = r->kept[ri];
        std::map<int, std::vector<int>> by_group;   // group -> entries, ascending column
        for (int k = r->rp[i]; k < r->rp[i + 1]; ++k) by_group[r->group[r->ci[k]]].push_back(k);
        for (const auto& gk : by_group) {
            double s = 0.0;
            bool var = false;
            for (int k : gk.second) {
                s += r->val[k] * r->weight[r->ci[k]];
                var = var || r->var_of[k] >= 0;
            }
            if (!var && s == 0.0) continue;
            if (var) {
                r->rvar_pos.push_back((int64_t)r->rci.size());
                for (int k : gk.second) {
                    const int v = r->var_of[k];
                    r->csrc.push_back(v);
                    r->cw.push_back(v >= 0 ? r->weight[r->ci[k]] : r->val[k] * r->weight[r->ci[k]]);
                }
                r->cptr.push_back((int)r->csrc.size());
            }
            r->rci.push_back(gk.first);
            r->rval.push_back(s);
        }
        r->rrp[ri + 1] = (int)r->rci.size();
    }
    return LPE_OK;
}

template <typename T>
int dupload(lpe_reduction* r, T** d, const std::vector<T>& h) {
    void* q = nullptr;
    RHIPCHK(hipMalloc(&q, std::max<size_t>(1, h.size()) * sizeof(T)));
    r->allocs.push_back(q);
    if (!h.empty()) RHIPCHK(hipMemcpy(q, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    *d = static_cast<T*>(q);
    return LPE_OK;
}

// Upload the plan at the first device call; later calls use that device.
int device_ready(lpe_reduction* r) {
    if (r->dev >= 0) {
        RHIPCHK(hipSetDevice(r->dev));
        return LPE_OK;
    }
    int dev = 0;
    RHIPCHK(hipGetDevice(&dev));
    PlanView p = r->host_view();
    int32_t *group, *rep, *gptr, *gmem, *gord, *rowmap, *elim, *cptr, *csrc, *lptr, *lcol;
    double *weight, *cmpw, *cw, *lval;
    RCHK(dupload(r, &group, r->group)); RCHK(dupload(r, &rep, r->rep)); RCHK(dupload(r, &gptr, r->gptr));
    RCHK(dupload(r, &gmem, r->gmem)); RCHK(dupload(r, &gord, r->gord)); RCHK(dupload(r, &rowmap, r->rowmap));
    RCHK(dupload(r, &elim, r->elim)); RCHK(dupload(r, &cptr, r->cptr)); RCHK(dupload(r, &csrc, r->csrc));
    RCHK(dupload(r, &lptr, r->lptr)); RCHK(dupload(r, &lcol, r->lcol)); RCHK(dupload(r, &weight, r->weight));
    RCHK(dupload(r, &cmpw, r->cmpw)); RCHK(dupload(r, &cw, r->cw)); RCHK(dupload(r, &lval, r->lval));
    RCHK(dupload(r, &r->d_tvar, r->tvar));
    p.group = group; p.rep = rep; p.gptr = gptr; p.gmem = gmem; p.gord = gord; p.rowmap = rowmap; p.elim = elim;
    p.cptr = cptr; p.csrc = csrc; p.lptr = lptr; p.lcol = lcol; p.weight = weight; p.cmpw = cmpw; p.cw = cw;
    p.lval = lval;
    r->dview = p;
    r->dev = dev;
    return LPE_OK;
}

int device_scratch(lpe_reduction* r, long count) {
    if (count <= r->scratch) return LPE_OK;
    void *a = nullptr, *b = nullptr;
    RHIPCHK(hipMalloc(&a, count * sizeof(double)));
    r->allocs.push_back(a);
    RHIPCHK(hipMalloc(&b, count * sizeof(double)));
    r->allocs.push_back(b);
    r->d_aty = static_cast<double*>(a);
    r->d_res = static_cast<double*>(b);
    r->scratch = count;
    return LPE_OK;
}

}  // namespace

extern "C" {

int lpe_reduction_create(const lpe_reduction_desc* d, lpe_reduction** out) {
    if (!d || !out) return set_error(LPE_EINVAL, "desc and out are required");
    *out = nullptr;
    if (d->m < 0 || d->n < 1 || d->neq < 0 || d->neq > d->m || d->nnz < 0 || d->nnz > INT32_MAX)
        return set_error(LPE_EINVAL, "invalid dimensions (need n >= 1, 0 <= neq <= m)");
    if (!(std::isfinite(d->max_scale_ratio) && d->max_scale_ratio >= 1.0))
        return set_error(LPE_EINVAL, "max_scale_ratio must be finite and at least one");
    if (!d->row_ptr || (d->nnz > 0 && (!d->col_idx || !d->values)) || (d->m > 0 && !d->rhs))
        return set_error(LPE_EINVAL, "CSR arrays and rhs are required");
    const int m = d->m, n = d->n;
    const long nnz = d->nnz;
    if (d->row_ptr[0] != 0 || d->row_ptr[m] != nnz) return set_error(LPE_EINVAL, "Invalid CSR row pointers");
    for (int i = 0; i < m; ++i) {
        if (d->row_ptr[i + 1] < d->row_ptr[i]) return set_error(LPE_EINVAL, "Invalid CSR row pointers");
        for (long k = d->row_ptr[i]; k < d->row_ptr[i + 1]; ++k) {
            const int c = d->col_idx[k];
            if (c < 0 || c >= n) return set_error(LPE_EINVAL, "Invalid CSR column indices");
            if (k > d->row_ptr[i] && c <= d->col_idx[k - 1])
                return set_error(LPE_EINVAL, "CSR rows must have sorted, unique column indices");
            if (!std::isfinite(d->values[k])) return set_error(LPE_EINVAL, "Matrix values must be finite");
        }
        if (!std::isfinite(d->rhs[i])) return set_error(LPE_EINVAL, "rhs contains invalid values");
    }
    if (d->n_var < 0 || (d->n_var > 0 && !d->var_pos)) return set_error(LPE_EINVAL, "invalid per-LP coefficient positions");
    lpe_reduction* r = new lpe_reduction();
    r->m = m; r->n = n; r->neq = d->neq; r->n_var = d->n_var;
    r->rp.assign(d->row_ptr, d->row_ptr + m + 1);
    r->ci.assign(d->col_idx, d->col_idx + nnz);
    r->val.assign(d->values, d->values + nnz);
    r->var_of.assign(nnz, -1);
    for (int v = 0; v < d->n_var; ++v) {
        const int64_t k = d->var_pos[v];
        if (k < 0 || k >= nnz || r->var_of[k] >= 0) {
            delete r;
            return set_error(LPE_EINVAL, "per-LP positions must be unique entries of the pattern");
        }
        r->var_of[k] = v;
        r->tvar.push_back(r->val[k]);
    }
    const int rc = build_plan(r, d->rhs, d->max_scale_ratio);
    if (rc != LPE_OK) {
        delete r;
        return rc;
    }
    *out = r;
    return LPE_OK;
}

int lpe_reduction_get_dims(const lpe_reduction* r, lpe_reduction_dims* dims) {
    if (!r || !dims) return set_error(LPE_EINVAL, "reduction and dims are required");
    dims->m = r->rm; dims->n = r->rn; dims->neq = r->rneq; dims->nnz = (int64_t)r->rci.size();
    dims->n_var = (int32_t)r->rvar_pos.size(); dims->n_eliminated = (int32_t)r->elim.size();
    dims->lift_nnz = (int64_t)r->lcol.size();
    return LPE_OK;
}

int lpe_reduction_get_plan(const lpe_reduction* r, int32_t* group, double* weight, int32_t* representative,
                           int32_t* kept_rows, int32_t* eliminated_rows) {
    if (!r) return set_error(LPE_EINVAL, "reduction is required");
    if (group) std::copy(r->group.begin(), r->group.end(), group);
    if (weight) std::copy(r->weight.begin(), r->weight.end(), weight);
    if (representative) std::copy(r->rep.begin(), r->rep.end(), representative);
    if (kept_rows) std::copy(r->kept.begin(), r->kept.end(), kept_rows);
    if (eliminated_rows) std::copy(r->elim.begin(), r->elim.end(), eliminated_rows);
    return LPE_OK;
}

int lpe_reduction_get_pattern(const lpe_reduction* r, int32_t* row_ptr, int32_t* col_idx, double* values,
                              int64_t* var_pos) {
    if (!r) return set_error(LPE_EINVAL, "reduction is required");
    if (row_ptr) std::copy(r->rrp.begin(), r->rrp.end(), row_ptr);
    if (col_idx) std::copy(r->rci.begin(), r->rci.end(), col_idx);
    if (values) std::copy(r->rval.begin(), r->rval.end(), values);
    if (var_pos) std::copy(r->rvar_pos.begin(), r->rvar_pos.end(), var_pos);
    return LPE_OK;
}

int lpe_reduction_get_lift_map(const lpe_reduction* r, int32_t* row_ptr, int32_t* col_idx, double* values) {
    if (!r) return set_error(LPE_EINVAL, "reduction is required");
    if (row_ptr) std::copy(r->lptr.begin(), r->lptr.end(), row_ptr);
    if (col_idx) std::copy(r->lcol.begin(), r->lcol.end(), col_idx);
    if (values) std::copy(r->lval.begin(), r->lval.end(), values);
    return LPE_OK;
}

int lpe_reduction_reduce(lpe_reduction* r, int32_t B, int32_t shared_c, const lpe_vectors* in, const lpe_reduced* out,
                         int32_t on_device, void* stream) {
    if (!r || !in || !out || B < 1) return set_error(LPE_EINVAL, "reduction, batch, inputs and outputs are required");
    if (!in->rhs || !in->lower || !in->upper || !out->rhs || !out->lower || !out->upper || !out->lower_witness ||
        !out->upper_witness || !out->status || (in->c && !out->c))
        return set_error(LPE_EINVAL, "rhs, bounds, witnesses and status arrays are required");
    const int rnv = (int)r->rvar_pos.size();
    if (rnv > 0 && !out->var_values) return set_error(LPE_EINVAL, "var_values output is required");
    const long var_ld = in->var_values ? r->n_var : 0;
    if (on_device) {
        RCHK(device_ready(r));
        hipStream_t s = (hipStream_t)stream;
        const PlanView& p = r->dview;
        hipLaunchKernelGGL(k_zero_status, dim3(blocks(B)), dim3(256), 0, s, out->status, B);
        hipLaunchKernelGGL(k_red_groups, dim3(blocks((long)B * r->rn)), dim3(256), 0, s, p, B, in->c, shared_c,
                           in->lower, in->upper, out->lower, out->upper, out->lower_witness, out->upper_witness, out->c,
                           out->status);
        if (r->m > 0)
            hipLaunchKernelGGL(k_red_rows, dim3(blocks((long)B * r->m)), dim3(256), 0, s, p, B, in->rhs, out->rhs,
                               out->status);
        if (rnv > 0)
            hipLaunchKernelGGL(k_red_values, dim3(blocks((long)B * rnv)), dim3(256), 0, s, p, B,
                               in->var_values ? in->var_values : r->d_tvar, var_ld, out->var_values);
        RHIPCHK(hipGetLastError());
        return LPE_OK;
    }
    const PlanView p = r->host_view();
    const double* var = in->var_values ? in->var_values : r->tvar.data();
    for (int b = 0; b < B; ++b) {
        int flags = 0;
        const long o = (long)b * r->n, ro = (long)b * r->rn;
        for (int g = 0; g < r->rn; ++g) {
            const bool want_c = in->c && (!shared_c || b == 0);
            double dummy;
            flags |= red_group(p, g, in->lower + o, in->upper + o, want_c ? (shared_c ? in->c : in->c + o) : nullptr,
                               out->lower + ro + g, out->upper + ro + g, out->lower_witness + ro + g,
                               out->upper_witness + ro + g,
                               want_c ? (shared_c ? out->c + g : out->c + ro + g) : &dummy);
        }
        for (int i = 0; i < r->m; ++i) {
            const double v = in->rhs[(long)b * r->m + i];
            if (r->rowmap[i] >= 0) out->rhs[(long)b * r->rm + r->rowmap[i]] = v;
            else if (v != 0.0) flags |= LPE_RED_INHOMOGENEOUS;
        }
        for (int q = 0; q < rnv; ++q) out->var_values[(long)b * rnv + q] = red_value(p, q, var + b * var_ld);
        out->status[b] = flags;
    }
    return LPE_OK;
}

int lpe_reduction_expand(lpe_reduction* r, int32_t B, const double* z, double* x, int32_t on_device, void* stream) {
    if (!r || !z || !x || B < 1) return set_error(LPE_EINVAL, "reduction, batch, z and x are required");
    if (on_device) {
        RCHK(device_ready(r));
        hipLaunchKernelGGL(k_expand, dim3(blocks((long)B * r->n)), dim3(256), 0, (hipStream_t)stream, r->dview, B, z, x);
        RHIPCHK(hipGetLastError());
        return LPE_OK;
    }
    for (int b = 0; b < B; ++b)
        for (int j = 0; j < r->n; ++j) x[(long)b * r->n + j] = r->weight[j] * z[(long)b * r->rn + r->group[j]];
    return LPE_OK;
}

int lpe_reduction_compress(lpe_reduction* r, int32_t B, const double* x, double* z, int32_t on_device, void* stream) {
    if (!r || !z || !x || B < 1) return set_error(LPE_EINVAL, "reduction, batch, x and z are required");
    if (on_device) {
        RCHK(device_ready(r));
        hipLaunchKernelGGL(k_compress, dim3(blocks((long)B * r->rn)), dim3(256), 0, (hipStream_t)stream, r->dview, B, x,
                           z);
        RHIPCHK(hipGetLastError());
        return LPE_OK;
    }
    const PlanView p = r->host_view();
    for (int b = 0; b < B; ++b)
        for (int g = 0; g < r->rn; ++g) z[(long)b * r->rn + g] = compress_group(p, g, x + (long)b * r->n);
    return LPE_OK;
}

int lpe_reduction_lift(lpe_reduction* r, lpe_handle* original, int32_t B, int32_t shared_c, const double* c,
                       const double* var_values, const double* z, const double* yr, const int32_t* lw,
                       const int32_t* uw, double* x, double* y, int32_t on_device, void* stream) {
    if (!r || !c || !z || !lw || !uw || !x || !y || B < 1 || (r->rm > 0 && !yr))
        return set_error(LPE_EINVAL, "reduction, batch, c, z, y_reduced, witnesses, x and y are required");
    if (on_device) {
        if (!original) return set_error(LPE_EINVAL, "the device lift needs the original engine handle");
        int om = 0, on = 0, ob = 0;
        RCHK(lpe::handle_dims(original, &om, &on, &ob));
        if (om != r->m || on != r->n || ob != B)
            return set_error(LPE_EINVAL, "original handle dimensions or batch differ from the reduction");
        RCHK(device_ready(r));
        RCHK(device_scratch(r, (long)B * r->n));
        hipStream_t s = (hipStream_t)stream;
        const PlanView& p = r->dview;
        hipLaunchKernelGGL(k_expand, dim3(blocks((long)B * r->n)), dim3(256), 0, s, p, B, z, x);
        if (r->m > 0) hipLaunchKernelGGL(k_lift_rows, dim3(blocks((long)B * r->m)), dim3(256), 0, s, p, B, yr, y);
        RHIPCHK(hipGetLastError());
        RCHK(lpe_spmv(original, 1, LPE_PREC_FP64, y, r->d_aty, stream));
        hipLaunchKernelGGL(k_lift_groups, dim3(blocks((long)B * r->rn)), dim3(256), 0, s, p, B, c, shared_c, r->d_aty,
                           lw, uw, r->d_res);
        if (!r->elim.empty())
            hipLaunchKernelGGL(k_lift_elim, dim3(blocks((long)B * r->elim.size())), dim3(256), 0, s, p, B, r->d_res, y);
        RHIPCHK(hipGetLastError());
        return LPE_OK;
    }
    const PlanView p = r->host_view();
    std::vector<double> aty(r->n), res(r->n);
    for (int b = 0; b < B; ++b) {
        const long o = (long)b * r->n, mo = (long)b * r->m;
        for (int j = 0; j < r->n; ++j) x[o + j] = r->weight[j] * z[(long)b * r->rn + r->group[j]];
        for (int i = 0; i < r->m; ++i) y[mo + i] = r->rowmap[i] >= 0 ? yr[(long)b * r->rm + r->rowmap[i]] : 0.0;
        std::fill(aty.begin(), aty.end(), 0.0);
        for (int i = 0; i < r->m; ++i)
            for (int k = r->rp[i]; k < r->rp[i + 1]; ++k) {
                const int v = r->var_of[k];
                const double a = v < 0 ? r->val[k] : (var_values ? var_values[(long)b * r->n_var + v] : r->tvar[v]);
                aty[r->ci[k]] += a * y[mo + i];
            }
        for (int g = 0; g < r->rn; ++g)
            lift_group(p, g, shared_c ? c : c + o, aty.data(), lw[(long)b * r->rn + g], uw[(long)b * r->rn + g],
                       res.data());
        for (int e = 0; e < (int)r->elim.size(); ++e) y[mo + r->elim[e]] = lift_row(p, e, res.data());
    }
    return LPE_OK;
}

int lpe_reduction_destroy(lpe_reduction* r) {
    if (!r) return set_error(LPE_EINVAL, "reduction is required");
    if (r->dev >= 0) (void)hipSetDevice(r->dev);
    delete r;
    return LPE_OK;
}

}  // extern "C"
