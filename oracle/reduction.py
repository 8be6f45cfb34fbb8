"""Exact homogeneous equality reduction on the CPU (test infrastructure; see
oracle/__init__.py).

Restates src/lp_equality_reduction.py:

* ``plan`` (:295-430): two-entry rows of the equality block with zero rhs and
  a coefficient magnitude ratio within ``max_scale_ratio`` are taken in row
  order while they join two different union-find trees (:330-347); each tree
  of the resulting forest gets null-space weights by a depth-first walk from
  its smallest column (:357-385), is dropped (its rows kept) when a weight is
  zero / non-finite or the weight magnitudes spread beyond the ratio
  (:386-392), and is normalised so the smallest column of largest |weight|
  (the representative) has weight exactly 1 (:393-400).  Groups are the valid
  trees plus singleton columns, ordered by smallest member (:403-409).
  T maps reduced z to x = T z (one entry per original column), the
  compression is z = (T^T T)^-1 T^T x (:433-444), and the dual-lift map is
  the tree solve of section ``_build_dual_lift_map`` (:463-517).
* ``reduce`` (:519-560): reduced bounds are the intersection of the bounds
  each member induces, the witness of an endpoint is the first member in
  (descending |weight|, ascending column) order that attains it (:127-160);
  reduced A = A[kept] T, c = T^T c, rhs = rhs[kept].
* ``expand`` / ``compress`` (:562-586), ``lift`` (the device lift of
  src/gpu_reduced_pdhg.py:81-103): y on kept rows from the reduced dual,
  q = c - A^T y, reduced cost T^T q, the bound witness chosen by its sign
  (>= 0: lower, else upper, missing: representative) takes reduced_cost /
  weight, the eliminated rows' duals are L (q - normal).

Pinned by: the reference's known answers (tests/test_lp_equality_reduction.py:
the equal-variable, scaled-relation, negative-weight, tree dual map, cycle,
near-zero and non-homogeneous, dynamic-vector, infeasible-intersection,
compression and missing-bound cases; tests/test_gpu_reduced_pdhg.py:74-111
witnesses), the algebraic identities E T = 0 and L E^T = I, and HiGHS: the
reduced LP has the original optimum and the lifted reduced HiGHS solution
passes the original certificate (tests/test_oracle_reduction.py).
"""

from __future__ import annotations

import hashlib
from dataclasses import dataclass

import numpy as np
from scipy.sparse import coo_matrix, csr_matrix


def _normalized(problem):
    a, rhs, lower, upper, c, neq = problem
    a = csr_matrix(a, dtype=np.float64, copy=True)
    a.sum_duplicates()
    a.eliminate_zeros()
    a.sort_indices()
    return (a, np.asarray(rhs, np.float64).copy(), np.asarray(lower, np.float64).copy(),
            np.asarray(upper, np.float64).copy(), np.asarray(c, np.float64).copy(), int(neq))


def equality_fingerprint(a: csr_matrix, neq: int) -> str:
    """sha256 over the equality block's shape, CSR arrays and neq
    (src/lp_equality_reduction.py:64-82)."""
    e = csr_matrix(a[:neq], copy=True)
    e.sum_duplicates()
    e.eliminate_zeros()
    e.sort_indices()
    h = hashlib.sha256()
    for name, value, dtype in (("shape", e.shape, "<i8"), ("indptr", e.indptr, "<i8"),
                               ("indices", e.indices, "<i8"), ("data", e.data, "<f8"), ("neq", [neq], "<i8")):
        arr = np.asarray(value, dtype=dtype)
        h.update(name.encode("ascii"))
        h.update(np.asarray(arr.shape, dtype="<i8").tobytes())
        h.update(arr.tobytes())
    return h.hexdigest()


@dataclass
class Plan:
    n: int
    m: int
    neq: int
    group: np.ndarray            # [n] reduced column of each original column
    weight: np.ndarray           # [n]
    representative: np.ndarray   # [r]
    kept_rows: np.ndarray
    eliminated_rows: np.ndarray
    T: csr_matrix                # [n, r]
    compression: csr_matrix      # [r, n]
    lift: csr_matrix             # [len(eliminated_rows), n]
    fingerprint: str

    @property
    def reduced_variables(self) -> int:
        return len(self.representative)


def plan(problem, max_scale_ratio: float = 1e8) -> Plan:
    """Reduction plan of one LP's equality block (src/lp_equality_reduction.py:295-461)."""
    if isinstance(max_scale_ratio, (bool, np.bool_)) or not np.isfinite(max_scale_ratio) or max_scale_ratio < 1.0:
        raise ValueError("max_scale_ratio must be finite and at least one")
    a, rhs, _, _, _, neq = _normalized(problem)
    m, n = a.shape
    root = list(range(n))

    def find(v):
        while root[v] != v:
            root[v] = root[root[v]]
            v = root[v]
        return v

    edges = {}       # row -> (left col, right col, left coef, right coef)
    for r in range(neq):
        lo, hi = a.indptr[r], a.indptr[r + 1]
        if hi - lo != 2 or rhs[r] != 0.0:
            continue
        mag = np.abs(a.data[lo:hi])
        if mag.min() == 0.0 or mag.max() / mag.min() > max_scale_ratio:
            continue
        left, right = int(a.indices[lo]), int(a.indices[lo + 1])
        rl, rr = find(left), find(right)
        if rl == rr:
            continue                     # a cycle: the row stays explicit
        if rl > rr:
            rl, rr = rr, rl
        root[rr] = rl
        edges[r] = (left, right, float(a.data[lo]), float(a.data[lo + 1]))

    nbrs = {}
    for r, (left, right, _, _) in edges.items():
        nbrs.setdefault(left, []).append((right, r))
        nbrs.setdefault(right, []).append((left, r))

    seen = set()
    trees = []
    for start in sorted(nbrs):
        if start in seen:
            continue
        alpha = {start: 1.0}
        order, rows, todo, ok = [], set(), [start], True
        while todo:
            v = todo.pop()
            if v in seen:
                continue
            seen.add(v)
            order.append(v)
            for u, r in nbrs[v]:
                rows.add(r)
                if u in alpha:
                    continue
                left, right, al, ar = edges[r]
                val = (-al * alpha[v]) / ar if v == left else (-ar * alpha[v]) / al
                if not np.isfinite(val) or val == 0.0:
                    ok = False
                alpha[u] = float(val)
                todo.append(u)
        if not ok or len(order) < 2:
            continue
        mags = np.array([abs(alpha[v]) for v in order])
        if not np.isfinite(mags).all() or mags.max() / mags.min() > max_scale_ratio:
            continue
        top = mags.max()
        rep = min(v for v in order if abs(alpha[v]) == top)
        s = alpha[rep]
        w = {v: alpha[v] / s for v in order}
        w[rep] = 1.0
        trees.append((sorted(order), w, sorted(rows)))

    in_tree = {v for nodes, _, _ in trees for v in nodes}
    comps = trees + [([v], {v: 1.0}, []) for v in range(n) if v not in in_tree]
    comps.sort(key=lambda t: t[0][0])
    group = np.empty(n, np.int64)
    weight = np.empty(n)
    reps, elim = [], []
    for g, (nodes, w, rows) in enumerate(comps):
        reps.append(next(v for v in nodes if w[v] == 1.0))
        elim.extend(rows)
        for v in nodes:
            group[v] = g
            weight[v] = w[v]
    r_count = len(comps)
    T = coo_matrix((weight, (np.arange(n), group)), shape=(n, r_count)).tocsr()
    denom = np.zeros(r_count)
    for v in range(n):                   # ascending column, as T's column sums accumulate
        denom[group[v]] += weight[v] * weight[v]
    compression = coo_matrix((weight / denom[group], (group, np.arange(n))), shape=(r_count, n)).tocsr()
    eliminated = np.array(sorted(elim), dtype=np.int64)
    gone = set(eliminated.tolist())
    kept = np.array([r for r in range(m) if r not in gone], dtype=np.int64)
    lift = _tree_dual_map(n, edges, comps, weight, eliminated)
    p = Plan(n=n, m=m, neq=neq, group=group, weight=weight, representative=np.array(reps, np.int64),
             kept_rows=kept, eliminated_rows=eliminated, T=T, compression=compression, lift=lift,
             fingerprint=equality_fingerprint(a, neq))
    E = a[eliminated]
    null = abs(E @ T)
    if null.nnz and null.max() > 1e-12:
        raise ArithmeticError("primal map is not in the equality null space")
    if len(eliminated):
        err = (lift @ E.T).toarray() - np.eye(len(eliminated))
        if np.abs(err).max() > 1e-10:
            raise ArithmeticError("tree dual map is not a left inverse of the eliminated transpose")
    return p


def _tree_dual_map(n, edges, comps, weight, eliminated):
    """Row e of L expresses the multiplier of eliminated row e through the
    stationarity residual d of the original columns: L E^T = I.  Solved leaf to
    root on a breadth-first tree from the representative
    (src/lp_equality_reduction.py:463-517)."""
    pos = {int(r): i for i, r in enumerate(eliminated)}
    expr = [None] * len(eliminated)
    for nodes, _, rows in comps:
        if not rows:
            continue
        rootv = next(v for v in nodes if weight[v] == 1.0)
        adj = {v: [] for v in nodes}
        for r in rows:
            left, right, _, _ = edges[r]
            adj[left].append((right, r))
            adj[right].append((left, r))
        parent, up = {rootv: None}, {}
        bfs = [rootv]
        for v in bfs:
            for u, r in adj[v]:
                if u in parent:
                    continue
                parent[u] = v
                up[u] = r
                bfs.append(u)
        kids = {v: [] for v in nodes}
        for v in bfs[1:]:
            kids[parent[v]].append(v)
        for v in reversed(bfs[1:]):
            r = up[v]
            left, right, al, ar = edges[r]
            coef = al if v == left else ar
            e = {v: 1.0 / coef}
            for k in kids[v]:
                kr = up[k]
                kl, krt, kal, kar = edges[kr]
                f = -(kal if v == kl else kar) / coef
                for col, val in expr[pos[kr]].items():
                    e[col] = e.get(col, 0.0) + f * val
            expr[pos[r]] = e
    rr, cc, vv = [], [], []
    for i, e in enumerate(expr):
        for col in sorted(e):
            if e[col] != 0.0:
                rr.append(i); cc.append(col); vv.append(e[col])
    return coo_matrix((vv, (rr, cc)), shape=(len(eliminated), n)).tocsr()


def reduce(p: Plan, problem):
    """Reduced LP and bound witnesses of one LP under a fixed plan
    (src/lp_equality_reduction.py:127-160, :519-560).  Returns
    (reduced problem tuple, lower witness, upper witness)."""
    a, rhs, lower, upper, c, neq = _normalized(problem)
    if a.shape != (p.m, p.n) or neq != p.neq:
        raise ValueError("LP dimensions/equality block changed from the reduction plan")
    if equality_fingerprint(a, neq) != p.fingerprint:
        raise ValueError("Equality fingerprint changed from the reduction plan")
    if len(p.eliminated_rows) and np.any(rhs[p.eliminated_rows] != 0.0):
        raise ValueError("An eliminated equality is no longer homogeneous")
    r = p.reduced_variables
    pos = p.weight > 0
    with np.errstate(over="raise", invalid="raise", divide="raise"):
        il = np.where(pos, lower, upper) / p.weight
        iu = np.where(pos, upper, lower) / p.weight
    rl = np.full(r, -np.inf)
    ru = np.full(r, np.inf)
    np.maximum.at(rl, p.group, il)
    np.minimum.at(ru, p.group, iu)
    if np.any(rl > ru):
        raise ValueError("Original bounds are infeasible under eliminated equalities")
    lw = np.full(r, -1, np.int64)
    uw = np.full(r, -1, np.int64)
    for g in range(r):
        mem = np.flatnonzero(p.group == g)
        mem = mem[np.lexsort((mem, -np.abs(p.weight[mem])))]
        for wit, ind, end in ((lw, il, rl[g]), (uw, iu, ru[g])):
            if not np.isfinite(end):
                continue
            hit = mem[ind[mem] == end]
            if hit.size:
                wit[g] = hit[0]
    ra = (a[p.kept_rows] @ p.T).tocsr()
    ra.sum_duplicates()
    ra.eliminate_zeros()
    ra.sort_indices()
    rc = np.asarray(p.T.T @ c).ravel()
    red = (ra, rhs[p.kept_rows].copy(), rl, ru, rc, p.neq - len(p.eliminated_rows))
    return red, lw, uw


def expand(p: Plan, z):
    return np.asarray(np.asarray(z, np.float64) @ p.T.T)


def compress(p: Plan, x):
    return np.asarray(np.asarray(x, np.float64) @ p.compression.T)


def lift(p: Plan, problem, z, y_reduced, lower_witness, upper_witness):
    """Original (x, y) from a reduced (z, y) (src/gpu_reduced_pdhg.py:81-103)."""
    a, _, _, _, c, _ = _normalized(problem)
    x = expand(p, z)
    y = np.zeros(p.m)
    y[p.kept_rows] = y_reduced
    q = c - a.T @ y
    rcost = np.asarray(p.T.T @ q).ravel()
    wit = np.where(rcost >= 0.0, lower_witness, upper_witness)
    wit = np.where(wit >= 0, wit, p.representative)
    normal = np.zeros(p.n)
    normal[wit] = rcost / p.weight[wit]
    if len(p.eliminated_rows):
        y[p.eliminated_rows] = p.lift @ (q - normal)
    return x, y
