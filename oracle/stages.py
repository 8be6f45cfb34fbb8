"""Cooperative lexicographic stages on the CPU (test infrastructure; see
oracle/__init__.py).

Restates, for one environment's joint LP, the stage sequence of the
reference's cooperative solver: stage 1 max-min growth
(src/community_solver.py:1712-1727), stage 2 with t in
[coexistence_fraction * t*, t*] and the aggregate live objective, or t fixed at
coexistence_fraction * t* (:1729-1763), stage 3 parsimonious exchange built as
src/community_solver.py:116-198 assembles it (existing rows, the optional
aggregate floor -g.v <= -objective_fraction * aggregate, then the rows
scale_i v_i - a_i <= 0 and -scale_i v_i - a_i <= 0 per exchange term,
objective sum a_i, a_i >= 0).  Every stage is solved by HiGHS (oracle/highs.py).

Pinned by: HiGHS optima; the stage-3 invariants (a_i = |scale_i v_i| at the
optimum, t at its fixed floor, the objective equals sum |scale_i v_i|) and a
brute-force check on a tiny joint LP (tests/test_oracle_stages.py).
"""

from __future__ import annotations

import numpy as np
from scipy.sparse import csr_matrix, hstack, vstack

from . import highs


def stage3_lp(a, rhs, lower, upper, neq, terms, biomass, g, t_col, t_bounds, aggregate=None,
              objective_fraction=0.99):
    """Stage-3 LP (A, rhs, lower, upper, c, neq) over [fluxes, t, a].

    terms: (species, column, stoich) per exchange term; g: live objective over
    the flux columns; aggregate: None for no aggregate floor row."""
    a = csr_matrix(a)
    m, n = a.shape
    n_aux = len(terms)
    ext = hstack([a, csr_matrix((m, n_aux))], format="csr")
    rows, b_extra = [], []
    if aggregate is not None:
        idx = np.flatnonzero(g)
        rows.append(csr_matrix((-g[idx], (np.zeros(idx.size, int), idx)), shape=(1, n + n_aux)))
        b_extra.append(-aggregate * objective_fraction)
    tr_r, tr_c, tr_v = [], [], []
    for i, (s, col, stoich) in enumerate(terms):
        scale = max(1e-12, float(biomass[s])) * stoich
        tr_r += [2 * i, 2 * i, 2 * i + 1, 2 * i + 1]
        tr_c += [col, n + i, col, n + i]
        tr_v += [scale, -1.0, -scale, -1.0]
    rows.append(csr_matrix((tr_v, (tr_r, tr_c)), shape=(2 * n_aux, n + n_aux)))
    b_extra += [0.0] * (2 * n_aux)
    A3 = vstack([ext, *rows], format="csr")
    A3.eliminate_zeros()
    lo = np.concatenate([lower, np.zeros(n_aux)])
    hi = np.concatenate([upper, np.full(n_aux, np.inf)])
    lo[t_col], hi[t_col] = t_bounds
    c = np.zeros(n + n_aux)
    c[n:] = 1.0
    return A3, np.concatenate([rhs, b_extra]), lo, hi, c, neq


def cooperative_chain(problem, terms, biomass, g, t_col, *, coexistence_fraction=0.99, objective_fraction=0.99,
                      optimize_live_objectives=False, parsimonious_exchange=True, t_star=None):
    """HiGHS stage sequence for one environment.  t_star: use this max-min
    growth instead of stage 1's (so a GPU chain's later stages can be checked
    on exactly the LP the GPU solved)."""
    a, rhs, lower, upper, c, neq = problem
    out = {}
    r1 = highs.solve(a, rhs, lower, upper, c, neq)
    out["stage1"] = r1
    if r1.status != "optimal":
        return out
    ts = max(0.0, float(r1.x[t_col])) if t_star is None else float(t_star)
    out["t_star"] = ts
    floor = ts * coexistence_fraction
    lo2, hi2 = lower.copy(), upper.copy()
    x = r1.x
    if optimize_live_objectives:
        lo2[t_col], hi2[t_col] = floor, ts
        c2 = np.concatenate([-g, [0.0]])
        r2 = highs.solve(a, rhs, lo2, hi2, c2, neq)
        out["stage2"] = r2
        if r2.status != "optimal":
            return out
        x = r2.x
    else:
        lo2[t_col], hi2[t_col] = floor, floor
    aggregate = float(g @ x[:t_col])
    out["aggregate"] = aggregate
    if parsimonious_exchange and terms:
        agg = aggregate if (optimize_live_objectives and aggregate > 1e-9) else None
        p3 = stage3_lp(a, rhs, lo2, hi2, neq, terms, biomass, g, t_col, (lo2[t_col], hi2[t_col]), agg,
                       objective_fraction)
        out["stage3_problem"] = p3
        out["stage3"] = highs.solve(*p3)
    return out
