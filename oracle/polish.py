"""FP64 active-set polish (test infrastructure; see oracle/__init__.py).

Restates the engine's polish (no reference counterpart: the reference fails
closed, src/gpu_pdhg_corrector.py:12-14).  From an iterate (x, y) in original
units: variables within tol*max(1, u-l) of a bound are fixed there; an
inequality row is active when its dual is below -tol_y or its slack is at most
tol_s*max(1,|b|).  With Dr, Dc the equilibration, A^ = Dr A Dc, R and D the 0/1
masks of active rows and free columns, both least-squares problems use
M = R A^ D A^T R and run `cg_iterations` conjugate-gradient steps:

    primal  M w = R Dr (b - A x_N)       x_F = x_N + Dc D A^T w  (then clipped)
    dual    min |D (c^ - A^T u)|, u0 = R y^,  y = R Dr u  (inequalities <= 0)

Pinned by: scipy.linalg.lstsq on the same active-set systems (the CG answer
converges to the least-squares solution) and by the certificate of HiGHS-
identified active sets (tests/test_oracle.py).
"""

from __future__ import annotations

import numpy as np
from scipy.sparse import csr_matrix, diags

from .spmv import csr_matvec_ordered, csr_rmatvec_ordered


def active_set(x, y, a, rhs, lower, upper, neq, dr, tol=1e-6, tol_y=1e-9, tol_s=1e-7, c=None,
               tol_far=1e-2, tol_r=1e-5):
    width = upper - lower
    width = np.where(np.isfinite(width) & (width > 1.0), width, 1.0)
    at_lo = (x - lower) <= tol * width
    at_hi = ~at_lo & ((upper - x) <= tol * width)
    if c is not None:
        # a reduced cost pointing at a nearby bound also fixes the variable there
        r = c - csr_rmatvec_ordered(a, y)
        thr = tol_r * (1.0 + np.abs(c))
        at_lo |= ~at_hi & (r > thr) & ((x - lower) <= tol_far * width)
        at_hi |= ~at_lo & (r < -thr) & ((upper - x) <= tol_far * width)
    xs = np.where(at_lo, lower, np.where(at_hi, upper, x))
    free = ~(at_lo | at_hi)
    m = a.shape[0]
    res = rhs - csr_matvec_ordered(a, xs)
    rows = np.ones(m, dtype=bool)
    if neq < m:
        yo = y[neq:]  # original-unit duals
        rows[neq:] = (yo < -tol_y) | (res[neq:] <= tol_s * np.maximum(1.0, np.abs(rhs[neq:])))
    return xs, free, rows, res


def polish(a, rhs, lower, upper, c, neq, x, y, dr, dc, *, cg_iterations=200, tol=1e-6):
    a = csr_matrix(a, dtype=np.float64)
    m, n = a.shape
    xs, free, rows, res = active_set(x, y, a, rhs, lower, upper, neq, dr, tol, c=c)
    As = diags(dr) @ a @ diags(dc)
    As = csr_matrix(As)
    Rm, Dm = rows.astype(float), free.astype(float)

    def M(v):
        return Rm * (As @ (Dm * (As.T @ (Rm * v))))

    rp = Rm * dr * res
    wp = np.zeros(m)
    pp = rp.copy()
    u = np.where(rows, y / dr, 0.0)
    q = Dm * (c * dc - As.T @ u)
    rd = Rm * (As @ q)
    pd = rd.copy()
    rho = np.array([rp @ rp, rd @ rd])
    W = [wp, u]; Rv = [rp, rd]; Pv = [pp, pd]
    for _ in range(cg_iterations):
        for s in range(2):
            S = M(Pv[s])
            pAp = Pv[s] @ S
            alpha = rho[s] / pAp if (pAp > 0 and np.isfinite(pAp) and rho[s] > 0) else 0.0
            W[s] = W[s] + alpha * Pv[s]
            Rv[s] = Rv[s] - alpha * S
            new = Rv[s] @ Rv[s]
            beta = new / rho[s] if (rho[s] > 0 and np.isfinite(new)) else 0.0
            rho[s] = new
            Pv[s] = Rv[s] + beta * Pv[s]
    xo = xs + np.where(free, dc * (As.T @ W[0]), 0.0)
    xo = np.where(free, np.minimum(np.maximum(xo, lower), upper), xs)
    yo = np.where(rows, dr * W[1], 0.0)
    yo[neq:] = np.minimum(yo[neq:], 0.0)
    return xo, yo, dict(free=int(free.sum()), active_rows=int(rows.sum()))
