"""Infeasibility certificates of an LP from a pair of rays (test
infrastructure; see oracle/__init__.py).

The LP is the canonical tuple (a, rhs, lower, upper, c, neq): min c.x subject
to a[:neq] x = rhs[:neq], a[neq:] x <= rhs[neq:], lower <= x <= upper.  The
restarted PDHG engine tests the iterate differences since the last restart
(dx, dy) as certificates, the test of PDLP (Applegate et al., "Practical
large-scale linear programming using primal-dual hybrid gradient", NeurIPS
2021, sec. 5; cuPDLP, Lu & Yang 2023); this module restates that test in
FP64 NumPy:

* dual ray dy (primal infeasibility, Farkas): with r = -a^T dy, the ray
  objective F = rhs.dy + sum_j (r_j > 0 ? r_j lower_j : r_j upper_j) over
  finite bounds must be positive, and its violations - dy_i > 0 on inequality
  rows, r_j > 0 with lower_j = -inf, r_j < 0 with upper_j = +inf - at most
  eps * F;
* primal ray dx (dual infeasibility, unbounded objective): c.dx < 0 and the
  violations |a_eq dx|, (a_ub dx)+, dx_j < 0 with lower_j finite, dx_j > 0
  with upper_j finite at most eps * |c.dx|.

Pinned by closed-form rays of small infeasible and unbounded LPs whose
status HiGHS confirms, by weak duality (no ray passes for a feasible,
bounded LP: F <= violation * (|slack|_1 + |x|_1) for any feasible x), and
by planted mistakes (a flipped sign of dy or a missing bound term changes
the verdict) in tests/test_oracle_infeasibility.py.
"""

from __future__ import annotations

import numpy as np
from scipy.sparse import csr_matrix


def dual_ray(a, rhs, lower, upper, neq, dy):
    """(F, violation) of a candidate dual ray dy."""
    a = csr_matrix(a)
    dy = np.asarray(dy, dtype=np.float64)
    r = -(a.T @ dy)
    lo_f, up_f = np.isfinite(lower), np.isfinite(upper)
    bt = np.where(r > 0, np.where(lo_f, r * np.where(lo_f, lower, 0.0), 0.0),
                  np.where(r < 0, np.where(up_f, r * np.where(up_f, upper, 0.0), 0.0), 0.0))
    F = float(np.asarray(rhs) @ dy + bt.sum())
    viol = [0.0]
    if dy.size > neq:
        viol.append(float(np.max(dy[neq:])))
    viol.append(float(np.max(np.where((r > 0) & ~lo_f, r, 0.0), initial=0.0)))
    viol.append(float(np.max(np.where((r < 0) & ~up_f, -r, 0.0), initial=0.0)))
    return F, max(viol)


def primal_ray(a, c, lower, upper, neq, dx):
    """(c.dx, violation) of a candidate primal ray dx."""
    a = csr_matrix(a)
    dx = np.asarray(dx, dtype=np.float64)
    adx = a @ dx
    viol = [0.0]
    if neq:
        viol.append(float(np.max(np.abs(adx[:neq]))))
    if adx.size > neq:
        viol.append(float(np.max(adx[neq:])))
    viol.append(float(np.max(np.where(np.isfinite(lower) & (dx < 0), -dx, 0.0), initial=0.0)))
    viol.append(float(np.max(np.where(np.isfinite(upper) & (dx > 0), dx, 0.0), initial=0.0)))
    return float(np.asarray(c) @ dx), max(viol)


def certificate_status(a, rhs, lower, upper, c, neq, dx, dy, eps=1e-8):
    """'primal_infeasible', 'dual_infeasible' or None for the ray pair."""
    F, dv = dual_ray(a, rhs, lower, upper, neq, dy)
    if np.isfinite(F) and np.isfinite(dv) and F > 0 and dv <= eps * F:
        return "primal_infeasible"
    cdx, pv = primal_ray(a, c, lower, upper, neq, dx)
    if np.isfinite(cdx) and np.isfinite(pv) and cdx < 0 and pv <= eps * (-cdx):
        return "dual_infeasible"
    return None
