"""Original-LP certificate (test infrastructure; see oracle/__init__.py).

Restates the host certificate src/cpu_repeated_lp.py:150-178 and the device
certificate src/gpu_block_lp.py:52-132 (without the optional box-dual and
direct-dual gates, which the engine's acceptance does not use:
src/gpu_pdhg_corrector.py:245-253 passes allow_box_dual=False).

Pinned by: the reference's known answers (tests/test_gpu_block_lp.py:39-62 -
an optimal (x, y) passes, x[2] += 1e-3 fails only its own block, a 1e-3
objective error is caught only when |c.x| is small, non-finite iterates are
rejected) and by HiGHS optima, which must pass (tests/test_oracle.py).
"""

from __future__ import annotations

import numpy as np

GATES = (1e-5, 1e-7, 1e-7)


def certificate(a, rhs, lower, upper, c, neq, x, y, gates=GATES) -> dict:
    """Gates of one LP: primal residual, dual violation, relative KKT gap."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    with np.errstate(invalid="ignore", over="ignore"):
        activity = a @ x
        reduced = c - a.T @ y
        error = activity - rhs
        # np.max propagates NaN like cp.max in the device certificate
        row_eq = np.max(np.abs(error[:neq])) if neq else 0.0
        row_ub = np.max(error[neq:]) if neq < len(rhs) else 0.0
        primal = np.max([0.0, row_eq, row_ub, np.max(lower - x), np.max(x - upper)])
        if np.isnan([row_eq, row_ub, np.max(lower - x), np.max(x - upper)]).any():
            primal = np.nan
        dual_rows = np.max(y[neq:]) if neq < len(rhs) else 0.0
        d1 = np.max(np.where(~np.isfinite(lower), reduced, 0.0))
        d2 = np.max(np.where(~np.isfinite(upper), -reduced, 0.0))
        dual = np.nan if np.isnan([dual_rows, d1, d2]).any() else max(0.0, dual_rows, d1, d2)
        target = np.where(reduced >= 0.0, lower, upper)
        bound_gap = np.sum(np.abs(reduced * (x - np.where(np.isfinite(target), target, x))))
        row_gap = np.sum(np.abs(y[neq:] * (rhs - activity)[neq:]))
        objective = float(np.sum(c * x))
        gap = (bound_gap + row_gap) / max(1.0, abs(objective))
    finite_primal = bool(np.isfinite(x).all() and np.isfinite(activity).all() and np.isfinite(objective))
    finite = finite_primal and bool(np.isfinite(y).all() and np.isfinite(reduced).all())
    passed = bool(finite and np.isfinite(gap) and primal <= gates[0] and dual <= gates[1] and gap <= gates[2])
    return dict(primal_residual=float(primal), dual_violation=float(dual), relative_kkt_gap=float(gap),
                objective=objective, certificate_passed=passed, finite=finite,
                bound_complementarity_gap=float(bound_gap), row_complementarity_gap=float(row_gap))


def relative_objective_error(objective: float, reference: float) -> float:
    """|c.x - c.x*| / max(1, |c.x*|), the north-star agreement measure."""
    return abs(objective - reference) / max(1.0, abs(reference))
