"""HiGHS FP64 through SciPy (test infrastructure; see oracle/__init__.py).

The reference's production path solves every LP with HiGHS dual simplex,
``threads=1, parallel='off', solver='simplex'`` (src/cpu_repeated_lp.py:287-296);
its offline teacher calls ``scipy.optimize.linprog(method='highs-ds')`` and
reads row duals from ``eqlin/ineqlin.marginals``
(src/offline_scipy_teacher.py:32-40).  highspy is not installed here; SciPy
vendors HiGHS (version reported by ``version()``).

Pinned by: closed-form known-answer LPs (tests/test_gpu_pdhg_corrector.py:29-95
answers) and brute-force vertex enumeration on tiny LPs (tests/test_oracle.py).
"""

from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np
from scipy.optimize import linprog


def version() -> str:
    try:
        from scipy.optimize._highspy import _core
        return "HiGHS " + str(_core._Highs().version())
    except Exception:  # pragma: no cover - SciPy without the private module
        import scipy
        return "HiGHS (scipy " + scipy.__version__ + ")"


@dataclass
class HighsResult:
    status: str            # "optimal" | "infeasible" | "unbounded" | "error"
    x: np.ndarray | None
    y: np.ndarray | None
    fun: float | None
    seconds: float
    iterations: int


def solve(a, rhs, lower, upper, c, neq, *, presolve=True, tol=1e-9) -> HighsResult:
    """Solve one canonical LP tuple with HiGHS dual simplex (FP64)."""
    m = a.shape[0]
    t0 = time.perf_counter()
    res = linprog(c, A_eq=a[:neq] if neq else None, b_eq=rhs[:neq] if neq else None,
                  A_ub=a[neq:] if neq < m else None, b_ub=rhs[neq:] if neq < m else None,
                  bounds=np.column_stack((lower, upper)), method="highs-ds",
                  options=dict(presolve=presolve, primal_feasibility_tolerance=tol,
                               dual_feasibility_tolerance=tol))
    dt = time.perf_counter() - t0
    status = {0: "optimal", 2: "infeasible", 3: "unbounded"}.get(res.status, "error")
    if res.status != 0:
        return HighsResult(status, None, None, None, dt, int(getattr(res, "nit", 0)))
    y = np.concatenate([np.asarray(res.eqlin.marginals) if neq else np.zeros(0),
                        np.asarray(res.ineqlin.marginals) if neq < m else np.zeros(0)])
    return HighsResult(status, np.asarray(res.x, dtype=np.float64), y, float(res.fun), dt, int(res.nit))
