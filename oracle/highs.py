"""HiGHS FP64 through SciPy (test infrastructure; see oracle/__init__.py).

The reference's production path solves every LP with HiGHS dual simplex,
``threads=1, parallel='off', solver='simplex'`` (src/cpu_repeated_lp.py:287-296);
its offline teacher calls ``scipy.optimize.linprog(method='highs-ds')`` and
reads row duals from ``eqlin/ineqlin.marginals``
(src/offline_scipy_teacher.py:32-40).  highspy is not installed here; SciPy
vendors HiGHS (version reported by ``version()``).

Pinned by: closed-form known-answer LPs (tests/test_gpu_pdhg_corrector.py:29-95
answers) and brute-force vertex enumeration on tiny LPs (tests/test_oracle.py).
``PersistentHighs`` (the warm, basis-reusing path) is pinned by the cold
``solve`` optimum of every updated LP and by the reference certificate gates.
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


class PersistentHighs:
    """One environment's persistent HiGHS model with hot updates and basis
    reuse: the reference's repeated-LP path (src/cpu_repeated_lp.py).

    * options ``threads=1, parallel='off', solver='simplex',
      simplex_strategy=1`` (dual simplex), no output (:287-296);
    * the model is passed once, row-wise, equality rows with
      ``row_lower = rhs`` and inequality rows with ``row_lower = -inf``
      (:298-313);
    * later LPs of the environment change only what differs: matrix values
      through ``changeCoeff``, costs, row bounds and column bounds (:565-584);
    * the basis saved after the previous solve is restored when the updates
      left HiGHS without a valid one (coefficient changes invalidate the
      factorization, not the combinatorial basis, :640-655).

    Used as the CPU baseline of a dFBA step whose LPs follow the previous
    step's: the timed solve starts from the previous step's basis.
    """

    def __init__(self, a, rhs, lower, upper, c, neq, *, presolve=True, tol=1e-9):
        from scipy.optimize._highspy import _core
        self._core = _core
        a = a.tocsr()
        a.sort_indices()
        self.neq = int(neq)
        self.m, self.n = a.shape
        self.indptr = a.indptr.astype(np.int32)
        self.indices = a.indices.astype(np.int32)
        self.rows = np.repeat(np.arange(self.m, dtype=np.int32), np.diff(self.indptr))
        h = _core._Highs()
        opts = dict(threads=1, parallel="off", solver="simplex", simplex_strategy=1, output_flag=False,
                    presolve="on" if presolve else "off", primal_feasibility_tolerance=tol,
                    dual_feasibility_tolerance=tol)
        for k, v in opts.items():
            h.setOptionValue(k, v)
        lp = _core.HighsLp()
        lp.num_col_, lp.num_row_ = self.n, self.m
        lp.col_cost_ = np.asarray(c, dtype=np.float64)
        lp.col_lower_ = np.asarray(lower, dtype=np.float64)
        lp.col_upper_ = np.asarray(upper, dtype=np.float64)
        rhs = np.asarray(rhs, dtype=np.float64)
        lp.row_lower_ = np.r_[rhs[:self.neq], np.full(self.m - self.neq, -np.inf)]
        lp.row_upper_ = rhs
        lp.a_matrix_.format_ = _core.MatrixFormat.kRowwise
        lp.a_matrix_.start_ = self.indptr
        lp.a_matrix_.index_ = self.indices
        lp.a_matrix_.value_ = a.data.astype(np.float64)
        h.passModel(lp)
        self.h = h
        self.values = a.data.astype(np.float64).copy()
        self.c = lp.col_cost_.copy()
        self.rhs = rhs.copy()
        self.lower = lp.col_lower_.copy()
        self.upper = lp.col_upper_.copy()
        self.basis = None

    def update(self, values, rhs, lower, upper, c=None) -> int:
        """Hot update to the next LP of the environment (same pattern);
        returns the number of changed matrix values."""
        values = np.asarray(values, dtype=np.float64)
        h = self.h
        changed = np.flatnonzero(values != self.values)
        for k in changed:
            h.changeCoeff(int(self.rows[k]), int(self.indices[k]), float(values[k]))
        if c is not None:
            c = np.asarray(c, dtype=np.float64)
            ids = np.flatnonzero(c != self.c).astype(np.int32)
            if ids.size:
                h.changeColsCost(ids.size, ids, c[ids])
            self.c = c.copy()
        rhs = np.asarray(rhs, dtype=np.float64)
        for i in np.flatnonzero(rhs != self.rhs):
            h.changeRowBounds(int(i), float(rhs[i]) if i < self.neq else -np.inf, float(rhs[i]))
        lower = np.asarray(lower, dtype=np.float64)
        upper = np.asarray(upper, dtype=np.float64)
        ids = np.flatnonzero((lower != self.lower) | (upper != self.upper)).astype(np.int32)
        if ids.size:
            h.changeColsBounds(ids.size, ids, lower[ids], upper[ids])
        if self.basis is not None and self.basis.valid and not h.getBasis().valid:
            h.setBasis(self.basis)
        self.values, self.rhs, self.lower, self.upper = values.copy(), rhs.copy(), lower.copy(), upper.copy()
        return int(changed.size)

    def solve(self) -> HighsResult:
        # HiGHS keeps one global task scheduler per process; a linprog call
        # before this one may have sized it for more threads than ``threads=1``
        self.h.resetGlobalScheduler(True)
        t0 = time.perf_counter()
        self.h.run()
        dt = time.perf_counter() - t0
        ms = self.h.getModelStatus()
        name = self.h.modelStatusToString(ms).lower()
        info = self.h.getInfo()
        its = int(getattr(info, "simplex_iteration_count", 0))
        if name != "optimal":
            status = "infeasible" if "infeasible" in name else ("unbounded" if "unbounded" in name else "error")
            return HighsResult(status, None, None, None, dt, its)
        sol = self.h.getSolution()
        self.basis = self.h.getBasis()
        x = np.asarray(sol.col_value, dtype=np.float64)
        y = np.asarray(sol.row_dual, dtype=np.float64)
        return HighsResult("optimal", x, y, float(info.objective_function_value), dt, its)
