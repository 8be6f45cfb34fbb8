"""Greenlet microbatching of per-environment LP calls (lpe.microbatch):
scheduling on the CPU with a HiGHS service (ragged call counts, stable IDs,
strict rejection, exceptions), and on the GPU through the batched backend."""

from types import SimpleNamespace

import numpy as np
import pytest
from scipy.optimize import linprog

from lpe.microbatch import YieldingLPBackend, drive_microbatch
from oracle.certificate import relative_objective_error


class HighsService:
    """solve_batch through SciPy HiGHS (the reference's exact CPU path)."""

    name, method = "highs_service", "highs-ds"

    def __init__(self):
        self.calls = []

    def solve_batch(self, requests, *, environment_ids=None):
        self.calls.append(list(environment_ids))
        out = []
        for c, kw in requests:
            r = linprog(c, method="highs-ds", **{k: v for k, v in kw.items() if not k.startswith("_")})
            out.append(SimpleNamespace(success=r.status == 0, x=r.x, fun=r.fun, diagnostics=dict(status=r.status)))
        return out


def _lp(scale, cap):
    """max x0 + scale x1 s.t. x0 + x1 <= cap, 0 <= x <= 1 (optimum closed form)."""
    return np.array([-1.0, -scale]), dict(A_ub=np.array([[1.0, 1.0]]), b_ub=np.array([cap]),
                                          bounds=[(0.0, 1.0), (0.0, 1.0)])


def _environment(k, calls):
    def run(backend):
        assert isinstance(backend, YieldingLPBackend)
        values = []
        for j in range(calls):
            c, kw = _lp(0.5 + 0.1 * k, 0.5 + 0.25 * j)
            r = backend.solve(c, **kw)
            values.append(r.fun)
        return values
    return run


def test_ragged_environments_batched_in_order():
    service = HighsService()
    calls = [3, 1, 2, 3]
    returns, record = drive_microbatch([_environment(k, n) for k, n in enumerate(calls)], service,
                                       environment_ids=["e0", "e1", "e2", "e3"])
    assert record["batches"] == [4, 3, 2] and record["lp_calls"] == sum(calls)
    assert service.calls == [["e0", "e1", "e2", "e3"], ["e0", "e2", "e3"], ["e0", "e3"]]
    for k, vals in enumerate(returns):
        assert len(vals) == calls[k]
        for j, v in enumerate(vals):
            cap, s = 0.5 + 0.25 * j, 0.5 + 0.1 * k
            best = -(max(s, 1.0) * min(cap, 1.0) + min(s, 1.0) * max(0.0, min(cap - 1.0, 1.0)))
            assert abs(v - best) <= 1e-9


def test_strict_rejection_and_environment_errors_unwind():
    service = HighsService()

    def infeasible(backend):
        return backend.solve(np.array([1.0]), A_ub=np.array([[1.0]]), b_ub=np.array([-1.0]), bounds=[(0.0, 1.0)])

    returns, _ = drive_microbatch([infeasible], service)
    assert returns[0].success is False                          # returned to the environment
    with pytest.raises(RuntimeError, match="rejected"):
        drive_microbatch([infeasible, _environment(0, 2)], service, strict=True)

    def broken(backend):
        backend.solve(*_lp(1.0, 1.0)[:1], **_lp(1.0, 1.0)[1])
        raise KeyError("environment failure")

    with pytest.raises(KeyError):
        drive_microbatch([broken, _environment(1, 3)], service)
    with pytest.raises(ValueError, match="unique"):
        drive_microbatch([_environment(0, 1), _environment(1, 1)], service, environment_ids=[1, 1])


@pytest.mark.gpu
def test_gpu_backend_microbatch_matches_highs(device):
    from lpe.backend import GpuLpBackend
    from lpe import workloads
    from oracle import highs
    ub = workloads.uptake_batch(batch=6, seed=5, m=60, n=90)
    backend = GpuLpBackend(precision="fp64")

    def env(b):
        def run(yb):
            out = []
            for step in range(2):
                a, rhs, lo, hi, c, neq = ub.problem(b)
                c = c * (1.0 + 0.5 * step)                  # the second call re-weights the objective
                r = yb.solve(c, A_eq=a[:neq].toarray(), b_eq=rhs[:neq], bounds=list(zip(lo, hi)))
                ref = highs.solve(a, rhs, lo, hi, c, neq)
                out.append((r.success, r.fun, ref.fun))
            return out
        return run

    returns, record = drive_microbatch([env(b) for b in range(ub.batch)], backend)
    assert record["batches"] == [ub.batch, ub.batch]
    for rows in returns:
        for ok, fun, ref in rows:
            assert ok
            assert relative_objective_error(fun, ref) <= 1e-6
    backend.close()
