"""Drop-in for the reference's FP64 PDHG corrector on the MI355X engine.

``GpuPdhgCorrector`` keeps the constructor, ``solve`` signature, result keys,
validation messages and fail-closed semantics of
src/gpu_pdhg_corrector.py:75-438: fixed-metric PDHG (Pock-Chambolle diagonal
metric, theta extrapolation, primal weight), certificate at iteration 0, at
every ``check_interval`` and at the last iteration, accepted LPs frozen, an
exhausted budget reported as ``accepted=False``.  The arithmetic runs in
liblpe's FP64 kernels (method "fixed"); the LPs may differ in every value of
their shared sparsity pattern.
"""

from __future__ import annotations

import time
from typing import Any, Iterable

import numpy as np
import torch
from scipy.sparse import csr_matrix

from .engine import BatchedLpEngine


def _validated_problem(problem: Any):
    """Copy and validate one canonical tuple (src/gpu_pdhg_corrector.py:35-72)."""
    if not isinstance(problem, (tuple, list)) or len(problem) != 6:
        raise ValueError("Each problem must be an (a, rhs, lower, upper, c, neq) tuple")
    supplied_a, supplied_rhs, supplied_lower, supplied_upper, supplied_c, neq = problem
    a = csr_matrix(supplied_a, dtype=np.float64, copy=True)
    a.sum_duplicates()
    a.eliminate_zeros()
    a.sort_indices()
    m, n = a.shape
    if n < 1 or not np.isfinite(a.data).all():
        raise ValueError("The LP matrix must have finite coefficients and columns")
    if isinstance(neq, (bool, np.bool_)) or not isinstance(neq, (int, np.integer)):
        raise ValueError("neq must be an integer")
    neq = int(neq)
    if not 0 <= neq <= m:
        raise ValueError("neq must identify an initial block of equality rows")
    vectors = []
    for name, supplied, shape, finite in (("rhs", supplied_rhs, (m,), True), ("lower", supplied_lower, (n,), False),
                                          ("upper", supplied_upper, (n,), False),
                                          ("objective", supplied_c, (n,), True)):
        value = np.asarray(supplied, dtype=np.float64)
        if value.shape != shape:
            raise ValueError(f"{name} has the wrong shape")
        if np.isnan(value).any() or (finite and not np.isfinite(value).all()):
            raise ValueError(f"{name} contains invalid values")
        vectors.append(value.copy())
    rhs, lower, upper, c = vectors
    if np.any(lower > upper):
        raise ValueError("Variable lower bounds cannot exceed upper bounds")
    return a, rhs, lower, upper, c, neq


def shared_pattern(matrices, all_entries: bool = False):
    """Union pattern of same-shape CSR matrices, the template values (LP 0)
    and the positions / per-LP values of every entry that differs (of every
    entry when ``all_entries``, so any later values fit the pattern)."""
    shape = matrices[0].shape
    union = abs(matrices[0])
    for a in matrices[1:]:
        union = union + abs(a)
    union = csr_matrix(union)
    union.sort_indices()
    union.data[:] = 1.0
    rows = np.repeat(np.arange(shape[0]), np.diff(union.indptr))
    vals = [np.asarray(csr_matrix(a)[rows, union.indices]).ravel() if union.nnz else np.zeros(0)
            for a in matrices]
    vals = np.stack(vals)
    varying = np.arange(vals.shape[1]) if all_entries else np.flatnonzero(np.any(vals != vals[0], axis=0))
    template = csr_matrix((vals[0].copy(), union.indices.copy(), union.indptr.copy()), shape=shape)
    return template, varying, vals[:, varying]


def pattern_values(template, matrices):
    """Values of each matrix at the template's entries ([len(matrices), nnz])."""
    rows = np.repeat(np.arange(template.shape[0]), np.diff(template.indptr))
    if not template.nnz:
        return np.zeros((len(matrices), 0))
    return np.stack([np.asarray(csr_matrix(a)[rows, template.indices]).ravel() for a in matrices])


class GpuPdhgCorrector:
    """Correct independent same-shape LP proposals on one MI355X.

    Parameters mirror src/gpu_pdhg_corrector.py:75-106.
    """

    name = "mi355x_fp64_batched_pdhg_original_lp_corrector"
    # per-LP coefficient positions of the engine: "varying" (entries that
    # differ in the constructor's batch) or "all" (values may change later)
    per_lp_entries = "varying"

    def __init__(self, problems: Iterable[Any], *, step_safety: float = 0.99, theta: float = 1.0,
                 primal_weight: float = 1.0) -> None:
        started = time.perf_counter()
        problems = list(problems)
        if not problems:
            raise ValueError("At least one independent LP is required")
        if not np.isfinite(step_safety) or not 0.0 < float(step_safety) < 1.0:
            raise ValueError("step_safety must be finite and strictly between zero and one")
        if not np.isfinite(theta) or not 0.0 <= float(theta) <= 1.0:
            raise ValueError("theta must be finite and between zero and one")
        if (isinstance(primal_weight, (bool, np.bool_)) or not np.isfinite(primal_weight)
                or float(primal_weight) <= 0.0):
            raise ValueError("primal_weight must be finite and strictly positive")
        normalized = tuple(_validated_problem(p) for p in problems)
        shape, neq = normalized[0][0].shape, normalized[0][-1]
        if any(p[0].shape != shape or p[-1] != neq for p in normalized):
            raise ValueError("A PDHG batch requires identical dimensions and equality counts")
        self.problems = normalized
        self.batch = len(normalized)
        self.m, self.n = shape
        self.neq = neq
        self.step_safety, self.theta, self.primal_weight = float(step_safety), float(theta), float(primal_weight)
        template, varying, values = shared_pattern([p[0] for p in normalized], self.per_lp_entries == "all")
        self._template = template
        host_assembly_seconds = time.perf_counter() - started
        t = time.perf_counter()
        self.engine = BatchedLpEngine(template, neq, self.batch, var_positions=varying, method="fixed",
                                      step_safety=step_safety, theta=theta, primal_weight=primal_weight)
        self.engine.update(np.stack([p[4] for p in normalized]), np.stack([p[1] for p in normalized]),
                           np.stack([p[2] for p in normalized]), np.stack([p[3] for p in normalized]),
                           var_values=values if varying.size else None)
        torch.cuda.synchronize()
        gpu_setup_seconds = time.perf_counter() - t
        # diagonal metric (src/gpu_pdhg_corrector.py:185-196), exposed flattened
        sig, tau = [], []
        for a, *_ in normalized:
            absolute = abs(a)
            rs = np.asarray(absolute.sum(axis=1)).ravel()
            cs = np.asarray(absolute.sum(axis=0)).ravel()
            sig.append(np.where(rs > 0.0, self.step_safety / np.where(rs > 0, rs, 1.0), 1.0) * self.primal_weight)
            tau.append(np.where(cs > 0.0, self.step_safety / np.where(cs > 0, cs, 1.0), 1.0) / self.primal_weight)
        self.sigma, self.tau = np.concatenate(sig), np.concatenate(tau)
        self.setup_timing = dict(host_assembly_seconds=host_assembly_seconds, gpu_transfer_seconds=0.0,
                                 gpu_setup_seconds=gpu_setup_seconds,
                                 setup_total_seconds=time.perf_counter() - started,
                                 primal_weight=self.primal_weight)
        self.history: list[dict[str, Any]] = []

    @staticmethod
    def _valid_budget(value: Any, name: str, *, allow_zero: bool) -> int:
        if isinstance(value, (bool, np.bool_)) or not isinstance(value, (int, np.integer)):
            raise ValueError(f"{name} must be an integer")
        value = int(value)
        if value < (0 if allow_zero else 1):
            raise ValueError(f"{name} must be {'non-negative' if allow_zero else 'positive'}")
        return value

    def _warm(self, supplied, shape, name):
        if supplied is None:
            return None
        value = supplied if isinstance(supplied, torch.Tensor) else np.asarray(supplied, dtype=np.float64)
        value = torch.as_tensor(value, dtype=torch.float64)
        flat = (int(np.prod(shape)),)
        if tuple(value.shape) == flat:
            value = value.reshape(shape)
        elif tuple(value.shape) != shape:
            raise ValueError(f"{name} must have shape {shape} or {flat}")
        if not bool(torch.isfinite(value).all()):
            raise ValueError(f"{name} must contain only finite values")
        return value

    def _certificate(self, x=None, y=None):
        metrics = self.engine.certify(x, y).cpu().numpy()
        return [dict(primal_residual=float(r[0]), dual_violation=float(r[1]), relative_kkt_gap=float(r[2]),
                     objective=float(r[3]), certificate_passed=bool(r[4]),
                     certificate_source="solver_row_dual", solver_dual_relative_kkt_gap=float(r[2]))
                for r in metrics]

    def _checkpoint(self, iteration, metrics, elapsed):
        accepted = [bool(r["certificate_passed"]) for r in metrics]
        return dict(iteration=int(iteration), accepted_count=int(sum(accepted)), batch=len(metrics),
                    primal_residual_maximum=float(max((r["primal_residual"] for r in metrics), default=np.inf)),
                    dual_violation_maximum=float(max((r["dual_violation"] for r in metrics), default=np.inf)),
                    relative_kkt_gap_maximum=float(max((r["relative_kkt_gap"] for r in metrics), default=np.inf)),
                    certificate_seconds=float(elapsed), primal_weight=self.primal_weight)

    def solve(self, initial_x=None, initial_y=None, *, iterations: int = 1000, check_interval: int = 50):
        iterations = self._valid_budget(iterations, "iterations", allow_zero=True)
        check_interval = self._valid_budget(check_interval, "check_interval", allow_zero=False)
        solve_started = time.perf_counter()
        x0 = self._warm(initial_x, (self.batch, self.n), "initial_x")
        y0 = self._warm(initial_y, (self.batch, self.m), "initial_y")
        t = time.perf_counter()
        self.engine.reset(x0, y0)
        torch.cuda.synchronize()
        warm_seconds = time.perf_counter() - t
        certificate_seconds = correction_seconds = 0.0
        t = time.perf_counter()
        metrics = self._certificate()
        elapsed = time.perf_counter() - t
        certificate_seconds += elapsed
        checkpoints = [self._checkpoint(0, metrics, elapsed)]
        accepted = np.array([r["certificate_passed"] for r in metrics], dtype=bool)
        accepted_iteration = np.where(accepted, 0, -1).astype(np.int64)
        if accepted.any():
            self.engine.set_accepted(accepted)
        done = 0
        budget = 0 if accepted.all() else iterations
        while done < budget:
            k = min(check_interval - done % check_interval, budget - done)
            t = time.perf_counter()
            self.engine.run(k, chunk=min(64, max(k, 1)))
            torch.cuda.synchronize()
            correction_seconds += time.perf_counter() - t
            done += k
            t = time.perf_counter()
            metrics = self._certificate()
            elapsed = time.perf_counter() - t
            certificate_seconds += elapsed
            checkpoints.append(self._checkpoint(done, metrics, elapsed))
            now = np.array([r["certificate_passed"] for r in metrics], dtype=bool)
            newly = ~accepted & now
            accepted_iteration[newly] = done
            accepted |= now
            if newly.any():
                self.engine.set_accepted(accepted)
            if accepted.all():
                break
        x, y, _ = self.engine.state()
        for i, row in enumerate(metrics):
            row["accepted_iteration"] = int(accepted_iteration[i])
            row["success"] = bool(accepted[i])
            row["solver"] = self.name
        total = time.perf_counter() - solve_started
        timing = dict(self.setup_timing)
        timing.update(warm_start_transfer_seconds=warm_seconds, correction_seconds=correction_seconds,
                      certificate_seconds=certificate_seconds,
                      host_control_seconds=max(0.0, total - warm_seconds - correction_seconds - certificate_seconds),
                      solve_total_seconds=total)
        result = dict(x=x, y=y, accepted=accepted, all_accepted=bool(accepted.all()), iterations_run=int(done),
                      metrics=metrics, checkpoints=checkpoints, timing=timing, cpu_lp_calls=0,
                      primal_weight=self.primal_weight,
                      scope="FP64 MI355X PDHG on unchanged original LPs; strict original-unit certificate; "
                            "fixed-budget failures remain rejected")
        self.history.append(dict(batch=self.batch, variables_per_problem=self.n, constraints_per_problem=self.m,
                                 nonzeros=int(sum(p[0].nnz for p in self.problems)), accepted=accepted.tolist(),
                                 all_accepted=bool(accepted.all()), iterations_run=int(done),
                                 metrics=[dict(r) for r in metrics], checkpoints=[dict(c) for c in checkpoints],
                                 timing=dict(timing), cpu_lp_calls=0, primal_weight=self.primal_weight))
        return result
