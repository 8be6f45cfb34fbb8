"""The reference's LP backend surfaces served by the MI355X engine.

* ``solve(c, **linprog_kwargs)`` - the per-LP backend called by the
  community solver (src/community_solver.py:1200-1205): returns
  ``SimpleNamespace(success, x, fun, message, ...)``; failure is
  ``success=False`` (the caller aborts that environment's step).
* ``solve_batch(requests, environment_ids=None)`` - the batched backend
  (src/cpu_repeated_lp.py:783-818): one unique stable environment ID per
  request, calls serialised, results in request order with ``diagnostics``
  holding the certificate (``certificate_passed``, ``primal_residual``,
  ``dual_violation``, ``relative_kkt_gap``, ``stage``, ...).

Requests of one stage are solved as one shared-pattern batch: LPs of other
sizes are padded to the group's largest row / column counts (empty rows,
columns fixed at zero) and the union of their patterns is the batch's
pattern, with per-LP coefficients where values differ (the reference's
corrector takes same-shape LPs only, src/gpu_pdhg_corrector.py:120-123).  Every answer is the engine's FP64 certificate; an LP that
does not certify within the budget is reported ``success=False`` (no CPU
fallback).  With ``warm_start`` each environment's previous primal/dual of
the same stage seeds its next solve (src/gpu_resident_reduced_pdhg.py:104-118).
"""

from __future__ import annotations

import time
from types import SimpleNamespace

import numpy as np

from . import _lib
from scipy.sparse import csr_matrix

from .corrector import shared_pattern
from .engine import BatchedLpEngine, SolveOptions
from .requests import normalize, pattern_key


class GpuLpBackend:
    name = "mi355x_batched_lp"

    def __init__(self, *, precision: str = "fp32", polish: bool = True, max_iterations: int = 20000,
                 check_interval: int = 64, warm_start: bool = True, launch: tuple | None = None,
                 pdhg: str = "restarted", mixed_shapes: bool = True):
        if pdhg not in ("restarted", "halpern"):
            raise ValueError("pdhg must be 'restarted' or 'halpern'")
        self.precision = precision
        self.pdhg = pdhg
        self.method = ("restarted_averaged_pdhg_" if pdhg == "restarted" else "restarted_halpern_pdhg_") + precision \
            + ("_polish" if polish else "")
        self.options = SolveOptions(max_iterations=max_iterations, check_interval=check_interval, polish=polish)
        if launch:
            self.options.rows_per_wave, self.options.lps_per_lane = launch
        self.warm_start = bool(warm_start)
        self.mixed_shapes = bool(mixed_shapes)
        self._engines: dict = {}
        self._warm: dict = {}
        self.history: list[dict] = []
        self.closed = False

    # ---------------------------------------------------------------- engines
    MAX_ENGINES = 4

    def _engine_for(self, key, template, varying, batch, shared_c):
        """The cached engine of this pattern and batch size; at most
        MAX_ENGINES stay resident (least recently used ones are closed, so a
        microbatched rollout whose pending batch shrinks does not keep one
        engine per size)."""
        cache_key = (key, tuple(varying.tolist()), batch, shared_c)
        eng = self._engines.pop(cache_key, None)
        if eng is not None:
            self._engines[cache_key] = eng            # most recently used last
        else:
            while len(self._engines) >= self.MAX_ENGINES:
                old = next(iter(self._engines))
                self._engines.pop(old).close()
            a, neq = template
            eng = BatchedLpEngine(a, neq, batch, var_positions=varying, precision=self.precision,
                                  shared_c=shared_c, method=self.pdhg)
            self._engines[cache_key] = eng
        return eng

    @staticmethod
    def _pad_to_common(problems):
        """LPs of different sizes as one shape: equality rows, inequality
        rows and columns each padded to the group's largest count (empty
        rows with zero right-hand side, columns fixed at 0 with zero cost),
        which leaves every LP's optimum, duals and certificate unchanged.
        Returns the padded problems and each LP's (neq, n_ineq, n)."""
        dims = [(p[5], p[0].shape[0] - p[5], p[0].shape[1]) for p in problems]
        E, U, N = (max(d[k] for d in dims) for k in range(3))
        out = []
        for (a, rhs, lo, hi, c, neq), (e, u, n) in zip(problems, dims):
            a = csr_matrix(a)
            rows = np.r_[np.arange(e), E + np.arange(u)]
            coo = a.tocoo()
            pa = csr_matrix((coo.data, (rows[coo.row], coo.col)), shape=(E + U, N))
            prhs = np.zeros(E + U)
            prhs[rows] = rhs
            plo, phi, pc = np.zeros(N), np.zeros(N), np.zeros(N)
            plo[:n], phi[:n], pc[:n] = lo, hi, c
            out.append((pa, prhs, plo, phi, pc, E))
        return out, dims

    def _solve_group(self, problems, env_ids, stage):
        dims = None
        if len({(p[0].shape, p[5]) for p in problems}) > 1:
            problems, dims = self._pad_to_common(problems)
        a0, _, _, _, c0, neq = problems[0]
        template, varying, values = shared_pattern([p[0] for p in problems])
        B = len(problems)
        shared_c = all(np.array_equal(p[4], c0) for p in problems)
        key = pattern_key(template, neq)
        eng = self._engine_for(key, (template, neq), varying, B, shared_c)
        c = c0 if shared_c else np.stack([p[4] for p in problems])
        eng.update(c, np.stack([p[1] for p in problems]), np.stack([p[2] for p in problems]),
                   np.stack([p[3] for p in problems]), var_values=values if varying.size else None)
        x0 = y0 = None
        if self.warm_start:
            prev = [self._warm.get((e, stage, key)) for e in env_ids]
            if all(p is not None for p in prev):
                x0 = np.stack([p[0] for p in prev]); y0 = np.stack([p[1] for p in prev])
        eng.reset(x0, y0)
        if self.precision == "fp32":
            # FP32 stragglers continue on an FP64 engine from their FP32 iterate
            info, x, y, status = eng.solve_with_fallback(self.options)
        else:
            info = eng.solve(self.options)
            x, y, status = eng.state()
        metrics = eng.certify(x, y).cpu().numpy()
        x, y = x.cpu().numpy(), y.cpu().numpy()
        xw, yw = x, y                     # the engine's layout, kept for warm starts
        if dims is not None:
            # back to each LP's own rows and columns
            E = neq
            x = [x[b, :d[2]] for b, d in enumerate(dims)]
            y = [np.r_[y[b, :d[0]], y[b, E:E + d[1]]] for b, d in enumerate(dims)]
        out = []
        for b, (p, e) in enumerate(zip(problems, env_ids)):
            ok = bool(status[b] == _lib.STATUS_OPTIMAL and metrics[b, 4] == 1.0)
            diag = dict(certificate_passed=ok, primal_residual=float(metrics[b, 0]),
                        dual_violation=float(metrics[b, 1]), relative_kkt_gap=float(metrics[b, 2]),
                        objective=float(metrics[b, 3]), status=_lib.STATUS_NAMES[int(status[b])],
                        stage=stage, iterations=int(info["iterations"]), warm_started=x0 is not None,
                        cpu_solver_runs=0, numerical_retry_count=0)
            if ok and self.warm_start:
                self._warm[(e, stage, key)] = (xw[b].copy(), yw[b].copy())
            if ok:
                message = "optimal"
            elif int(status[b]) in (_lib.STATUS_PRIMAL_INFEASIBLE, _lib.STATUS_DUAL_INFEASIBLE):
                message = "model status " + diag["status"] + " (ray certificate)"
            else:
                message = "original LP certificate failed: " + diag["status"]
            out.append(SimpleNamespace(success=ok, x=x[b].copy() if ok else None,
                                       fun=float(metrics[b, 3]) if ok else None, nit=int(info["iterations"]),
                                       diagnostics=diag, message=message))
        return out, info

    # ------------------------------------------------------------- interfaces
    def solve_batch(self, requests, *, environment_ids=None):
        if self.closed:
            raise RuntimeError("The batched LP backend is closed")
        requests = list(requests)
        if environment_ids is None:
            environment_ids = list(range(len(requests)))
        environment_ids = list(environment_ids)
        if len(environment_ids) != len(requests) or len(set(environment_ids)) != len(environment_ids):
            raise ValueError("One unique stable environment ID is required per request")
        started = time.perf_counter()
        problems = [normalize(c, kw) for c, kw in requests]      # validate everything first
        stages = [kw.get("_stage") for _, kw in requests]
        groups: dict = {}
        for i, (p, st) in enumerate(zip(problems, stages)):
            # LPs of one stage share a launch; sizes are padded to a common shape
            groups.setdefault(st if self.mixed_shapes else (p[0].shape, p[5], st), []).append(i)
        results = [None] * len(requests)
        infos = []
        for key, idx in groups.items():
            stage = key if self.mixed_shapes else key[2]
            res, info = self._solve_group([problems[i] for i in idx], [environment_ids[i] for i in idx], stage)
            infos.append(info)
            for i, r in zip(idx, res):
                results[i] = r
        self.history.append(dict(batch=len(requests), groups=len(groups), cpu_lp_calls=0,
                                 seconds=time.perf_counter() - started, solve_info=infos,
                                 rows=[r.diagnostics for r in results]))
        return results

    def solve(self, c, **kwargs):
        """Per-LP backend (src/community_solver.py:1200-1205)."""
        return self.solve_batch([(c, kwargs)], environment_ids=[kwargs.pop("_environment_id", 0)])[0]

    def clear_models(self):
        for eng in self._engines.values():
            eng.close()
        self._engines.clear()
        self._warm.clear()

    def close(self):
        if not self.closed:
            self.clear_models()
            self.closed = True
