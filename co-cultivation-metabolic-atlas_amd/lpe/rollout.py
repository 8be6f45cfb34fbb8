"""Device-resident dFBA rollout of many environments on the batched engine.

Each step mirrors the reference's dFBA step for the cooperative consortium
(src/dfba_simulator.py:1618 step -> :646-666 uptake bounds -> community LP ->
:1326-1435 Euler medium update), restated on synthetic data:

* uptake bounds  lb = max(-Vmax*C/(Km+C), -1000) on each species' medium
  exchanges (src/dfba_simulator.py:646-666);
* medium rows    sum_s -biomass_s * v_ex(s,i) <= C_i/dt
  (src/community_solver.py:1120-1149); growth rows t - mu_k <= 0 (:1114-1119);
* Euler update   C_i <- max(0, C_i + sum_s v_ex(s,i) * biomass_s * dt)
  (src/dfba_simulator.py:1352,1425-1427);
* growth         mu_s = clip(v_bio_s / FLUX_UNIT, -0.5, 2) per hour and
  X_s <- clip(X_s + mu_s X_s cap dt, 0, K), cap = max(0, 1 - sum_s X_s / K)
  when mu_s > 0 (else 1), K = 20 g/L (src/dfba_simulator.py:1317-1321,
  carrying_capacity default :104).  The reference reads the biomass flux
  of a genome-scale model as mu in 1/h; the synthetic generator's fluxes
  live on the flux-bound scale (|v| <= 1000, optimal biomass fluxes of
  150-700), so they are read in units of FLUX_UNIT = 1000, which gives
  mu = 0.15-0.7 per hour.  Every species is updated from the same pre-step
  total biomass (the reference updates species one after another).

All per-environment state is a torch tensor on the GPU; the LPs of one step
are one batched solve.  Warm starts: "previous" (the last step's primal and
dual), "none" (cold) or a callable (conc, lower, biomass, x_prev, y_prev) ->
(x0, y0) such as the learned model of ``lpe.warmstart`` (which also receives
the engine, already updated to the step's LPs, to score candidate starts).
"""

from __future__ import annotations

import dataclasses
import time
from typing import Optional

import numpy as np
import torch

from . import _lib
from .engine import BatchedLpEngine, SolveOptions
from .workloads import BOUND, UNLIMITED_SUPPLY, Consortium, consortium

FLUX_UNIT = 1000.0           # synthetic flux per unit of growth rate (1/h)
MU_CLIP = (-0.5, 2.0)        # src/dfba_simulator.py:1317
CARRYING_CAPACITY = 20.0     # g/L, src/dfba_simulator.py:104


def active_columns(x: torch.Tensor, lower: torch.Tensor, upper: torch.Tensor, tol: float = 1e-6) -> torch.Tensor:
    """[B, n] bool: variables at a finite bound (within tol * max(1, width))."""
    width = torch.clamp(upper - lower, min=1.0)
    at_lo = torch.isfinite(lower) & ((x - lower) <= tol * width)
    at_up = torch.isfinite(upper) & ((upper - x) <= tol * width)
    return at_lo | at_up


class ConsortiumRollout:
    def __init__(self, model: Optional[Consortium] = None, n_env: int = 4096, *, seed: int = 2,
                 precision: str = "fp32", options: Optional[SolveOptions] = None,
                 warm_start="previous", dt: float = 1.0, conc_range=(1e-2, 10.0),
                 autotune: bool = True, fallback: bool = True, device: Optional[torch.device] = None,
                 method: str = "restarted", reduce: bool = False, flux_unit: float = FLUX_UNIT,
                 capacity: float = CARRYING_CAPACITY, engine=None, tuning: Optional[dict] = None):
        self.model = model if model is not None else consortium(seed)
        self.flux_unit = float(flux_unit)
        self.capacity = float(capacity)
        md = self.model
        self.n_env = int(n_env)
        self.dt = float(dt)
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        self.options = options or SolveOptions()
        self.warm_start = warm_start
        if engine is not None:
            # an existing engine of this model and batch (its graphs and launch
            # shape carry over, e.g. from a warm-up rollout)
            if engine.batch != self.n_env or tuple(engine.A.shape) != tuple(md.shape):
                raise ValueError("the shared engine must match the model's shape and the number of environments")
            self.engine = engine
        elif reduce:
            # exact equality reduction: the engine iterates on the reduced LP,
            # accepts on the original certificate (the real-GEM path)
            from .reduction import ReducedLpEngine
            rhs_template = np.zeros(md.shape[0])
            rhs_template[:md.neq] = md.rhs_species
            self.engine = ReducedLpEngine(md.A, md.neq, self.n_env, var_positions=md.var_positions,
                                          rhs_template=rhs_template, precision=precision, shared_c=True,
                                          device=self.device, method=method)
        else:
            self.engine = BatchedLpEngine(md.A, md.neq, self.n_env, var_positions=md.var_positions,
                                          precision=precision, shared_c=True, device=self.device, method=method)
        rng = np.random.default_rng(seed + 7)
        lo, hi = conc_range
        dev = self.device
        self.conc = torch.as_tensor(np.exp(rng.uniform(np.log(lo), np.log(hi), size=(self.n_env, md.n_medium))),
                                    device=dev)
        self.biomass = torch.as_tensor(rng.uniform(0.05, 0.5, size=(self.n_env, md.n_species)), device=dev)
        self._lower0 = torch.as_tensor(md.lower, device=dev)
        self._upper0 = torch.as_tensor(md.upper, device=dev)
        self._rhs_species = torch.as_tensor(md.rhs_species, device=dev)
        ts, ti = md.terms()
        self._term_species = torch.as_tensor(ts, device=dev, dtype=torch.long)
        self._term_medium = torch.as_tensor(ti, device=dev, dtype=torch.long)
        self._term_cols = torch.as_tensor(md.term_cols, device=dev, dtype=torch.long)
        unl = md.unlimited()
        self._mm_terms = torch.as_tensor(np.flatnonzero(~unl[ti]), device=dev, dtype=torch.long)
        self._unlimited = torch.as_tensor(unl, device=dev)
        self._biomass_cols = torch.as_tensor(md.biomass_cols, device=dev, dtype=torch.long)
        self._vmax = torch.as_tensor(md.vmax, device=dev)
        self._km = torch.as_tensor(md.km, device=dev)
        self.c = md.c.copy()
        self.x = None
        self.y = None
        self.omega = None
        self.autotune = autotune
        self.fallback = fallback
        self.tuning = tuning
        if tuning is not None:
            self.options.check_interval = int(tuning["chunk"])
        self.step_index = 0
        self.log: list[dict] = []
        self.track_active = False   # record per-step active-set changes in the log
        self._active = None
        self._inputs = None         # (rhs, lower, upper) reused by step()

    def lp_inputs(self, conc=None, biomass=None):
        """(rhs [B,m], lower [B,n], upper [B,n], coupling values [B,3*K]) on the device."""
        md = self.model
        conc = self.conc if conc is None else conc
        biomass = self.biomass if biomass is None else biomass
        B = conc.shape[0]
        lower = self._lower0.expand(B, -1).clone()
        t = self._mm_terms
        s, i = self._term_species[t], self._term_medium[t]
        ci = conc[:, i]
        mm = -self._vmax[s, i] * ci / (self._km[s, i] + ci)
        lower[:, self._term_cols[t]] = torch.clamp(mm, min=-BOUND)
        upper = self._upper0.expand(B, -1).contiguous()
        rhs = torch.zeros((B, md.shape[0]), dtype=torch.float64, device=self.device)
        rhs[:, :md.neq] = self._rhs_species
        rhs[:, md.neq:md.neq + conc.shape[1]] = torch.where(self._unlimited, UNLIMITED_SUPPLY, conc / self.dt)
        coupling = -torch.clamp(biomass, min=1e-12)[:, self._term_species]
        return rhs, lower, upper, coupling.contiguous()

    def _step_inputs(self):
        """``lp_inputs()`` for the step, into buffers kept across steps: only
        the Michaelis-Menten uptake bounds and the medium rows change, so they
        are rewritten in place instead of rebuilding the [B, n] and [B, m]
        arrays (the engine copies its inputs at ``update``; ``lp_inputs``
        still returns fresh arrays to other callers)."""
        if self._inputs is None:
            rhs, lower, upper, coupling = self.lp_inputs()
            self._inputs = (rhs, lower, upper)
            return rhs, lower, upper, coupling
        rhs, lower, upper = self._inputs
        md = self.model
        t = self._mm_terms
        s, i = self._term_species[t], self._term_medium[t]
        ci = self.conc[:, i]
        mm = -self._vmax[s, i] * ci / (self._km[s, i] + ci)
        lower[:, self._term_cols[t]] = torch.clamp(mm, min=-BOUND)
        rhs[:, md.neq:md.neq + self.conc.shape[1]] = torch.where(self._unlimited, UNLIMITED_SUPPLY,
                                                                  self.conc / self.dt)
        coupling = -torch.clamp(self.biomass, min=1e-12)[:, self._term_species]
        return rhs, lower, upper, coupling.contiguous()

    def euler(self, x: torch.Tensor, mask: Optional[torch.Tensor] = None):
        """Medium and biomass update from the solved fluxes (skipped for LPs
        that did not certify, whose environment keeps its state)."""
        md = self.model
        terms = torch.zeros((x.shape[0], md.n_species, md.n_medium), dtype=x.dtype, device=x.device)
        terms[:, self._term_species, self._term_medium] = x[:, self._term_cols] * self.biomass[:, self._term_species]
        delta = terms.sum(dim=1) * self.dt                   # species summed in order (deterministic)
        conc = torch.clamp(self.conc + delta, min=0.0)
        mu = torch.clamp(x[:, self._biomass_cols] / self.flux_unit, MU_CLIP[0], MU_CLIP[1])
        cap = torch.clamp(1.0 - self.biomass.sum(dim=1, keepdim=True) / self.capacity, min=0.0)
        dx = mu * self.biomass * torch.where(mu > 0, cap, torch.ones_like(cap)) * self.dt
        biomass = torch.clamp(self.biomass + dx, min=0.0, max=self.capacity)
        if mask is not None:
            conc = torch.where(mask[:, None], conc, self.conc)
            biomass = torch.where(mask[:, None], biomass, self.biomass)
        self.conc, self.biomass = conc, biomass

    def step(self, *, advance: bool = True) -> dict:
        rhs, lower, upper, coupling = self._step_inputs()
        eng = self.engine
        t0 = time.perf_counter()
        eng.update(self.c, rhs, lower, upper, var_values=coupling)
        if self.autotune and self.tuning is None:
            from .autotune import autotune
            self.tuning = autotune(eng)
            self.options.check_interval = int(self.tuning["chunk"])
        x0 = y0 = None
        start_mlp = None
        if callable(self.warm_start):
            if getattr(self.warm_start, "wants_engine", False):
                x0, y0 = self.warm_start(self.conc, lower, self.biomass, self.x, self.y, engine=eng)
                start_mlp = getattr(self.warm_start, "last_choice", None)
            else:
                x0, y0 = self.warm_start(self.conc, lower, self.biomass, self.x, self.y)
        elif self.warm_start == "previous" and self.x is not None:
            x0, y0 = self.x, self.y
        eng.reset(x0, y0)
        if x0 is not None and self.omega is not None:
            eng.set_primal_weight(self.omega)       # warm start of the primal weight
        opts = self.options
        if x0 is not None and not opts.first_check_interval:
            # a warm start is checked (and possibly polished) after half a chunk
            opts = dataclasses.replace(opts, first_check_interval=max(1, opts.check_interval // 2))
        if x0 is not None and not opts.polish_kkt_start:
            # ... and polished at once when it is already as close as an
            # iterate the polish would be tried from: between steps the active
            # set rarely changes, so one polish from the previous solution
            # certifies where 32 PDHG iterations and up to two polishes were
            # spent otherwise (4096 consortium LPs, steps 2-8: 277 instead of
            # 419 ms)
            opts = dataclasses.replace(opts, polish_kkt_start=opts.polish_kkt)
        if self.fallback and eng.precision == "fp32":
            info, x, y, status = eng.solve_with_fallback(opts)
        else:
            info = eng.solve(opts)
            x, y, status = eng.state()
        certified = status == _lib.STATUS_OPTIMAL
        # the engine's work is joined to the caller's stream; waiting on that
        # stream alone lets rollouts on other streams keep the device busy
        torch.cuda.current_stream(self.device).synchronize()
        info["wall_seconds"] = time.perf_counter() - t0
        info["step"] = self.step_index
        info["start_mlp"] = start_mlp
        if self.track_active:
            self._record_active(info, x, y, lower, upper, certified)
        self.log.append(info)
        if advance:
            ok = torch.as_tensor(certified, device=self.device)
            if self.x is None or bool(np.all(certified)):
                self.x, self.y = x, y
            else:
                self.x = torch.where(ok[:, None], x, self.x)
                self.y = torch.where(ok[:, None], y, self.y)
            self.euler(x, ok)
            self.omega = eng.primal_weight()
            self.step_index += 1
        return dict(info, certified_mask=certified, x=x, y=y)

    def _record_active(self, info, x, y, lower, upper, certified):
        """Active set of each certified LP (columns at a bound, inequality rows
        with a nonzero dual) against the previous step's: mean changed
        entries per environment and the share of environments whose set
        changed (environments certified at both steps)."""
        md = self.model
        cols = active_columns(x, lower, upper)
        rows = y[:, md.neq:] < -1e-9
        ok = torch.as_tensor(certified, device=self.device)
        if self._active is not None:
            pc, pr, pok = self._active
            both = ok & pok
            nb = int(both.sum())
            if nb:
                dc = (cols != pc)[both].sum(dim=1)
                dr = (rows != pr)[both].sum(dim=1)
                info["active_changed_columns"] = float(dc.double().mean())
                info["active_changed_rows"] = float(dr.double().mean())
                info["active_changed_share"] = float(((dc + dr) > 0).double().mean())
        info["active_columns"] = float(cols[ok].sum(dim=1).double().mean()) if bool(ok.any()) else None
        self._active = (cols, rows, ok)

    def snapshot(self):
        return self.conc.clone(), self.biomass.clone()

    def restore(self, snap):
        self.conc, self.biomass = snap[0].clone(), snap[1].clone()


class CooperativeRollout:
    """dFBA rollout whose step is the reference's default cooperative step:
    max-min growth, then parsimonious exchange (src/community_solver.py:
    1712-1800, parsimonious_exchange=True by default, :554), with the medium
    and biomass advanced from the chosen stage-3 fluxes (the same Euler and
    growth updates as ``ConsortiumRollout``, src/dfba_simulator.py:1317-1321,
    :1352, :1425-1427).  Stage 1 starts from the previous step's stage-1
    solution and stage 3 from the previous step's stage-3 solution (projected
    onto the new bounds), each with its primal weights; ``warm=False`` starts
    every step cold.  An environment whose step does not certify keeps its
    state (the reference aborts that environment's step)."""

    def __init__(self, model: Optional[Consortium] = None, n_env: int = 4096, *, seed: int = 2,
                 precision: str = "fp32", options: Optional[SolveOptions] = None, stages=None,
                 dt: float = 1.0, conc_range=(1e-2, 10.0), device: Optional[torch.device] = None,
                 fallback: bool = True, warm: bool = True):
        from .stages import CooperativeStages
        model = model if model is not None else consortium(seed)
        self.options = options or SolveOptions()
        self.driver = CooperativeStages(model, n_env, precision=precision, options=self.options, stages=stages,
                                        device=device, fallback=fallback)
        # the state and its dynamics; the stage driver's joint engine is shared
        self.state = ConsortiumRollout(model, n_env, seed=seed, precision=precision, options=self.options,
                                       warm_start="none", dt=dt, conc_range=conc_range, autotune=False,
                                       fallback=fallback, device=device, engine=self.driver.joint)
        self.model = model
        self.n_env = int(n_env)
        self.warm = bool(warm)
        self._prev = None
        self.step_index = 0
        self.log: list[dict] = []

    def lp_inputs(self, conc=None, biomass=None):
        return self.state.lp_inputs(conc, biomass)

    def snapshot(self):
        return self.state.snapshot()

    def step(self) -> dict:
        st = self.state
        rhs, lower, upper, coupling = st.lp_inputs()
        kw = {}
        if self.warm and self._prev is not None:
            (x1, y1), (x3, y3), omega = self._prev
            kw = dict(x0=x1, y0=y1, x3_0=x3, y3_0=y3, omega=omega)
            # a warm start is checked (and possibly polished) at once
            self.driver.options = dataclasses.replace(self.options, polish_kkt_start=self.options.polish_kkt,
                                                      first_check_interval=max(1, self.options.check_interval // 2))
        else:
            self.driver.options = self.options
        t0 = time.perf_counter()
        out = self.driver.solve(rhs, lower, upper, coupling, st.biomass, **kw)
        ok = out["certified"]
        it = out["iterates"]
        if "parsimonious_exchange" in it:
            self._prev = (it["max_min"], it["parsimonious_exchange"], out["omega"])
        st.euler(out["x"], ok)
        torch.cuda.synchronize()
        info = {"step": self.step_index, "certified": int(ok.sum()), "wall_seconds": time.perf_counter() - t0}
        for name, si in out["stage_info"].items():
            info[name] = {k: si.get(k) for k in ("certified", "iterations", "polish_attempts", "cg_steps")}
        self.log.append(info)
        self.step_index += 1
        return dict(info, certified_mask=ok, x=out["x"], t_star=out["t_star"], transfer=out.get("transfer"))
