"""Cooperative lexicographic dFBA stages on the batched engine.

The reference's cooperative community step (src/community_solver.py:1207-1844)
solves, per environment, up to three LPs in sequence:

1. max-min growth: minimise -t subject to the species stoichiometry
   (equalities), the shared-medium rows sum_s biomass_s * stoich * v <= C/dt
   and the growth rows t - mu_k <= 0 (:1114-1149, :1712-1727);
2. optionally (``optimize_live_objectives``) the aggregate live objective
   with t held in [f_c t*, t*] (:1729-1757), f_c = ``coexistence_fraction``;
   otherwise t is fixed at f_c t* and stage 1's answer is kept (:1758-1763);
3. parsimonious exchange (``parsimonious_exchange``): one auxiliary column
   a_i >= 0 per (species, exchange) term with the two rows
   +-scale_i v_i - a_i <= 0, scale_i = max(1e-12, biomass_s) * stoich_i,
   an optional aggregate floor -g.v <= -f_o * aggregate, and the objective
   sum_i a_i (:80-198, :1765-1800).

Here every stage is one batched solve over all environments: stage 1 and 2
share the joint LP's pattern (one engine, updated between stages), stage 3
has its own extended pattern whose per-environment coefficients (the medium
rows and the transfer rows) are per-LP matrix values.  Each stage is
warm-started from the previous one (x extended by |scale_i v_i|, duals of
new rows 0).  An environment whose stage fails keeps no answer (the
reference aborts that environment's step, src/community_solver.py:1747-1749).
"""

from __future__ import annotations

import dataclasses
import time
from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch
from scipy.sparse import coo_matrix

from . import _lib
from .engine import BatchedLpEngine, SolveOptions
from .workloads import Consortium

TRANSFER_STOICH = -1.0   # exchange reactions consume the medium metabolite


@dataclass
class StageOptions:
    """Stage controls with the reference's defaults (src/community_solver.py:550-554)."""

    coexistence_fraction: float = 0.99
    objective_fraction: float = 0.99
    optimize_live_objectives: bool = False
    parsimonious_exchange: bool = True
    # the aggregate and parsimonious-exchange LPs (objectives over the optimal
    # face of stage 1) do not reach the polish threshold from FP32 iterates;
    # they run in FP64
    later_precision: str = "fp64"
    # iteration budget of stages 2 and 3 (None: the solve options' budget)
    later_max_iterations: Optional[int] = None
    # active-set polish in stages 2 and 3 (None: as the solve options say)
    later_polish: Optional[bool] = None
    # PDHG method of stage 1 and of the later (FP64) stages ("restarted" or
    # "halpern"; Halpern iterates need FP64 to reach the polish threshold)
    method: str = "restarted"
    later_method: str = "restarted"
    # exact homogeneous-equality reduction of every stage LP (iterate on the
    # reduced LP, accept on the original LP's certificate): the real-GEM path
    reduce: bool = False


def exchange_terms(model: Consortium):
    """(species, joint column, stoich) of every medium exchange term, in
    medium-metabolite order as the reference iterates ``_exchange_terms``."""
    mask = model.medium_mask if model.medium_mask is not None else np.ones(model.medium_cols.shape, bool)
    return [(s, int(model.medium_cols[s, i]), TRANSFER_STOICH)
            for i in range(model.n_medium) for s in range(model.n_species) if mask[s, i]]


def live_objective(model: Consortium) -> np.ndarray:
    """Sum of the members' live objectives (their biomass fluxes, maximised),
    over the flux columns (src/community_solver.py:1066-1082)."""
    g = np.zeros(model.shape[1] - 1)
    g[model.biomass_cols] = 1.0
    return g


def stage3_pattern(model: Consortium, with_floor: bool):
    """Extended stage-3 matrix (template values), its equality count and the
    CSR positions of the per-environment values: the medium-row coefficients
    (as in the joint LP) followed by the two transfer coefficients of every
    exchange term."""
    A = model.A.tocoo()
    m, n = model.shape
    terms = exchange_terms(model)
    n_aux = len(terms)
    rows, cols, vals = [A.row], [A.col], [A.data]
    r = m
    if with_floor:
        g = live_objective(model)
        idx = np.flatnonzero(g)
        rows.append(np.full(idx.size, r)); cols.append(idx); vals.append(-g[idx])
        r += 1
    tr0 = r
    for i, (_, col, _) in enumerate(terms):
        rows.append(np.array([r, r, r + 1, r + 1]))
        cols.append(np.array([col, n + i, col, n + i]))
        vals.append(np.array([1.0, -1.0, -1.0, -1.0]))
        r += 2
    A3 = coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                    shape=(r, n + n_aux)).tocsr()
    A3.sum_duplicates()
    A3.sort_indices()

    def pos(row, col):
        seg = slice(A3.indptr[row], A3.indptr[row + 1])
        k = np.flatnonzero(A3.indices[seg] == col)
        assert k.size == 1
        return A3.indptr[row] + k[0]

    coupling = np.array([pos(model.medium_rows[i], model.medium_cols[s, i]) for s, i in zip(*model.terms())],
                        dtype=np.int64)
    transfer = np.array([[pos(tr0 + 2 * i, col), pos(tr0 + 2 * i + 1, col)] for i, (_, col, _) in enumerate(terms)],
                        dtype=np.int64)
    return A3, model.neq, coupling, transfer, tr0


class CooperativeStages:
    """All environments' stage LPs as batched solves on MI355X."""

    def __init__(self, model: Consortium, n_env: int, *, precision: str = "fp32",
                 options: Optional[SolveOptions] = None, stages: Optional[StageOptions] = None,
                 device: Optional[torch.device] = None, fallback: bool = True):
        self.model = model
        self.n_env = int(n_env)
        self.precision = precision
        self.options = options or SolveOptions()
        self.stages = stages or StageOptions()
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        self.fallback = fallback
        m, n = model.shape
        self.n_flux = n - 1
        self.g = live_objective(model)
        self.terms = exchange_terms(model)
        self.n_aux = len(self.terms)
        if self.stages.reduce:
            from .reduction import ReducedLpEngine as engine_class
        else:
            engine_class = BatchedLpEngine

        def engine(A, neq, var_positions, prec, method):
            kw = {}
            if self.stages.reduce:
                # species rows keep their rhs in every stage; the rest start homogeneous
                kw["rhs_template"] = np.zeros(A.shape[0])
                kw["rhs_template"][:model.neq] = model.rhs_species
            return engine_class(A, neq, self.n_env, var_positions=var_positions, precision=prec, shared_c=True,
                                device=self.device, method=method, **kw)

        self.joint = engine(model.A, model.neq, model.var_positions, precision, self.stages.method)
        self.joint2 = self.joint
        if self.stages.optimize_live_objectives and self.stages.later_precision != precision:
            self.joint2 = engine(model.A, model.neq, model.var_positions, self.stages.later_precision,
                                 self.stages.later_method)
        self.stage3 = None
        if self.stages.parsimonious_exchange and self.n_aux:
            A3, neq3, coupling3, transfer3, tr0 = stage3_pattern(model, self.stages.optimize_live_objectives)
            self.A3, self.tr0 = A3, tr0
            self.var3 = np.concatenate([coupling3.ravel(), transfer3.ravel()])
            self.stage3 = engine(A3, neq3, self.var3, self.stages.later_precision, self.stages.later_method)
            self.c3 = np.zeros(A3.shape[1])
            self.c3[n:] = 1.0
        self.term_species = torch.as_tensor([s for s, _, _ in self.terms], device=self.device)
        self.term_cols = torch.as_tensor([c for _, c, _ in self.terms], device=self.device)
        self.log: list[dict] = []

    def _solve(self, eng, later: bool = False):
        opts = self.options
        if later and self.stages.later_max_iterations is not None:
            opts = dataclasses.replace(opts, max_iterations=int(self.stages.later_max_iterations))
        if later and self.stages.later_polish is not None:
            opts = dataclasses.replace(opts, polish=bool(self.stages.later_polish))
        if self.fallback and eng.precision == "fp32":
            info, x, y, status = eng.solve_with_fallback(opts)
        else:
            info = eng.solve(opts)
            x, y, status = eng.state()
        return info, x, y, status

    def solve(self, rhs, lower, upper, coupling, biomass, x0=None, y0=None, x3_0=None, y3_0=None,
              omega=None) -> dict:
        """One cooperative step for every environment.

        rhs [B, m], lower/upper [B, n], coupling [B, T] (the present medium
        terms' coefficients -biomass_s), biomass [B, S] are device tensors (as
        ``ConsortiumRollout.lp_inputs`` returns them).  (x0, y0) start stage 1;
        (x3_0, y3_0) start stage 3 instead of stage 1's answer (a previous
        dFBA step's stage-3 solution, projected onto the new bounds); omega
        maps stage names to primal weights to start from.  Returns the chosen
        fluxes x [B, n_flux], t*, the aggregate objective, per-stage status,
        the final iterates and primal weights of each stage ("iterates",
        "omega") and the mask of environments whose every stage certified."""
        t0 = time.perf_counter()
        md, st, dev = self.model, self.stages, self.device
        m, n = md.shape
        B = self.n_env
        out = {"stage_info": {}}
        # stage 1: max-min growth
        omega = omega or {}
        out["iterates"], out["omega"] = {}, {}
        eng = self.joint
        eng.update(md.c, rhs, lower, upper, var_values=coupling)
        eng.reset(x0, y0)
        if x0 is not None and omega.get("max_min") is not None:
            eng.set_primal_weight(omega["max_min"])
        info1, x1, y1, s1 = self._solve(eng)
        out["iterates"]["max_min"] = (x1, y1)
        out["omega"]["max_min"] = eng.primal_weight()
        ok = torch.as_tensor(s1 == _lib.STATUS_OPTIMAL, device=dev)
        t_star = torch.clamp(x1[:, md.t_col], min=0.0)
        floor = t_star * st.coexistence_fraction
        out["stage_info"]["max_min"] = info1
        x, y = x1, y1
        lo2, up2 = lower.clone(), upper.clone()
        if st.optimize_live_objectives:
            lo2[:, md.t_col] = floor
            up2[:, md.t_col] = t_star
            c2 = np.concatenate([-self.g, [0.0]])
            e2 = self.joint2
            e2.update(c2, rhs, lo2, up2, var_values=coupling)
            e2.reset(x1, y1)
            info2, x2, y2, s2 = self._solve(e2, later=True)
            ok &= torch.as_tensor(s2 == _lib.STATUS_OPTIMAL, device=dev)
            x, y = x2, y2
            out["stage_info"]["aggregate_objective"] = info2
        else:
            lo2[:, md.t_col] = floor
            up2[:, md.t_col] = floor
        aggregate = x[:, :self.n_flux] @ torch.as_tensor(self.g, device=dev)
        chosen = x
        if self.stage3 is not None:
            bm = torch.clamp(biomass, min=1e-12)
            scale = bm[:, self.term_species] * TRANSFER_STOICH                  # [B, n_aux]
            transfer = torch.stack([scale, -scale], dim=2).reshape(B, -1)
            var3 = torch.cat([coupling, transfer], dim=1).contiguous()
            extra = self.A3.shape[0] - m
            rhs3 = torch.zeros((B, self.A3.shape[0]), dtype=torch.float64, device=dev)
            rhs3[:, :m] = rhs
            if st.optimize_live_objectives:
                use = aggregate > 1e-9
                # rows whose environment has no aggregate to retain are made vacuous
                rhs3[:, m] = torch.where(use, -aggregate * st.objective_fraction,
                                         torch.full_like(aggregate, 1e6))
            lo3 = torch.cat([lo2, torch.zeros((B, self.n_aux), dtype=torch.float64, device=dev)], dim=1)
            up3 = torch.cat([up2, torch.full((B, self.n_aux), np.inf, dtype=torch.float64, device=dev)], dim=1)
            if x3_0 is not None:
                x3 = torch.minimum(torch.maximum(x3_0, lo3), up3)
                y3 = y3_0
            else:
                xs = torch.minimum(torch.maximum(x, lo2), up2)
                a0 = torch.abs(scale * xs[:, self.term_cols])
                x3 = torch.cat([xs, a0], dim=1)
                y3 = torch.cat([y, torch.zeros((B, extra), dtype=torch.float64, device=dev)], dim=1)
            e3 = self.stage3
            e3.update(self.c3, rhs3, lo3.contiguous(), up3.contiguous(), var_values=var3)
            e3.reset(x3.contiguous(), y3.contiguous())
            if x3_0 is not None and omega.get("parsimonious_exchange") is not None:
                e3.set_primal_weight(omega["parsimonious_exchange"])
            info3, x3, y3, s3 = self._solve(e3, later=True)
            out["iterates"]["parsimonious_exchange"] = (x3, y3)
            out["omega"]["parsimonious_exchange"] = e3.primal_weight()
            ok &= torch.as_tensor(s3 == _lib.STATUS_OPTIMAL, device=dev)
            chosen = x3[:, :n]
            out["stage_info"]["parsimonious_exchange"] = info3
            out["transfer"] = x3[:, n:]
        torch.cuda.synchronize()
        out.update(x=chosen[:, :self.n_flux], t=chosen[:, md.t_col], t_star=t_star, aggregate=aggregate,
                   certified=ok, seconds=time.perf_counter() - t0)
        self.log.append({k: v for k, v in out.items() if k in ("seconds",)})
        return out
