"""The consortium dFBA rollout over its whole 48-step horizon (dt = 1 h),
the horizon of the reference's rollouts (src/dfba_simulator.py:1326-1435:
uptake bounds, community LP, Euler medium update, growth), on the product
path: FP32 PDHG with FP64 polish and certificate, previous-step starts,
FP64 continuation of stragglers.

Checked at every step: every environment's LP certifies (the FP64
certificate on the device), the medium stays finite and non-negative, and
biomass stays within [0, K].  At steps 0, 24 and 47 1 % of the environments, spread over the batch, are solved
by HiGHS on the LP the step solved: objective within 1e-6 and the device
certificate equal to the host certificate of the returned (x, y).  The
growth-rate LP alone secretes the limited medium metabolites (its optimal
face is left at a vertex), so the medium accumulates over the horizon."""

import numpy as np
import pytest
import torch

from lpe import workloads
from oracle import certificate as cert, highs

pytestmark = pytest.mark.gpu

ENVS = 1024
STEPS = 48
CHECK_STEPS = (0, 24, 47)


def _highs_fun(p):
    r = highs.solve(*p)
    return r.status, r.fun


def _check(problems, x, y, metrics, sample):
    import multiprocessing as mp
    import os
    probs = [problems(b) for b in sample]
    for b, p in zip(sample, probs):
        host = cert.certificate(*p, x[b], y[b])
        assert host["certificate_passed"], (b, host)
        np.testing.assert_allclose(metrics[b, 0], host["primal_residual"], rtol=1e-6, atol=1e-12)
    workers = max(1, min(16, len(os.sched_getaffinity(0)), len(probs)))
    with mp.get_context("spawn").Pool(workers) as pool:
        refs = pool.map(_highs_fun, probs, chunksize=1)
    worst = 0.0
    for b, p, (st, fun) in zip(sample, probs, refs):
        assert st == "optimal", b
        err = cert.relative_objective_error(float(p[4] @ x[b]), fun)
        assert err <= 1e-6, (b, err)
        worst = max(worst, err)
    return worst


def test_consortium_rollout_48_steps(device):
    from lpe.engine import SolveOptions
    from lpe.rollout import CARRYING_CAPACITY, ConsortiumRollout
    ro = ConsortiumRollout(workloads.consortium(2), ENVS, seed=5, precision="fp32", autotune=False,
                           warm_start="previous", device=device, options=SolveOptions(max_iterations=6000))
    ro.track_active = True
    sample = np.unique(np.linspace(0, ENVS - 1, max(3, ENVS // 100)).round().astype(int)).tolist()
    limited = ~ro._unlimited
    conc0 = ro.conc[:, limited].clone()
    per_step = []
    for step in range(STEPS):
        snap = ro.snapshot()
        out = ro.step()
        assert out["certified"] == ENVS, (step, {k: out[k] for k in ("certified", "iterations", "polish_attempts")})
        assert bool(torch.isfinite(ro.conc).all()) and float(ro.conc.min()) >= 0.0, step
        assert float(ro.biomass.min()) >= 0.0 and float(ro.biomass.max()) <= CARRYING_CAPACITY, step
        per_step.append((out["iterations"], out["polish_attempts"], round(out["wall_seconds"] * 1e3, 1)))
        if step in CHECK_STEPS:
            cb = workloads.ConsortiumBatch(model=ro.model, conc=snap[0].cpu().numpy(), biomass=snap[1].cpu().numpy())
            x, y = out["x"], out["y"]
            metrics = ro.engine.certify(x, y).cpu().numpy()
            worst = _check(cb.problem, x.cpu().numpy(), y.cpu().numpy(), metrics, sample)
            print(f"step {step}: HiGHS max rel objective error {worst:.2e} over {len(sample)} LPs", flush=True)
    conc_end = ro.conc[:, limited]
    grown = float((conc_end > conc0).double().mean())
    print("per step (PDHG iterations, polish attempts, ms):", per_step, flush=True)
    print(f"limited medium entries (environment x metabolite) above their initial value after "
          f"{STEPS} steps: {grown:.3f}", flush=True)
    # The growth-rate LP alone leaves the exchange fluxes at a vertex of its
    # optimal face: the synthetic species secrete every medium metabolite up
    # to their flux bounds, so the limited medium accumulates over the horizon
    # and the uptake bounds -Vmax C/(Km + C) saturate (the LPs change less
    # from step to step as the rollout proceeds).
    assert grown > 0.5, grown
    ro.engine.close()
