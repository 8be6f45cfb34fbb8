"""The workload configurations at their full sizes on the GPU: the 256-LP
uptake batch on the 1000 x 1500 species LP, and two steps of the
4096-environment consortium dFBA rollout (3663 x 5301 joint LP).  Properties
checked at full size: every LP certifies; for at least 1 % of each batch
(spread over it) the device certificate equals the oracle's host certificate
of the returned (x, y) and the objective is within 1e-6 of HiGHS; the
rollout's second step (warm started from the first) stays certified."""

import numpy as np
import pytest
import torch

from lpe import workloads
from oracle import certificate as cert, highs

pytestmark = pytest.mark.gpu

REL_OBJ_TOL = 1e-6


def _highs_fun(p):
    r = highs.solve(*p)
    return r.status, r.fun


def _sample(batch):
    """At least 1 % of the batch, spread over it (first and last included)."""
    k = max(3, -(-batch // 100))
    return np.unique(np.linspace(0, batch - 1, k).round().astype(int)).tolist()


def _check_sample(problems, x, y, metrics, sample):
    import multiprocessing as mp
    import os
    probs = [problems(b) for b in sample]
    for b, p in zip(sample, probs):
        host = cert.certificate(*p, x[b], y[b])
        assert host["certificate_passed"], (b, host)
        np.testing.assert_allclose(metrics[b, 0], host["primal_residual"], rtol=1e-6, atol=1e-12)
        np.testing.assert_allclose(metrics[b, 2], host["relative_kkt_gap"], rtol=1e-6, atol=1e-13)
    try:
        workers = max(1, min(16, len(os.sched_getaffinity(0)), len(probs)))
    except AttributeError:
        workers = 1
    with mp.get_context("spawn").Pool(workers) as pool:
        refs = pool.map(_highs_fun, probs, chunksize=1)
    for b, p, (st, fun) in zip(sample, probs, refs):
        assert st == "optimal", b
        assert cert.relative_objective_error(float(p[4] @ x[b]), fun) <= REL_OBJ_TOL, b


def test_uptake_batch_256_full_size(device):
    from lpe.engine import BatchedLpEngine, SolveOptions
    ub = workloads.uptake_batch(batch=256, seed=1)
    assert ub.S.shape == (1000, 1500)
    eng = BatchedLpEngine(ub.S, ub.neq, ub.batch, precision="fp32", shared_c=True, device=device)
    eng.update(ub.c, np.tile(ub.rhs, (ub.batch, 1)), ub.lower, ub.upper)
    eng.reset()
    info = eng.solve(SolveOptions(max_iterations=20000))
    assert info["certified"] == ub.batch, info
    x, y, status = eng.state()
    metrics = eng.certify(x, y).cpu().numpy()
    assert (metrics[:, 4] > 0.5).all()
    _check_sample(ub.problem, x.cpu().numpy(), y.cpu().numpy(), metrics, sample=_sample(ub.batch))
    eng.close()


def test_consortium_rollout_4096_two_steps(device):
    from lpe.engine import SolveOptions
    from lpe.rollout import ConsortiumRollout
    ro = ConsortiumRollout(workloads.consortium(2), 4096, seed=2, precision="fp32", autotune=False,
                           warm_start="previous", device=device, options=SolveOptions(max_iterations=6000))
    assert ro.model.shape == (3663, 5301)
    for step in range(2):
        snap = ro.snapshot()
        out = ro.step()
        assert out["certified"] == 4096, (step, out)
        cb = workloads.ConsortiumBatch(model=ro.model, conc=snap[0].cpu().numpy(), biomass=snap[1].cpu().numpy())
        x, y = out["x"], out["y"]
        metrics = ro.engine.certify(x, y).cpu().numpy()
        assert (metrics[:, 4] > 0.5).all()
        if step == 1:
            _check_sample(cb.problem, x.cpu().numpy(), y.cpu().numpy(), metrics, sample=_sample(4096))
    ro.engine.close()
