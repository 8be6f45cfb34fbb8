"""The workload configurations at their full sizes on the GPU: the 256-LP
uptake batch on the 1000 x 1500 species LP, and two steps of the
4096-environment consortium dFBA rollout (3663 x 5301 joint LP).  Properties
checked at full size: every LP certifies, the device certificate of sampled
LPs equals the oracle's host certificate of the returned (x, y), sampled
objectives are within 1e-6 of HiGHS, and the rollout's second step (warm
started from the first) stays certified."""

import numpy as np
import pytest
import torch

from lpe import workloads
from oracle import certificate as cert, highs

pytestmark = pytest.mark.gpu

REL_OBJ_TOL = 1e-6


def _check_sample(problems, x, y, metrics, sample):
    for b in sample:
        p = problems(b)
        host = cert.certificate(*p, x[b], y[b])
        assert host["certificate_passed"], (b, host)
        np.testing.assert_allclose(metrics[b, 0], host["primal_residual"], rtol=1e-6, atol=1e-12)
        np.testing.assert_allclose(metrics[b, 2], host["relative_kkt_gap"], rtol=1e-6, atol=1e-13)
        ref = highs.solve(*p)
        assert cert.relative_objective_error(float(p[4] @ x[b]), ref.fun) <= REL_OBJ_TOL, b


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
    _check_sample(ub.problem, x.cpu().numpy(), y.cpu().numpy(), metrics, sample=[0, 97, 255])
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
            _check_sample(cb.problem, x.cpu().numpy(), y.cpu().numpy(), metrics, sample=[0, 4095])
    ro.engine.close()
