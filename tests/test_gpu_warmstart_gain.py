"""What the trained warm-start network gains, at the consortium's full LP
size (3663 x 5301, 1024 environments, 12 dFBA steps of one rollout): the
same steps solved from the learned (hybrid) start, from the previous step's
solution and cold.  The network is trained by the library's own kernels on
certified solves of two other seeded rollouts.  Printed per step: PDHG
iterations, polish rounds, wall time and the LPs started from the network."""

import numpy as np
import pytest
import torch

from lpe import workloads

pytestmark = pytest.mark.gpu

ENVS, STEPS = 1024, 12


def _opts():
    from lpe.engine import SolveOptions
    return SolveOptions(max_iterations=6000, polish=True, polish_cg_iterations=200)


@pytest.fixture(scope="module")
def trained(device):
    from lpe.rollout import ConsortiumRollout
    from lpe.warmstart import LearnedWarmStart
    md = workloads.consortium(2)
    lws = LearnedWarmStart(md, device=device)
    for seed in (11, 12):
        ro = ConsortiumRollout(md, ENVS, seed=seed, options=_opts(), autotune=False, device=device)
        for _ in range(STEPS):
            conc, bio = ro.conc.clone(), ro.biomass.clone()
            _, lower, upper, _ = ro.lp_inputs()
            out = ro.step()
            lws.add_solutions(conc, lower, upper, bio, out["x"], out["y"], out["certified_mask"])
        ro.engine.close()
    hist = lws.train(epochs=15, batch=1024, lr=1e-3)
    assert hist[-1] < hist[0]
    return md, lws


def _run(md, device, warm_start):
    from lpe.rollout import ConsortiumRollout
    ro = ConsortiumRollout(md, ENVS, seed=2, options=_opts(), autotune=False, device=device, warm_start=warm_start)
    rows = []
    for _ in range(STEPS):
        ro.step()
        l = ro.log[-1]
        rows.append((l["iterations"], l["polish_attempts"], l["wall_seconds"], l["certified"], l.get("start_mlp")))
    ro.engine.close()
    return rows


def test_learned_start_gain_over_previous_and_cold(device, trained):
    md, lws = trained
    runs = {"learned": _run(md, device, lws), "previous": _run(md, device, "previous"),
            "cold": _run(md, device, "none")}
    for name, rows in runs.items():
        print(f"\n{name:9s} " + " ".join(f"[{it}it {pol}p {w * 1e3:.0f}ms{'' if s is None else f' mlp{s}'}]"
                                        for it, pol, w, _, s in rows))
        assert all(r[3] == ENVS for r in rows), (name, [r[3] for r in rows])
    lr, pr, cr = runs["learned"], runs["previous"], runs["cold"]
    # step 0 has no previous solution: the network's start against a cold start
    assert lr[0][4] == ENVS
    assert lr[0][0] * 3 <= cr[0][0], (lr[0], cr[0])
    assert lr[0][2] < cr[0][2]
    # step 1: the previous solution is stale (biomass grew by tens of percent);
    # the LPs whose previous solution scores far from optimal start from the network
    assert lr[1][4] >= ENVS // 2, lr[1]
    assert lr[1][0] < pr[1][0], (lr[1], pr[1])
    assert lr[1][2] < pr[1][2], (lr[1], pr[1])
    # once the rollout settles the previous solutions are kept
    assert all(r[4] == 0 for r in lr[3:]), [r[4] for r in lr]
    # over the rollout the learned start is never worse than the previous-step start
    # by more than the cost of scoring the previous solutions (one certificate per step)
    wl, wp = sum(r[2] for r in lr[1:]), sum(r[2] for r in pr[1:])
    assert wl <= 1.15 * wp, (wl, wp)
    assert sum(r[0] for r in lr[1:]) <= sum(r[0] for r in pr[1:])
    # no step past the fifth is slower than the slowest of steps 2-5
    assert max(r[2] for r in lr[6:]) <= 1.25 * max(r[2] for r in lr[2:6]), [r[2] for r in lr]
    # and far fewer iterations than solving every step cold
    assert sum(r[0] for r in lr) * 5 <= sum(r[0] for r in cr)
