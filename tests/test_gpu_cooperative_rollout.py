"""The reference's default dFBA step - max-min growth, then parsimonious
exchange, with the medium and biomass advanced from the stage-3 fluxes
(src/community_solver.py:1712-1800; Euler and growth, src/dfba_simulator.py:
1317-1321, :1352, :1425-1427) - as a device-resident rollout of the synthetic
consortium (lpe.rollout.CooperativeRollout), each stage warm-started from the
previous step's.

Checked: every environment certifies in every stage at every step; at the
first and last step 1 % of the environments (spread over the batch) agree
with the HiGHS stage chain of the oracle on the step's LPs (t* and the
stage-3 objective on the LP the device solved, 1e-6 relative); warm steps
need far fewer stage-3 iterations than the cold first step; and the
parsimonious fluxes draw part of the limited medium down (the growth LP
alone secretes every medium metabolite)."""

import numpy as np
import pytest
import torch

from lpe import workloads
from oracle import stages as ostages
from oracle.certificate import relative_objective_error

pytestmark = pytest.mark.gpu

ENVS = 512
STEPS = 12


def _chain(args):
    problem, terms, biomass, g, t_col, t_gpu = args
    ref = ostages.cooperative_chain(problem, terms, biomass, g, t_col)
    same = ostages.cooperative_chain(problem, terms, biomass, g, t_col, t_star=t_gpu)
    return ref.get("t_star"), (same["stage3"].fun if "stage3" in same and same["stage3"].status == "optimal" else None)


def test_cooperative_rollout_warm_steps(device):
    import multiprocessing as mp
    import os
    from lpe.engine import SolveOptions
    from lpe.rollout import CooperativeRollout
    from lpe.stages import exchange_terms, live_objective
    md = workloads.consortium(2)
    ro = CooperativeRollout(md, ENVS, seed=4, device=device, options=SolveOptions(max_iterations=20000))
    limited = ~ro.state._unlimited
    conc0 = ro.state.conc[:, limited].clone()
    sample = np.unique(np.linspace(0, ENVS - 1, max(3, ENVS // 100)).round().astype(int)).tolist()
    terms, g = exchange_terms(md), live_objective(md)
    stage3_iters = []
    for step in range(STEPS):
        conc, bio = (v.cpu().numpy() for v in ro.snapshot())
        out = ro.step()
        assert out["certified"] == ENVS, (step, out)
        stage3_iters.append(out["parsimonious_exchange"]["iterations"])
        if step in (0, STEPS - 1):
            cb = workloads.ConsortiumBatch(model=md, conc=conc, biomass=bio)
            t_gpu = out["t_star"].cpu().numpy()
            obj3 = out["transfer"].sum(dim=1).cpu().numpy()
            jobs = [(cb.problem(b), terms, bio[b], g, md.t_col, float(t_gpu[b])) for b in sample]
            workers = max(1, min(16, len(os.sched_getaffinity(0)), len(jobs)))
            with mp.get_context("spawn").Pool(workers) as pool:
                refs = pool.map(_chain, jobs, chunksize=1)
            for b, (t_ref, f3) in zip(sample, refs):
                assert t_ref is not None and f3 is not None, (step, b)
                assert relative_objective_error(t_gpu[b], t_ref) <= 1e-6, (step, b, t_gpu[b], t_ref)
                assert relative_objective_error(obj3[b], f3) <= 1e-6, (step, b, obj3[b], f3)
    print("stage-3 PDHG iterations per step:", stage3_iters, flush=True)
    # warm starts from the previous step's stage-3 solution
    assert max(stage3_iters[2:]) < 0.5 * stage3_iters[0], stage3_iters
    drawn = float((ro.state.conc[:, limited] < conc0).double().mean())
    print(f"limited medium entries below their initial value after {STEPS} steps: {drawn:.3f}", flush=True)
    assert drawn > 0.2, drawn
    ro.driver.joint.close()
    ro.driver.stage3.close()
