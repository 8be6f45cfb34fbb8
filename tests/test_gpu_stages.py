"""The cooperative stage driver (lpe/stages.py) on MI355X against the HiGHS
stage chain of oracle/stages.py (src/community_solver.py:1207-1844)."""

import numpy as np
import pytest
import torch
from scipy.sparse import csr_matrix

from oracle import stages as ostages
from oracle.certificate import relative_objective_error
from lpe import workloads

pytestmark = pytest.mark.gpu


def tiny_model():
    """The two-species closed-form LP of tests/test_oracle_stages.py as a
    consortium structure (biomass 1 in the template medium row)."""
    dense = np.array([
        [1.0, 1.0, 0.0, 0.0, 0.0],
        [0.0, 0.0, 0.5, 1.0, 0.0],
        [-1.0, 0.0, -1.0, 0.0, 0.0],
        [0.0, -1.0, 0.0, 0.0, 1.0],
        [0.0, 0.0, 0.0, -1.0, 1.0],
    ])
    a = csr_matrix(dense)
    a.sort_indices()
    coupling = np.array([[a.indptr[2] + 0], [a.indptr[2] + 1]], dtype=np.int64)   # (2,0) and (2,2)
    return workloads.Consortium(
        A=a, rhs_species=np.zeros(2), lower=np.array([-10.0, 0.0, -10.0, 0.0, 0.0]),
        upper=np.array([0.0, 10.0, 0.0, 10.0, 10.0]), c=np.array([0.0, 0.0, 0.0, 0.0, -1.0]), neq=2,
        col_offsets=np.array([0, 2, 4]), row_offsets=np.array([0, 1, 2]), biomass_cols=np.array([1, 3]),
        medium_cols=np.array([[0], [2]]), coupling_nnz=coupling, medium_rows=np.array([2]),
        vmax=np.ones((2, 1)), km=np.ones((2, 1)), t_col=4)


def _inputs(md, conc, biomass, device):
    B = conc.shape[0]
    cb = workloads.ConsortiumBatch(model=md, conc=conc, biomass=biomass)
    rhs = torch.as_tensor(cb.rhs(), device=device)
    lower = torch.as_tensor(np.tile(md.lower, (B, 1)), device=device)   # no uptake limits
    upper = torch.as_tensor(cb.upper(), device=device)
    coupling = torch.as_tensor(cb.coupling_values().reshape(B, -1), device=device)
    return cb, rhs, lower, upper, coupling, torch.as_tensor(biomass, device=device)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_tiny_closed_form(device, precision):
    from lpe.stages import CooperativeStages
    md = tiny_model()
    conc = np.array([[1.2], [0.6], [2.0], [1.0]])
    biomass = np.array([[0.3, 0.2], [0.3, 0.2], [0.1, 0.4], [0.5, 0.5]])
    _, rhs, lower, upper, coupling, bm = _inputs(md, conc, biomass, device)
    drv = CooperativeStages(md, 4, precision=precision, device=device)
    out = drv.solve(rhs, lower, upper, coupling, bm)
    assert bool(out["certified"].all())
    t_star = conc[:, 0] / (biomass[:, 0] + 2 * biomass[:, 1])
    np.testing.assert_allclose(out["t_star"].cpu().numpy(), t_star, rtol=1e-7)
    np.testing.assert_allclose(out["t"].cpu().numpy(), 0.99 * t_star, rtol=1e-7)
    # the stage-3 objective to the path's bar (1e-6 relative; the certificate's gap gate is 1e-7)
    np.testing.assert_allclose(out["transfer"].sum(dim=1).cpu().numpy(), 0.99 * conc[:, 0], rtol=1e-6)


@pytest.mark.parametrize("live", [False, True])
def test_generated_consortium_against_highs_chain(device, live):
    from lpe.stages import CooperativeStages, StageOptions, exchange_terms, live_objective
    model = workloads.consortium(seed=2)
    cb = workloads.consortium_batch(batch=2, seed=4, model=model, conc_range=(1e-2, 10.0))
    B = cb.batch
    rhs = torch.as_tensor(cb.rhs(), device=device)
    lower = torch.as_tensor(cb.lower(), device=device)
    upper = torch.as_tensor(cb.upper(), device=device)
    coupling = torch.as_tensor(cb.coupling_values().reshape(B, -1), device=device)
    bm = torch.as_tensor(cb.biomass, device=device)
    drv = CooperativeStages(model, B, precision="fp32", device=device,
                            stages=StageOptions(optimize_live_objectives=live))
    out = drv.solve(rhs, lower, upper, coupling, bm)
    assert bool(out["certified"].all()), out["stage_info"]
    terms, g = exchange_terms(model), live_objective(model)
    t_gpu = out["t_star"].cpu().numpy()
    obj3 = out["transfer"].sum(dim=1).cpu().numpy()
    for b in range(B):
        ref = ostages.cooperative_chain(cb.problem(b), terms, cb.biomass[b], g, model.t_col,
                                        optimize_live_objectives=live)
        assert relative_objective_error(t_gpu[b], ref["t_star"]) <= 1e-6
        # the later stages on exactly the LP the GPU solved (its own t*)
        same = ostages.cooperative_chain(cb.problem(b), terms, cb.biomass[b], g, model.t_col,
                                         optimize_live_objectives=live, t_star=t_gpu[b])
        assert same["stage3"].status == "optimal"
        assert relative_objective_error(obj3[b], same["stage3"].fun) <= 1e-6, (b, obj3[b], same["stage3"].fun)
