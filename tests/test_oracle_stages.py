"""Pins of the cooperative stage oracle (oracle/stages.py): a two-species joint
LP with closed-form stage answers, and the stage-3 invariants on a generated
consortium."""

import numpy as np
import pytest
from scipy.sparse import csr_matrix

from oracle import stages
from lpe import workloads


def tiny(biomass=(0.3, 0.2), supply=1.2):
    """Species A: v0 exchange (<= 0), v1 growth, v1 = -v0.  Species B: v2
    exchange, v3 growth, v3 = -v2 / 2.  Medium: -bA v0 - bB v2 <= C.  Growth
    rows t - v1 <= 0, t - v3 <= 0.  Max-min growth t* = C / (bA + 2 bB)."""
    bA, bB = biomass
    dense = np.array([
        [1.0, 1.0, 0.0, 0.0, 0.0],      # v1 + v0 = 0
        [0.0, 0.0, 0.5, 1.0, 0.0],      # v3 + v2/2 = 0
        [-bA, 0.0, -bB, 0.0, 0.0],      # medium
        [0.0, -1.0, 0.0, 0.0, 1.0],     # t - v1 <= 0
        [0.0, 0.0, 0.0, -1.0, 1.0],     # t - v3 <= 0
    ])
    a = csr_matrix(dense)
    rhs = np.array([0.0, 0.0, supply, 0.0, 0.0])
    lower = np.array([-10.0, 0.0, -10.0, 0.0, 0.0])
    upper = np.array([0.0, 10.0, 0.0, 10.0, 10.0])
    c = np.array([0.0, 0.0, 0.0, 0.0, -1.0])
    terms = [(0, 0, -1.0), (1, 2, -1.0)]
    g = np.array([0.0, 1.0, 0.0, 1.0])
    return (a, rhs, lower, upper, c, 2), terms, np.array(biomass), g, 4


def test_closed_form_max_min_and_parsimonious_exchange():
    problem, terms, biomass, g, t_col = tiny()
    out = stages.cooperative_chain(problem, terms, biomass, g, t_col)
    t_star = 1.2 / (0.3 + 2 * 0.2)
    assert out["t_star"] == pytest.approx(t_star, rel=1e-9)
    r3 = out["stage3"]
    assert r3.status == "optimal"
    t = 0.99 * t_star
    # v0 = -t, v2 = -2t, a = (bA t, 2 bB t), objective 0.99 C
    np.testing.assert_allclose(r3.x[:5], [-t, t, -2 * t, t, t], atol=1e-9)
    np.testing.assert_allclose(r3.x[5:], [0.3 * t, 0.2 * 2 * t], atol=1e-9)
    assert r3.fun == pytest.approx(0.99 * 1.2, rel=1e-9)


def test_closed_form_aggregate_stage():
    """With live objectives, stage 2 maximises v1 + v3 with t in [0.99 t*, t*]:
    all supply goes to species A beyond B's share of the floor."""
    problem, terms, biomass, g, t_col = tiny()
    out = stages.cooperative_chain(problem, terms, biomass, g, t_col, optimize_live_objectives=True)
    t_star = 1.2 / 0.7
    floor = 0.99 * t_star
    # B at the floor (v3 = floor, uses 2 bB floor), A takes the rest: v1 = (C - 2 bB floor) / bA
    v1 = (1.2 - 2 * 0.2 * floor) / 0.3
    assert out["stage2"].fun == pytest.approx(-(v1 + floor), rel=1e-9)
    assert out["aggregate"] == pytest.approx(v1 + floor, rel=1e-9)
    r3 = out["stage3"]
    assert r3.status == "optimal"
    # the aggregate floor row keeps g.v >= 0.99 aggregate
    assert g @ r3.x[:4] >= 0.99 * (v1 + floor) - 1e-7


def test_stage3_invariants_on_generated_consortium():
    model = workloads.consortium(seed=4, shapes=((120, 180), (150, 220), (90, 130)), n_medium=20)
    cb = workloads.consortium_batch(batch=2, seed=4, model=model)
    from lpe.stages import exchange_terms, live_objective
    terms = exchange_terms(model)
    g = live_objective(model)
    out = stages.cooperative_chain(cb.problem(0), terms, cb.biomass[0], g, model.t_col)
    r3 = out["stage3"]
    assert r3.status == "optimal"
    n = model.shape[1]
    v = r3.x[:n]
    scale = np.array([max(1e-12, cb.biomass[0][s]) * st for s, _, st in terms])
    flux = np.array([v[col] for _, col, _ in terms])
    np.testing.assert_allclose(r3.x[n:], np.abs(scale * flux), atol=1e-7)
    assert r3.fun == pytest.approx(np.abs(scale * flux).sum(), rel=1e-7, abs=1e-9)
    assert v[model.t_col] == pytest.approx(0.99 * out["t_star"], rel=1e-9, abs=1e-12)


@pytest.mark.parametrize("with_floor", [False, True])
def test_engine_stage3_pattern_matches_oracle_assembly(with_floor):
    """The batched stage-3 pattern with one environment's per-LP values equals
    the oracle's assembly of that environment's stage-3 matrix."""
    from lpe.stages import exchange_terms, live_objective, stage3_pattern
    model = workloads.consortium(seed=4, shapes=((120, 180), (150, 220), (90, 130)), n_medium=20)
    cb = workloads.consortium_batch(batch=2, seed=4, model=model)
    A3, neq3, coupling, transfer, tr0 = stage3_pattern(model, with_floor)
    terms, g = exchange_terms(model), live_objective(model)
    b = 1
    vals = A3.data.copy()
    vals[coupling.ravel()] = cb.coupling_values()[b].ravel()
    scale = np.maximum(cb.biomass[b], 1e-12)[[s for s, _, _ in terms]] * -1.0
    vals[transfer[:, 0]] = scale
    vals[transfer[:, 1]] = -scale
    mine = csr_matrix((vals, A3.indices, A3.indptr), shape=A3.shape).toarray()
    a, rhs, lo, hi, c, neq = cb.problem(b)
    ref = stages.stage3_lp(a, rhs, lo, hi, neq, terms, cb.biomass[b], g, model.t_col, (0.0, 1.0),
                           aggregate=1.0 if with_floor else None)[0].toarray()
    assert neq3 == neq
    np.testing.assert_array_equal(mine, ref)
