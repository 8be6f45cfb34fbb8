"""The reference's plugin surfaces on the MI355X engine: GpuPdhgCorrector on
the reference's own known-answer LPs (tests/test_gpu_pdhg_corrector.py:29-112)
and the batched / per-LP backend (src/cpu_repeated_lp.py:783-818,
src/community_solver.py:1200-1205) against HiGHS."""

import json
import os

import numpy as np
import pytest
from scipy.sparse import csr_matrix

from conftest import ROOT
from oracle import highs
from oracle.certificate import relative_objective_error
from lpe import workloads
from lpe.requests import problem_request

GOLDEN = os.path.join(ROOT, "tests", "golden", "reference_known_answers.json")


def _known():
    with open(GOLDEN) as f:
        return json.load(f)


def _bounds(values):
    return np.array([-np.inf if v is None else v for v in values], dtype=float)


def _problem(lp):
    lower = _bounds(lp["lower"])
    upper = np.array([np.inf if v is None else v for v in lp["upper"]], dtype=float)
    return (csr_matrix(np.array(lp["a"], dtype=float)), np.array(lp["rhs"], dtype=float), lower, upper,
            np.array(lp["c"], dtype=float), lp["neq"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["inequality_dual_sign", "box_only_infinite_unused_bound"])
def test_corrector_reference_known_answers(device, name):
    from lpe.corrector import GpuPdhgCorrector
    lp = next(l for l in _known()["lps"] if l["name"] == name)
    corr = GpuPdhgCorrector([_problem(lp)], step_safety=lp["step_safety"])
    res = corr.solve(iterations=lp["iterations"], check_interval=lp["check_interval"])
    assert res["all_accepted"], res["metrics"]
    np.testing.assert_allclose(res["x"].cpu().numpy()[0], lp["x"], atol=lp["atol_x"])
    if lp["y"] is not None:
        np.testing.assert_allclose(res["y"].cpu().numpy()[0], lp["y"], atol=lp["atol_y"])
    assert res["metrics"][0]["certificate_passed"] and res["metrics"][0]["success"]


@pytest.mark.gpu
def test_corrector_independent_blocks_known_answer(device):
    """Two LPs with different right-hand sides in one batch (:66-82)."""
    from lpe.corrector import GpuPdhgCorrector
    lps = [l for l in _known()["lps"] if l["name"].startswith("block_independent")]
    corr = GpuPdhgCorrector([_problem(l) for l in lps], step_safety=lps[0]["step_safety"])
    res = corr.solve(iterations=lps[0]["iterations"], check_interval=lps[0]["check_interval"])
    assert res["all_accepted"]
    for b, lp in enumerate(lps):
        np.testing.assert_allclose(res["x"].cpu().numpy()[b], lp["x"], atol=lp["atol_x"])
        np.testing.assert_allclose(res["y"].cpu().numpy()[b], lp["y"], atol=lp["atol_y"])


@pytest.mark.gpu
def test_corrector_infeasible_known_answer_rejected(device):
    from lpe.corrector import GpuPdhgCorrector
    k = _known()["infeasible"]
    corr = GpuPdhgCorrector([_problem(k)])
    res = corr.solve(iterations=k["iterations"], check_interval=k["check_interval"])
    assert not res["accepted"].any()
    assert res["metrics"][0]["success"] is False


def _uptake_requests(batch, seed=3):
    ub = workloads.uptake_batch(batch=batch, seed=seed, m=120, n=180)
    return [problem_request(ub.problem(b)) for b in range(batch)], ub


@pytest.mark.gpu
def test_backend_solve_batch_against_highs_and_warm_reuse(device):
    from lpe.backend import GpuLpBackend
    reqs, ub = _uptake_requests(6)
    backend = GpuLpBackend(precision="fp32", polish=True)
    try:
        res = backend.solve_batch(reqs, environment_ids=[10, 11, 12, 13, 14, 15])
        assert len(res) == 6
        for b, r in enumerate(res):
            assert r.success, r.diagnostics
            ref = highs.solve(*ub.problem(b))
            assert relative_objective_error(r.fun, ref.fun) <= 1e-6
            assert r.diagnostics["certificate_passed"] and r.diagnostics["primal_residual"] <= 1e-5
            assert not r.diagnostics["warm_started"]
        first_iterations = res[0].nit
        again = backend.solve_batch(reqs, environment_ids=[10, 11, 12, 13, 14, 15])
        assert all(r.success and r.diagnostics["warm_started"] for r in again)
        assert again[0].nit <= first_iterations
        one = backend.solve(reqs[2][0], **dict(reqs[2][1], _environment_id=99))
        assert one.success
        assert relative_objective_error(one.fun, res[2].fun) <= 1e-6
    finally:
        backend.close()


@pytest.mark.gpu
@pytest.mark.parametrize("pdhg", ["restarted", "halpern"])
def test_backend_mixed_shapes_and_stages_grouped(device, pdhg):
    from lpe.backend import GpuLpBackend
    reqs_a, ub_a = _uptake_requests(2, seed=5)
    sp = workloads.species_lp(seed=6, m=80, n=120)
    c_b, kw_b = problem_request(sp)
    kw_b2 = dict(kw_b, _stage="aggregate")
    for mixed in (True, False):
        _grouping_run(pdhg, mixed, reqs_a, c_b, kw_b, kw_b2, sp)


def _grouping_run(pdhg, mixed, reqs_a, c_b, kw_b, kw_b2, sp):
    from lpe.backend import GpuLpBackend
    backend = GpuLpBackend(precision="fp64", polish=True, pdhg=pdhg, mixed_shapes=mixed)
    try:
        res = backend.solve_batch(reqs_a + [(c_b, kw_b), (c_b, kw_b2)], environment_ids=["a", "b", "c", "d"])
        assert [r.success for r in res] == [True] * 4
        assert res[3].diagnostics["stage"] == "aggregate"
        ref = highs.solve(*sp)
        for r in res[2:]:
            assert relative_objective_error(r.fun, ref.fun) <= 1e-6
        # one launch per stage (the two shapes of the default stage share one),
        # or per shape and stage without padding
        assert backend.history[-1]["groups"] == (2 if mixed else 3)
    finally:
        backend.close()


@pytest.mark.gpu
def test_backend_infeasible_request_fails_closed(device):
    from lpe.backend import GpuLpBackend
    backend = GpuLpBackend(precision="fp64", polish=True, max_iterations=2000)
    try:
        r = backend.solve([0.0, 0.0], A_ub=[[1.0, 0.0], [-1.0, 0.0]], b_ub=[0.0, -1.0], bounds=[(None, None)] * 2)
        assert r.success is False and r.x is None and r.fun is None
        # x1 <= 0 and x1 >= 1: certified primal infeasible by a dual ray before the budget
        assert r.diagnostics["status"] == "primal_infeasible" and "primal_infeasible" in r.message
        assert r.nit < 2000
    finally:
        backend.close()


# ---------------------------------------------------------------- CPU-side
def test_backend_rejects_duplicate_or_missing_ids_before_any_work():
    from lpe.backend import GpuLpBackend
    backend = GpuLpBackend()
    reqs = [([1.0], dict(A_ub=[[1.0]], b_ub=[1.0])), ([1.0], dict(A_ub=[[1.0]], b_ub=[1.0]))]
    with pytest.raises(ValueError):
        backend.solve_batch(reqs, environment_ids=[1, 1])
    with pytest.raises(ValueError):
        backend.solve_batch(reqs, environment_ids=[1])
    with pytest.raises(ValueError):
        backend.solve_batch([([1.0], dict(integrality=[1]))])
    backend.close()
    with pytest.raises(RuntimeError):
        backend.solve_batch(reqs)


def test_corrector_validation_before_device_work():
    from lpe.corrector import GpuPdhgCorrector
    p = (csr_matrix(np.eye(1)), np.ones(1), np.zeros(1), np.ones(1), np.ones(1), 1)
    for kwargs in (dict(step_safety=1.0), dict(step_safety=0.0), dict(theta=1.5), dict(primal_weight=0.0),
                   dict(primal_weight=True)):
        with pytest.raises(ValueError):
            GpuPdhgCorrector([p], **kwargs)
    with pytest.raises(ValueError):
        GpuPdhgCorrector([])
    q = (csr_matrix(np.eye(2)), np.ones(2), np.zeros(2), np.ones(2), np.ones(2), 2)
    with pytest.raises(ValueError):
        GpuPdhgCorrector([p, q])


@pytest.mark.gpu
def test_backend_keeps_a_bounded_number_of_engines(device):
    """Batches of shrinking size (a microbatched rollout) reuse at most
    MAX_ENGINES resident engines instead of one per size."""
    from lpe.backend import GpuLpBackend
    ub = workloads.uptake_batch(batch=6, seed=9, m=60, n=90)
    be = GpuLpBackend(precision="fp64")
    for size in (6, 5, 4, 3, 2, 1, 6):
        res = be.solve_batch([problem_request(ub.problem(b)) for b in range(size)],
                             environment_ids=list(range(size)))
        assert all(r.success for r in res), size
        assert len(be._engines) <= be.MAX_ENGINES
    be.close()


@pytest.mark.gpu
def test_backend_solves_lps_of_different_sizes_in_one_batch(device):
    """Requests of three sizes (and inequality rows in some) share one engine
    launch: each is padded to the common shape and answered in its own rows
    and columns, certified and within 1e-6 of HiGHS; a warm-started repeat
    of the batch reuses the padded layout."""
    from lpe.backend import GpuLpBackend
    probs = []
    for seed, (m, n) in enumerate(((60, 90), (40, 70), (80, 100))):
        ub = workloads.uptake_batch(batch=2, seed=seed + 20, m=m, n=n)
        for b in range(2):
            a, rhs, lo, hi, c, neq = ub.problem(b)
            if seed == 1:          # last 5 equality rows as <= rows
                neq = m - 5
            probs.append((a, rhs, lo, hi, c, neq))
    be = GpuLpBackend(precision="fp32")
    try:
        for _ in range(2):
            res = be.solve_batch([problem_request(p) for p in probs], environment_ids=list(range(len(probs))))
            assert be.history[-1]["groups"] == 1
            for p, r in zip(probs, res):
                ref = highs.solve(*p)
                assert r.success, r.message
                assert r.x.shape == (p[0].shape[1],)
                assert relative_objective_error(r.fun, ref.fun) <= 1e-6
    finally:
        be.close()
