"""The resident reduced service (lpe.reduction.GpuResidentReducedPdhg) on the
reference's cases (tests/test_gpu_resident_reduced_pdhg.py): graph = loop =
non-resident corrector, value updates through try_update, rebuild requests
and invalid input without mutation, owned results, final-certificate
authority."""

import numpy as np
import pytest
import torch
from scipy.sparse import csr_matrix

pytestmark = pytest.mark.gpu


def _problem(*, inequality=(1.0, 0.5), rhs=0.8, lower=(0.0, -1.0), upper=(1.0, 2.0), c=(-1.0, -0.3)):
    return (csr_matrix(np.array([[1.0, -1.0], inequality], dtype=np.float64)), np.array([0.0, rhs]),
            np.array(lower), np.array(upper), np.array(c), 1)


def _service(problems=None, **kwargs):
    from lpe.reduction import GpuResidentReducedPdhg
    problems = problems if problems is not None else [_problem(), _problem(rhs=0.9)]
    return GpuResidentReducedPdhg(problems, chunk_size=4, **kwargs)


def _np(t):
    return t.cpu().numpy() if isinstance(t, torch.Tensor) else np.asarray(t)


def _same_result(left, right):
    for name in ("x", "y", "reduced_x", "reduced_y"):
        np.testing.assert_allclose(_np(left[name]), _np(right[name]), atol=1e-10, rtol=1e-10)
    assert left["accepted"].tolist() == right["accepted"].tolist()
    assert left["iterations_run"] == right["iterations_run"]
    for lrow, rrow in zip(left["metrics"], right["metrics"]):
        for name in ("primal_residual", "dual_violation", "relative_kkt_gap", "objective"):
            np.testing.assert_allclose(lrow[name], rrow[name], atol=1e-10, rtol=1e-10)
        assert lrow["certificate_passed"] == rrow["certificate_passed"]


@pytest.mark.parametrize("iterations", [0, 3, 4, 17])
def test_graph_loop_and_corrector_equivalence(device, iterations):
    from lpe.reduction import GpuReducedPdhgCorrector
    problems = [_problem(), _problem(rhs=0.9)]
    graph, loop = _service(problems), _service(problems, use_graph=False)
    original = GpuReducedPdhgCorrector(problems)
    args = dict(iterations=iterations, check_interval=17)
    result = graph.solve(**args)
    _same_result(result, loop.solve(**args))
    _same_result(result, original.solve(**args))
    assert result["cpu_lp_calls"] == 0
    assert "host input/update/control" in result["scope"]


def test_value_updates_refresh_witnesses(device):
    service = _service([_problem(rhs=3.0), _problem(rhs=3.0)])
    assert service.upper_witness.ravel().tolist() == [0, 0]
    assert service.lower_witness.ravel().tolist() == [0, 0]
    changed = [_problem(inequality=(0.25, 1.5), rhs=4.0, lower=(-1.0, 0.0), upper=(2.0, 0.7), c=(-0.4, -0.8)),
               _problem(inequality=(2.0, 0.25), rhs=5.0, lower=(-1.0, 0.0), upper=(2.0, 0.9), c=(-0.6, -0.4))]
    assert service.try_update(changed)
    assert service.upper_witness.ravel().tolist() == [1, 1]
    assert service.lower_witness.ravel().tolist() == [1, 1]
    assert service.last_update_timing["reused"]
    result = service.solve(iterations=12, check_interval=12)
    rebuilt = _service(changed).solve(iterations=12, check_interval=12)
    _same_result(result, rebuilt)
    assert result["all_accepted"]
    np.testing.assert_allclose(_np(result["x"]), [[0.7, 0.7], [0.9, 0.9]], atol=1e-12)


def test_returned_arrays_are_owned(device):
    service = _service([_problem(rhs=3.0)])
    first = service.solve(iterations=8, check_interval=8)
    snap = {k: _np(first[k]).copy() for k in ("x", "y", "reduced_x", "reduced_y")}
    assert service.try_update([_problem(rhs=3.0, upper=(0.3, 2.0))])
    second = service.solve(iterations=8, check_interval=8)
    for k, v in snap.items():
        np.testing.assert_array_equal(_np(first[k]), v)
        assert first[k].data_ptr() != second[k].data_ptr() or first[k].numel() == 0


@pytest.mark.parametrize("change,reason", [("order", "batch_or_environment_order_changed"),
                                           ("batch", "batch_or_environment_order_changed"),
                                           ("original_pattern", "original_csr_pattern_changed"),
                                           ("reduced_pattern", "reduced_csr_pattern_changed")])
def test_rebuild_requests_do_not_mutate(device, change, reason):
    service = _service(environment_ids=[11, 29])
    old = service.original_problems
    problems, ids = [_problem(), _problem()], [11, 29]
    if change == "order":
        ids.reverse()
    elif change == "batch":
        problems, ids = problems[:1], ids[:1]
    elif change == "original_pattern":
        problems[0] = _problem(inequality=(1.0, 0.0))
    else:
        problems[0] = _problem(inequality=(1.0, -1.0))
    assert not service.try_update(problems, environment_ids=ids)
    assert service.last_update_timing["reason"] == reason
    assert service.original_problems is old
    service.solve(iterations=0)


@pytest.mark.parametrize("bad", ["equality", "equality_rhs", "nan", "bounds"])
def test_invalid_problem_rejected_before_mutation(device, bad):
    service = _service()
    old = service.original_problems
    problem = list(_problem())
    if bad == "equality":
        problem[0] = csr_matrix([[1.0, -2.0], [1.0, 0.5]])
    elif bad == "equality_rhs":
        problem[1] = np.array([1.0, 0.8])
    elif bad == "nan":
        problem[4] = np.array([np.nan, -0.3])
    else:
        problem[2], problem[3] = np.array([1.0, 0.0]), np.array([2.0, 0.5])
    with pytest.raises(ValueError):
        service.try_update([tuple(problem), _problem()])
    assert service.original_problems is old and service._valid


def test_interrupted_update_invalidates(device, monkeypatch):
    service = _service()

    def fail(*args, **kwargs):
        raise RuntimeError("simulated device update failure")

    monkeypatch.setattr(service.engine, "update", fail)
    with pytest.raises(RuntimeError, match="simulated"):
        service.try_update([_problem(rhs=2.0), _problem(rhs=2.0)])
    with pytest.raises(RuntimeError, match="invalidated"):
        service.solve(iterations=1)


def test_final_certificate_is_the_authority(device, monkeypatch):
    service = _service([_problem(rhs=3.0, c=(1.0, 0.3)), _problem()])
    original = service._certificate
    calls = 0

    def certificate(x, y):
        nonlocal calls
        rows = original(x, y)
        if calls:
            rows[0]["certificate_passed"] = False
            rows[0]["relative_kkt_gap"] = 1.0
        rows[1]["certificate_passed"] = False
        calls += 1
        return rows

    monkeypatch.setattr(service, "_certificate", certificate)
    result = service.solve(iterations=4, check_interval=4)
    assert result["checkpoints"][0]["accepted_count"] == 1
    assert not result["all_accepted"] and result["accepted"].tolist() == [False, False]
    assert result["metrics"][0]["accepted_iteration"] == -1 and not result["metrics"][0]["success"]
    assert not service.history[-1]["accepted"][0]


def test_certified_environment_stays_frozen(device):
    service = _service([_problem(rhs=3.0, c=(1.0, 0.3)), _problem(rhs=3.0)])
    result = service.solve(iterations=8, check_interval=8)
    assert result["all_accepted"]
    assert result["metrics"][0]["accepted_iteration"] == 0
    np.testing.assert_array_equal(_np(result["x"])[0], [0.0, 0.0])
    assert result["cpu_lp_calls"] == 0
