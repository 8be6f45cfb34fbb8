"""The resident reduced service, lpe.reduction.GpuResidentReducedPdhg
(src/gpu_resident_reduced_pdhg.py), on the reference's service cases - kept
as data in tests/golden/resident_reduction_cases.json - and checked against
this repository's oracle: the service's witnesses and reduced problems are
oracle.reduction's, and every answer it accepts is HiGHS's optimum of the
original LP.  Contract points: the graph, loop and non-resident drivers
agree; a same-pattern update reuses the resident buffers; pattern / batch /
order changes ask for a rebuild (with the reference's reason strings) and
invalid input raises, both without mutation; a failed device update
invalidates the service; results are owned; the final original-LP
certificate decides acceptance; a certified LP stays frozen."""

import json
import os

import numpy as np
import pytest
import torch
from scipy.sparse import csr_matrix

from conftest import ROOT
from oracle import highs, reduction as ored
from oracle.certificate import relative_objective_error

pytestmark = pytest.mark.gpu

with open(os.path.join(ROOT, "tests", "golden", "resident_reduction_cases.json")) as _f:
    CASES = json.load(_f)


def lp(d):
    """LP tuple from a fixture record ("nan" strings become NaN)."""
    f = lambda v: np.array([np.nan if x == "nan" else x for x in v], dtype=np.float64)  # noqa: E731
    rec = dict(CASES["base"], **d)
    return (csr_matrix(np.array(rec["a"], dtype=np.float64)), f(rec["rhs"]), f(rec["lower"]), f(rec["upper"]),
            f(rec["c"]), int(rec["neq"]))


def lps(records):
    return [lp(r) for r in records]


def make_service(problems, **kwargs):
    from lpe.reduction import GpuResidentReducedPdhg
    kwargs.setdefault("chunk_size", CASES["equivalence"]["chunk_size"])
    return GpuResidentReducedPdhg(problems, **kwargs)


def host(t):
    return t.detach().cpu().numpy().copy() if isinstance(t, torch.Tensor) else np.array(t, copy=True)


def summary(result):
    """Arrays that identify a service result."""
    metr = np.array([[r[k] for k in ("primal_residual", "dual_violation", "relative_kkt_gap", "objective")]
                     for r in result["metrics"]])
    return dict(x=host(result["x"]), y=host(result["y"]), rx=host(result["reduced_x"]),
                ry=host(result["reduced_y"]), accepted=np.array(result["accepted"], dtype=bool),
                passed=np.array([r["certificate_passed"] for r in result["metrics"]]), metrics=metr,
                iterations=int(result["iterations_run"]))


def assert_same(a, b, tol=1e-10):
    assert a["iterations"] == b["iterations"]
    np.testing.assert_array_equal(a["accepted"], b["accepted"])
    np.testing.assert_array_equal(a["passed"], b["passed"])
    for k in ("x", "y", "rx", "ry", "metrics"):
        np.testing.assert_allclose(a[k], b[k], rtol=tol, atol=tol, err_msg=k)


def assert_accepted_are_optimal(problems, s):
    for b, p in enumerate(problems):
        if s["accepted"][b]:
            ref = highs.solve(*p)
            assert ref.status == "optimal"
            assert relative_objective_error(float(p[4] @ s["x"][b]), ref.fun) <= 1e-6, b


def oracle_witnesses(problems):
    pl = ored.plan(problems[0])
    red = [ored.reduce(pl, p) for p in problems]
    return np.array([r[1] for r in red]).ravel().tolist(), np.array([r[2] for r in red]).ravel().tolist()


@pytest.mark.parametrize("iterations", CASES["equivalence"]["iterations"])
def test_graph_loop_and_one_shot_drivers_agree(device, iterations):
    from lpe.reduction import GpuReducedPdhgCorrector
    eq = CASES["equivalence"]
    problems = lps(eq["problems"])
    args = dict(iterations=iterations, check_interval=eq["check_interval"])
    runs = [make_service(problems).solve(**args), make_service(problems, use_graph=False).solve(**args),
            GpuReducedPdhgCorrector(problems).solve(**args)]
    sums = [summary(r) for r in runs]
    assert_same(sums[0], sums[1])
    assert_same(sums[0], sums[2])
    assert_accepted_are_optimal(problems, sums[0])
    assert runs[0]["cpu_lp_calls"] == 0 and "host input/update/control" in runs[0]["scope"]


def test_value_update_reuses_buffers_and_refreshes_witnesses(device):
    vu = CASES["value_update"]
    before, after = lps(vu["before"]), lps(vu["after"])
    svc = make_service(before)
    lw, uw = oracle_witnesses(before)
    assert svc.lower_witness.ravel().tolist() == vu["witnesses_before"]["lower"] == lw
    assert svc.upper_witness.ravel().tolist() == vu["witnesses_before"]["upper"] == uw
    assert svc.try_update(after)
    assert svc.last_update_timing["reused"] and svc.last_update_timing["reason"] == "same_pattern_value_update"
    lw, uw = oracle_witnesses(after)
    assert svc.lower_witness.ravel().tolist() == vu["witnesses_after"]["lower"] == lw
    assert svc.upper_witness.ravel().tolist() == vu["witnesses_after"]["upper"] == uw
    args = dict(iterations=vu["iterations"], check_interval=vu["iterations"])
    updated = summary(svc.solve(**args))
    assert updated["accepted"].all()
    assert_same(updated, summary(make_service(after).solve(**args)))
    np.testing.assert_allclose(updated["x"], vu["x_after"], atol=1e-12)
    for b, p in enumerate(after):
        np.testing.assert_allclose(updated["x"][b], highs.solve(*p).x, atol=1e-9)


def test_results_are_owned_copies(device):
    own = CASES["ownership"]
    svc = make_service(lps(own["first"]))
    args = dict(iterations=own["iterations"], check_interval=own["iterations"])
    first = svc.solve(**args)
    kept = {k: host(first[k]) for k in ("x", "y", "reduced_x", "reduced_y")}
    for k in kept:                                  # the caller scribbles on its result ...
        first[k].add_(1.0)
    again = summary(svc.solve(**args))              # ... which the service never reads back
    np.testing.assert_array_equal(again["x"], kept["x"])
    np.testing.assert_array_equal(again["ry"], kept["reduced_y"])
    assert svc.try_update(lps(own["second"]))
    svc.solve(**args)                               # a later solve leaves the first result alone
    for k, v in kept.items():
        np.testing.assert_array_equal(host(first[k]), v + 1.0)


@pytest.mark.parametrize("case", CASES["rebuild"]["cases"], ids=lambda c: c["name"])
def test_rebuild_requests_leave_the_service_unchanged(device, case):
    ids = CASES["rebuild"]["environment_ids"]
    svc = make_service(lps([{}, {}]), environment_ids=ids)
    problems_before = svc.original_problems
    assert not svc.try_update(lps(case["problems"]), environment_ids=case["ids"])
    assert svc.last_update_timing["reason"] == case["reason"]
    assert svc.original_problems is problems_before
    assert svc.environment_ids == tuple(ids)
    svc.solve(iterations=0)


@pytest.mark.parametrize("case", CASES["invalid"], ids=lambda c: c["name"])
def test_invalid_input_raises_before_any_change(device, case):
    svc = make_service(lps([{}, {}]))
    problems_before = svc.original_problems
    bad = {k: v for k, v in case.items() if k != "name"}
    with pytest.raises(ValueError):
        svc.try_update(lps([bad, {}]))
    assert svc.original_problems is problems_before and svc._valid
    svc.solve(iterations=0)


class _EngineWithFailingUpdate:
    """The service's engine with a device update that fails part-way."""

    def __init__(self, engine):
        self._engine = engine

    def update(self, *args, **kwargs):
        raise OSError("device update failed")

    def __getattr__(self, name):
        return getattr(self._engine, name)


def test_failed_device_update_invalidates_the_service(device):
    svc = make_service(lps([{}, {}]))
    svc.engine = _EngineWithFailingUpdate(svc.engine)
    with pytest.raises(OSError):
        svc.try_update(lps([{"rhs": [0.0, 2.0]}, {"rhs": [0.0, 2.0]}]))
    with pytest.raises(RuntimeError, match="invalidated"):
        svc.solve(iterations=1)


class _ScheduledCertificate:
    """The service's certificate with failures forced on a schedule:
    LP 1 always, LP 0 from the second call on."""

    def __init__(self, certificate):
        self._certificate = certificate
        self.calls = 0

    def __call__(self, x, y):
        rows = self._certificate(x, y)
        rows[1]["certificate_passed"] = False
        if self.calls > 0:
            rows[0]["certificate_passed"] = False
            rows[0]["relative_kkt_gap"] = 1.0
        self.calls += 1
        return rows


def test_final_certificate_decides_acceptance(device):
    auth = CASES["authority"]
    svc = make_service(lps(auth["problems"]))
    svc._certificate = _ScheduledCertificate(svc._certificate)
    result = svc.solve(iterations=auth["iterations"], check_interval=auth["iterations"])
    assert result["checkpoints"][0]["accepted_count"] == auth["accepted_at_first_checkpoint"]
    assert list(result["accepted"]) == auth["accepted_final"] and not result["all_accepted"]
    assert result["metrics"][0]["accepted_iteration"] == -1 and not result["metrics"][0]["success"]
    assert not svc.history[-1]["accepted"][0]


def test_certified_lp_stays_frozen(device):
    fr = CASES["frozen"]
    problems = lps(fr["problems"])
    result = make_service(problems).solve(iterations=fr["iterations"], check_interval=fr["iterations"])
    s = summary(result)
    assert result["all_accepted"] and result["cpu_lp_calls"] == 0
    assert result["metrics"][0]["accepted_iteration"] == fr["lp0_accepted_iteration"]
    np.testing.assert_array_equal(s["x"][0], fr["lp0_x"])
    assert_accepted_are_optimal(problems, s)
