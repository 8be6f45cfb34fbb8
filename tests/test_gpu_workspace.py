"""The fixed-pattern workspace (lpe.workspace.GpuPdhgWorkspace) on the
reference's cases (tests/test_gpu_pdhg_workspace.py): graph and loop routes
against the oracle's fixed-metric PDHG with chunks and tail, projected warm
starts, dynamic values / vectors / metrics / theta, frozen environments,
fail-closed validation without mutation, owned inputs, a box-only problem,
budget validation; plus rows whose equality / inequality order is not
equality-first."""

import numpy as np
import pytest
import torch
from scipy.sparse import block_diag, csr_matrix

from oracle import pdhg as opdhg

pytestmark = pytest.mark.gpu


def _problem(scale=1., *, rhs=1., upper=2., cost=-1.):
    return (csr_matrix(np.array([[scale, -1.], [1., 1.]], dtype=np.float64)),
            np.array([0., rhs]), np.array([0., 0.]), np.array([upper, upper]), np.array([cost, -.25]), 1)


class _Batch:
    """Block data of a list of canonical LPs with the reference metric."""

    def __init__(self, problems, *, theta=1.0, primal_weight=1.0, mask=None):
        self.problems = problems
        self.batch = len(problems)
        self.m, self.n = problems[0][0].shape
        self.a = csr_matrix(block_diag([p[0] for p in problems], format="csr"))
        self.a.sort_indices()
        self.rhs = np.concatenate([p[1] for p in problems])
        self.lower = np.concatenate([p[2] for p in problems])
        self.upper = np.concatenate([p[3] for p in problems])
        self.c = np.concatenate([p[4] for p in problems])
        if mask is None:
            mask = np.arange(self.m) >= problems[0][5]
        self.row_mask = np.asarray(mask, dtype=bool)
        self.inequality_mask = np.tile(self.row_mask, self.batch)
        metrics = [opdhg.reference_metric(p[0], 0.99, primal_weight) for p in problems]
        self.sigma = np.concatenate([s for s, _ in metrics])
        self.tau = np.concatenate([t for _, t in metrics])
        self.theta = theta

    def oracle(self, iterations, x=None, y=None):
        xs, ys = [], []
        for b, (a, rhs, lo, up, c, neq) in enumerate(self.problems):
            sl_n, sl_m = slice(b * self.n, (b + 1) * self.n), slice(b * self.m, (b + 1) * self.m)
            xo, yo = opdhg.fixed_pdhg(a, rhs, lo, up, c, neq, iterations=iterations, sigma=self.sigma[sl_m],
                                      tau=self.tau[sl_n], theta=self.theta,
                                      x0=None if x is None else np.asarray(x).reshape(self.batch, self.n)[b],
                                      y0=None if y is None else np.asarray(y).reshape(self.batch, self.m)[b],
                                      inequality_mask=self.row_mask)
            xs.append(xo)
            ys.append(yo)
        return np.stack(xs), np.stack(ys)


def _workspace(data=None, **kwargs):
    from lpe.workspace import GpuPdhgWorkspace
    data = data or _Batch([_problem(), _problem(rhs=1.4)])
    args = dict(batch_size=data.batch, n_variables=data.n, n_constraints=data.m, rhs=data.rhs, c=data.c,
                lower=data.lower, upper=data.upper, tau=data.tau, sigma=data.sigma,
                inequality_mask=data.inequality_mask, theta=data.theta, env_ids=("a", "b")[:data.batch],
                chunk_size=8)
    args.update(kwargs)
    return GpuPdhgWorkspace(data.a, **args), data


def _np(t):
    return t.cpu().numpy() if isinstance(t, torch.Tensor) else np.asarray(t)


@pytest.mark.parametrize("iterations", [0, 1, 7, 8, 17, 64])
def test_graph_and_loop_match_oracle_with_chunks_and_tail(device, iterations):
    ws, data = _workspace()
    ref_x, ref_y = data.oracle(iterations)
    ws.run(iterations)
    graph_x, graph_y = [_np(v) for v in ws.state()]
    assert ws.run_stats["graph_chunks"] == iterations // 8
    assert ws.run_stats["python_tail_iterations"] == iterations % 8
    ws.reset()
    ws.run(iterations, use_graph=False)
    loop_x, loop_y = [_np(v) for v in ws.state()]
    np.testing.assert_array_equal(graph_x, loop_x)
    np.testing.assert_array_equal(graph_y, loop_y)
    np.testing.assert_allclose(graph_x, ref_x, atol=1e-12, rtol=1e-12)
    np.testing.assert_allclose(graph_y, ref_y, atol=1e-12, rtol=1e-12)
    assert ws.setup_timing["kernels_per_iteration"] == 4


def test_warm_starts_are_projected_and_match_oracle(device):
    ws, data = _workspace()
    x = np.array([[-2., 9.], [.5, .7]])
    y = np.array([[3., 4.], [-1., -.3]])
    ws.reset(x, y)
    ws.run(27)
    for actual, expected in zip(ws.state(), data.oracle(27, x, y)):
        np.testing.assert_allclose(_np(actual), expected, atol=1e-12, rtol=1e-12)


def test_dynamic_values_vectors_metrics_and_theta(device):
    ws, _ = _workspace()
    updated = _Batch([_problem(.7, rhs=.8, upper=.9, cost=-.6), _problem(1.3, rhs=1.8, upper=1.4, cost=-.4)],
                     theta=.5, primal_weight=2.)
    ws.update_problem(a=updated.a, rhs=updated.rhs, c=updated.c, lower=updated.lower, upper=updated.upper,
                      tau=updated.tau, sigma=updated.sigma, inequality_mask=updated.inequality_mask,
                      theta=updated.theta, env_ids=iter(["a", "b"]))
    with pytest.raises(RuntimeError, match="reset"):
        ws.run(1)
    ws.reset()
    ws.run(49)
    for actual, expected in zip(ws.state(), updated.oracle(49)):
        np.testing.assert_allclose(_np(actual), expected, atol=1e-12, rtol=1e-12)
    assert ws.update_timing["matrix_values_updated"]
    assert ws.theta == .5


def test_accepted_environments_freeze_and_reset_clears_mask(device):
    ws, _ = _workspace()
    ws.run(3)
    before = [_np(v) for v in ws.state()]
    ws.set_accepted(np.array([True, False]))
    ws.run(19)
    after = [_np(v) for v in ws.state()]
    for left, right in zip(before, after):
        np.testing.assert_array_equal(left[0], right[0])
    assert not np.array_equal(before[0][1], after[0][1])
    with pytest.raises(ValueError, match="unfreeze"):
        ws.set_accepted(np.array([False, False]))
    ws.set_accepted(np.array([True, True]))
    all_frozen = [_np(v).copy() for v in ws.state()]
    ws.run(8)
    for expected, actual in zip(all_frozen, ws.state()):
        np.testing.assert_array_equal(expected, _np(actual))
    ws.reset()
    assert ws.accepted.tolist() == [False, False]


@pytest.mark.parametrize("bad", [["b", "a"], ["a", "a"], [True, "b"], ["a"]])
def test_env_id_changes_fail_closed_without_mutating_problem(device, bad):
    ws, _ = _workspace()
    old = ws.rhs
    with pytest.raises(ValueError, match="env_ids"):
        ws.update_problem(rhs=np.array([0., 9., 0., 9.]), env_ids=bad)
    np.testing.assert_array_equal(ws.rhs, old)
    ws.run(1)


def test_pattern_changes_and_cross_environment_edges_fail_closed(device):
    ws, data = _workspace()
    changed = data.a.copy()
    changed.indices[1] = 0          # duplicate column is not canonical
    with pytest.raises(ValueError, match="sorted, unique"):
        ws.update_problem(a=changed, tau=data.tau, sigma=data.sigma)
    changed = data.a.copy()
    changed.indices[1] = 2
    with pytest.raises(ValueError, match="couple"):
        ws.update_problem(a=changed, tau=data.tau, sigma=data.sigma)
    with pytest.raises(ValueError, match="pattern"):
        ws.update_problem(a=csr_matrix(np.eye(4)), tau=data.tau, sigma=data.sigma)
    ws.run(1)


@pytest.mark.parametrize("field,value", [
    ("rhs", [0., np.nan, 0., 1.]), ("c", [0., np.inf, 0., 1.]),
    ("lower", [0., np.nan, 0., 1.]), ("upper", [0., -np.inf, 0., 1.]),
    ("tau", [0., 1., 1., 1.]), ("sigma", [-1., 1., 1., 1.]),
    ("inequality_mask", [0, 1, 0, 1]), ("rhs", [0., 1.]),
])
def test_invalid_updates_do_not_partially_mutate(device, field, value):
    ws, _ = _workspace()
    before = ws.c
    update = dict(c=np.array([-5., -5., -5., -5.]))
    update[field] = np.asarray(value)
    with pytest.raises(ValueError):
        ws.update_problem(**update)
    np.testing.assert_array_equal(ws.c, before)
    ws.run(1)


def test_mask_shape_and_warm_start_fail_before_state_mutation(device):
    ws, _ = _workspace()
    before = [_np(v) for v in ws.state()]
    with pytest.raises(ValueError, match="y"):
        ws.reset(np.ones(4), np.full(4, np.nan))
    for old, new in zip(before, ws.state()):
        np.testing.assert_array_equal(old, _np(new))
    with pytest.raises(ValueError, match="shape"):
        ws.set_accepted(np.array([True]))


def test_matrix_update_requires_new_metric_and_owns_input_buffers(device):
    data = _Batch([_problem(), _problem(rhs=1.4)])
    ws, _ = _workspace(data)
    original = ws.rhs
    data.rhs.fill(99.)
    np.testing.assert_array_equal(ws.rhs, original)
    with pytest.raises(ValueError, match="tau and sigma"):
        ws.update_problem(a=data.a)


def test_zero_row_box_problem(device):
    problem = (csr_matrix((0, 1)), np.empty(0), np.array([0.]), np.array([2.]), np.array([-1.]), 0)
    ws, data = _workspace(_Batch([problem]))
    ws.run(17)
    x, y = ws.state()
    assert tuple(y.shape) == (1, 0)
    np.testing.assert_array_equal(_np(x), [[2.]])
    np.testing.assert_allclose(_np(x), data.oracle(17)[0], atol=1e-12)


def test_explicit_loop_only_and_budget_validation(device):
    ws, _ = _workspace(use_graph=False)
    ws.run(3)
    assert ws.run_stats["graph_chunks"] == 0
    with pytest.raises(RuntimeError, match="no captured graph"):
        ws.run(8, use_graph=True)
    for value in (-1, True, 1.5):
        with pytest.raises(ValueError, match="iterations"):
            ws.run(value)


@pytest.mark.parametrize("name", ["rhs", "c", "lower", "upper", "tau", "sigma", "inequality_mask", "theta"])
def test_constructor_rejects_missing_initial_numeric_buffers(device, name):
    with pytest.raises(ValueError, match="required"):
        _workspace(**{name: None})


def test_rows_in_caller_order_with_inequality_first(device):
    """An inequality row ahead of the equality row: the engine keeps equality
    rows first internally, inputs and outputs stay in the caller's order."""
    flip = lambda p: (csr_matrix(p[0].toarray()[::-1]), p[1][::-1].copy(), p[2], p[3], p[4], p[5])  # noqa: E731
    problems = [flip(_problem()), flip(_problem(1.2, rhs=1.4))]
    data = _Batch(problems, mask=[True, False])
    ws, _ = _workspace(data)
    y0 = np.array([[2., -1.], [-.5, .3]])
    ws.reset(None, y0)
    ws.run(23)
    for actual, expected in zip(ws.state(), data.oracle(23, None, y0)):
        np.testing.assert_allclose(_np(actual), expected, atol=1e-12, rtol=1e-12)
    # the inequality dual (row 0) stays non-positive
    assert (_np(ws.state()[1])[:, 0] <= 0).all()


def test_mask_must_match_across_environments(device):
    data = _Batch([_problem(), _problem(rhs=1.4)])
    with pytest.raises(ValueError, match="same in every environment"):
        _workspace(data, inequality_mask=np.array([False, True, True, False]))
