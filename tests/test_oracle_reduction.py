"""Pins oracle/reduction.py: the reference's known answers
(tests/test_lp_equality_reduction.py, tests/test_gpu_reduced_pdhg.py:74-111),
the identities E T = 0 and L E^T = I, and HiGHS on random forest LPs."""

import numpy as np
import pytest
from scipy.sparse import csr_matrix, random as sparse_random, vstack

from oracle import highs
from oracle import reduction as R
from oracle.certificate import certificate, relative_objective_error


def _problem(a, rhs, lower, upper, c, neq):
    return (csr_matrix(np.asarray(a, dtype=np.float64)), np.asarray(rhs, np.float64), np.asarray(lower, np.float64),
            np.asarray(upper, np.float64), np.asarray(c, np.float64), neq)


def _passes(problem, x, y):
    return certificate(*problem, x, y)["certificate_passed"]


def test_equal_variables_known_answer():
    # min -x0-x1, x0 = x1, 0 <= x <= 1  ->  min -2z, 0 <= z <= 1
    p = _problem([[1.0, -1.0]], [0.0], [0.0, 0.0], [1.0, 1.0], [-1.0, -1.0], 1)
    plan = R.plan(p)
    red, lw, uw = R.reduce(plan, p)
    assert plan.T.shape == (2, 1)
    np.testing.assert_array_equal(plan.eliminated_rows, [0])
    np.testing.assert_allclose(plan.T.toarray(), [[1.0], [1.0]])
    assert red[0].shape == (0, 1)
    np.testing.assert_allclose(red[2:5], ([0.0], [1.0], [-2.0]))
    x, y = R.lift(plan, p, [1.0], np.empty(0), lw, uw)
    np.testing.assert_allclose(x, [1.0, 1.0])
    np.testing.assert_allclose(y, [1.0])
    assert _passes(p, x, y)


def test_scaled_relation_known_answer():
    # -2 x0 + x1 = 0: x = [0.5, 1] z, the representative is the large-weight column
    p = _problem([[-2.0, 1.0]], [0.0], [0.0, 0.0], [1.0, 2.0], [-1.0, 0.0], 1)
    plan = R.plan(p)
    red, lw, uw = R.reduce(plan, p)
    np.testing.assert_allclose(plan.T.toarray(), [[0.5], [1.0]])
    np.testing.assert_array_equal(plan.representative, [1])
    assert uw.tolist() == [1]
    np.testing.assert_allclose(red[3], [2.0])
    np.testing.assert_allclose(red[4], [-0.5])
    x, y = R.lift(plan, p, [2.0], np.empty(0), lw, uw)
    np.testing.assert_allclose(y, [0.5])
    assert _passes(p, x, y)


def test_negative_weight_maps_lower_to_original_upper():
    p = _problem([[1.0, 1.0]], [0.0], [0.0, -1.0], [1.0, 0.0], [0.0, -1.0], 1)
    plan = R.plan(p)
    _, lw, uw = R.reduce(plan, p)
    np.testing.assert_allclose(plan.T.toarray(), [[1.0], [-1.0]])
    assert lw.tolist() == [0]
    x, y = R.lift(plan, p, [0.0], np.empty(0), lw, uw)
    np.testing.assert_allclose(y, [-1.0])
    assert _passes(p, [0.0, 0.0], y)


def test_tree_dual_map_known_answer():
    p = _problem([[1.0, -1.0, 0.0], [0.0, 1.0, -1.0]], [0.0, 0.0], [-np.inf] * 3, [np.inf] * 3, [0.0] * 3, 2)
    plan = R.plan(p)
    d = np.array([1.0, 0.0, -1.0])
    np.testing.assert_allclose(plan.T.toarray(), [[1.0], [1.0], [1.0]])
    np.testing.assert_allclose(plan.lift @ d, [1.0, 1.0])
    np.testing.assert_allclose(p[0].T @ (plan.lift @ d), d)


def test_cycle_row_is_kept():
    p = _problem([[1.0, -1.0, 0.0], [0.0, 1.0, -1.0], [-1.0, 0.0, 1.0]], [0.0] * 3, [0.0] * 3, [1.0] * 3,
                 [0.0] * 3, 3)
    plan = R.plan(p)
    red, _, _ = R.reduce(plan, p)
    np.testing.assert_array_equal(plan.eliminated_rows, [0, 1])
    np.testing.assert_array_equal(plan.kept_rows, [2])
    assert red[5] == 1 and red[0].shape == (1, 1) and red[0].nnz == 0


def test_near_zero_and_nonhomogeneous_rows_are_kept():
    near = _problem([[1.0, -1e-12]], [0.0], [0.0, 0.0], [1.0, 1.0], [0.0, 0.0], 1)
    inhom = _problem([[1.0, -1.0]], [1.0], [0.0, 0.0], [2.0, 2.0], [0.0, 0.0], 1)
    assert R.plan(near).eliminated_rows.size == 0
    assert R.plan(inhom).eliminated_rows.size == 0


def test_fixed_plan_dynamic_vectors_and_stale_equalities():
    initial = _problem([[1.0, -1.0], [2.0, 3.0]], [0.0, 8.0], [0.0, 0.0], [4.0, 4.0], [1.0, 2.0], 1)
    changed = _problem([[1.0, -1.0], [4.0, -1.0]], [0.0, 7.0], [-1.0, 0.0], [3.0, 2.0], [-2.0, 5.0], 1)
    plan = R.plan(initial)
    red, _, _ = R.reduce(plan, changed)
    np.testing.assert_allclose(red[0].toarray(), [[3.0]])
    np.testing.assert_allclose(red[1], [7.0])
    np.testing.assert_allclose(red[2], [0.0])
    np.testing.assert_allclose(red[3], [2.0])
    np.testing.assert_allclose(red[4], [3.0])
    stale = list(changed)
    stale[0] = csr_matrix([[1.0, -2.0], [4.0, -1.0]])
    with pytest.raises(ValueError, match="fingerprint"):
        R.reduce(plan, tuple(stale))
    inhom = list(changed)
    inhom[1] = np.array([0.25, 7.0])
    with pytest.raises(ValueError, match="homogeneous"):
        R.reduce(plan, tuple(inhom))


def test_infeasible_bound_intersection():
    p = _problem([[1.0, -1.0]], [0.0], [1.0, 0.0], [2.0, 0.5], [0.0, 0.0], 1)
    with pytest.raises(ValueError, match="infeasible"):
        R.reduce(R.plan(p), p)


def test_compression_known_answer():
    p = _problem([[1.0, -1.0]], [0.0], [-5.0, -5.0], [5.0, 5.0], [0.0, 0.0], 1)
    plan = R.plan(p)
    np.testing.assert_allclose(R.compress(plan, [2.0, 2.0]), [2.0])
    np.testing.assert_allclose(R.compress(plan, [[1.0, 1.0], [2.0, 2.0]]), [[1.0], [2.0]])
    np.testing.assert_allclose(R.expand(plan, [[1.0], [2.0]]), [[1.0, 1.0], [2.0, 2.0]])
    np.testing.assert_allclose(R.compress(plan, [1.0, 2.0]), [1.5])


def test_missing_bound_lifts_but_fails_original_dual_gate():
    p = _problem([[1.0, -1.0]], [0.0], [-np.inf, -np.inf], [np.inf, np.inf], [1.0, 0.0], 1)
    plan = R.plan(p)
    _, lw, uw = R.reduce(plan, p)
    x, y = R.lift(plan, p, [0.0], np.empty(0), lw, uw)
    c = certificate(*p, x, y)
    assert not c["certificate_passed"]
    assert c["dual_violation"] == pytest.approx(1.0)


@pytest.mark.parametrize("value", [True, 0.5, np.nan, np.inf])
def test_scale_ratio_validation(value):
    p = _problem([[1.0, -1.0]], [0.0], [0.0, 0.0], [1.0, 1.0], [0.0, 0.0], 1)
    with pytest.raises(ValueError, match="max_scale_ratio"):
        R.plan(p, max_scale_ratio=value)


def test_negative_weight_witnesses_per_environment():
    # tests/test_gpu_reduced_pdhg.py:74-111: x1 = -x0 in two environments
    lower_opt = _problem([[1.0, 1.0]], [0.0], [0.0, -3.0], [2.0, 0.0], [1.0, 0.0], 1)
    upper_neg = _problem([[1.0, 1.0]], [0.0], [0.0, -2.0], [3.0, 0.0], [-1.0, 0.0], 1)
    plan = R.plan(lower_opt)
    np.testing.assert_allclose(plan.weight, [1.0, -1.0])
    assert R.reduce(plan, lower_opt)[1].tolist() == [0]
    red, lw, uw = R.reduce(plan, upper_neg)
    assert uw.tolist() == [1]
    x, y = R.lift(plan, upper_neg, [2.0], np.empty(0), lw, uw)
    np.testing.assert_allclose(x, [2.0, -2.0])
    assert _passes(upper_neg, x, y)


def forest_lp(seed, n_chain=30, n_free=12, m_extra=10):
    """Random LP whose equality block holds chains of two-entry rows (with a
    cycle and a badly scaled row), dense-ish kept equalities and inequalities."""
    rng = np.random.default_rng(seed)
    n = n_chain + n_free
    rows = []
    for j in range(0, n_chain - 1):
        if j % 6 == 5:
            continue                                   # break into several trees
        r = np.zeros(n)
        r[j], r[j + 1] = rng.choice([-1, 1]) * rng.uniform(0.5, 2.0), rng.choice([-1, 1]) * rng.uniform(0.5, 2.0)
        rows.append(r)
    r = np.zeros(n); r[0], r[2] = 1.0, -1.0                # closes a cycle in the first tree
    rows.append(r)
    r = np.zeros(n); r[n_chain], r[n_chain + 1] = 1.0, 1e-10   # badly scaled: kept
    rows.append(r)
    eq = np.array(rows)
    kept_eq = sparse_random(3, n, density=0.3, random_state=seed).toarray()
    ub = sparse_random(m_extra, n, density=0.4, random_state=seed + 1).toarray()
    ub[ub != 0] -= 0.5
    a = csr_matrix(np.vstack([eq, kept_eq, ub]))
    x0 = rng.uniform(-1.0, 1.0, n)
    # a feasible point: the equality block's null space through the plan's own T
    neq = eq.shape[0] + 3
    rhs = np.zeros(a.shape[0])
    lower, upper = np.full(n, -5.0), np.full(n, 5.0)
    lower[::7] = -np.inf
    c = rng.normal(size=n)
    p0 = (a, rhs, lower, upper, c, neq)
    plan = R.plan(p0)
    x_feas = R.expand(plan, R.compress(plan, x0) * 0.2)
    act = a @ x_feas
    rhs[eq.shape[0]:neq] = act[eq.shape[0]:neq]
    rhs[neq:] = act[neq:] + rng.uniform(0.1, 1.0, m_extra)
    return (a, rhs, lower, upper, c, neq)


@pytest.mark.parametrize("seed", range(6))
def test_identities_and_highs_optimum(seed):
    p = forest_lp(seed)
    plan = R.plan(p)
    a = p[0]
    assert len(plan.eliminated_rows) >= 15
    assert 0 not in plan.kept_rows and a.shape[0] - 1 >= plan.kept_rows[-1]
    E = a[plan.eliminated_rows]
    assert abs(E @ plan.T).max() <= 1e-12
    np.testing.assert_allclose((plan.lift @ E.T).toarray(), np.eye(len(plan.eliminated_rows)), atol=1e-10)
    # every original column has exactly one T entry; the representatives carry weight 1
    assert (np.diff(plan.T.indptr) == 1).all()
    np.testing.assert_array_equal(plan.weight[plan.representative], 1.0)
    assert (np.abs(plan.weight) <= 1.0).all()
    # compression inverts expansion on the null space
    z = np.random.default_rng(seed).normal(size=plan.reduced_variables)
    np.testing.assert_allclose(R.compress(plan, R.expand(plan, z)), z, rtol=1e-13, atol=1e-14)
    red, lw, uw = R.reduce(plan, p)
    full = highs.solve(*p)
    sub = highs.solve(*red)
    assert full.status == "optimal" and sub.status == "optimal"
    assert relative_objective_error(sub.fun, full.fun) <= 1e-9
    x, y = R.lift(plan, p, sub.x, sub.y, lw, uw)
    c = certificate(*p, x, y)
    assert c["certificate_passed"], c


def test_brute_force_tiny_nullspace():
    # the reduced LP of a 3-cycle plus a pendant edge equals the LP with the
    # null space parameterised by hand: x1 = x0, x2 = x1, x3 = -2 x2
    p = _problem([[1, -1, 0, 0], [0, 1, -1, 0], [1, 0, -1, 0], [0, 0, 2, 1], [1, 1, 1, 1]],
                 [0, 0, 0, 0, 3], [-1, -1, -1, -4], [1, 1, 1, 4], [1, 2, 3, -1], 4)
    plan = R.plan(p)
    np.testing.assert_array_equal(plan.eliminated_rows, [0, 1, 3])
    np.testing.assert_array_equal(plan.kept_rows, [2, 4])
    red, _, _ = R.reduce(plan, p)
    # weights normalised to the largest |alpha| (x3 = -2 x0): x = [-0.5, -0.5, -0.5, 1] z
    np.testing.assert_allclose(plan.weight, [-0.5, -0.5, -0.5, 1.0])
    # the kept cycle row x0 - x2 cancels to an empty row; the inequality sums the weights
    np.testing.assert_allclose(red[0].toarray(), [[0.0], [-0.5]])
    assert red[0].nnz == 1 and red[5] == 1
    np.testing.assert_allclose(red[4], [-0.5 * 6 - 1])
    np.testing.assert_allclose(red[2], [-2.0])
    np.testing.assert_allclose(red[3], [2.0])
    ref = highs.solve(*p)
    sub = highs.solve(*red)
    assert relative_objective_error(sub.fun, ref.fun) <= 1e-12
