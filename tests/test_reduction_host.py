"""The native reduction plan and its host path (liblpe, no device) against
oracle/reduction.py: plan arrays bit-exact, reduced data bit-exact, the lift
to 1e-12; and the reference's interface (src/lp_equality_reduction.py) on the
known answers of tests/test_lp_equality_reduction.py."""

import os

import numpy as np
import pytest
from scipy.sparse import csr_matrix

from conftest import ROOT
from lpe import sbml, workloads
from lpe.reduction import HomogeneousEqualityReduction, _NativePlan
from oracle import highs
from oracle import reduction as R
from oracle.certificate import certificate, relative_objective_error
from test_oracle_reduction import _problem, forest_lp

FIXTURE = os.path.join(ROOT, "tests", "golden", "gem_consortium.npz")


def known_lps():
    return [
        _problem([[1.0, -1.0]], [0.0], [0.0, 0.0], [1.0, 1.0], [-1.0, -1.0], 1),
        _problem([[-2.0, 1.0]], [0.0], [0.0, 0.0], [1.0, 2.0], [-1.0, 0.0], 1),
        _problem([[1.0, 1.0]], [0.0], [0.0, -1.0], [1.0, 0.0], [0.0, -1.0], 1),
        _problem([[1.0, -1.0, 0.0], [0.0, 1.0, -1.0]], [0.0, 0.0], [-np.inf] * 3, [np.inf] * 3, [0.0] * 3, 2),
        _problem([[1.0, -1.0, 0.0], [0.0, 1.0, -1.0], [-1.0, 0.0, 1.0]], [0.0] * 3, [0.0] * 3, [1.0] * 3,
                 [0.0] * 3, 3),
        _problem([[1.0, -1e-12]], [0.0], [0.0, 0.0], [1.0, 1.0], [0.0, 0.0], 1),
        _problem([[1, -1, 0, 0], [0, 1, -1, 0], [1, 0, -1, 0], [0, 0, 2, 1], [1, 1, 1, 1]],
                 [0, 0, 0, 0, 3], [-1, -1, -1, -4], [1, 1, 1, 4], [1, 2, 3, -1], 4),
    ]


def gem_problem():
    md = sbml.consortium_from_gems(sbml.load_gems(FIXTURE), seed=2)
    cb = workloads.consortium_batch(batch=1, seed=3, model=md)
    return md, cb.problem(0)


def cases():
    out = [(f"known{i}", p) for i, p in enumerate(known_lps())]
    out += [(f"forest{s}", forest_lp(s)) for s in range(4)]
    return out


def _assert_plan_equal(native, ref):
    np.testing.assert_array_equal(native.group, ref.group)
    np.testing.assert_array_equal(native.weight, ref.weight)            # bit-exact
    np.testing.assert_array_equal(native.representative, ref.representative)
    np.testing.assert_array_equal(native.kept, ref.kept_rows)
    np.testing.assert_array_equal(native.eliminated, ref.eliminated_rows)
    assert native.lift_map.shape == ref.lift.shape
    np.testing.assert_array_equal(native.lift_map.indptr, ref.lift.indptr)
    np.testing.assert_array_equal(native.lift_map.indices, ref.lift.indices)
    np.testing.assert_array_equal(native.lift_map.data, ref.lift.data)


@pytest.mark.parametrize("name", [c[0] for c in cases()])
def test_native_plan_equals_oracle(name):
    p = dict(cases())[name]
    ref = R.plan(p)
    nat = _NativePlan(p[0], p[5], p[1])
    _assert_plan_equal(nat, ref)
    red, lw, uw = R.reduce(ref, p)
    # the native reduced pattern keeps exactly the reference's nonzeros
    np.testing.assert_array_equal(nat.reduced_matrix.toarray(), red[0].toarray())
    assert nat.reduced_matrix.nnz == red[0].nnz
    assert nat.rneq == red[5]


def test_native_plan_equals_oracle_on_the_real_gem_consortium():
    md, p = gem_problem()
    ref = R.plan(p)
    nat = _NativePlan(p[0], p[5], p[1], md.var_positions)
    _assert_plan_equal(nat, ref)
    assert nat.n_eliminated == 2411 and nat.rn == 4323
    # per-LP coupling entries stay per-LP in the reduced pattern
    assert nat.r_n_var == md.var_positions.size


@pytest.mark.parametrize("name", [c[0] for c in cases()])
def test_host_reduce_expand_compress_lift_equal_oracle(name):
    p = dict(cases())[name]
    ref = R.plan(p)
    plan = HomogeneousEqualityReduction(p)
    red = plan.reduce(p)
    o_red, o_lw, o_uw = R.reduce(ref, p)
    for k in (1, 2, 3, 4):
        np.testing.assert_array_equal(red.problem[k], o_red[k])
    np.testing.assert_array_equal(red.lower_witness, o_lw)
    np.testing.assert_array_equal(red.upper_witness, o_uw)
    np.testing.assert_array_equal(red.problem[0].toarray(), o_red[0].toarray())
    rng = np.random.default_rng(7)
    z = rng.normal(size=(3, plan.reduced_variables))
    np.testing.assert_array_equal(plan.expand_primal(z), R.expand(ref, z))
    x = rng.normal(size=(3, plan.original_variables))
    np.testing.assert_array_equal(plan.compress_primal(x), R.compress(ref, x))
    # the native host lift (the device kernels' arithmetic) against the oracle
    nat = plan._native
    yr = rng.normal(size=(1, len(ref.kept_rows)))
    import ctypes as C
    from lpe import _lib
    xo = np.zeros((1, ref.n)); yo = np.zeros((1, ref.m))
    c = np.ascontiguousarray(p[4]); zz = np.ascontiguousarray(z[:1])
    lw = np.ascontiguousarray(o_lw, dtype=np.int32); uw = np.ascontiguousarray(o_uw, dtype=np.int32)
    vp = lambda a: a.ctypes.data_as(C.c_void_p)   # noqa: E731
    _lib.check(nat._lib.lpe_reduction_lift(nat._h, None, 1, 1, vp(c), None, vp(zz), vp(yr) if yr.size else None,
                                           vp(lw), vp(uw), vp(xo), vp(yo), 0, None))
    x_ref, y_ref = R.lift(ref, p, zz[0], yr[0], o_lw, o_uw)
    np.testing.assert_array_equal(xo[0], x_ref)
    np.testing.assert_allclose(yo[0], y_ref, rtol=1e-12, atol=1e-12)


def test_reference_interface_known_answers():
    # src/lp_equality_reduction.py interface on tests/test_lp_equality_reduction.py's cases
    p = _problem([[1.0, -1.0]], [0.0], [0.0, 0.0], [1.0, 1.0], [-1.0, -1.0], 1)
    plan = HomogeneousEqualityReduction.from_problem(p)
    reduced = plan.reduce(p)
    assert plan.transform.shape == (2, 1)
    np.testing.assert_array_equal(plan.eliminated_rows, [0])
    assert reduced.problem[0].shape == (0, 1)
    np.testing.assert_allclose(reduced.problem[2:5], ([0.0], [1.0], [-2.0]))
    x = reduced.expand_primal([1.0])
    y = reduced.lift_dual(np.empty(0), np.asarray([-2.0]))
    np.testing.assert_allclose(x, [1.0, 1.0])
    np.testing.assert_allclose(y, [1.0])
    assert certificate(*p, x, y)["certificate_passed"]

    p = _problem([[-2.0, 1.0]], [0.0], [0.0, 0.0], [1.0, 2.0], [-1.0, 0.0], 1)
    plan = HomogeneousEqualityReduction(p)
    reduced = plan.reduce(p)
    np.testing.assert_allclose(plan.transform.toarray(), [[0.5], [1.0]])
    np.testing.assert_array_equal(plan.representatives, [1])
    assert reduced.upper_witness.tolist() == [1]
    np.testing.assert_allclose(reduced.allocate_bound_reduced_cost([-0.5]), [0.0, -0.5])
    y = reduced.lift_dual([], [-0.5])
    np.testing.assert_allclose(y, [0.5])
    assert certificate(*p, reduced.expand_primal([2.0]), y)["certificate_passed"]

    p = _problem([[1.0, -1.0]], [0.0], [-5.0, -5.0], [5.0, 5.0], [0.0, 0.0], 1)
    plan = HomogeneousEqualityReduction(p)
    np.testing.assert_allclose(plan.compress_primal([2.0, 2.0], check=True), [2.0])
    np.testing.assert_allclose(plan.compress_primal([[1.0, 1.0], [2.0, 2.0]]), [[1.0], [2.0]])
    np.testing.assert_allclose(plan.compress_primal([1.0, 2.0]), [1.5])
    with pytest.raises(ValueError, match="does not satisfy"):
        plan.compress_primal([1.0, 2.0], check=True)

    initial = _problem([[1.0, -1.0], [2.0, 3.0]], [0.0, 8.0], [0.0, 0.0], [4.0, 4.0], [1.0, 2.0], 1)
    changed = _problem([[1.0, -1.0], [4.0, -1.0]], [0.0, 7.0], [-1.0, 0.0], [3.0, 2.0], [-2.0, 5.0], 1)
    reduced = HomogeneousEqualityReduction(initial).reduce(changed).problem
    np.testing.assert_allclose(reduced[0].toarray(), [[3.0]])
    np.testing.assert_allclose([reduced[1], reduced[2], reduced[3], reduced[4]], [[7.0], [0.0], [2.0], [3.0]])
    stale = list(changed)
    stale[0] = csr_matrix([[1.0, -2.0], [4.0, -1.0]])
    with pytest.raises(ValueError, match="fingerprint"):
        HomogeneousEqualityReduction(initial).reduce(tuple(stale))
    inhom = list(changed)
    inhom[1] = np.array([0.25, 7.0])
    with pytest.raises(ValueError, match="homogeneous"):
        HomogeneousEqualityReduction(initial).reduce(tuple(inhom))

    bad = _problem([[1.0, -1.0]], [0.0], [1.0, 0.0], [2.0, 0.5], [0.0, 0.0], 1)
    with pytest.raises(ValueError, match="infeasible"):
        HomogeneousEqualityReduction(bad).reduce(bad)

    free = _problem([[1.0, -1.0]], [0.0], [-np.inf, -np.inf], [np.inf, np.inf], [1.0, 0.0], 1)
    y = HomogeneousEqualityReduction(free).reduce(free).lift_dual([], [1.0])
    c = certificate(*free, [0.0, 0.0], y)
    assert not c["certificate_passed"] and c["dual_violation"] == pytest.approx(1.0)
    for value in (True, 0.5, np.nan, np.inf):
        with pytest.raises(ValueError, match="max_scale_ratio"):
            HomogeneousEqualityReduction(p, max_scale_ratio=value)


def test_rows_with_per_lp_entries_are_never_eliminated():
    p = _problem([[1.0, -1.0, 0.0], [0.0, 1.0, -1.0]], [0.0, 0.0], [0.0] * 3, [1.0] * 3, [0.0] * 3, 2)
    a = p[0]
    nat = _NativePlan(a, 2, p[1], var_positions=np.array([3]))      # row 1, column 2
    np.testing.assert_array_equal(nat.eliminated, [0])
    np.testing.assert_array_equal(nat.kept, [1])
    # reduced row 0 = [x1 (group of x0, x1), x2]: the per-LP entry is the second
    np.testing.assert_array_equal(nat.reduced_matrix.toarray(), [[1.0, -1.0]])
    assert nat.r_n_var == 1 and nat.reduced_var_positions.tolist() == [1]


def test_reduced_gem_lp_has_the_original_optimum():
    md, p = gem_problem()
    plan = HomogeneousEqualityReduction(p)
    red = plan.reduce(p)
    full = highs.solve(*p)
    sub = highs.solve(*red.problem)
    assert full.status == "optimal" and sub.status == "optimal"
    assert relative_objective_error(sub.fun, full.fun) <= 1e-9
    ref = R.plan(p)
    _, lw, uw = R.reduce(ref, p)
    x, y = R.lift(ref, p, sub.x, sub.y, lw, uw)
    assert relative_objective_error(float(p[4] @ x), full.fun) <= 1e-9
    assert certificate(*p, x, y)["primal_residual"] <= 1e-7
