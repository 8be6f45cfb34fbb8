"""The infeasibility-certificate restatement (oracle/infeasibility.py) against
closed-form rays whose LP status HiGHS confirms, weak duality on feasible LPs,
and planted mistakes."""

import numpy as np
import pytest
from scipy.sparse import csr_matrix

from oracle import highs, infeasibility as inf
from lpe import workloads


def _lp(a, rhs, lo, hi, c, neq):
    return csr_matrix(np.array(a, dtype=float)), np.array(rhs, float), np.array(lo, float), np.array(hi, float), \
        np.array(c, float), neq


def test_farkas_ray_of_a_row_range_infeasible_lp():
    # x1 + x2 = 3 with 0 <= x <= 1: dy = 1 gives r = (-1, -1), F = 3 - 2 = 1
    p = _lp([[1, 1]], [3], [0, 0], [1, 1], [1, 1], 1)
    assert highs.solve(*p).status == "infeasible"
    assert inf.certificate_status(*p, dx=np.zeros(2), dy=np.array([1.0])) == "primal_infeasible"
    assert inf.dual_ray(p[0], p[1], p[2], p[3], 1, [1.0]) == (1.0, 0.0)
    # the opposite sign is no certificate (F = -3 - 0 < 0)
    assert inf.certificate_status(*p, dx=np.zeros(2), dy=np.array([-1.0])) is None


def test_farkas_ray_needs_the_bound_terms_and_the_inequality_sign():
    # x1 - x2 <= -2, 0 <= x <= 1: dy = -1 (y_ub <= 0), r = (1, -1), F = 2 + 0 - 1 = 1
    p = _lp([[1, -1]], [-2], [0, 0], [1, 1], [0, 0], 0)
    assert highs.solve(*p).status == "infeasible"
    assert inf.certificate_status(*p, dx=np.zeros(2), dy=np.array([-1.0])) == "primal_infeasible"
    # a positive inequality dual violates the cone
    F, viol = inf.dual_ray(p[0], p[1], p[2], p[3], 0, [1.0])
    assert viol == 1.0
    # with x2 unbounded above the same ray is infeasible to use (r_2 < 0 with u = inf)
    q = _lp([[1, -1]], [-2], [0, 0], [1, np.inf], [0, 0], 0)
    assert highs.solve(*q).status == "optimal"
    assert inf.certificate_status(*q, dx=np.zeros(2), dy=np.array([-1.0])) is None


def test_primal_ray_of_an_unbounded_lp():
    # min -x1 subject to x1 - x2 = 0, x >= 0: dx = (1, 1) keeps A x, lowers c.x
    p = _lp([[1, -1]], [0], [0, 0], [np.inf, np.inf], [-1, 0], 1)
    assert highs.solve(*p).status == "unbounded"
    assert inf.certificate_status(*p, dx=np.array([1.0, 1.0]), dy=np.zeros(1)) == "dual_infeasible"
    # a ray leaving the equality or a finite bound is no certificate
    assert inf.certificate_status(*p, dx=np.array([1.0, 0.5]), dy=np.zeros(1)) is None
    r = _lp([[1, -1]], [0], [0, 0], [np.inf, 5.0], [-1, 0], 1)
    assert highs.solve(*r).status == "optimal"
    assert inf.certificate_status(*r, dx=np.array([1.0, 1.0]), dy=np.zeros(1)) is None


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_no_ray_certifies_a_feasible_bounded_lp(seed):
    """Weak duality: for a feasible x, F <= violation * (|slack|_1 + |x|_1), so
    random rays, the optimal dual and iterate-like differences never pass."""
    rng = np.random.default_rng(seed)
    sb = workloads.status_batch(n_each=1, seed=seed, m=40, n=60)
    a, rhs, lo, hi, c, neq = sb.problem(0)
    ref = highs.solve(a, rhs, lo, hi, c, neq)
    assert ref.status == "optimal"
    for _ in range(200):
        dy = rng.normal(size=a.shape[0]) * 10 ** rng.uniform(-3, 3)
        dx = rng.normal(size=a.shape[1]) * 10 ** rng.uniform(-3, 3)
        assert inf.certificate_status(a, rhs, lo, hi, c, neq, dx, dy) is None
    assert inf.certificate_status(a, rhs, lo, hi, c, neq, ref.x, ref.y) is None


def test_status_batch_statuses_agree_with_highs():
    sb = workloads.status_batch(n_each=2, seed=5, m=60, n=90)
    for b in range(sb.batch):
        assert highs.solve(*sb.problem(b)).status == sb.kind[b], b


def test_known_optimum_batch_is_the_highs_optimum():
    """The constructed optimum is HiGHS's (unique): same fluxes, duals and
    active sets, and it passes the certificate."""
    from oracle import certificate as cert
    kb = workloads.known_optimum_batch(batch=3, seed=7, m=60, n=100, n_ineq=15)
    for b in range(kb.batch):
        p = kb.problem(b)
        ref = highs.solve(*p)
        assert ref.status == "optimal"
        np.testing.assert_allclose(ref.x, kb.x_star[b], atol=1e-7)
        np.testing.assert_allclose(ref.y, kb.y_star[b], atol=1e-7)
        assert cert.certificate(*p, kb.x_star[b], kb.y_star[b])["certificate_passed"]
