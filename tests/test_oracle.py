"""CPU checks that pin the oracle before it is trusted as the GPU checker.

Each check is chosen so a plausible mistake (dropped term, wrong sign or
index, transposed operand) fails it: reference known answers held by the
reference's tests (tests/golden/reference_known_answers.json), library
routines (scipy.sparse products, HiGHS), brute-force vertex enumeration on
tiny LPs, closed forms and invariants.
"""

import itertools
import json
import os

import numpy as np
import pytest
from scipy.sparse import csr_matrix

from lpe import workloads
from oracle import certificate as cert, highs, pdhg, polish, spmv

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _known():
    with open(os.path.join(GOLDEN, "reference_known_answers.json")) as f:
        return json.load(f)


def _arr(v, none):
    return np.array([none if e is None else e for e in v], dtype=float)


# ----------------------------------------------------------------- products
def test_ordered_spmv_matches_scipy():
    rng = np.random.default_rng(0)
    A = workloads.species_model(200, 300, seed=1).S
    x = rng.standard_normal((3, 300))
    np.testing.assert_allclose(spmv.csr_matvec_ordered(A, x), (A @ x.T).T, rtol=1e-14, atol=1e-12)
    y = rng.standard_normal((3, 200))
    np.testing.assert_allclose(spmv.csr_rmatvec_ordered(A, y), (A.T @ y.T).T, rtol=1e-14, atol=1e-12)
    Ai = csr_matrix((np.round(A.data), A.indices, A.indptr), shape=A.shape)
    xi = rng.integers(-9, 9, size=(2, 300)).astype(float)
    np.testing.assert_array_equal(spmv.csr_matvec_ordered(Ai, xi), (Ai @ xi.T).T)
    # a transposed operand is caught: on a square non-symmetric matrix A x != A^T x
    Q = workloads.species_model(80, 80, seed=3, n_exchange=5).S
    z = rng.standard_normal(80)
    assert not np.allclose(spmv.csr_matvec_ordered(Q, z), spmv.csr_rmatvec_ordered(Q, z))
    np.testing.assert_allclose(spmv.csr_rmatvec_ordered(Q, z), Q.T @ z, rtol=1e-14, atol=1e-12)


def test_ordered_spmv_float32_stays_float32():
    A = workloads.species_model(50, 80, seed=2).S
    A32 = csr_matrix((A.data.astype(np.float32), A.indices, A.indptr), shape=A.shape)
    x = np.ones((1, 80), dtype=np.float32)
    assert spmv.csr_matvec_ordered(A32, x).dtype == np.float32


# ------------------------------------------------------------- certificate
def test_certificate_reference_known_answers():
    k = _known()["certificate_blocks"]
    A = csr_matrix(np.array(k["a_ub"]))
    lo = np.array([b[0] for b in k["bounds"]]); hi = np.array([b[1] for b in k["bounds"]])
    c = np.array(k["c"])
    x = np.array(k["x"]); y = np.array(k["y"])
    probs = [(A, np.array([b]), lo, hi, c, 0) for b in k["b_ub"]]
    rows = [cert.certificate(*p, x[2 * i:2 * i + 2], y[i:i + 1]) for i, p in enumerate(probs)]
    assert all(r["certificate_passed"] for r in rows)
    x[k["perturb_index"]] += k["perturb"]
    rows = [cert.certificate(*p, x[2 * i:2 * i + 2], y[i:i + 1]) for i, p in enumerate(probs)]
    assert [r["certificate_passed"] for r in rows] == k["passed_after_perturb"]


def test_certificate_scale_aware_gap_known_answer():
    k = _known()["certificate_scale"]
    got = []
    for ci, xi in zip(k["c"], k["x"]):
        p = (csr_matrix((0, 1)), np.zeros(0), np.array([k["bounds"][0]]), np.array([k["bounds"][1]]),
             np.array([ci]), 0)
        got.append(cert.certificate(*p, np.array([xi]), np.zeros(0))["certificate_passed"])
    assert got == k["passed"]


def test_certificate_rejects_nonfinite_and_wrong_sign_dual():
    p = (csr_matrix(np.array([[1.0]])), np.array([1.0]), np.array([0.0]), np.array([2.0]), np.array([-1.0]), 0)
    assert cert.certificate(*p, np.array([1.0]), np.array([-1.0]))["certificate_passed"]
    assert not cert.certificate(*p, np.array([np.nan]), np.array([-1.0]))["certificate_passed"]
    # an inequality dual of the wrong sign is a dual violation
    r = cert.certificate(*p, np.array([1.0]), np.array([1.0]))
    assert not r["certificate_passed"] and r["dual_violation"] == 1.0


# ------------------------------------------------------------------ HiGHS
def _brute_force(a, rhs, lo, hi, c, neq):
    """Enumerate vertices of a tiny LP: fix n - rank active constraints."""
    a = a.toarray()
    m, n = a.shape
    best = None
    cons = [(a[i], rhs[i], "eq" if i < neq else "ub") for i in range(m)]
    bound_rows = [(np.eye(n)[j], lo[j]) for j in range(n)] + [(np.eye(n)[j], hi[j]) for j in range(n)]
    candidates = [(r, v) for r, v, _ in cons] + bound_rows
    for subset in itertools.combinations(range(len(candidates)), n):
        M = np.array([candidates[i][0] for i in subset]); v = np.array([candidates[i][1] for i in subset])
        if abs(np.linalg.det(M)) < 1e-12:
            continue
        x = np.linalg.solve(M, v)
        if np.any(x < lo - 1e-9) or np.any(x > hi + 1e-9):
            continue
        if neq and np.max(np.abs(a[:neq] @ x - rhs[:neq])) > 1e-9:
            continue
        if neq < m and np.max(a[neq:] @ x - rhs[neq:]) > 1e-9:
            continue
        f = c @ x
        best = f if best is None or f < best else best
    return best


@pytest.mark.parametrize("seed", range(6))
def test_highs_matches_vertex_enumeration(seed):
    rng = np.random.default_rng(seed)
    n, m = 4, 3
    a = csr_matrix(rng.integers(-3, 4, size=(m, n)).astype(float))
    lo = -rng.uniform(0, 2, n); hi = rng.uniform(1, 3, n)
    x0 = lo + rng.uniform(0.1, 0.9, n) * (hi - lo)
    rhs = a @ x0 + np.r_[0.0, rng.uniform(0, 1, m - 1)]
    c = rng.standard_normal(n)
    r = highs.solve(a, rhs, lo, hi, c, 1)
    assert r.status == "optimal"
    np.testing.assert_allclose(r.fun, _brute_force(a, rhs, lo, hi, c, 1), atol=1e-8)
    assert cert.certificate(a, rhs, lo, hi, c, 1, r.x, r.y)["certificate_passed"]


# ------------------------------------------------------------------ PDHG
def test_persistent_highs_hot_updates_match_cold_solves():
    """The warm CPU path (persistent model, hot updates, previous basis)
    solves each updated LP to the cold linprog optimum; the updates cover
    bounds, rhs and per-LP matrix values, and the warm solves need far fewer
    simplex iterations than a cold start."""
    md = workloads.consortium(2, shapes=((120, 180), (150, 220), (90, 130)), n_medium=12)
    cb = workloads.consortium_batch(batch=3, seed=3, model=md, conc_range=(1e-2, 10.0))
    nxt = workloads.ConsortiumBatch(model=md, conc=cb.conc * 0.7, biomass=cb.biomass * 1.05)
    for b in range(cb.batch):
        p0, p1 = cb.problem(b), nxt.problem(b)
        model = highs.PersistentHighs(*p0)
        r0 = model.solve()
        ref0 = highs.solve(*p0)
        assert r0.status == ref0.status == "optimal"
        assert cert.relative_objective_error(r0.fun, ref0.fun) <= 1e-9
        changed = model.update(p1[0].data, p1[1], p1[2], p1[3])
        assert changed == md.var_positions.size       # every biomass coefficient moved
        r1 = model.solve()
        ref1 = highs.solve(*p1)
        assert r1.status == "optimal"
        assert cert.relative_objective_error(r1.fun, ref1.fun) <= 1e-9
        assert np.abs(r1.x - r0.x).max() > 1e-6       # the update moved the solution
        assert r1.iterations < max(ref1.iterations, 1) / 4
        gates = cert.certificate(*p1, r1.x, r1.y)
        assert gates["certificate_passed"], gates


def test_reference_metric_known_answer():
    k = _known()["metric"]
    s, t = pdhg.reference_metric(csr_matrix(np.array(k["a"])), k["step_safety"], 1.0)
    np.testing.assert_allclose(s, k["sigma"]); np.testing.assert_allclose(t, k["tau"])
    s2, t2 = pdhg.reference_metric(csr_matrix(np.array(k["a"])), k["step_safety"], 2.0)
    np.testing.assert_allclose(s2, 2 * s); np.testing.assert_allclose(t2, t / 2)
    np.testing.assert_allclose(np.outer(s2, t2), np.outer(s, t))


@pytest.mark.parametrize("name", ["inequality_dual_sign", "block_independent_0", "block_independent_1",
                                  "box_only_infinite_unused_bound"])
def test_fixed_pdhg_reference_known_answers(name):
    lp = next(l for l in _known()["lps"] if l["name"] == name)
    a = csr_matrix(np.array(lp["a"])); a.eliminate_zeros()
    p = (a, np.array(lp["rhs"]), _arr(lp["lower"], -np.inf), _arr(lp["upper"], np.inf), np.array(lp["c"]),
         lp["neq"])
    x, y = pdhg.fixed_pdhg(*p, iterations=lp["iterations"], step_safety=lp["step_safety"])
    np.testing.assert_allclose(x, lp["x"], atol=lp["atol_x"])
    if lp["y"] is not None:
        np.testing.assert_allclose(y, lp["y"], atol=lp["atol_y"])


def test_fixed_pdhg_infeasible_never_certifies():
    k = _known()["infeasible"]
    p = (csr_matrix(np.array(k["a"])), np.array(k["rhs"]), _arr(k["lower"], -np.inf), _arr(k["upper"], np.inf),
         np.array(k["c"]), k["neq"])
    x, y = pdhg.fixed_pdhg(*p, iterations=k["iterations"])
    assert not cert.certificate(*p, x, y)["certificate_passed"]


def test_fixed_pdhg_optimum_is_fixed_point():
    ub = workloads.uptake_batch(batch=1, seed=3, m=40, n=60)
    p = ub.problem(0)
    r = highs.solve(*p)
    x, y = pdhg.fixed_pdhg(*p, iterations=5, x0=r.x, y0=r.y)
    np.testing.assert_allclose(x, r.x, atol=1e-7)


def test_restarted_pdhg_certifies_and_matches_highs():
    ub = workloads.uptake_batch(batch=1, seed=3, m=120, n=180)
    p = ub.problem(0)
    ref = highs.solve(*p)
    r = pdhg.restarted_pdhg(*p, iterations=20000)
    assert r["status"] == "optimal"
    assert r["certificate"]["certificate_passed"]
    assert cert.relative_objective_error(r["certificate"]["objective"], ref.fun) <= 1e-6
    a, rhs, lo, hi, c, neq = p
    assert np.all(r["x"] >= lo - 1e-12) and np.all(r["x"] <= hi + 1e-12)


def test_halpern_pdhg_certifies_and_matches_highs():
    # restarted reflected Halpern PDHG (Lu & Yang 2024) with the engine's checks
    for seed in (3, 4):
        ub = workloads.uptake_batch(batch=1, seed=seed, m=120, n=180)
        p = ub.problem(0)
        ref = highs.solve(*p)
        r = pdhg.restarted_pdhg(*p, iterations=20000, halpern=True)
        assert r["status"] == "optimal" and r["certificate"]["certificate_passed"]
        assert cert.relative_objective_error(r["certificate"]["objective"], ref.fun) <= 1e-6
        a, rhs, lo, hi, c, neq = p
        assert np.all(r["x"] >= lo - 1e-12) and np.all(r["x"] <= hi + 1e-12)


def test_halpern_steps_known_answers():
    # box-only LP (no rows): T(x) = clip(x - c, l, u) (unit metric and weight).
    # r = 1: z1 = T(z0) = [-1, 2]; z2 = 2/3 (2 T(z1) - z1) + 1/3 z0 = [-2, 4].
    # r = 0: z1 = (T(z0) + z0) / 2 = [-0.5, 1]; z2 = 2/3 T(z1) + 1/3 z0 = [-1, 2].
    p = (csr_matrix((0, 2)), np.zeros(0), np.array([-5.0, -5.0]), np.array([5.0, 5.0]), np.array([1.0, -2.0]), 0)
    for refl, answers in ((1.0, ([-1.0, 2.0], [-2.0, 4.0])), (0.0, ([-0.5, 1.0], [-1.0, 2.0]))):
        for k, want in zip((1, 2), answers):
            r = pdhg.restarted_pdhg(*p, iterations=k, check_interval=k, certify=False, halpern=True, reflection=refl)
            np.testing.assert_allclose(r["x"], want, rtol=0, atol=1e-15)


def test_restarted_pdhg_average_is_running_mean_and_box_invariant():
    """The average the first check evaluates (and may restart to) is the
    mean of iterates 1..64, each obtained by a separate run stopped there; a
    planted off-by-one (iterates 0..63) does not match."""
    ub = workloads.uptake_batch(batch=1, seed=4, m=40, n=60)
    p = ub.problem(0)
    K = 64
    r = pdhg.restarted_pdhg(*p, iterations=K, check_interval=K, certify=False)
    assert r["history"][0]["iteration"] == K
    ls = p[2] / r["col_scale"]; us = p[3] / r["col_scale"]
    assert np.all(r["x_scaled"] >= ls - 1e-12) and np.all(r["x_scaled"] <= us + 1e-12)
    xs = [pdhg.restarted_pdhg(*p, iterations=j, check_interval=j, certify=False)["x_scaled"] for j in range(1, K + 1)]
    ys = [pdhg.restarted_pdhg(*p, iterations=j, check_interval=j, certify=False)["y_scaled"] for j in range(1, K + 1)]
    x0 = np.clip(np.zeros(p[0].shape[1]), ls, us)
    np.testing.assert_allclose(r["x_candidate_scaled"], np.mean(xs, axis=0), rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(r["y_candidate_scaled"], np.mean(ys, axis=0), rtol=1e-13, atol=1e-15)
    assert np.max(np.abs(r["x_candidate_scaled"] - np.mean([x0] + xs[:-1], axis=0))) > 1e-6
    assert np.all(r["x_candidate_scaled"] >= ls - 1e-12) and np.all(r["x_candidate_scaled"] <= us + 1e-12)


def test_ruiz_scaling_equilibrates():
    A = workloads.species_model(100, 150, seed=7).S
    dr, dc = pdhg.ruiz_scaling(A, 20)
    As = csr_matrix(A.multiply(dr[:, None]).multiply(dc[None, :]))
    rmax = abs(As).max(axis=1).toarray().ravel(); cmax = abs(As).max(axis=0).toarray().ravel()
    nz_r = np.diff(A.indptr) > 0
    assert np.all(np.abs(rmax[nz_r] - 1) < 1e-3) and np.all(np.abs(cmax - 1) < 1e-3)


# ---------------------------------------------------------------- polish
def test_polish_equals_least_squares_on_highs_active_set():
    ub = workloads.uptake_batch(batch=1, seed=5, m=60, n=90)
    p = ub.problem(0)
    a, rhs, lo, hi, c, neq = p
    ref = highs.solve(*p)
    dr, dc = pdhg.ruiz_scaling(a)
    rng = np.random.default_rng(1)
    xn = np.clip(ref.x + 1e-9 * rng.standard_normal(ref.x.size), lo, hi)
    xo, yo, info = polish.polish(*p, xn, ref.y, dr, dc, cg_iterations=400)
    c1 = cert.certificate(*p, xo, yo)
    assert c1["certificate_passed"], c1
    assert cert.relative_objective_error(c1["objective"], ref.fun) <= 1e-9
    # dual least squares via a dense solver on the same active set
    xs, free, rows, _ = polish.active_set(xn, ref.y, csr_matrix(a), rhs, lo, hi, neq, dr, c=c)
    As = (csr_matrix(a).multiply(dr[:, None]).multiply(dc[None, :])).toarray()
    u, *_ = np.linalg.lstsq(As[np.ix_(rows, free)].T, (c * dc)[free], rcond=None)
    np.testing.assert_allclose(yo[rows], dr[rows] * u, rtol=1e-6, atol=1e-8)


def test_polish_with_wrong_active_set_fails_closed():
    ub = workloads.uptake_batch(batch=1, seed=5, m=60, n=90)
    p = ub.problem(0)
    a, rhs, lo, hi, c, neq = p
    ref = highs.solve(*p)
    dr, dc = pdhg.ruiz_scaling(a)
    # a point far from the optimum: the identified set is wrong and the certificate must fail
    x_bad = np.clip(0.5 * (lo + hi), lo, hi)
    xo, yo, _ = polish.polish(*p, x_bad, np.zeros_like(ref.y), dr, dc, cg_iterations=50)
    assert not cert.certificate(*p, xo, yo)["certificate_passed"]
