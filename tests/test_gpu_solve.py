"""GPU parity of the restarted-averaged PDHG, polish and certificate.

The oracle is HiGHS FP64 (the reference's exact CPU path) plus the CPU
restatements in oracle/.  Bar: every certified LP passes the reference's
certificate (primal <= 1e-5, dual <= 1e-7, relative KKT gap <= 1e-7,
src/gpu_block_lp.py:84) and is within 1e-6 relative objective error of HiGHS.
"""

import numpy as np
import pytest

from lpe import workloads
from oracle import highs, certificate as cert, pdhg as opdhg

pytestmark = pytest.mark.gpu

REL_OBJ_TOL = 1e-6


def _engine(A, neq, batch, **kw):
    from lpe.engine import BatchedLpEngine
    return BatchedLpEngine(A, neq, batch, **kw)


def _uptake(batch=8, m=120, n=180, seed=3):
    return workloads.uptake_batch(batch=batch, seed=seed, m=m, n=n)


def _load_uptake(eng, ub):
    eng.update(ub.c, np.tile(ub.rhs, (ub.batch, 1)), ub.lower, ub.upper)


@pytest.mark.parametrize("method", ["restarted", "halpern"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_restarted_iterates_match_oracle_before_restarts(device, precision, method):
    ub = _uptake(batch=3)
    eng = _engine(ub.S, ub.neq, ub.batch, precision=precision, shared_c=True, method=method)
    _load_uptake(eng, ub)
    eng.reset()
    eng.run(50)
    x, y, _ = eng.state()
    tol = 1e-12 if precision == "fp64" else 1e-5
    for b in range(ub.batch):
        r = opdhg.restarted_pdhg(*ub.problem(b), iterations=50, check_interval=50, precision=precision,
                                 certify=False, halpern=method == "halpern")
        np.testing.assert_allclose(x.cpu().numpy()[b], r["x"], rtol=tol, atol=tol * 1e3)
        np.testing.assert_allclose(y.cpu().numpy()[b], r["y"], rtol=tol, atol=tol * 1e3)


@pytest.mark.parametrize("method", ["restarted", "halpern"])
def test_restarted_fp64_bit_exact_single_chunk(device, method):
    ub = _uptake(batch=2)
    eng = _engine(ub.S, ub.neq, ub.batch, precision="fp64", shared_c=True, method=method)
    _load_uptake(eng, ub)
    eng.reset()
    eng.run(30, chunk=10, use_graph=True)
    gx, gy, _ = eng.state()
    eng.reset()
    eng.run(30, use_graph=False)
    lx, ly, _ = eng.state()
    assert bool((gx == lx).all()) and bool((gy == ly).all())
    r = opdhg.restarted_pdhg(*ub.problem(0), iterations=30, check_interval=30, certify=False,
                             halpern=method == "halpern")
    np.testing.assert_allclose(gx.cpu().numpy()[0], r["x"], rtol=1e-13, atol=1e-10)


def _no_accept_options(iterations):
    """Checks, restarts and primal-weight updates on the device, but no LP
    accepted (gates that nothing passes) and no polish or infeasibility test:
    the device's decisions alone, as the oracle makes them."""
    from lpe.engine import SolveOptions
    return SolveOptions(max_iterations=iterations, check_interval=64, polish=False, eps_primal=1e-300,
                        eps_dual=1e-300, eps_gap=1e-300, infeasibility_interval=-1)


def test_fp64_iterates_match_oracle_across_restarts(device):
    """1024 iterations with 16 device checks: every LP restarts 6 times and
    updates its primal weight at each restart, as the oracle decides; the
    iterates agree to 1e-11 (relative) and the weights to 1e-9."""
    ub = _uptake(batch=4)
    eng = _engine(ub.S, ub.neq, ub.batch, precision="fp64", shared_c=True)
    _load_uptake(eng, ub)
    eng.reset()
    info = eng.solve(_no_accept_options(1024))
    x, y, status = eng.state()
    x, y, om = x.cpu().numpy(), y.cpu().numpy(), eng.primal_weight()
    total = 0
    for b in range(ub.batch):
        r = opdhg.restarted_pdhg(*ub.problem(b), iterations=1024, check_interval=64, certify=False)
        assert r["restarts"] >= 2
        total += r["restarts"]
        np.testing.assert_allclose(x[b], r["x"], rtol=1e-11, atol=1e-11)
        np.testing.assert_allclose(y[b], r["y"], rtol=1e-11, atol=1e-11)
        np.testing.assert_allclose(om[b], r["omega"], rtol=1e-9)
    assert info["restarts"] == total
    assert (status == 2).all()


def test_fp32_iterates_bit_exact_before_restarts_and_within_gamma_bound(device):
    """FP32 iterates: identical to the FP32 oracle (same operations in the same
    order) before the first restart, and within 32 k u (1 + |x|) of the FP64
    oracle after k iterations (u = 2^-24, a gamma_k-type forward error bound;
    measured constant <= 16.4)."""
    ub = _uptake(batch=3)
    u = 2.0 ** -24
    for k in (10, 50, 200):
        eng = _engine(ub.S, ub.neq, ub.batch, precision="fp32", shared_c=True)
        _load_uptake(eng, ub)
        eng.reset()
        eng.run(k)
        x, y, _ = eng.state()
        x, y = x.cpu().numpy(), y.cpu().numpy()
        for b in range(ub.batch):
            r32 = opdhg.restarted_pdhg(*ub.problem(b), iterations=k, check_interval=k, precision="fp32",
                                       certify=False)
            np.testing.assert_array_equal(x[b], r32["x"])
            np.testing.assert_array_equal(y[b], r32["y"])
            r64 = opdhg.restarted_pdhg(*ub.problem(b), iterations=k, check_interval=k, certify=False)
            sx, sy = r64["x_scaled"], r64["y_scaled"]
            ex = np.abs(x[b] / r64["col_scale"] - sx)
            ey = np.abs(y[b] / r64["row_scale"] - sy)
            assert np.all(ex <= 32 * k * u * (1 + np.abs(sx))), (k, b, float(np.max(ex / (k * u * (1 + np.abs(sx))))))
            assert np.all(ey <= 32 * k * u * (1 + np.abs(sy))), (k, b)
        eng.close()


def test_fp32_restart_decisions_match_oracle(device):
    """FP32 iterates over 1024 iterations: the device restarts each LP as often
    as the FP32 oracle; the primal weights then differ only by the last bits
    of their FP64 reductions fed back through FP32 (within 2 %)."""
    ub = _uptake(batch=4)
    eng = _engine(ub.S, ub.neq, ub.batch, precision="fp32", shared_c=True)
    _load_uptake(eng, ub)
    eng.reset()
    info = eng.solve(_no_accept_options(1024))
    om = eng.primal_weight()
    total = 0
    for b in range(ub.batch):
        r = opdhg.restarted_pdhg(*ub.problem(b), iterations=1024, check_interval=64, precision="fp32",
                                 certify=False)
        total += r["restarts"]
        np.testing.assert_allclose(om[b], r["omega"], rtol=2e-2)
    assert info["restarts"] == total


@pytest.mark.parametrize("precision,polish,method", [("fp64", False, "restarted"), ("fp64", True, "restarted"),
                                                     ("fp32", True, "restarted"), ("fp64", False, "halpern"),
                                                     ("fp64", True, "halpern"), ("fp32", True, "halpern")])
def test_solve_small_batch_against_highs(device, precision, polish, method):
    ub = _uptake(batch=8)
    eng = _engine(ub.S, ub.neq, ub.batch, precision=precision, shared_c=True, method=method)
    _load_uptake(eng, ub)
    eng.reset()
    info = eng.solve(max_iterations=40000, check_interval=64, polish=polish)
    x, y, status = eng.state()
    x, y = x.cpu().numpy(), y.cpu().numpy()
    assert info["certified"] == ub.batch, info
    for b in range(ub.batch):
        p = ub.problem(b)
        ref = highs.solve(*p)
        c = cert.certificate(*p, x[b], y[b])
        assert status[b] == 1
        assert c["certificate_passed"], c
        assert cert.relative_objective_error(c["objective"], ref.fun) <= REL_OBJ_TOL


def test_species_lp_full_size_fp32_polish(device):
    p = workloads.species_lp(0)
    A, rhs, lo, hi, c, neq = p
    eng = _engine(A, neq, 2, precision="fp32", shared_c=True)
    eng.update(c, np.tile(rhs, (2, 1)), np.tile(lo, (2, 1)), np.tile(hi, (2, 1)))
    eng.reset()
    info = eng.solve(max_iterations=20000, check_interval=64, polish=True)
    x, y, status = eng.state()
    ref = highs.solve(*p)
    for b in range(2):
        cc = cert.certificate(*p, x.cpu().numpy()[b], y.cpu().numpy()[b])
        assert cc["certificate_passed"], (cc, info)
        assert cert.relative_objective_error(cc["objective"], ref.fun) <= REL_OBJ_TOL


def test_consortium_batch_per_lp_coefficients(device):
    model = workloads.consortium(seed=4, shapes=((120, 180), (150, 220), (90, 130)), n_medium=20)
    cb = workloads.consortium_batch(batch=6, seed=4, model=model)
    md = cb.model
    eng = _engine(md.A, md.neq, cb.batch, precision="fp32", shared_c=True,
                  var_positions=md.coupling_nnz.ravel())
    eng.update(md.c, cb.rhs(), cb.lower(), cb.upper(), var_values=cb.coupling_values().reshape(cb.batch, -1))
    eng.reset()
    # FP32 stalls on some of these small consortium LPs: the FP64 fallback finishes them
    from lpe.engine import SolveOptions
    info, x, y, status = eng.solve_with_fallback(SolveOptions(max_iterations=8000, polish=True),
                                                 SolveOptions(max_iterations=40000, polish=True), capacity=64)
    assert info["certified"] == cb.batch, info
    for b in range(cb.batch):
        p = cb.problem(b)
        ref = highs.solve(*p)
        c = cert.certificate(*p, x.cpu().numpy()[b], y.cpu().numpy()[b])
        assert c["certificate_passed"], c
        assert cert.relative_objective_error(c["objective"], ref.fun) <= REL_OBJ_TOL


def test_device_certificate_matches_oracle(device):
    ub = _uptake(batch=4)
    eng = _engine(ub.S, ub.neq, ub.batch, shared_c=True)
    _load_uptake(eng, ub)
    xs, ys = [], []
    for b in range(ub.batch):
        r = highs.solve(*ub.problem(b))
        xs.append(r.x); ys.append(r.y)
    xs, ys = np.array(xs), np.array(ys)
    xs[1, 5] += 1e-3  # breaks only LP 1 (tests/test_gpu_block_lp.py:51-52)
    metrics = eng.certify(xs, ys).cpu().numpy()
    for b in range(ub.batch):
        ref = cert.certificate(*ub.problem(b), xs[b], ys[b])
        assert bool(metrics[b, 4]) == ref["certificate_passed"]
        assert metrics[b, 0] == ref["primal_residual"]
        assert metrics[b, 1] == ref["dual_violation"]
        np.testing.assert_allclose(metrics[b, 2], ref["relative_kkt_gap"], rtol=1e-10, atol=1e-18)
        np.testing.assert_allclose(metrics[b, 3], ref["objective"], rtol=1e-13)
    assert metrics[0, 4] == 1.0 and metrics[1, 4] == 0.0


def test_nonfinite_and_infeasible_fail_closed(device):
    from scipy.sparse import csr_matrix
    # x <= 0 and x >= 1 (tests/test_gpu_pdhg_corrector.py:98-112)
    A = csr_matrix(np.array([[1.0], [-1.0]]))
    eng = _engine(A, 0, 1, shared_c=True)
    eng.update(np.zeros(1), np.array([[0.0, -1.0]]), np.full((1, 1), -np.inf), np.full((1, 1), np.inf))
    eng.reset()
    info = eng.solve(max_iterations=640, check_interval=64, polish=True)
    _, _, status = eng.state()
    assert info["certified"] == 0 and status[0] != 1
    m = eng.certify(np.full((1, 1), np.nan), np.full((1, 2), -1.0)).cpu().numpy()
    assert m[0, 4] == 0.0


def test_polish_matches_oracle_active_set_least_squares(device):
    """One polish call from the same FP64 PDHG iterate: the engine's active-set
    polish (Jacobi-PCG on both systems) and oracle/polish.py (plain CG on the
    same systems) converge to the same least-squares answers, so the LPs the
    oracle's polish certifies are certified by the engine with the same
    objective (1e-9 relative).  Catches a wrong operator, mask, preconditioner
    or CG scalar in either system."""
    from lpe import _lib
    from oracle import polish as opolish
    model = workloads.consortium(seed=4, shapes=((120, 180), (150, 220), (90, 130)), n_medium=20)
    cb = workloads.consortium_batch(batch=4, seed=4, model=model)
    md = cb.model
    eng = _engine(md.A, md.neq, cb.batch, precision="fp64", shared_c=True, var_positions=md.coupling_nnz.ravel())
    eng.update(md.c, cb.rhs(), cb.lower(), cb.upper(), var_values=cb.coupling_values().reshape(cb.batch, -1))
    eng.reset()
    eng.solve(max_iterations=3000, check_interval=64, polish=False)
    x, y, _ = eng.state()
    x, y = x.cpu().numpy(), y.cpu().numpy()
    sc = eng.scaling()
    oracle_ok = []
    for b in range(cb.batch):
        p = cb.problem(b)
        xo, yo = opolish.polish(*p, x[b], y[b], sc["row_scale"], sc["col_scale"], cg_iterations=400)[:2]
        c = cert.certificate(*p, xo, yo)
        oracle_ok.append((c["certificate_passed"], c["objective"]))
    assert sum(ok for ok, _ in oracle_ok) >= 2, oracle_ok
    eng.polish(np.ones(cb.batch, np.uint8), cg_iterations=200)
    xg, yg, status = eng.state()
    xg = xg.cpu().numpy()
    for b, (ok, obj) in enumerate(oracle_ok):
        if not ok:
            continue
        assert status[b] == _lib.STATUS_OPTIMAL, (b, status)
        got = float(md.c @ xg[b])
        assert abs(got - obj) <= 1e-9 * max(1.0, abs(obj)), (b, got, obj)


def test_status_equals_a_fresh_certificate_after_whole_batch_polish(device):
    """A solve whose LPs are all accepted by one polish of the whole batch
    does not certify its answers a second time (that polish certificate is
    the same computation over the same data).  The statuses it returns must
    equal a fresh device certificate of the returned (x, y) and the host
    certificate - also when some LPs are not started near their optimum and
    the final certificate does run."""
    from lpe.engine import SolveOptions
    # 48 LPs in a 64-wide layout: polishing more than half of it runs on the
    # whole batch (fewer LPs run on a compacted view and are certified again)
    ub = _uptake(batch=48)
    refs = [highs.solve(*ub.problem(b)) for b in range(ub.batch)]
    x0 = np.array([r.x for r in refs])
    y0 = np.array([r.y for r in refs])
    for far in (False, True):
        eng = _engine(ub.S, ub.neq, ub.batch, precision="fp32", shared_c=True)
        _load_uptake(eng, ub)
        xs, ys = x0.copy(), y0.copy()
        if far:
            xs[1::3] = 0.0          # a third of the batch starts at the origin
            ys[1::3] = 0.0
        eng.reset(xs, ys)
        opts = SolveOptions(max_iterations=20000, check_interval=64, polish=True, polish_kkt=1e-4,
                            polish_kkt_start=1e-4)
        info = eng.solve(opts)
        x, y, status = eng.state()
        assert info["certified"] == ub.batch, (far, info)
        fresh = eng.certify(x, y).cpu().numpy()
        xn, yn = x.cpu().numpy(), y.cpu().numpy()
        for b in range(ub.batch):
            host = cert.certificate(*ub.problem(b), xn[b], yn[b])
            assert (status[b] == 1) == bool(fresh[b, 4]) == host["certificate_passed"], (far, b)
            assert cert.relative_objective_error(host["objective"], refs[b].fun) <= REL_OBJ_TOL
        eng.close()


def test_primal_only_cg_passes_give_the_same_answers(device, monkeypatch):
    """Once every dual system has stopped, the polish's remaining CG steps run
    the primal-only passes on a contiguous copy of p1.  They carry the same
    arithmetic as the primal half of the two-system passes, and each LP's CG
    stops at the same step whatever the block lengths, so the answers are
    bit-identical to the two-system passes (LPE_CG_SPLIT=0)."""
    from lpe.engine import SolveOptions
    model = workloads.consortium(seed=4, shapes=((120, 180), (150, 220), (90, 130)), n_medium=20)
    cb = workloads.consortium_batch(batch=96, seed=4, model=model)
    md = cb.model
    answers = []
    for split in ("0", "1"):
        monkeypatch.setenv("LPE_CG_SPLIT", split)
        eng = _engine(md.A, md.neq, cb.batch, precision="fp32", shared_c=True,
                      var_positions=md.coupling_nnz.ravel())
        eng.update(md.c, cb.rhs(), cb.lower(), cb.upper(), var_values=cb.coupling_values().reshape(cb.batch, -1))
        eng.reset()
        info = eng.solve(SolveOptions(max_iterations=4000, check_interval=64, polish=True))
        x, y, status = eng.state()
        answers.append((x.cpu().numpy(), y.cpu().numpy(), status.copy(), info["cg_steps"]))
        eng.close()
    (x0, y0, s0, c0), (x1, y1, s1, c1) = answers
    assert (s0 == s1).all()
    assert int((s1 == 1).sum()) >= cb.batch // 2, s1
    np.testing.assert_array_equal(x0, x1)
    np.testing.assert_array_equal(y0, y1)
