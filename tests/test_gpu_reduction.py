"""Equality reduction on MI355X: the device kernels against liblpe's host
path (same arithmetic: bit-exact) and the oracle, the reduced engine against
HiGHS, and the reference's reduced-corrector known answers
(tests/test_gpu_reduced_pdhg.py)."""

import ctypes as C

import numpy as np
import pytest
import torch

from lpe import _lib
from oracle import highs
from oracle import reduction as R
from oracle.certificate import certificate, relative_objective_error
from test_oracle_reduction import _problem, forest_lp
from test_reduction_host import gem_problem

pytestmark = pytest.mark.gpu


def _batch_of(p, B, seed):
    """B LPs on p's pattern: bounds, kept rhs and c perturbed (eliminated rows stay homogeneous)."""
    rng = np.random.default_rng(seed)
    a, rhs, lo, up, c, neq = p
    plan = R.plan(p)
    rhs_b = np.tile(rhs, (B, 1))
    kept_eq = [r for r in plan.kept_rows if r < neq]
    ineq = np.arange(neq, a.shape[0])
    rhs_b[:, ineq] += rng.uniform(0.0, 0.2, size=(B, ineq.size))
    lo_b = np.tile(lo, (B, 1))
    up_b = np.tile(up, (B, 1))
    fin = np.isfinite(lo)
    lo_b[:, fin] -= rng.uniform(0.0, 0.5, size=(B, fin.sum()))
    up_b += rng.uniform(0.0, 0.5, size=up_b.shape)
    c_b = np.tile(c, (B, 1)) + rng.normal(scale=0.1, size=(B, c.size))
    del kept_eq
    return rhs_b, lo_b, up_b, c_b


def _vp(t):
    return None if t is None else t.data_ptr()


def test_device_reduce_expand_compress_lift_match_host_path(device):
    from lpe.engine import BatchedLpEngine
    from lpe.reduction import _NativePlan
    md, p = gem_problem()
    B = 6
    a = p[0]
    plan = _NativePlan(a, p[5], p[1], md.var_positions)
    rng = np.random.default_rng(5)
    rhs, lo, up, c = np.tile(p[1], (B, 1)), np.tile(p[2], (B, 1)), np.tile(p[3], (B, 1)), np.tile(p[4], (B, 1))
    lo[:, np.isfinite(p[2])] -= rng.uniform(0.0, 1.0, size=(B, int(np.isfinite(p[2]).sum())))
    c += rng.normal(scale=0.1, size=c.shape)
    var = np.tile(a.data[md.var_positions], (B, 1)) * rng.uniform(0.5, 2.0, size=(B, md.var_positions.size))
    outs = {}
    for on_device in (0, 1):
        dev = device if on_device else torch.device("cpu")
        t = lambda x: torch.as_tensor(np.ascontiguousarray(x), device=dev)   # noqa: E731
        ins = [t(c), t(rhs), t(lo), t(up), t(var)]
        o = dict(c=torch.zeros((B, plan.rn), dtype=torch.float64, device=dev),
                 rhs=torch.zeros((B, plan.rm), dtype=torch.float64, device=dev),
                 lower=torch.zeros((B, plan.rn), dtype=torch.float64, device=dev),
                 upper=torch.zeros((B, plan.rn), dtype=torch.float64, device=dev),
                 var=torch.zeros((B, plan.r_n_var), dtype=torch.float64, device=dev),
                 lw=torch.zeros((B, plan.rn), dtype=torch.int32, device=dev),
                 uw=torch.zeros((B, plan.rn), dtype=torch.int32, device=dev),
                 status=torch.full((B,), 7, dtype=torch.int32, device=dev))
        vin = _lib.LpeVectors()
        vin.c, vin.rhs, vin.lower, vin.upper, vin.var_values = [x.data_ptr() for x in ins]
        vout = _lib.LpeReduced()
        vout.c, vout.rhs, vout.lower, vout.upper = o["c"].data_ptr(), o["rhs"].data_ptr(), o["lower"].data_ptr(), o["upper"].data_ptr()
        vout.var_values, vout.lower_witness, vout.upper_witness = o["var"].data_ptr(), o["lw"].data_ptr(), o["uw"].data_ptr()
        vout.status = o["status"].data_ptr()
        stream = torch.cuda.current_stream().cuda_stream if on_device else None
        _lib.check(plan._lib.lpe_reduction_reduce(plan._h, B, 0, C.byref(vin), C.byref(vout), on_device, stream))
        z = t(rng.normal(size=(B, plan.rn))) if not on_device else t(outs[0]["z"].numpy())
        x = torch.zeros((B, plan.n), dtype=torch.float64, device=dev)
        _lib.check(plan._lib.lpe_reduction_expand(plan._h, B, z.data_ptr(), x.data_ptr(), on_device, stream))
        zc = torch.zeros((B, plan.rn), dtype=torch.float64, device=dev)
        xin = t(outs[0]["xin"].numpy()) if on_device else t(rng.normal(size=(B, plan.n)))
        _lib.check(plan._lib.lpe_reduction_compress(plan._h, B, xin.data_ptr(), zc.data_ptr(), on_device, stream))
        yr = t(outs[0]["yr"].numpy()) if on_device else t(rng.normal(size=(B, plan.rm)))
        xl = torch.zeros((B, plan.n), dtype=torch.float64, device=dev)
        yl = torch.zeros((B, plan.m), dtype=torch.float64, device=dev)
        orig = None
        if on_device:
            orig = BatchedLpEngine(a, p[5], B, var_positions=md.var_positions, method="fixed", device=device)
            orig.update(c, rhs, lo, up, var_values=var)
        _lib.check(plan._lib.lpe_reduction_lift(plan._h, None if orig is None else orig._h, B, 0, ins[0].data_ptr(),
                                                ins[4].data_ptr(), z.data_ptr(), yr.data_ptr(), o["lw"].data_ptr(),
                                                o["uw"].data_ptr(), xl.data_ptr(), yl.data_ptr(), on_device, stream))
        if on_device:
            torch.cuda.synchronize()
            orig.close()
        outs[on_device] = {k: v.cpu() for k, v in o.items()}
        outs[on_device].update(z=z.cpu(), x=x.cpu(), xin=xin.cpu(), zc=zc.cpu(), yr=yr.cpu(), xl=xl.cpu(), yl=yl.cpu())
    h, d = outs[0], outs[1]
    for k in ("c", "rhs", "lower", "upper", "var", "lw", "uw", "status", "x", "zc", "xl"):
        assert torch.equal(h[k], d[k]), k                                   # bit-exact
    assert int(h["status"].abs().sum()) == 0
    # the device lift sums A^T y in another order: agreement to 1e-12
    np.testing.assert_allclose(d["yl"].numpy(), h["yl"].numpy(), rtol=1e-10, atol=1e-10)
    # and the host path against the oracle on LP 0
    ref = R.plan(p)
    p0 = list(p)
    a0 = a.copy(); a0.data[md.var_positions] = var[0]
    p0[0], p0[2], p0[4] = a0, lo[0], c[0]
    red, lw, uw = R.reduce(ref, tuple(p0))
    np.testing.assert_array_equal(h["lower"][0].numpy(), red[2])
    np.testing.assert_array_equal(h["c"][0].numpy(), red[4])
    np.testing.assert_array_equal(h["lw"][0].numpy(), lw)
    x_ref, y_ref = R.lift(ref, tuple(p0), h["z"][0].numpy(), h["yr"][0].numpy(), lw, uw)
    np.testing.assert_allclose(d["yl"][0].numpy(), y_ref, rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_reduced_engine_against_highs(device, precision):
    from lpe.engine import SolveOptions
    from lpe.reduction import ReducedLpEngine
    p = forest_lp(3)
    B = 8
    rhs, lo, up, c = _batch_of(p, B, seed=11)
    eng = ReducedLpEngine(p[0], p[5], B, precision=precision, device=device)
    assert eng.reduced_shape[1] < p[0].shape[1]
    eng.update(c, rhs, lo, up)
    eng.reset()
    info = eng.solve(SolveOptions(max_iterations=20000), fallback=precision == "fp32")
    x, y, status = eng.state()
    assert info["certified"] == B and info["original_rejected"] == 0, info
    x, y = x.cpu().numpy(), y.cpu().numpy()
    for b in range(B):
        pb = (p[0], rhs[b], lo[b], up[b], c[b], p[5])
        ref = highs.solve(*pb)
        cert = certificate(*pb, x[b], y[b])
        assert cert["certificate_passed"], (b, cert)
        assert relative_objective_error(cert["objective"], ref.fun) <= 1e-6, b
    eng.close()


def _equal_box(upper=1.0):
    return _problem([[1.0, -1.0]], [0.0], [0.0, 0.0], [upper, upper], [-1.0, -1.0], 1)


def test_corrector_box_only_reduction(device):
    from lpe.reduction import GpuReducedPdhgCorrector
    corr = GpuReducedPdhgCorrector([_equal_box()], step_safety=0.9)
    assert corr.reduced_problems[0][0].shape == (0, 1)
    assert corr.plan.T.shape == (2, 1)
    res = corr.solve(iterations=3, check_interval=1)
    assert res["all_accepted"] and res["iterations_run"] == 1 and res["cpu_lp_calls"] == 0
    assert tuple(res["x"].shape) == (1, 2) and tuple(res["y"].shape) == (1, 1)
    assert tuple(res["reduced_x"].shape) == (1, 1) and tuple(res["reduced_y"].shape) == (1, 0)
    np.testing.assert_allclose(res["x"].cpu().numpy(), [[1.0, 1.0]], atol=1e-12)
    assert res["metrics"][0]["certificate_passed"]
    assert "ORIGINAL-LP certificate" in res["scope"]


def test_corrector_negative_weight_witnesses(device):
    from lpe.reduction import GpuReducedPdhgCorrector
    lower_opt = _problem([[1.0, 1.0]], [0.0], [0.0, -3.0], [2.0, 0.0], [1.0, 0.0], 1)
    upper_neg = _problem([[1.0, 1.0]], [0.0], [0.0, -2.0], [3.0, 0.0], [-1.0, 0.0], 1)
    corr = GpuReducedPdhgCorrector([lower_opt, upper_neg], step_safety=0.9)
    np.testing.assert_allclose(corr.plan.weights, [1.0, -1.0])
    assert corr.reductions[0].lower_witness.tolist() == [0]
    assert corr.reductions[1].upper_witness.tolist() == [1]
    res = corr.solve(iterations=4, check_interval=1)
    assert res["all_accepted"]
    assert [r["accepted_iteration"] for r in res["metrics"]] == [0, 2]
    np.testing.assert_allclose(res["x"].cpu().numpy(), [[0.0, 0.0], [2.0, -2.0]], atol=1e-12)
    assert all(r["certificate_passed"] for r in res["metrics"])
    np.testing.assert_allclose(res["reduced_x"].cpu().numpy().ravel(), [0.0, 2.0])


def test_corrector_fixed_plan_and_stale_equalities(device):
    from lpe.reduction import GpuReducedPdhgCorrector
    first = GpuReducedPdhgCorrector([_equal_box(1.0)])
    reused = GpuReducedPdhgCorrector([_equal_box(2.0)], plan=first.plan)
    assert reused.plan is first.plan
    np.testing.assert_allclose(reused.reduced_problems[0][3], [2.0])
    stale = list(_equal_box(2.0))
    from scipy.sparse import csr_matrix
    stale[0] = csr_matrix([[1.0, -2.0]])
    with pytest.raises(ValueError, match="fingerprint"):
        GpuReducedPdhgCorrector([tuple(stale)], plan=first.plan)
    stale_rhs = list(_equal_box(2.0))
    stale_rhs[1] = np.asarray([0.25])
    with pytest.raises(ValueError, match="homogeneous"):
        GpuReducedPdhgCorrector([tuple(stale_rhs)], plan=first.plan)


def test_corrector_warm_start_contract_and_zero_budget(device):
    from lpe.reduction import GpuReducedPdhgCorrector
    p = _problem([[1.0, -1.0], [1.0, 0.0]], [0.0, 0.5], [0.0, 0.0], [2.0, 2.0], [-1.0, -1.0], 1)
    corr = GpuReducedPdhgCorrector([p])
    with pytest.raises(ValueError, match="initial_x"):
        corr.solve(initial_x=np.zeros((1, 1)), iterations=0)
    with pytest.raises(ValueError, match="initial_y"):
        corr.solve(initial_y=np.zeros((1, 1)), iterations=0)
    res = corr.solve(initial_x=np.zeros((1, 2)), initial_y=np.zeros((1, 2)), iterations=0, check_interval=1)
    assert not res["all_accepted"] and res["accepted"].tolist() == [False]
    assert res["iterations_run"] == 0 and res["metrics"][0]["accepted_iteration"] == -1
    assert not res["metrics"][0]["certificate_passed"] and res["cpu_lp_calls"] == 0
    assert tuple(res["x"].shape) == (1, 2) and tuple(res["y"].shape) == (1, 2)


def test_corrector_converges_on_a_forest_lp(device):
    from lpe.reduction import GpuReducedPdhgCorrector
    p = forest_lp(1)
    corr = GpuReducedPdhgCorrector([p, p])
    res = corr.solve(iterations=200000, check_interval=500)
    ref = highs.solve(*p)
    for b in range(2):
        assert res["metrics"][b]["certificate_passed"], res["metrics"][b]
        assert relative_objective_error(res["metrics"][b]["objective"], ref.fun) <= 1e-5
