"""lpe_certify_noalloc (the handle-free certificate for callers that cannot
allocate) against the handle-based lpe_certify on the same iterates: two
calls in a row on one device, the second on a smaller problem and batch, so
stale scratch contents of the first call would show.  Max-based metrics,
flags and the activity / reduced-cost products are the same arithmetic in
the same order (exact); sums differ only in summation order (1e-12)."""

import ctypes as C

import numpy as np
import pytest
import torch

from lpe import _lib, workloads

SUM_RTOL = 1e-12


def _compare(ref: np.ndarray, got: np.ndarray, batch: int):
    assert got.shape == (batch, 8)
    for k in (_lib.METRIC_NAMES.index("primal_residual"), _lib.METRIC_NAMES.index("dual_violation"),
              _lib.METRIC_NAMES.index("certificate_passed"), _lib.METRIC_NAMES.index("finite")):
        np.testing.assert_array_equal(got[:, k], ref[:, k], err_msg=_lib.METRIC_NAMES[k])
    for k in (2, 3, 6, 7):           # gap, objective, bound and row complementarity sums
        fin = np.isfinite(ref[:, k])
        np.testing.assert_array_equal(np.isfinite(got[:, k]), fin)
        np.testing.assert_allclose(got[fin, k], ref[fin, k], rtol=SUM_RTOL, atol=1e-300,
                                   err_msg=_lib.METRIC_NAMES[k])


def _engine_iterates(eng, iterations):
    """Iterates of a short solve: some LPs certified, others not."""
    from lpe.engine import SolveOptions
    eng.reset()
    eng.solve(SolveOptions(max_iterations=iterations, check_interval=64, polish=True))
    x, y, _ = eng.state()
    return x, y


@pytest.mark.gpu
def test_noalloc_certificate_matches_lpe_certify_twice(device):
    from lpe.engine import BatchedLpEngine
    from lpe.noalloc import DeviceBatch

    # call 1: consortium LPs (3663 x 5301) with per-environment coupling coefficients
    B1 = 48
    cb = workloads.consortium_batch(batch=B1, seed=3, conc_range=(1e-2, 10.0))
    md = cb.model
    eng = BatchedLpEngine(md.A, md.neq, B1, var_positions=md.var_positions, precision="fp32", shared_c=True,
                          device=device)
    eng.update(md.c, cb.rhs(), cb.lower(), cb.upper(), var_values=cb.coupling_values())
    x1, y1 = _engine_iterates(eng, 6000)
    x1[1::3, 11] += 1e-3                        # every third LP moved off its certified point
    x1[5, 7] = float("nan")                     # a non-finite primal fails closed in both
    y1[9, -1] = 3.0                             # a wrong-signed inequality dual
    ref1 = eng.certify(x1, y1).cpu().numpy()
    values = np.tile(md.A.data, (B1, 1))
    values[:, md.var_positions] = cb.coupling_values()
    db1 = DeviceBatch(md.A, md.neq, cb.rhs(), cb.lower(), cb.upper(), md.c, values=values, device=device)
    met1 = torch.full((B1, 8), -7.0, dtype=torch.float64, device=device)
    db1.certify(x1, y1, met1)
    torch.cuda.synchronize()
    got1 = met1.cpu().numpy()
    _compare(ref1, got1, B1)
    assert 0 < int(got1[:, 4].sum()) < B1, "the batch must mix certified and rejected LPs"
    assert got1[5, 5] == 0.0 and got1[5, 4] == 0.0
    eng.close()

    # call 2: smaller LPs (1000 x 1500 uptake batch) and a smaller batch, per-LP c
    B2 = 20
    ub = workloads.uptake_batch(batch=B2, seed=4)
    eng2 = BatchedLpEngine(ub.S, ub.neq, B2, precision="fp64", shared_c=False, device=device)
    c2 = np.tile(ub.c, (B2, 1))
    c2[:, 3] = np.linspace(-0.5, 0.5, B2)
    rhs2 = np.tile(ub.rhs, (B2, 1))
    eng2.update(c2, rhs2, ub.lower, ub.upper)
    x2, y2 = _engine_iterates(eng2, 2048)
    ref2 = eng2.certify(x2, y2).cpu().numpy()
    db2 = DeviceBatch(ub.S, ub.neq, rhs2, ub.lower, ub.upper, c2, device=device)
    met2 = torch.full((B2, 8), -7.0, dtype=torch.float64, device=device)
    db2.certify(x2, y2, met2)
    torch.cuda.synchronize()
    _compare(ref2, met2.cpu().numpy(), B2)
    # a prefix of the batch (the same problem, fewer LPs)
    met3 = torch.full((B2, 8), -7.0, dtype=torch.float64, device=device)
    db2.certify(x2, y2, met3, batch=7)
    torch.cuda.synchronize()
    got3 = met3.cpu().numpy()
    _compare(ref2[:7], got3[:7], 7)
    assert (got3[7:] == -7.0).all(), "LPs past the batch must be left untouched"
    eng2.close()


@pytest.mark.gpu
def test_noalloc_certificate_rejects_batches_beyond_the_scratch(device):
    from lpe.noalloc import DeviceBatch, max_batch
    ub = workloads.uptake_batch(batch=2, seed=4)
    db = DeviceBatch(ub.S, ub.neq, np.tile(ub.rhs, (2, 1)), ub.lower, ub.upper, ub.c, device=device)
    limit = max_batch(*ub.S.shape)
    assert limit > 2
    x = torch.zeros((2, ub.S.shape[1]), dtype=torch.float64, device=device)
    y = torch.zeros((2, ub.S.shape[0]), dtype=torch.float64, device=device)
    met = torch.zeros((2, 8), dtype=torch.float64, device=device)
    p = _lib.LpeBatchProblem.from_buffer_copy(db._p)
    p.batch = limit + 1
    rc = _lib.load().lpe_certify_noalloc(C.byref(p), x.data_ptr(), y.data_ptr(), met.data_ptr(), 0.0, 0.0, 0.0,
                                         torch.cuda.current_stream().cuda_stream)
    assert rc == _lib.LPE_EINVAL
    assert "static scratch" in _lib.last_error()
