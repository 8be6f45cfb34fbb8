"""GPU parity of the sparse kernels and the fixed-metric PDHG operator.

Restates tests/test_gpu_pdhg_workspace.py:35-104 (graph == loop bit-exact;
CPU operator at 1e-12; projected warm starts) against the oracle's ordered
restatement, which the engine matches bit for bit.
"""

import numpy as np
import pytest
from scipy.sparse import csr_matrix

from lpe import workloads
from oracle import spmv as ospmv, pdhg as opdhg

pytestmark = pytest.mark.gpu


def _engine(A, neq, batch, **kw):
    from lpe.engine import BatchedLpEngine
    return BatchedLpEngine(A, neq, batch, **kw)


def test_spmv_fp64_bit_exact_against_ordered_host(device):
    rng = np.random.default_rng(0)
    sp = workloads.species_model(300, 450, seed=5)
    A = sp.S
    B = 70  # not a multiple of 64: exercises padding
    eng = _engine(A, A.shape[0], B)
    x = rng.standard_normal((B, A.shape[1])) * 100
    y = rng.standard_normal((B, A.shape[0]))
    ax = eng.spmv(x).cpu().numpy()
    aty = eng.spmv(y, transpose=True).cpu().numpy()
    np.testing.assert_array_equal(ax, ospmv.csr_matvec_ordered(A, x))
    np.testing.assert_array_equal(aty, ospmv.csr_rmatvec_ordered(A, y))
    # library routine: exact on integer-valued data
    xi = rng.integers(-50, 50, size=(B, A.shape[1])).astype(float)
    Ai = csr_matrix((np.round(A.data), A.indices, A.indptr), shape=A.shape)
    eng2 = _engine(Ai, A.shape[0], B)
    np.testing.assert_array_equal(eng2.spmv(xi).cpu().numpy(), (Ai @ xi.T).T)


def test_spmv_fp32_matches_float32_host(device):
    rng = np.random.default_rng(1)
    A = workloads.species_model(200, 300, seed=6).S
    B = 64
    eng = _engine(A, A.shape[0], B)
    x = rng.standard_normal((B, A.shape[1]))
    got = eng.spmv(x, precision="fp32").cpu().numpy()
    A32 = csr_matrix((A.data.astype(np.float32), A.indices, A.indptr), shape=A.shape)
    ref = ospmv.csr_matvec_ordered(A32, x.astype(np.float32)).astype(np.float64)
    np.testing.assert_array_equal(got, ref)


def _two_problems():
    # tests/test_gpu_pdhg_workspace.py:18-21
    def prob(scale=1., rhs=1., upper=2., cost=-1.):
        return (csr_matrix(np.array([[scale, -1.], [1., 1.]])), np.array([0., rhs]),
                np.array([0., 0.]), np.array([upper, upper]), np.array([cost, -.25]), 1)
    return [prob(), prob(rhs=1.4)]


def _fixed_engine(problems, **kw):
    A = problems[0][0]
    eng = _engine(A, problems[0][5], len(problems), method="fixed", **kw)
    eng.update(np.stack([p[4] for p in problems]), np.stack([p[1] for p in problems]),
               np.stack([p[2] for p in problems]), np.stack([p[3] for p in problems]))
    eng.reset()
    return eng


@pytest.mark.parametrize("iterations", [0, 1, 7, 8, 17, 64])
def test_fixed_pdhg_graph_loop_and_cpu_operator(device, iterations):
    probs = _two_problems()
    eng = _fixed_engine(probs)
    eng.run(iterations, chunk=8, use_graph=True)
    gx, gy, _ = [v.cpu().numpy() if hasattr(v, "cpu") else v for v in eng.state()]
    eng.reset()
    eng.run(iterations, chunk=8, use_graph=False)
    lx, ly, _ = [v.cpu().numpy() if hasattr(v, "cpu") else v for v in eng.state()]
    np.testing.assert_array_equal(gx, lx)
    np.testing.assert_array_equal(gy, ly)
    for b, p in enumerate(probs):
        rx, ry = opdhg.fixed_pdhg(*p, iterations=iterations)
        np.testing.assert_array_equal(gx[b], rx)
        np.testing.assert_array_equal(gy[b], ry)


def test_fixed_pdhg_warm_start_projection(device):
    probs = _two_problems()
    eng = _fixed_engine(probs)
    x = np.array([[-2., 9.], [.5, .7]])
    y = np.array([[3., 4.], [-1., -.3]])
    eng.reset(x, y)
    eng.run(27)
    gx, gy, _ = eng.state()
    for b, p in enumerate(probs):
        rx, ry = opdhg.fixed_pdhg(*p, iterations=27, x0=x[b], y0=y[b])
        np.testing.assert_allclose(gx.cpu().numpy()[b], rx, atol=1e-12, rtol=1e-12)
        np.testing.assert_allclose(gy.cpu().numpy()[b], ry, atol=1e-12, rtol=1e-12)


def test_fixed_metric_known_answer(device):
    # tests/test_gpu_pdhg_corrector.py:115-120
    A = csr_matrix(np.array([[2.0, -1.0]]))
    eng = _engine(A, 1, 1, method="fixed", step_safety=0.9)
    eng.update(np.zeros((1, 2)), np.ones((1, 1)), np.zeros((1, 2)), np.full((1, 2), 2.0))
    sc = eng.scaling()
    np.testing.assert_allclose(sc["sigma"], [0.3])
    np.testing.assert_allclose(sc["tau"], [0.45, 0.9])


def test_fixed_pdhg_consortium_batch_per_lp_coefficients(device):
    cb = workloads.consortium_batch(batch=5, seed=4,
                                    model=workloads.consortium(seed=4, shapes=((60, 90), (80, 110), (50, 70)), n_medium=12))
    md = cb.model
    eng = _engine(md.A, md.neq, cb.batch, method="fixed", var_positions=md.coupling_nnz.ravel())
    eng.update(np.tile(md.c, (cb.batch, 1)), cb.rhs(), cb.lower(), cb.upper(),
               var_values=cb.coupling_values().reshape(cb.batch, -1))
    eng.reset()
    eng.run(40)
    gx, gy, _ = eng.state()
    for b in range(cb.batch):
        rx, ry = opdhg.fixed_pdhg(*cb.problem(b), iterations=40)
        np.testing.assert_array_equal(gx.cpu().numpy()[b], rx)
        np.testing.assert_array_equal(gy.cpu().numpy()[b], ry)
