"""Status and active-set parity with HiGHS (src/cpu_repeated_lp.py:341-345,
:694-699: infeasible and unbounded LPs are reported as such and never pass
as success).

* A batch mixing feasible, primal infeasible and unbounded LPs of one
  pattern: on the product path (FP32 iterates, FP64 continuation of
  uncertified LPs) every LP's status equals HiGHS's (optimal / infeasible /
  unbounded -> LPE_STATUS_OPTIMAL / PRIMAL_INFEASIBLE / DUAL_INFEASIBLE).
* LPs built around a unique, primal- and dual-nondegenerate optimum: the
  engine's active set (columns at their lower / upper bound, inequality rows
  with a nonzero dual) equals HiGHS's index for index, and its fluxes and
  duals equal HiGHS's to 1e-5.
"""

import numpy as np
import pytest

from lpe import _lib, workloads
from oracle import highs

pytestmark = pytest.mark.gpu

STATUS_OF = {"optimal": _lib.STATUS_OPTIMAL, "infeasible": _lib.STATUS_PRIMAL_INFEASIBLE,
             "unbounded": _lib.STATUS_DUAL_INFEASIBLE}


def _status_run(device, precision, budget=20000):
    from lpe.engine import BatchedLpEngine, SolveOptions
    sb = workloads.status_batch(n_each=4, seed=5)
    ref = [highs.solve(*sb.problem(b)) for b in range(sb.batch)]
    assert [r.status for r in ref] == sb.kind
    eng = BatchedLpEngine(sb.A, sb.neq, sb.batch, precision=precision, shared_c=True, device=device)
    eng.update(sb.c, sb.rhs, sb.lower, sb.upper)
    eng.reset()
    opts = SolveOptions(max_iterations=budget, check_interval=64)
    if precision == "fp32":
        # the product path: FP32 iterates, LPs left uncertified continue in FP64
        info, x, _, status = eng.solve_with_fallback(opts)
    else:
        info = eng.solve(opts)
        x, _, status = eng.state()
    print(f"\n{precision}: statuses {list(status)} after {info['iterations']} iterations")
    x = x.cpu().numpy()
    for b, r in enumerate(ref):
        if r.status == "optimal":
            assert status[b] == _lib.STATUS_OPTIMAL
            assert abs(sb.c @ x[b] - r.fun) <= 1e-6 * max(1.0, abs(r.fun)), b
    eng.close()
    return sb, ref, status


def test_status_matches_highs_on_the_product_path(device):
    """FP32 iterates with the FP64 continuation of uncertified LPs: optimal,
    primal infeasible and unbounded LPs all end with HiGHS's status."""
    _, ref, status = _status_run(device, "fp32")
    assert list(status) == [STATUS_OF[r.status] for r in ref]


def test_primal_infeasibility_is_certified_in_fp64(device):
    """FP64 iterates: every primal infeasible LP ends PRIMAL_INFEASIBLE (a dual
    ray certificate, tolerance 1e-8); an unbounded LP is never reported
    optimal or primal infeasible (its primal-ray candidate, the last step,
    converges to the ray only as O(1/k) and may meet 1e-8 after the budget)."""
    _, ref, status = _status_run(device, "fp64")
    for b, r in enumerate(ref):
        if r.status == "infeasible":
            assert status[b] == _lib.STATUS_PRIMAL_INFEASIBLE, b
        elif r.status == "unbounded":
            assert status[b] in (_lib.STATUS_DUAL_INFEASIBLE, _lib.STATUS_LIMIT), b


def test_backend_reports_infeasible_and_unbounded_as_failures(device):
    """solve_batch (src/cpu_repeated_lp.py:783-818 contract): success only for
    certified optima; the message names the status."""
    from lpe.backend import GpuLpBackend
    from lpe.requests import problem_request
    sb = workloads.status_batch(n_each=1, seed=6)
    be = GpuLpBackend(precision="fp32")
    res = be.solve_batch([problem_request(sb.problem(b)) for b in range(sb.batch)],
                         environment_ids=list(range(sb.batch)))
    for b, r in enumerate(res):
        assert r.success == (sb.kind[b] == "optimal"), (b, r.message)
        if sb.kind[b] == "infeasible":
            assert "primal_infeasible" in r.message
        elif sb.kind[b] == "unbounded":
            assert "dual_infeasible" in r.message
    be.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_active_set_and_fluxes_match_highs_at_a_unique_optimum(device, precision):
    from lpe.engine import BatchedLpEngine, SolveOptions
    kb = workloads.known_optimum_batch(batch=8, seed=7)
    eng = BatchedLpEngine(kb.A, kb.neq, kb.batch, precision=precision, shared_c=False, device=device)
    eng.update(kb.c, kb.rhs, kb.lower, kb.upper)
    eng.reset()
    opts = SolveOptions(max_iterations=20000, check_interval=64)
    if precision == "fp32":
        info, x, y, status = eng.solve_with_fallback(opts)     # the product path
    else:
        info = eng.solve(opts)
        x, y, status = eng.state()
    B = kb.batch
    assert int((status == _lib.STATUS_OPTIMAL).sum()) == B, (info, list(status))
    x, y = x.cpu().numpy(), y.cpu().numpy()
    for b in range(kb.batch):
        p = kb.problem(b)
        ref = highs.solve(*p)
        lo, up = p[2], p[3]
        width = np.maximum(up - lo, 1.0)
        # HiGHS's active set is the constructed one
        np.testing.assert_array_equal(np.abs(ref.x - lo) <= 1e-9 * width, kb.at_lower[b])
        np.testing.assert_array_equal(np.abs(up - ref.x) <= 1e-9 * width, kb.at_upper[b])
        np.testing.assert_array_equal(ref.y[kb.neq:] < -1e-9, kb.active_rows[b])
        # the engine's, index for index
        np.testing.assert_array_equal(np.abs(x[b] - lo) <= 1e-9 * width, kb.at_lower[b], err_msg=f"LP {b}")
        np.testing.assert_array_equal(np.abs(up - x[b]) <= 1e-9 * width, kb.at_upper[b], err_msg=f"LP {b}")
        np.testing.assert_array_equal(y[b, kb.neq:] < -1e-9, kb.active_rows[b], err_msg=f"LP {b}")
        np.testing.assert_array_equal(y[b, kb.neq:][~kb.active_rows[b]], 0.0)
        # fluxes and duals of the unique optimum, to the certificate's primal
        # gate (1e-5) - the dual within its gap gate at these magnitudes
        np.testing.assert_allclose(x[b], ref.x, atol=1e-5, rtol=0)
        np.testing.assert_allclose(y[b], ref.y, atol=1e-5, rtol=0)
    eng.close()
