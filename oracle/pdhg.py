"""PDHG restatements (test infrastructure; see oracle/__init__.py).

Two methods:

* ``fixed_pdhg`` - the reference's fixed-metric PDHG recurrence
  (tests/test_gpu_pdhg_workspace.py:35-50; src/gpu_pdhg_corrector.py:340-358;
  kernels src/gpu_pdhg_workspace.py:122-134) with its diagonal metric
  ``reference_metric`` (src/gpu_pdhg_corrector.py:185-196).
  Pinned by: the reference's known-answer metric (sigma = [0.3],
  tau = [0.45, 0.9], tests/test_gpu_pdhg_corrector.py:115-120), the
  reciprocal primal-weight product (:140-164), and convergence of its
  known-answer LPs (:29-95) to the closed-form optimum.

* ``restarted_pdhg`` - restarted-averaged PDHG on the Ruiz-equilibrated
  problem with a Pock-Chambolle diagonal metric, KKT-based restarts and
  primal-weight smoothing (Applegate et al., "Practical large-scale linear
  programming using primal-dual hybrid gradient", NeurIPS 2021; Lu & Yang,
  cuPDLP.jl 2023).  The reference has no restarted method
  (src/gpu_pdhg_corrector.py:86-94 is plain theta=1 PDHG).
  Pinned by: invariants (iterates stay in the box and dual cone; an optimal
  pair is a fixed point; the average is the running mean) and by HiGHS
  optima on the synthetic workloads (tests/test_oracle.py).
  With ``halpern=True`` the iterations are restarted reflected Halpern PDHG
  (Lu & Yang, "Restarted Halpern PDHG for linear programming", 2024), pinned
  by closed-form iterates on a box-only LP (plain and reflected) and by HiGHS
  optima (tests/test_oracle.py).

Both use the engine's summation order (oracle.spmv) and separate multiply /
add roundings; ``precision='fp32'`` runs the iterate arithmetic in float32.
"""

from __future__ import annotations

import numpy as np
from scipy.sparse import csr_matrix

from .spmv import csr_matvec_ordered, csr_rmatvec_ordered
from .certificate import certificate, GATES


def reference_metric(a, step_safety=0.99, primal_weight=1.0):
    """sigma_i = w*eta/sum_j|a_ij|, tau_j = eta/(w*sum_i|a_ij|); empty -> 1.
    The absolute sums run in storage order (rows) and row order (columns)."""
    absolute = abs(csr_matrix(a))
    absolute.sort_indices()
    row_sum = csr_matvec_ordered(absolute, np.ones(absolute.shape[1]))
    col_sum = csr_rmatvec_ordered(absolute, np.ones(absolute.shape[0]))
    base_sigma = np.where(row_sum > 0.0, step_safety / np.where(row_sum > 0, row_sum, 1.0), 1.0)
    base_tau = np.where(col_sum > 0.0, step_safety / np.where(col_sum > 0, col_sum, 1.0), 1.0)
    return base_sigma * primal_weight, base_tau / primal_weight


def _dual_step(y, rhs, ax, sigma, ineq, frozen=None):
    delta = sigma * (rhs - ax)
    cand = y + delta
    out = np.where(ineq & ~np.isnan(cand), np.minimum(cand, 0.0), cand)
    if frozen is not None:
        out = np.where(frozen, y, out)
    return out


def _primal_step(x, aty, c, tau, lo, hi, theta, frozen=None):
    step = tau * (c - aty)
    cand = x - step
    nxt = np.where(np.isnan(cand), cand, np.minimum(np.maximum(cand, lo), hi))
    if frozen is not None:
        nxt = np.where(frozen, x, nxt)
    ext = theta * (nxt - x)
    return nxt, nxt + ext


def fixed_pdhg(a, rhs, lower, upper, c, neq, *, iterations, sigma=None, tau=None, theta=1.0,
               step_safety=0.99, primal_weight=1.0, x0=None, y0=None, inequality_mask=None):
    """Fixed-metric PDHG, x and y in original units (one LP).  Inequality
    rows are rows >= neq, or the rows ``inequality_mask`` marks (the
    workspace's ``dual_upper`` = 0 rows, src/gpu_pdhg_workspace.py:255-257)."""
    a = csr_matrix(a)
    m, n = a.shape
    if sigma is None or tau is None:
        sigma, tau = reference_metric(a, step_safety, primal_weight)
    ineq = np.arange(m) >= neq if inequality_mask is None else np.asarray(inequality_mask, dtype=bool)
    x = np.zeros(n) if x0 is None else np.asarray(x0, dtype=np.float64).copy()
    y = np.zeros(m) if y0 is None else np.asarray(y0, dtype=np.float64).copy()
    x = np.minimum(np.maximum(x, lower), upper)
    y = np.where(ineq, np.minimum(y, 0.0), y)
    xbar = x.copy()
    for _ in range(iterations):
        if m:
            y = _dual_step(y, rhs, csr_matvec_ordered(a, xbar), sigma, ineq)
        x, xbar = _primal_step(x, csr_rmatvec_ordered(a, y) if m else np.zeros(n), c, tau, lower, upper, theta)
    return x, y


def ruiz_scaling(a, iterations=10):
    """Simultaneous row/column inf-norm equilibration (as lpe's ruiz_equilibrate)."""
    a = csr_matrix(a)
    m, n = a.shape
    dr = np.ones(m); dc = np.ones(n)
    coo = a.tocoo()
    for _ in range(iterations):
        v = np.abs(coo.data * dr[coo.row] * dc[coo.col])
        rmax = np.zeros(m); cmax = np.zeros(n)
        np.maximum.at(rmax, coo.row, v)
        np.maximum.at(cmax, coo.col, v)
        dr = np.where(rmax > 0, dr / np.sqrt(np.where(rmax > 0, rmax, 1.0)), dr)
        dc = np.where(cmax > 0, dc / np.sqrt(np.where(cmax > 0, cmax, 1.0)), dc)
    return dr, dc


def shared_metric(a_scaled, eta=0.99):
    absolute = abs(csr_matrix(a_scaled))
    rs = np.asarray(absolute.sum(axis=1)).ravel()
    cs = np.asarray(absolute.sum(axis=0)).ravel()
    return (np.where(rs > 0, eta / np.where(rs > 0, rs, 1.0), 1.0),
            np.where(cs > 0, eta / np.where(cs > 0, cs, 1.0), 1.0))


def _kkt_values(As, bs, cs, ls, us, ineq, x, y, dr, dc, neq):
    """Scaled-space KKT pieces and original-unit certificate values (FP64)."""
    ax = csr_matvec_ordered(As, x)
    aty = csr_rmatvec_ordered(As, y)
    res = bs - ax
    rp = np.where(ineq, np.minimum(res, 0.0), res)
    rc = cs - aty
    lfin, ufin = np.isfinite(ls), np.isfinite(us)
    rdl = np.where(~lfin & (rc > 0), rc, 0.0)
    rdu = np.where(~ufin & (rc < 0), -rc, 0.0)
    bt = np.where((rc > 0) & lfin, ls * rc, np.where((rc < 0) & ufin, us * rc, 0.0))
    pobj = float(np.sum(cs * x))
    dobj = float(np.sum(bs * y) + np.sum(bt))
    rpn = float(np.sqrt(np.sum(rp * rp)))
    rdn = float(np.sqrt(np.sum(rdl * rdl + rdu * rdu)))
    gap = abs(pobj - dobj)
    perr = np.where(ineq, -res, np.abs(res)) / dr
    tgt = np.where(rc >= 0, ls, us)
    bgap = np.sum(np.where(np.isfinite(tgt), np.abs(rc * (x - tgt)), 0.0))
    rgap = np.sum(np.where(ineq, np.abs(y * res), 0.0))
    bviol = np.maximum(ls - x, x - us) * dc
    dviol = np.maximum(rdl, rdu) / dc
    yorig = np.where(ineq, y * dr, -np.inf)
    primal = max(0.0, float(np.max(perr, initial=-np.inf)), float(np.max(bviol, initial=-np.inf)))
    dual = max(0.0, float(np.max(yorig, initial=-np.inf)), float(np.max(dviol, initial=-np.inf)))
    cgap = (bgap + rgap) / max(1.0, abs(pobj))
    finite = bool(np.isfinite(ax).all() and np.isfinite(y).all() and np.isfinite(rc).all() and np.isfinite(x).all())
    return dict(rp=rpn, rd=rdn, pobj=pobj, dobj=dobj, gap=gap, primal=primal, dual=dual, cgap=cgap, finite=finite)


def restarted_pdhg(a, rhs, lower, upper, c, neq, *, iterations, check_interval=64, precision="fp64",
                   x0=None, y0=None, ruiz_iterations=10, eta=0.99, theta=1.0, omega0=None,
                   certify=True, gates=GATES, halpern=False, reflection=1.0):
    """Restarted-averaged PDHG for one LP, decisions as in the engine's k_check_decide.

    ``halpern=True``: restarted reflected Halpern PDHG (Lu & Yang, "Restarted
    Halpern PDHG for linear programming", 2024) with the same checks: each
    iteration applies the primal-first PDHG operator T to z = (x, y) and moves
    to lam ((1 + r) T(z) - r z) + (1 - lam) z0, lam = (k + 1) / (k + 2), k the
    iterations since the last restart and z0 the restart point; at a check the
    candidates are z and T(z) of the chunk's last iteration (instead of the
    current iterate and the average), and a restart moves z and z0 to the
    chosen one."""
    a = csr_matrix(a, dtype=np.float64)
    a.sort_indices()
    m, n = a.shape
    T = np.float32 if precision == "fp32" else np.float64
    dr, dc = ruiz_scaling(a, ruiz_iterations)
    As = csr_matrix((a.data * dr[np.repeat(np.arange(m), np.diff(a.indptr))] * dc[a.indices], a.indices, a.indptr),
                    shape=a.shape)
    sig, tau = shared_metric(As, eta)
    bs, cs = rhs * dr, c * dc
    ls, us = lower * (1.0 / dc), upper * (1.0 / dc)
    ineq = np.arange(m) >= neq
    bnorm, cnorm = float(np.sqrt(np.sum(bs * bs))), float(np.sqrt(np.sum(cs * cs)))
    AsT = As.astype(T)
    x = (np.zeros(n) if x0 is None else np.asarray(x0, dtype=np.float64) * (1.0 / dc)).astype(T)
    y = (np.zeros(m) if y0 is None else np.asarray(y0, dtype=np.float64) * (1.0 / dr)).astype(T)
    lsT, usT, bsT, csT = ls.astype(T), us.astype(T), bs.astype(T), cs.astype(T)
    sigT, tauT = sig.astype(T), tau.astype(T)
    x = np.minimum(np.maximum(x, lsT), usT)
    y = np.where(ineq, np.minimum(y, T(0)), y).astype(T)
    xbar, xlast, ylast = x.copy(), x.copy(), y.copy()
    xsum, ysum = np.zeros(n, T), np.zeros(m, T)
    if omega0 is None or omega0 <= 0:   # PDLP default ||c^|| / ||b^||
        omega0 = cnorm / bnorm if (bnorm > 0 and cnorm > 0) else 1.0
        omega0 = min(max(omega0, 1e-8), 1e8)
    omega, kkt_last, kkt_prev, cnt, iters, restarts = float(omega0), np.inf, np.inf, 0, 0, 0
    xa = ya = None
    status, accepted_at = "running", -1
    history = []

    def evaluate(xx, yy):
        return _kkt_values(As, bs, cs, ls, us, ineq, xx.astype(np.float64), yy.astype(np.float64), dr, dc, neq)

    def kkt_w(v):
        return float(np.sqrt(omega * omega * v["rp"] ** 2 + v["rd"] ** 2 / (omega * omega) + v["gap"] ** 2))

    def rel(v):
        return max(v["rp"] / (1 + bnorm), v["rd"] / (1 + cnorm), v["gap"] / (1 + abs(v["pobj"]) + abs(v["dobj"])))

    def passes(v):
        return v["finite"] and np.isfinite(v["cgap"]) and v["primal"] <= gates[0] and v["dual"] <= gates[1] and v["cgap"] <= gates[2]

    cur = evaluate(x, y)
    if certify and passes(cur):
        status, accepted_at = "optimal", 0
    refl = T(reflection)
    while status == "running" and iters < iterations:
        k = min(check_interval, iterations - iters)
        om = T(omega); iom = T(1.0 / omega)
        for _ in range(k):
            if halpern:
                lam = T(cnt + 1) / T(cnt + 2)
                aty = csr_rmatvec_ordered(AsT, y).astype(T) if m else np.zeros(n, T)
                xp, xbar = _primal_step(x, aty, csT, tauT * iom, lsT, usT, T(theta))
                xp, xbar = xp.astype(T), xbar.astype(T)
                xn = (lam * ((T(1) + refl) * xp - refl * x) + (T(1) - lam) * xlast).astype(T)
                yp = y
                if m:
                    ax = csr_matvec_ordered(AsT, xbar).astype(T)
                    yp = _dual_step(y, bsT, ax, sigT * om, ineq).astype(T)
                    y = (lam * ((T(1) + refl) * yp - refl * y) + (T(1) - lam) * ylast).astype(T)
                x = xn
                cnt += 1
                continue
            if m:
                ax = csr_matvec_ordered(AsT, xbar).astype(T)
                y = _dual_step(y, bsT, ax, sigT * om, ineq).astype(T)
                ysum = (ysum + y).astype(T)
            aty = csr_rmatvec_ordered(AsT, y).astype(T) if m else np.zeros(n, T)
            x, xbar = _primal_step(x, aty, csT, tauT * iom, lsT, usT, T(theta))
            x, xbar = x.astype(T), xbar.astype(T)
            xsum = (xsum + x).astype(T)
        iters += k
        div = cnt if halpern else cnt + k
        fp = None
        if halpern:
            # probe: T(z) of the current iterate (the second candidate and the
            # restart point); fixed-point residual |z - T(z)| in the w-norm
            aty = csr_rmatvec_ordered(AsT, y).astype(T) if m else np.zeros(n, T)
            xp, xbar_p = _primal_step(x, aty, csT, tauT * iom, lsT, usT, T(theta))
            xp, xbar_p = xp.astype(T), xbar_p.astype(T)
            yp = y
            if m:
                ax = csr_matvec_ordered(AsT, xbar_p).astype(T)
                yp = _dual_step(y, bsT, ax, sigT * om, ineq).astype(T)
            xa, ya = xp.astype(np.float64), yp.astype(np.float64)
            fx = float(np.sum((x.astype(np.float64) - xa) ** 2))
            fy = float(np.sum((y.astype(np.float64) - ya) ** 2))
            fp = float(np.sqrt(omega * fx + fy / omega))
        else:
            xa, ya = (xsum.astype(np.float64) / div), (ysum.astype(np.float64) / div)
        vc, va = evaluate(x, y), evaluate(xa, ya)
        kc, ka = kkt_w(vc), kkt_w(va)
        best = 1 if ka < kc else 0
        history.append(dict(iteration=iters, kkt_cur=kc, kkt_avg=ka,
                            rel=min(rel(vc), rel(va)) if halpern else rel(va if best else vc), omega=omega))
        if certify and (passes(vc) or passes(va)):
            if not passes(vc):
                x, y = xa.astype(T), ya.astype(T)
            status, accepted_at = "optimal", iters
            break
        if halpern:
            best = 1                         # Halpern restarts to T(z), on the fixed-point residual
            kc = ka = fp
        kcand = ka if best else kc
        restart = False
        if not np.isfinite(kkt_last):
            kkt_last = kcand
        else:
            restart = kcand <= 0.2 * kkt_last or (kcand <= 0.8 * kkt_last and kcand > kkt_prev) or div >= 0.36 * iters
        kkt_prev = kcand
        if restart:
            cx, cy = (xa.astype(T), ya.astype(T)) if best else (x, y)
            dx = float(np.sqrt(np.sum((cx.astype(np.float64) - xlast.astype(np.float64)) ** 2)))
            dy = float(np.sqrt(np.sum((cy.astype(np.float64) - ylast.astype(np.float64)) ** 2)))
            if dx > 1e-10 and dy > 1e-10 and np.isfinite(dx) and np.isfinite(dy):
                omega = float(np.exp(0.5 * np.log(dy / dx) + 0.5 * np.log(omega)))
                omega = min(max(omega, 1e-8), 1e8)
            x, y = cx.copy(), cy.copy()
            xbar, xlast, ylast = x.copy(), x.copy(), y.copy()
            xsum = np.zeros(n, T); ysum = np.zeros(m, T)
            kkt_last, kkt_prev, cnt = kcand, np.inf, 0
            if halpern:
                # the restart reference is the fixed-point residual at the new point
                om2, iom2 = T(omega), T(1.0 / omega)
                aty = csr_rmatvec_ordered(AsT, y).astype(T) if m else np.zeros(n, T)
                xq, xbq = _primal_step(x, aty, csT, tauT * iom2, lsT, usT, T(theta))
                xq, xbq = xq.astype(T), xbq.astype(T)
                yq = y
                if m:
                    ax = csr_matvec_ordered(AsT, xbq).astype(T)
                    yq = _dual_step(y, bsT, ax, sigT * om2, ineq).astype(T)
                fx = float(np.sum((x.astype(np.float64) - xq.astype(np.float64)) ** 2))
                fy = float(np.sum((y.astype(np.float64) - yq.astype(np.float64)) ** 2))
                kkt_last = float(np.sqrt(omega * fx + fy / omega))
            restarts += 1
        elif not halpern:
            cnt += k
    xo = x.astype(np.float64) * dc
    yo = y.astype(np.float64) * dr
    return dict(x=xo, y=yo, x_scaled=x, y_scaled=y, x_candidate_scaled=xa, y_candidate_scaled=ya,
                status=status if status != "running" else "iteration_limit",
                iterations=iters, accepted_iteration=accepted_at, restarts=restarts, omega=omega,
                history=history, row_scale=dr, col_scale=dc, sigma=sig, tau=tau,
                certificate=certificate(a, rhs, lower, upper, c, neq, xo, yo))
