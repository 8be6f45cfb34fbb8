"""Python handle over liblpe: one batch of shared-pattern LPs resident on the GPU.

``BatchedLpEngine`` is the host side of the C ABI in include/lpe.h.  Device
arrays are torch tensors on the current HIP device; every call enqueues on
torch's current stream.  Per-LP arrays are batch-major ``[B, len]`` FP64, the
layout the reference's services accept (src/gpu_pdhg_workspace.py:169-186).
"""

from __future__ import annotations

import ctypes as C
import dataclasses
from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch
from scipy.sparse import csr_matrix

from . import _lib
from ._lib import check


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def _device_array(value, shape, name, device, *, allow_flat=True) -> torch.Tensor:
    """FP64 contiguous device tensor of ``shape`` (flat inputs are reshaped)."""
    if isinstance(value, torch.Tensor):
        t = value.to(device=device, dtype=torch.float64)
    else:
        t = torch.as_tensor(np.asarray(value, dtype=np.float64), device=device)
    flat = (int(np.prod(shape)),)
    if tuple(t.shape) == flat and allow_flat:
        t = t.reshape(shape)
    elif tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} must have shape {tuple(shape)} or {flat}")
    return t.contiguous()


@dataclass
class SolveOptions:
    """Controls of ``BatchedLpEngine.solve`` (mirrors lpe_solve_opts)."""

    max_iterations: int = 20000
    check_interval: int = 64
    polish: bool = True
    polish_kkt: float = 1e-4
    polish_cg_iterations: int = 200
    use_graph: bool = True
    rows_per_wave: int = 0
    lps_per_lane: int = 0
    eps_primal: float = 1e-5
    eps_dual: float = 1e-7
    eps_gap: float = 1e-7
    stop_after_failed_polishes: int = 0
    polish_kkt_start: float = 0.0
    polish_stable_checks: int = 0
    first_check_interval: int = 0
    infeasibility_interval: int = 0
    eps_infeasible: float = 0.0

    def to_c(self) -> _lib.LpeSolveOpts:
        o = _lib.LpeSolveOpts()
        o.max_iterations = int(self.max_iterations)
        o.check_interval = int(self.check_interval)
        o.polish = 1 if self.polish else 0
        o.polish_kkt = float(self.polish_kkt)
        o.polish_cg_iterations = int(self.polish_cg_iterations)
        o.use_graph = 1 if self.use_graph else 0
        o.rows_per_wave = int(self.rows_per_wave)
        o.lps_per_lane = int(self.lps_per_lane)
        o.eps_certify_primal = float(self.eps_primal)
        o.eps_certify_dual = float(self.eps_dual)
        o.eps_certify_gap = float(self.eps_gap)
        o.stop_after_failed_polishes = int(self.stop_after_failed_polishes)
        o.polish_kkt_start = float(self.polish_kkt_start)
        o.polish_stable_checks = int(self.polish_stable_checks)
        o.first_check_interval = int(self.first_check_interval)
        o.infeasibility_interval = int(self.infeasibility_interval)
        o.eps_infeasible = float(self.eps_infeasible)
        return o


class BatchedLpEngine:
    """B LPs sharing the sparsity pattern of ``A`` (m x n CSR, rows [A_eq; A_ub]).

    Parameters
    ----------
    A:
        Template matrix.  Its values are the coefficients of every LP except
        at ``var_positions`` (indices into ``A.data`` after canonicalisation),
        whose per-LP values are given to :meth:`update`.
    neq:
        Number of leading equality rows.
    precision:
        ``"fp64"`` or ``"fp32"`` (iterate arithmetic of the PDHG bulk work;
        checks, polish and certificates are FP64 either way).
    method:
        ``"restarted"`` (restarted-averaged PDHG on the equilibrated problem),
        ``"halpern"`` (restarted reflected Halpern PDHG on the equilibrated
        problem, same checks, restarts and polish) or ``"fixed"`` (the
        reference's fixed-metric PDHG, src/gpu_pdhg_corrector.py:185-196 and
        :340-358).
    """

    def __init__(self, A, neq: int, batch: int, *, var_positions=None, precision: str = "fp64",
                 method: str = "restarted", step_safety: float = 0.99, theta: float = 1.0,
                 primal_weight: Optional[float] = None, ruiz_iterations: int = 10, shared_c: bool = False,
                 device: Optional[torch.device] = None):
        self._lib = _lib.load()
        if not torch.cuda.is_available():
            raise RuntimeError("BatchedLpEngine requires a HIP device")
        A = csr_matrix(A, dtype=np.float64, copy=True)
        A.sort_indices()
        self.A = A
        self.m, self.n = A.shape
        self.neq = int(neq)
        self.batch = int(batch)
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.precision = precision
        self.method = method
        self.shared_c = bool(shared_c)
        var_positions = np.zeros(0, dtype=np.int64) if var_positions is None else np.asarray(var_positions, dtype=np.int64).ravel()
        self.var_positions = var_positions
        self.n_var = int(var_positions.size)
        self._rp = np.ascontiguousarray(A.indptr, dtype=np.int32)
        self._ci = np.ascontiguousarray(A.indices, dtype=np.int32)
        self._val = np.ascontiguousarray(A.data, dtype=np.float64)
        self._vp = np.ascontiguousarray(var_positions, dtype=np.int64)
        d = _lib.LpeDesc()
        d.m, d.n, d.neq, d.batch, d.nnz = self.m, self.n, self.neq, self.batch, int(A.nnz)
        d.row_ptr = self._rp.ctypes.data_as(C.POINTER(C.c_int32))
        d.col_idx = self._ci.ctypes.data_as(C.POINTER(C.c_int32))
        d.values = self._val.ctypes.data_as(C.POINTER(C.c_double))
        d.n_var = self.n_var
        d.var_pos = self._vp.ctypes.data_as(C.POINTER(C.c_int64))
        d.precision = {"fp64": _lib.PREC_FP64, "fp32": _lib.PREC_FP32}[precision]
        d.method = {"fixed": _lib.METHOD_FIXED, "restarted": _lib.METHOD_RESTARTED,
                    "halpern": _lib.METHOD_HALPERN}[method]
        d.step_safety = float(step_safety)
        d.theta = float(theta)
        # restarted / Halpern: None -> ||c^||/||b^|| per LP (PDLP default); fixed: 1.0
        d.primal_weight = float(primal_weight) if primal_weight is not None else (0.0 if method != "fixed" else 1.0)
        d.ruiz_iterations = int(ruiz_iterations)
        d.shared_c = 1 if shared_c else 0
        handle = C.c_void_p()
        with torch.cuda.device(self.device):
            check(self._lib.lpe_create(C.byref(d), C.byref(handle)))
        self._h = handle
        self._vectors = None
        self.last_info: dict = {}

    # ------------------------------------------------------------------ setup
    def update(self, c, rhs, lower, upper, var_values=None) -> None:
        """Validated all-or-nothing copy of the per-LP vectors (lpe_update)."""
        dev = self.device
        B, m, n = self.batch, self.m, self.n
        c_t = _device_array(c, (n,) if self.shared_c else (B, n), "c", dev)
        rhs_t = _device_array(rhs, (B, m), "rhs", dev)
        lo_t = _device_array(lower, (B, n), "lower", dev)
        up_t = _device_array(upper, (B, n), "upper", dev)
        vv_t = None
        if var_values is not None:
            vv_t = _device_array(var_values, (B, self.n_var), "var_values", dev)
        v = _lib.LpeVectors()
        v.c = c_t.data_ptr(); v.rhs = rhs_t.data_ptr(); v.lower = lo_t.data_ptr(); v.upper = up_t.data_ptr()
        v.var_values = vv_t.data_ptr() if vv_t is not None else None
        with torch.cuda.device(dev):
            check(self._lib.lpe_update(self._h, C.byref(v), _stream()))
        self._vectors = (c_t, rhs_t, lo_t, up_t, vv_t)

    def reset(self, x0=None, y0=None) -> None:
        """Cold or warm start, projected into the box / dual cone (lpe_reset)."""
        x_t = None if x0 is None else _device_array(x0, (self.batch, self.n), "x", self.device)
        y_t = None if y0 is None else _device_array(y0, (self.batch, self.m), "y", self.device)
        with torch.cuda.device(self.device):
            check(self._lib.lpe_reset(self._h, x_t.data_ptr() if x_t is not None else None,
                                      y_t.data_ptr() if y_t is not None else None, _stream()))

    # -------------------------------------------------------------- iteration
    def run(self, iterations: int, chunk: int = 64, use_graph: bool = True) -> None:
        with torch.cuda.device(self.device):
            check(self._lib.lpe_run(self._h, int(iterations), int(chunk), 1 if use_graph else 0, _stream()))

    def set_accepted(self, mask) -> None:
        m = np.ascontiguousarray(np.asarray(mask, dtype=np.uint8).reshape(self.batch))
        with torch.cuda.device(self.device):
            check(self._lib.lpe_set_accepted(self._h, m.ctypes.data_as(C.c_void_p), _stream()))

    def solve(self, options: Optional[SolveOptions] = None, **kwargs) -> dict:
        opts = options or SolveOptions(**kwargs)
        info = _lib.LpeSolveInfo()
        with torch.cuda.device(self.device):
            check(self._lib.lpe_solve(self._h, C.byref(opts.to_c()), C.byref(info), _stream()))
        self.last_info = info.as_dict()
        return dict(self.last_info)

    def solve_with_fallback(self, options: Optional[SolveOptions] = None,
                            fallback: Optional[SolveOptions] = None, capacity: int = 256):
        """Solve, then hand the LPs that did not certify to an FP64 engine.

        FP32 iterates stall at about single-precision accuracy; an LP whose
        active set is not identifiable there is continued in FP64 (warm-started
        from its FP32 primal, dual and primal weight) on a second engine of
        ``capacity`` LPs.  Returns (info, x [B,n], y [B,m], status [B]) with
        the fallback's results merged in.
        """
        capacity = max(1, min(int(capacity), self.batch))
        options = options or SolveOptions()
        if options.polish and not options.stop_after_failed_polishes:
            # LPs whose FP32 polish failed twice go to FP64 instead of iterating on
            options = dataclasses.replace(options, stop_after_failed_polishes=2)
        info = self.solve(options)
        x, y, status = self.state()
        bad = np.flatnonzero((status != _lib.STATUS_OPTIMAL) & (status != _lib.STATUS_PRIMAL_INFEASIBLE)
                             & (status != _lib.STATUS_DUAL_INFEASIBLE))
        info["fallback_lps"] = int(bad.size)
        info["fallback_certified"] = 0
        if bad.size == 0 or self._vectors is None:
            return info, x, y, status
        fb_opts = fallback or SolveOptions(max_iterations=20000, polish=True)
        if getattr(self, "_fallback", None) is not None and self._fallback.batch != capacity:
            self._fallback.close()
            self._fallback = None
        if getattr(self, "_fallback", None) is None:
            self._fallback = BatchedLpEngine(self.A, self.neq, capacity, var_positions=self.var_positions,
                                             precision="fp64", shared_c=self.shared_c, device=self.device,
                                             method="restarted" if self.method == "fixed" else self.method)
        fb = self._fallback
        omega = self.primal_weight()
        c_t, rhs_t, lo_t, up_t, vv_t = self._vectors
        for start in range(0, bad.size, capacity):
            part = bad[start:start + capacity]
            pad = np.concatenate([part, np.full(capacity - part.size, part[0])])
            idx = torch.as_tensor(pad, device=self.device)
            fb.update(c_t if self.shared_c else c_t[idx], rhs_t[idx], lo_t[idx], up_t[idx],
                      var_values=None if vv_t is None else vv_t[idx])
            fb.reset(x[idx], y[idx])
            fb.set_primal_weight(omega[pad]) if self.method != "fixed" else None
            fb_info = fb.solve(fb_opts)
            fx, fy, fs = fb.state()
            k = part.size
            sel = torch.as_tensor(part, device=self.device)
            x[sel] = fx[:k]
            y[sel] = fy[:k]
            status[part] = fs[:k]
            info["fallback_certified"] += int((fs[:k] == _lib.STATUS_OPTIMAL).sum())
            info["fallback_seconds"] = info.get("fallback_seconds", 0.0) + fb_info["seconds"]
        info["certified"] = int((status == _lib.STATUS_OPTIMAL).sum())
        return info, x, y, status

    def polish(self, mask=None, cg_iterations: int = 200) -> None:
        m = None if mask is None else np.ascontiguousarray(np.asarray(mask, dtype=np.uint8).reshape(self.batch))
        with torch.cuda.device(self.device):
            check(self._lib.lpe_polish(self._h, None if m is None else m.ctypes.data_as(C.c_void_p),
                                       int(cg_iterations), _stream()))

    # ---------------------------------------------------------------- results
    def state(self):
        """Owned copies: x [B, n], y [B, m] (original units) and status [B]."""
        x = torch.empty((self.batch, self.n), dtype=torch.float64, device=self.device)
        y = torch.empty((self.batch, self.m), dtype=torch.float64, device=self.device)
        status = np.zeros(self.batch, dtype=np.int32)
        with torch.cuda.device(self.device):
            check(self._lib.lpe_state(self._h, x.data_ptr(), y.data_ptr() if self.m else None,
                                      status.ctypes.data_as(C.c_void_p), _stream()))
        return x, y, status

    def certify(self, x=None, y=None, eps=(1e-5, 1e-7, 1e-7)) -> torch.Tensor:
        """FP64 certificate [B, 8] of (x, y), or of the current iterate."""
        metrics = torch.empty((self.batch, _lib.NMETRIC), dtype=torch.float64, device=self.device)
        x_t = y_t = None
        if x is not None:
            x_t = _device_array(x, (self.batch, self.n), "x", self.device)
            y_t = _device_array(y, (self.batch, self.m), "y", self.device)
        with torch.cuda.device(self.device):
            check(self._lib.lpe_certify(self._h, None if x_t is None else x_t.data_ptr(),
                                        None if y_t is None or self.m == 0 else y_t.data_ptr(),
                                        metrics.data_ptr(), float(eps[0]), float(eps[1]), float(eps[2]),
                                        _stream()))
        return metrics

    def spmv(self, x, transpose: bool = False, precision: str = "fp64") -> torch.Tensor:
        lin, lout = (self.m, self.n) if transpose else (self.n, self.m)
        x_t = _device_array(x, (self.batch, lin), "x", self.device)
        out = torch.empty((self.batch, lout), dtype=torch.float64, device=self.device)
        with torch.cuda.device(self.device):
            check(self._lib.lpe_spmv(self._h, 1 if transpose else 0,
                                     {"fp64": _lib.PREC_FP64, "fp32": _lib.PREC_FP32}[precision],
                                     x_t.data_ptr(), out.data_ptr(), _stream()))
        return out

    def scaling(self) -> dict:
        dr = np.zeros(self.m); dc = np.zeros(self.n); sg = np.zeros(self.m); tu = np.zeros(self.n)
        p = lambda a: a.ctypes.data_as(C.c_void_p)
        check(self._lib.lpe_get_scaling(self._h, p(dr), p(dc), p(sg), p(tu)))
        return {"row_scale": dr, "col_scale": dc, "sigma": sg, "tau": tu}

    def set_primal_weight(self, omega) -> None:
        w = np.ascontiguousarray(np.asarray(omega, dtype=np.float64).reshape(self.batch))
        with torch.cuda.device(self.device):
            check(self._lib.lpe_set_primal_weight(self._h, w.ctypes.data_as(C.c_void_p), _stream()))

    def set_metric(self, tau, sigma, theta: float) -> None:
        """Caller-supplied diagonal metric of the fixed-metric method: tau
        [B, n], sigma [B, m] (strictly positive), theta in [0, 1]; a reset is
        required afterwards and the next ``update`` restores the default metric."""
        tau_t = _device_array(tau, (self.batch, self.n), "tau", self.device)
        sig_t = _device_array(sigma, (self.batch, self.m), "sigma", self.device) if self.m else None
        with torch.cuda.device(self.device):
            check(self._lib.lpe_set_metric(self._h, tau_t.data_ptr(), None if sig_t is None else sig_t.data_ptr(),
                                           float(theta), _stream()))

    def primal_weight(self) -> np.ndarray:
        w = np.zeros(self.batch)
        check(self._lib.lpe_get_primal_weight(self._h, w.ctypes.data_as(C.c_void_p)))
        return w

    def set_launch(self, rows_per_wave: int, lps_per_lane: int) -> None:
        check(self._lib.lpe_set_launch(self._h, int(rows_per_wave), int(lps_per_lane)))

    def time_iterations(self, iterations: int, chunk: int = 64, use_graph: bool = True) -> float:
        ms = C.c_double(0.0)
        with torch.cuda.device(self.device):
            check(self._lib.lpe_time_iterations(self._h, int(iterations), int(chunk), 1 if use_graph else 0,
                                                C.byref(ms), _stream()))
        return float(ms.value)

    def time_kernel(self, kind: str, reps: int = 20) -> float:
        """Average ms of one launch of the 'dual' or 'primal' step kernel, or
        of the polish's CG passes with every LP active: 'cg_cols' (q = D A^T p
        and |q|^2), 'cg_rows' (s = R A q with the CG update) or
        'cg_direction'."""
        kinds = {"dual": 0, "primal": 1, "cg_cols": 3, "cg_rows": 4, "cg_direction": 5}
        ms = C.c_double(0.0)
        with torch.cuda.device(self.device):
            check(self._lib.lpe_time_kernel(self._h, kinds[kind], int(reps), C.byref(ms), _stream()))
        return float(ms.value)

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.lpe_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
