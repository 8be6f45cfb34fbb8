"""Exact homogeneous equality reduction on the MI355X engine.

``HomogeneousEqualityReduction`` and ``ReducedEqualityLP`` keep the interface
of src/lp_equality_reduction.py:165-586 (plan attributes, ``reduce``,
``expand_primal``, ``compress_primal``, ``allocate_bound_reduced_cost``,
``lift_dual``, error messages); the plan itself is built by liblpe
(``lpe_reduction_create``: union-find forest over the two-entry homogeneous
equality rows, null-space weights, dual-lift map) and the bound / objective
reduction and primal maps run through its native host path.

``ReducedLpEngine`` is the batched device path: one plan for a batch that
shares a pattern (rows holding per-LP coefficients are never eliminated),
reduced per-LP data built by liblpe kernels, PDHG + polish on the reduced LP
(``BatchedLpEngine``), then x = T z and the dual lift on the device and the
ORIGINAL-LP certificate as the only acceptance (src/gpu_reduced_pdhg.py:19-103).
``GpuReducedPdhgCorrector`` mirrors src/gpu_reduced_pdhg.py:19-158 on top of
the fixed-metric corrector.
"""

from __future__ import annotations

import ctypes as C
import hashlib
import time
from typing import Optional

import numpy as np
import torch
from scipy.sparse import coo_matrix, csr_matrix

from . import _lib
from ._lib import check
from .corrector import GpuPdhgCorrector, _validated_problem, pattern_values, shared_pattern
from .engine import BatchedLpEngine, SolveOptions, _device_array, _stream


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def _vp(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def equality_fingerprint(a: csr_matrix, neq: int) -> str:
    """sha256 over the equality block (src/lp_equality_reduction.py:64-82)."""
    e = csr_matrix(a[:neq], copy=True)
    e.sum_duplicates()
    e.eliminate_zeros()
    e.sort_indices()
    digest = hashlib.sha256()
    for name, value, dtype in (("shape", e.shape, "<i8"), ("indptr", e.indptr, "<i8"), ("indices", e.indices, "<i8"),
                               ("data", e.data, "<f8"), ("neq", [neq], "<i8")):
        arr = np.asarray(value, dtype=dtype)
        digest.update(name.encode("ascii"))
        digest.update(np.asarray(arr.shape, dtype="<i8").tobytes())
        digest.update(arr.tobytes())
    return digest.hexdigest()


class _NativePlan:
    """Owner of one ``lpe_reduction`` handle and its host arrays."""

    def __init__(self, a: csr_matrix, neq: int, rhs: np.ndarray, var_positions=None, max_scale_ratio: float = 1e8):
        self._lib = _lib.load()
        a = csr_matrix(a, dtype=np.float64, copy=True)
        a.sort_indices()
        self.a = a
        self.m, self.n = a.shape
        self.neq = int(neq)
        rp = np.ascontiguousarray(a.indptr, dtype=np.int32)
        ci = np.ascontiguousarray(a.indices, dtype=np.int32)
        val = np.ascontiguousarray(a.data, dtype=np.float64)
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        vp = np.ascontiguousarray(np.zeros(0) if var_positions is None else var_positions, dtype=np.int64).ravel()
        self.var_positions = vp
        d = _lib.LpeReductionDesc()
        d.m, d.n, d.neq, d.nnz = self.m, self.n, self.neq, int(a.nnz)
        d.row_ptr, d.col_idx, d.values = _ptr(rp, C.c_int32), _ptr(ci, C.c_int32), _ptr(val, C.c_double)
        d.rhs = _ptr(rhs, C.c_double)
        d.n_var, d.var_pos = int(vp.size), _ptr(vp, C.c_int64)
        d.max_scale_ratio = float(max_scale_ratio)
        h = C.c_void_p()
        check(self._lib.lpe_reduction_create(C.byref(d), C.byref(h)))
        self._h = h
        dims = _lib.LpeReductionDims()
        check(self._lib.lpe_reduction_get_dims(h, C.byref(dims)))
        self.rm, self.rn, self.rneq = dims.m, dims.n, dims.neq
        self.n_eliminated, self.r_n_var = dims.n_eliminated, dims.n_var
        self.group = np.zeros(self.n, np.int32)
        self.weight = np.zeros(self.n)
        self.representative = np.zeros(self.rn, np.int32)
        self.kept = np.zeros(self.rm, np.int32)
        self.eliminated = np.zeros(self.n_eliminated, np.int32)
        check(self._lib.lpe_reduction_get_plan(h, _vp(self.group), _vp(self.weight), _vp(self.representative),
                                               _vp(self.kept), _vp(self.eliminated)))
        rrp = np.zeros(self.rm + 1, np.int32)
        rci = np.zeros(dims.nnz, np.int32)
        rval = np.zeros(dims.nnz)
        self.reduced_var_positions = np.zeros(self.r_n_var, np.int64)
        check(self._lib.lpe_reduction_get_pattern(h, _vp(rrp), _vp(rci), _vp(rval), _vp(self.reduced_var_positions)))
        self.reduced_matrix = csr_matrix((rval, rci, rrp), shape=(self.rm, self.rn))
        lp = np.zeros(self.n_eliminated + 1, np.int32)
        lc = np.zeros(dims.lift_nnz, np.int32)
        lv = np.zeros(dims.lift_nnz)
        check(self._lib.lpe_reduction_get_lift_map(h, _vp(lp), _vp(lc), _vp(lv)))
        self.lift_map = csr_matrix((lv, lc, lp), shape=(self.n_eliminated, self.n))

    def close(self):
        if getattr(self, "_h", None):
            self._lib.lpe_reduction_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ReducedEqualityLP:
    """One LP reduced by a fixed plan (src/lp_equality_reduction.py:165-268)."""

    def __init__(self, plan, problem, original_problem, kept_matrix, lower_witness, upper_witness):
        self.plan = plan
        self.problem = problem
        self.original_problem = original_problem
        self.kept_matrix = kept_matrix
        self.lower_witness = lower_witness
        self.upper_witness = upper_witness
        self.fallback_witness = plan.representatives.copy()

    def expand_primal(self, reduced_x):
        return self.plan.expand_primal(reduced_x)

    def compress_primal(self, original_x, *, check: bool = False, atol: float = 1e-9):
        return self.plan.compress_primal(original_x, check=check, atol=atol)

    def allocate_bound_reduced_cost(self, reduced_cost):
        """Each group's reduced cost on its witness bound: lower for positive,
        upper for negative, the representative when that endpoint has none
        (src/lp_equality_reduction.py:190-218)."""
        reduced_cost = np.asarray(reduced_cost, dtype=np.float64)
        if reduced_cost.shape != (self.plan.reduced_variables,) or not np.isfinite(reduced_cost).all():
            raise ValueError("reduced_cost has the wrong shape or non-finite values")
        original = np.zeros(self.plan.original_variables)
        for g in np.flatnonzero(reduced_cost):
            w = self.lower_witness[g] if reduced_cost[g] > 0.0 else self.upper_witness[g]
            if w < 0:
                w = self.fallback_witness[g]
            original[w] = reduced_cost[g] / self.plan.weights[w]
        return original

    def lift_dual(self, reduced_y, reduced_cost):
        """All original row duals from the kept rows' duals and the reduced
        costs (src/lp_equality_reduction.py:220-260)."""
        reduced_y = np.asarray(reduced_y, dtype=np.float64)
        if reduced_y.shape != (len(self.plan.kept_rows),) or not np.isfinite(reduced_y).all():
            raise ValueError("reduced_y has the wrong shape or non-finite values")
        reduced_cost = np.asarray(reduced_cost, dtype=np.float64)
        if reduced_cost.shape != (self.plan.reduced_variables,) or not np.isfinite(reduced_cost).all():
            raise ValueError("reduced_cost has the wrong shape or non-finite values")
        a, _, _, _, c, _ = self.original_problem
        q = c - np.asarray(self.kept_matrix.T @ reduced_y).ravel()
        expected = np.asarray(self.plan.transform.T @ q).ravel()
        scale = max(1.0, float(np.max(np.abs(expected), initial=0.0)))
        if float(np.max(np.abs(expected - reduced_cost), initial=0.0)) > 1e-10 * scale:
            raise ValueError("reduced_cost is inconsistent with the reduced row dual")
        normal = self.allocate_bound_reduced_cost(reduced_cost)
        y = np.empty(a.shape[0])
        y[self.plan.kept_rows] = reduced_y
        y[self.plan.eliminated_rows] = np.asarray(self.plan.dual_lift_map @ (q - normal)).ravel()
        recovered = c - np.asarray(a.T @ y).ravel()
        if float(np.max(np.abs(recovered - normal), initial=0.0)) > 1e-8:
            raise ArithmeticError("Eliminated equality dual lift lost original stationarity")
        return y


class HomogeneousEqualityReduction:
    """Fixed map x = T z eliminating independent two-entry homogeneous
    equalities (src/lp_equality_reduction.py:295-586), built by liblpe."""

    def __init__(self, problem, *, max_scale_ratio: float = 1e8, var_positions=None):
        if (isinstance(max_scale_ratio, (bool, np.bool_)) or not np.isfinite(max_scale_ratio)
                or float(max_scale_ratio) < 1.0):
            raise ValueError("max_scale_ratio must be finite and at least one")
        a, rhs, _lower, _upper, _c, neq = _validated_problem(problem)
        self._native = _NativePlan(a, neq, rhs, var_positions, max_scale_ratio)
        nat = self._native
        self.original_shape = a.shape
        self.original_variables = a.shape[1]
        self.original_equalities = neq
        self.max_scale_ratio = float(max_scale_ratio)
        self.equality_fingerprint = equality_fingerprint(a, neq)
        self.reduced_variables = nat.rn
        self.original_to_reduced = nat.group.astype(np.int64)
        self.weights = nat.weight.copy()
        self.representatives = nat.representative.astype(np.int64)
        self.kept_rows = nat.kept.astype(np.int64)
        self.eliminated_rows = nat.eliminated.astype(np.int64)
        self.reduced_equalities = nat.rneq
        self.transform = coo_matrix((self.weights, (np.arange(self.original_variables), self.original_to_reduced)),
                                    shape=(self.original_variables, self.reduced_variables)).tocsr()
        self.T = self.transform
        denom = np.zeros(self.reduced_variables)
        np.add.at(denom, self.original_to_reduced, self.weights ** 2)
        self.compression = coo_matrix((self.weights / denom[self.original_to_reduced],
                                       (self.original_to_reduced, np.arange(self.original_variables))),
                                      shape=(self.reduced_variables, self.original_variables)).tocsr()
        self.dual_lift_map = nat.lift_map
        order = np.argsort(self.original_to_reduced, kind="stable")
        splits = np.flatnonzero(np.diff(self.original_to_reduced[order])) + 1
        self.component_members = tuple(np.split(order, splits))

    @classmethod
    def from_problem(cls, problem, **kwargs):
        return cls(problem, **kwargs)

    def reduce(self, problem) -> ReducedEqualityLP:
        """Reduced LP and bound witnesses of one LP (src/lp_equality_reduction.py:519-560)."""
        original = _validated_problem(problem)
        a, rhs, lower, upper, c, neq = original
        if a.shape != self.original_shape or neq != self.original_equalities:
            raise ValueError("LP dimensions/equality block changed from the reduction plan")
        if equality_fingerprint(a, neq) != self.equality_fingerprint:
            raise ValueError("Equality fingerprint changed from the reduction plan")
        nat = self._native
        r, rm = self.reduced_variables, len(self.kept_rows)
        out = dict(c=np.zeros(r), rhs=np.zeros(rm), lower=np.zeros(r), upper=np.zeros(r),
                   lw=np.zeros(r, np.int32), uw=np.zeros(r, np.int32), status=np.zeros(1, np.int32))
        vin = _lib.LpeVectors()
        vin.c, vin.rhs, vin.lower, vin.upper, vin.var_values = _vp(c), _vp(rhs), _vp(lower), _vp(upper), None
        vout = _lib.LpeReduced()
        vout.c, vout.rhs, vout.lower, vout.upper = _vp(out["c"]), _vp(out["rhs"]), _vp(out["lower"]), _vp(out["upper"])
        vout.var_values = None
        vout.lower_witness, vout.upper_witness, vout.status = _vp(out["lw"]), _vp(out["uw"]), _vp(out["status"])
        if nat.r_n_var:
            raise ValueError("a single-LP plan has no per-LP coefficient positions")
        check(nat._lib.lpe_reduction_reduce(nat._h, 1, 0, C.byref(vin), C.byref(vout), 0, None))
        st = int(out["status"][0])
        if st & _lib.RED_INHOMOGENEOUS:
            raise ValueError("An eliminated equality is no longer homogeneous")
        if st & _lib.RED_OVERFLOW:
            raise ValueError("Forest bound transformation overflowed or became invalid")
        if st & _lib.RED_INFEASIBLE:
            raise ValueError("Original bounds are infeasible under eliminated equalities")
        kept_matrix = a[self.kept_rows].tocsr()
        reduced_a = (kept_matrix @ self.transform).tocsr()
        reduced_a.sum_duplicates()
        reduced_a.eliminate_zeros()
        reduced_a.sort_indices()
        reduced = (reduced_a, out["rhs"], out["lower"], out["upper"], out["c"], self.reduced_equalities)
        return ReducedEqualityLP(self, reduced, original, kept_matrix, out["lw"].astype(np.int64),
                                 out["uw"].astype(np.int64))

    def _primal_map(self, fn, value, width, out_width, name):
        value = np.asarray(value, dtype=np.float64)
        if value.shape[-1:] != (width,) or not np.isfinite(value).all():
            raise ValueError(f"{name} has the wrong final dimension or non-finite values")
        flat = np.ascontiguousarray(value.reshape(-1, width))
        out = np.zeros((flat.shape[0], out_width))
        if flat.shape[0]:
            check(fn(self._native._h, flat.shape[0], _vp(flat), _vp(out), 0, None))
        return out.reshape(value.shape[:-1] + (out_width,))

    def expand_primal(self, reduced_x):
        return self._primal_map(self._native._lib.lpe_reduction_expand, reduced_x, self.reduced_variables,
                                self.original_variables, "reduced_x")

    def compress_primal(self, original_x, *, check: bool = False, atol: float = 1e-9):
        reduced = self._primal_map(self._native._lib.lpe_reduction_compress, original_x, self.original_variables,
                                   self.reduced_variables, "original_x")
        if check and not np.allclose(self.expand_primal(reduced), original_x, rtol=0.0, atol=atol):
            raise ValueError("original_x does not satisfy the eliminated equalities")
        return reduced


EqualityReductionPlan = HomogeneousEqualityReduction


class ReducedLpEngine:
    """B LPs of one pattern solved in the equality-reduced space on MI355X,
    accepted only by the original-LP certificate.

    ``update``/``reset``/``solve``/``state``/``certify`` take and return
    original coordinates, like ``BatchedLpEngine``.  The reduced per-LP data
    are built on the device at every ``update``; a reduced LP that certifies
    but whose lifted original solution fails the original certificate is
    reported as not optimal (``STATUS_NUMERICAL``), never accepted.
    """

    def __init__(self, A, neq: int, batch: int, *, rhs_template=None, var_positions=None, precision: str = "fp32",
                 method: str = "restarted", shared_c: bool = False, device: Optional[torch.device] = None,
                 max_scale_ratio: float = 1e8, **engine_kwargs):
        A = csr_matrix(A, dtype=np.float64, copy=True)
        A.sort_indices()
        self.A = A
        self.m, self.n = A.shape
        self.neq, self.batch = int(neq), int(batch)
        self.shared_c = bool(shared_c)
        self.precision, self.method = precision, method
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        rhs_template = np.zeros(self.m) if rhs_template is None else np.asarray(rhs_template, dtype=np.float64)
        self.plan = _NativePlan(A, self.neq, rhs_template, var_positions, max_scale_ratio)
        p = self.plan
        self.var_positions = p.var_positions
        # a box-only reduced LP keeps one empty inequality row for the engine
        self._pad = 1 if p.rm == 0 else 0
        red_a = p.reduced_matrix
        if self._pad:
            red_a = csr_matrix((np.zeros(0), np.zeros(0, np.int32), np.zeros(2, np.int32)), shape=(1, p.rn))
        self.reduced = BatchedLpEngine(red_a, p.rneq, self.batch, var_positions=p.reduced_var_positions,
                                       precision=precision, method=method, shared_c=shared_c, device=self.device,
                                       **engine_kwargs)
        self.original = BatchedLpEngine(A, self.neq, self.batch, var_positions=p.var_positions, precision="fp64",
                                        method="fixed", shared_c=shared_c, device=self.device)
        B, dev = self.batch, self.device
        f64 = dict(dtype=torch.float64, device=dev)
        self._rc = torch.zeros((p.rn,) if shared_c else (B, p.rn), **f64)
        self._rrhs = torch.zeros((B, p.rm + self._pad), **f64)
        self._rlo = torch.zeros((B, p.rn), **f64)
        self._rup = torch.zeros((B, p.rn), **f64)
        self._rvar = torch.zeros((B, max(1, p.r_n_var)), **f64)
        self._lw = torch.zeros((B, p.rn), dtype=torch.int32, device=dev)
        self._uw = torch.zeros((B, p.rn), dtype=torch.int32, device=dev)
        self._status = torch.zeros(B, dtype=torch.int32, device=dev)
        self._c = None
        self._result = None
        self.last_info: dict = {}

    @property
    def reduced_shape(self):
        return self.plan.rm, self.plan.rn

    def update(self, c, rhs, lower, upper, var_values=None) -> None:
        """Validate and reduce the per-LP data on the device; raises (without
        changing the engines) when an LP's bounds are infeasible under the
        eliminated equalities or an eliminated row is no longer homogeneous."""
        p, B, dev = self.plan, self.batch, self.device
        c_t = _device_array(c, (self.n,) if self.shared_c else (B, self.n), "c", dev)
        rhs_t = _device_array(rhs, (B, self.m), "rhs", dev)
        lo_t = _device_array(lower, (B, self.n), "lower", dev)
        up_t = _device_array(upper, (B, self.n), "upper", dev)
        vv_t = None if var_values is None else _device_array(var_values, (B, p.var_positions.size), "var_values", dev)
        vin = _lib.LpeVectors()
        vin.c, vin.rhs, vin.lower, vin.upper = c_t.data_ptr(), rhs_t.data_ptr(), lo_t.data_ptr(), up_t.data_ptr()
        vin.var_values = None if vv_t is None else vv_t.data_ptr()
        vout = _lib.LpeReduced()
        vout.c, vout.rhs = self._rc.data_ptr(), self._rrhs.data_ptr()
        vout.lower, vout.upper, vout.var_values = self._rlo.data_ptr(), self._rup.data_ptr(), self._rvar.data_ptr()
        vout.lower_witness, vout.upper_witness = self._lw.data_ptr(), self._uw.data_ptr()
        vout.status = self._status.data_ptr()
        if self._pad:
            # the padded row has no entries: the reduced rhs is written for rm = 0 rows only
            self._rrhs.zero_()
        with torch.cuda.device(dev):
            check(p._lib.lpe_reduction_reduce(p._h, B, 1 if self.shared_c else 0, C.byref(vin), C.byref(vout), 1,
                                              _stream()))
        st = self._status.cpu().numpy()
        if (st & _lib.RED_INHOMOGENEOUS).any():
            raise ValueError("An eliminated equality is no longer homogeneous")
        if (st & _lib.RED_OVERFLOW).any():
            raise ValueError("Forest bound transformation overflowed or became invalid")
        if (st & _lib.RED_INFEASIBLE).any():
            raise ValueError("Original bounds are infeasible under eliminated equalities")
        self.original.update(c_t, rhs_t, lo_t, up_t, var_values=vv_t)
        self.reduced.update(self._rc, self._rrhs, self._rlo, self._rup,
                            var_values=self._rvar[:, :p.r_n_var] if p.r_n_var else None)
        self._c = c_t
        self._result = None

    def compress(self, x) -> torch.Tensor:
        x_t = _device_array(x, (self.batch, self.n), "x", self.device)
        z = torch.empty((self.batch, self.plan.rn), dtype=torch.float64, device=self.device)
        with torch.cuda.device(self.device):
            check(self.plan._lib.lpe_reduction_compress(self.plan._h, self.batch, x_t.data_ptr(), z.data_ptr(), 1,
                                                        _stream()))
        return z

    def expand(self, z) -> torch.Tensor:
        z_t = _device_array(z, (self.batch, self.plan.rn), "z", self.device)
        x = torch.empty((self.batch, self.n), dtype=torch.float64, device=self.device)
        with torch.cuda.device(self.device):
            check(self.plan._lib.lpe_reduction_expand(self.plan._h, self.batch, z_t.data_ptr(), x.data_ptr(), 1,
                                                      _stream()))
        return x

    def reset(self, x0=None, y0=None) -> None:
        """Warm start in original coordinates: z = compression x, y' = y[kept]."""
        z = None if x0 is None else self.compress(x0)
        yr = None
        if y0 is not None:
            y_t = _device_array(y0, (self.batch, self.m), "y", self.device)
            yr = torch.zeros((self.batch, self.plan.rm + self._pad), dtype=torch.float64, device=self.device)
            if self.plan.rm:
                yr[:, :self.plan.rm] = y_t[:, torch.as_tensor(self.plan.kept, device=self.device, dtype=torch.long)]
        self.reduced.reset(z, yr)
        self._result = None

    def lift(self, z, y_reduced):
        """Original (x, y) of reduced iterates (src/gpu_reduced_pdhg.py:81-103)."""
        p, B = self.plan, self.batch
        z_t = _device_array(z, (B, p.rn), "z", self.device)
        yr = _device_array(y_reduced, (B, p.rm + self._pad), "y_reduced", self.device)[:, :p.rm].contiguous()
        x = torch.empty((B, self.n), dtype=torch.float64, device=self.device)
        y = torch.empty((B, self.m), dtype=torch.float64, device=self.device)
        if self.m == 0:
            with torch.cuda.device(self.device):
                check(p._lib.lpe_reduction_expand(p._h, B, z_t.data_ptr(), x.data_ptr(), 1, _stream()))
            return x, y
        with torch.cuda.device(self.device):
            check(p._lib.lpe_reduction_lift(p._h, self.original._h, B, 1 if self.shared_c else 0, self._c.data_ptr(),
                                            None, z_t.data_ptr(), yr.data_ptr() if p.rm else None,
                                            self._lw.data_ptr(), self._uw.data_ptr(), x.data_ptr(), y.data_ptr(), 1,
                                            _stream()))
        return x, y

    def solve(self, options: Optional[SolveOptions] = None, *, fallback: bool = False, **kwargs) -> dict:
        """Reduced solve, lift, original certificate.  ``fallback`` routes LPs
        the FP32 solve leaves uncertified through an FP64 reduced engine."""
        opts = options or SolveOptions(**kwargs)
        if fallback:
            info, z, yr, status = self.reduced.solve_with_fallback(opts)
        else:
            info = self.reduced.solve(opts)
            z, yr, status = self.reduced.state()
        x, y = self.lift(z, yr)
        metrics = self.original.certify(x, y, eps=(opts.eps_primal, opts.eps_dual, opts.eps_gap))
        passed = metrics[:, 4].cpu().numpy() > 0.5
        optimal = status == _lib.STATUS_OPTIMAL
        rejected = optimal & ~passed
        status = status.copy()
        status[rejected] = _lib.STATUS_NUMERICAL
        info = dict(info)
        info["reduced_certified"] = int(optimal.sum())
        info["original_rejected"] = int(rejected.sum())
        info["certified"] = int((status == _lib.STATUS_OPTIMAL).sum())
        self._result = (x, y, status, metrics)
        self.last_info = info
        return dict(info)

    def state(self):
        """x [B, n], y [B, m] in original coordinates and status [B]."""
        if self._result is None:
            z, yr, status = self.reduced.state()
            x, y = self.lift(z, yr)
            return x, y, status
        x, y, status, _ = self._result
        return x.clone(), y.clone(), status.copy()

    def certify(self, x=None, y=None, eps=(1e-5, 1e-7, 1e-7)) -> torch.Tensor:
        if x is None:
            x, y, _ = self.state()
        return self.original.certify(x, y, eps=eps)

    def set_primal_weight(self, omega) -> None:
        self.reduced.set_primal_weight(omega)

    def primal_weight(self) -> np.ndarray:
        return self.reduced.primal_weight()

    def solve_with_fallback(self, options: Optional[SolveOptions] = None):
        """``BatchedLpEngine.solve_with_fallback``'s interface: (info, x, y, status)."""
        info = self.solve(options, fallback=True)
        x, y, status = self.state()
        return info, x, y, status

    def close(self) -> None:
        for part in ("reduced", "original"):
            eng = getattr(self, part, None)
            if eng is not None:
                eng.close()
        if getattr(self, "plan", None) is not None:
            self.plan.close()


class GpuReducedPdhgCorrector(GpuPdhgCorrector):
    """Iterate in the exact reduced space; accept only a lifted original LP
    (src/gpu_reduced_pdhg.py:19-158).  Warm starts and returned x / y use
    original coordinates; ``reduced_x`` / ``reduced_y`` are the iterates."""

    name = "mi355x_fp64_equality_reduced_pdhg_original_certificate"

    def __init__(self, problems, *, plan: Optional[HomogeneousEqualityReduction] = None, **kwargs):
        started = time.perf_counter()
        original = tuple(_validated_problem(p) for p in problems)
        if not original:
            raise ValueError("At least one independent LP is required")
        shape, neq = original[0][0].shape, original[0][-1]
        if any(p[0].shape != shape or p[-1] != neq for p in original):
            raise ValueError("Original LP batch requires identical dimensions and equality counts")
        before = time.perf_counter()
        self.plan = HomogeneousEqualityReduction.from_problem(original[0]) if plan is None else plan
        plan_seconds = time.perf_counter() - before if plan is None else 0.0
        before = time.perf_counter()
        self.reductions = tuple(self.plan.reduce(p) for p in original)
        self.reduced_problems = tuple(r.problem for r in self.reductions)
        host_seconds = time.perf_counter() - before
        self._pad = 1 if self.reduced_problems[0][0].shape[0] == 0 else 0
        solve_problems = self.reduced_problems
        if self._pad:
            # a box-only reduced LP keeps one empty inequality row for the engine
            empty = csr_matrix((1, self.plan.reduced_variables))
            solve_problems = tuple((empty, np.zeros(1), p[2], p[3], p[4], 0) for p in self.reduced_problems)
        super().__init__(solve_problems, **kwargs)
        self.original_problems = original
        self.original_m, self.original_n = shape
        self.original_neq = neq
        before = time.perf_counter()
        template, varying, values = shared_pattern([p[0] for p in original], self.per_lp_entries == "all")
        self._original_template = template
        self._original = BatchedLpEngine(template, neq, self.batch, var_positions=varying, method="fixed",
                                         device=self.engine.device)
        self._original.update(np.stack([p[4] for p in original]), np.stack([p[1] for p in original]),
                              np.stack([p[2] for p in original]), np.stack([p[3] for p in original]),
                              var_values=values if varying.size else None)
        dev = self.engine.device
        self._kept = torch.as_tensor(self.plan.kept_rows, device=dev)
        self._lw = np.stack([r.lower_witness for r in self.reductions])
        self._uw = np.stack([r.upper_witness for r in self.reductions])
        self._c = np.stack([p[4] for p in original])
        torch.cuda.synchronize()
        self.setup_timing.update(equality_plan_build_seconds=plan_seconds, host_equality_reduction_seconds=host_seconds,
                                 original_lift_setup_seconds=time.perf_counter() - before,
                                 setup_total_seconds=time.perf_counter() - started,
                                 original_variables=self.original_n, original_rows=self.original_m,
                                 reduced_variables=self.n, reduced_rows=self.m - self._pad,
                                 eliminated_equalities=len(self.plan.eliminated_rows),
                                 equality_fingerprint=self.plan.equality_fingerprint)
        self._last_lift = None

    def expand_and_lift(self, z, y_reduced):
        """Device-only expansion and dual postsolve through liblpe
        (src/gpu_reduced_pdhg.py:81-103)."""
        # every LP of the batch shares the equality block (the plan's
        # fingerprint), so the per-LP entries lie in kept rows and the single
        # plan's groups, witnesses and lift map serve the whole batch; the
        # original engine holds each LP's values for A^T y
        B = self.batch
        lp = self.plan._native
        dev = self.engine.device
        z_t = _device_array(z, (B, self.n), "z", dev)
        yr = _device_array(y_reduced, (B, self.m), "y_reduced", dev)[:, :self.m - self._pad].contiguous()
        c_t = torch.as_tensor(self._c, device=dev)
        lw = torch.as_tensor(self._lw.astype(np.int32), device=dev)
        uw = torch.as_tensor(self._uw.astype(np.int32), device=dev)
        x = torch.empty((B, self.original_n), dtype=torch.float64, device=dev)
        y = torch.empty((B, self.original_m), dtype=torch.float64, device=dev)
        with torch.cuda.device(dev):
            if self.original_m == 0:
                check(lp._lib.lpe_reduction_expand(lp._h, B, z_t.data_ptr(), x.data_ptr(), 1, _stream()))
            else:
                check(lp._lib.lpe_reduction_lift(lp._h, self._original._h, B, 0, c_t.data_ptr(), None, z_t.data_ptr(),
                                                 yr.data_ptr() if yr.shape[1] else None, lw.data_ptr(), uw.data_ptr(),
                                                 x.data_ptr(), y.data_ptr(), 1, _stream()))
        return x, y

    def _certificate(self, x=None, y=None):
        if x is None:
            x, y, _ = self.engine.state()
        ox, oy = self.expand_and_lift(x, y)
        self._last_lift = (ox, oy)
        metrics = self._original.certify(ox, oy).cpu().numpy()
        return [dict(primal_residual=float(r[0]), dual_violation=float(r[1]), relative_kkt_gap=float(r[2]),
                     objective=float(r[3]), certificate_passed=bool(r[4]), certificate_source="solver_row_dual",
                     solver_dual_relative_kkt_gap=float(r[2]))
                for r in metrics]

    def solve(self, initial_x=None, initial_y=None, *, iterations: int = 1000, check_interval: int = 50):
        started = time.perf_counter()
        dev = self.engine.device
        before = time.perf_counter()
        full_x = self._warm(initial_x, (self.batch, self.original_n), "initial_x")
        full_y = self._warm(initial_y, (self.batch, self.original_m), "initial_y")
        z = None
        if full_x is not None:
            z = torch.as_tensor(self.plan.compress_primal(full_x.cpu().numpy()), device=dev)
        yr = None
        if full_y is not None:
            yr = torch.zeros((self.batch, self.m), dtype=torch.float64, device=dev)
            if self.m - self._pad:
                yr[:, :self.m - self._pad] = full_y.to(dev)[:, self._kept]
        compression_seconds = time.perf_counter() - before
        result = super().solve(initial_x=z, initial_y=yr, iterations=iterations, check_interval=check_interval)
        result["reduced_x"], result["reduced_y"] = result["x"], result["y"][:, :self.m - self._pad]
        result["x"], result["y"] = self._last_lift
        result["timing"]["original_warm_compression_seconds"] = compression_seconds
        result["timing"]["solve_total_seconds"] = time.perf_counter() - started
        result["equality_reduction"] = True
        result["scope"] = ("Exact homogeneous equality reduction; MI355X reduced-space PDHG, device primal/dual "
                           "restoration, strict ORIGINAL-LP certificate. No CPU LP fallback.")
        self.history[-1]["timing"] = dict(result["timing"])
        self.history[-1]["equality_reduction"] = True
        return result


def _same_pattern(left, right) -> bool:
    return (left.shape == right.shape and np.array_equal(left.indptr, right.indptr)
            and np.array_equal(left.indices, right.indices))


class GpuResidentReducedPdhg(GpuReducedPdhgCorrector):
    """Persistent fixed-pattern reduced service (src/gpu_resident_reduced_pdhg.py:23-177).

    ``try_update`` refreshes every value of a same-pattern batch in the
    resident engines and returns False (without mutation) for batch / order /
    CSR-pattern changes; invalid numeric input raises before any mutation.
    ``solve`` certifies the lifted original LP at every checkpoint and lets
    the final certificate decide acceptance; results are owned copies.
    """

    name = "mi355x_resident_reduced_pdhg_original_certificate"
    per_lp_entries = "all"

    def __init__(self, problems, *, environment_ids=None, chunk_size: int = 64, use_graph: bool = True, plan=None,
                 **kwargs):
        started = time.perf_counter()
        super().__init__(problems, plan=plan, **kwargs)
        self.environment_ids = tuple(range(self.batch)) if environment_ids is None else tuple(environment_ids)
        if len(self.environment_ids) != self.batch:
            raise ValueError("environment_ids must name every LP of the batch")
        self.chunk_size = int(chunk_size)
        self.use_graph = bool(use_graph)
        self.setup_timing.update(setup_total_seconds=time.perf_counter() - started)
        self.last_update_timing = dict(reused=False, reason="initial_setup", total_seconds=0.0)
        self._valid = True

    def try_update(self, problems, *, environment_ids=None) -> bool:
        started = time.perf_counter()
        original = tuple(_validated_problem(p) for p in problems)
        ids = tuple(range(len(original))) if environment_ids is None else tuple(environment_ids)

        def rebuild(reason):
            self.last_update_timing = dict(reused=False, reason=reason, total_seconds=time.perf_counter() - started)
            return False

        if len(original) != self.batch or ids != self.environment_ids:
            return rebuild("batch_or_environment_order_changed")
        # fingerprint, homogeneity, dimensions and bound intersections are
        # validated before any resident buffer changes
        reductions = tuple(self.plan.reduce(p) for p in original)
        reduced = tuple(r.problem for r in reductions)
        if not all(_same_pattern(p[0], q[0]) for p, q in zip(original, self.original_problems)):
            return rebuild("original_csr_pattern_changed")
        if not all(_same_pattern(p[0], q[0]) for p, q in zip(reduced, self.reduced_problems)):
            return rebuild("reduced_csr_pattern_changed")
        preparation = time.perf_counter() - started
        solve_problems = reduced
        if self._pad:
            empty = csr_matrix((1, self.plan.reduced_variables))
            solve_problems = tuple((empty, np.zeros(1), p[2], p[3], p[4], 0) for p in reduced)
        torch.cuda.synchronize()
        self._valid = False
        before = time.perf_counter()
        self.engine.update(np.stack([p[4] for p in solve_problems]), np.stack([p[1] for p in solve_problems]),
                           np.stack([p[2] for p in solve_problems]), np.stack([p[3] for p in solve_problems]),
                           var_values=pattern_values(self._template, [p[0] for p in solve_problems]))
        self._original.update(np.stack([p[4] for p in original]), np.stack([p[1] for p in original]),
                              np.stack([p[2] for p in original]), np.stack([p[3] for p in original]),
                              var_values=pattern_values(self._original_template, [p[0] for p in original]))
        torch.cuda.synchronize()
        self.original_problems = original
        self.reductions, self.reduced_problems = reductions, reduced
        self.problems = solve_problems
        self._lw = np.stack([r.lower_witness for r in reductions])
        self._uw = np.stack([r.upper_witness for r in reductions])
        self._c = np.stack([p[4] for p in original])
        self._last_lift = None
        self._valid = True
        self.last_update_timing = dict(reused=True, reason="same_pattern_value_update",
                                       host_preparation_seconds=preparation,
                                       device_update_seconds=time.perf_counter() - before,
                                       total_seconds=time.perf_counter() - started)
        return True

    @property
    def lower_witness(self) -> np.ndarray:
        return self._lw.copy()

    @property
    def upper_witness(self) -> np.ndarray:
        return self._uw.copy()

    def solve(self, initial_x=None, initial_y=None, *, iterations: int = 1000, check_interval: int = 64):
        if not self._valid:
            raise RuntimeError("An interrupted update invalidated this solver; rebuild it")
        iterations = self._valid_budget(iterations, "iterations", allow_zero=True)
        check_interval = self._valid_budget(check_interval, "check_interval", allow_zero=False)
        started = time.perf_counter()
        dev = self.engine.device
        before = time.perf_counter()
        full_x = self._warm(initial_x, (self.batch, self.original_n), "initial_x")
        full_y = self._warm(initial_y, (self.batch, self.original_m), "initial_y")
        z = None if full_x is None else torch.as_tensor(self.plan.compress_primal(full_x.cpu().numpy()), device=dev)
        yr = None
        if full_y is not None:
            yr = torch.zeros((self.batch, self.m), dtype=torch.float64, device=dev)
            if self.m - self._pad:
                yr[:, :self.m - self._pad] = full_y.to(dev)[:, self._kept]
        self.engine.reset(z, yr)
        torch.cuda.synchronize()
        warm_seconds = time.perf_counter() - before
        correction_seconds = certificate_seconds = control_seconds = 0.0
        accepted = np.zeros(self.batch, dtype=bool)
        accepted_iteration = np.full(self.batch, -1, dtype=np.int64)
        checkpoints = []
        completed = 0
        while True:
            zs, ys, _ = self.engine.state()
            before = time.perf_counter()
            metrics = self._certificate(zs, ys)
            elapsed = time.perf_counter() - before
            certificate_seconds += elapsed
            checkpoints.append(self._checkpoint(completed, metrics, elapsed))
            now = np.array([r["certificate_passed"] for r in metrics], dtype=bool)
            accepted_iteration[~accepted & now] = completed
            accepted |= now
            if accepted.all() or completed == iterations:
                break
            before = time.perf_counter()
            self.engine.set_accepted(accepted)
            control_seconds += time.perf_counter() - before
            count = min(check_interval, iterations - completed)
            before = time.perf_counter()
            self.engine.run(count, chunk=self.chunk_size, use_graph=self.use_graph)
            torch.cuda.synchronize()
            correction_seconds += time.perf_counter() - before
            completed += count
        # the final lifted certificate is the authority for the returned candidate
        accepted &= now
        accepted_iteration[~accepted] = -1
        for i, row in enumerate(metrics):
            row.update(accepted_iteration=int(accepted_iteration[i]), success=bool(accepted[i]), solver=self.name)
        timing = dict(warm_start_transfer_seconds=warm_seconds, correction_seconds=correction_seconds,
                      certificate_seconds=certificate_seconds, host_control_seconds=control_seconds,
                      solve_total_seconds=time.perf_counter() - started)
        x, y = self._last_lift
        result = dict(x=x.clone(), y=y.clone(), reduced_x=zs.clone(), reduced_y=ys[:, :self.m - self._pad].clone(),
                      accepted=accepted, all_accepted=bool(accepted.all()), iterations_run=completed, metrics=metrics,
                      checkpoints=checkpoints, timing=timing, cpu_lp_calls=0, primal_weight=self.primal_weight,
                      equality_reduction=True,
                      scope="Resident FP64 reduced MI355X iterations; original-LP certification; "
                            "host input/update/control remains; no CPU optimizer or fallback")
        self.history.append(dict(batch=self.batch, iterations_run=completed, all_accepted=bool(accepted.all()),
                                 accepted=accepted.tolist(), metrics=[dict(r) for r in metrics], timing=dict(timing),
                                 cpu_lp_calls=0, update=dict(self.last_update_timing),
                                 execution="graph" if self.use_graph else "loop"))
        return result
