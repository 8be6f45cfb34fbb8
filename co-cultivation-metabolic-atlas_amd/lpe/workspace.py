"""Drop-in for the reference's fixed-pattern PDHG workspace on the MI355X engine.

``GpuPdhgWorkspace`` keeps the constructor, methods, statistics and
validation behaviour of src/gpu_pdhg_workspace.py:55-384: a block-diagonal CSR
of ``batch`` environments with immutable pattern and environment order;
``update_problem`` validates a complete update before anything changes
(matrix values require new tau / sigma); ``reset`` projects a cold or warm
start and is required after every update; ``set_accepted`` freezes
environments (only ``reset`` unfreezes them); ``run`` enqueues fixed work,
full chunks as graph replays and the tail as single launches; ``state``
returns x [B, n] and y [B, m].

The iteration is liblpe's fixed-metric FP64 PDHG with the caller's diagonal
metric (``lpe_set_metric``).  Differences from the reference: inputs are
SciPy CSR / NumPy / torch (not CuPy), ``state`` returns owned device tensors,
graph chunks are captured inside liblpe (no graph object is exposed), and
the inequality mask must be the same in every environment (the engine keeps
equality rows first internally; inputs and outputs stay in the caller's row
order).
"""

from __future__ import annotations

import time
from typing import Optional

import numpy as np
import torch
from scipy.sparse import csr_matrix

from .engine import BatchedLpEngine


def _integer(value, name, minimum=1):
    if isinstance(value, (bool, np.bool_)) or not isinstance(value, (int, np.integer)) or value < minimum:
        raise ValueError(f"{name} must be an integer >= {minimum}")
    return int(value)


def _ids(values, count):
    ids = tuple(range(count)) if values is None else tuple(values)
    if len(ids) != count or any(isinstance(v, (bool, np.bool_)) or not isinstance(v, (str, int, np.integer))
                                for v in ids):
        raise ValueError("env_ids must contain one integer/string identifier per environment")
    if len(set(ids)) != count:
        raise ValueError("env_ids must be unique")
    return ids


def _host(value):
    if isinstance(value, torch.Tensor):
        return value.detach().cpu().numpy()
    return np.asarray(value)


class GpuPdhgWorkspace:
    """Fixed-pattern batched FP64 PDHG service (src/gpu_pdhg_workspace.py:55-384)."""

    def __init__(self, a, *, batch_size, n_variables, n_constraints, rhs, c, lower, upper, tau, sigma,
                 inequality_mask, theta=1.0, env_ids=None, chunk_size=64, use_graph=True,
                 stream: Optional[torch.cuda.Stream] = None):
        started = time.perf_counter()
        self.batch = _integer(batch_size, "batch_size")
        self.n = _integer(n_variables, "n_variables")
        self.m = _integer(n_constraints, "n_constraints", minimum=0)
        self.chunk_size = _integer(chunk_size, "chunk_size")
        if max(self.batch * self.n, self.batch * self.m) >= np.iinfo(np.int32).max:
            raise ValueError("Block dimensions exceed int32 CSR capacity")
        if any(value is None for value in (rhs, c, lower, upper, tau, sigma, inequality_mask, theta)):
            raise ValueError("Initial problem vectors, metrics, mask and theta are required")
        if not isinstance(use_graph, (bool, np.bool_)):
            raise ValueError("use_graph must be boolean")
        self.use_graph = bool(use_graph)
        self.env_ids = _ids(env_ids, self.batch)
        self.device = torch.device("cuda", torch.cuda.current_device())
        self.stream = stream if stream is not None else torch.cuda.current_stream(self.device)
        self.iterations_run = 0
        self.update_timing: dict = {}
        self.run_stats: dict = {}
        self._ready = False
        self.engine = None
        a = self._matrix_pattern(a)
        self._indptr, self._indices = a.indptr.copy(), a.indices.copy()
        vectors = self._validated_vectors(dict(rhs=rhs, c=c, lower=lower, upper=upper, tau=tau, sigma=sigma,
                                               inequality_mask=inequality_mask), full=True)
        theta = self._theta(theta)
        self._apply(a, vectors, theta, rebuild=True)
        self.reset()
        self.setup_seconds = time.perf_counter() - started
        self.setup_timing = dict(setup_total_seconds=self.setup_seconds, chunk_size=self.chunk_size,
                                 graph_enabled=self.use_graph, kernels_per_iteration=4 if self.m else 2,
                                 spmv_backend="liblpe_fixed_metric_fp64")

    # ------------------------------------------------------------------ validation
    def _matrix_pattern(self, a) -> csr_matrix:
        """Canonical block CSR of the configured shape (src/gpu_pdhg_workspace.py:170-188)."""
        if getattr(a, "format", None) != "csr" or a.shape != (self.batch * self.m, self.batch * self.n):
            raise ValueError("a must be a block CSR matrix of the configured shape")
        ptr, cols = np.asarray(a.indptr), np.asarray(a.indices)
        data = np.asarray(a.data)
        if data.dtype.kind != "f":
            raise ValueError("a requires floating-point values")
        if ptr.shape != (a.shape[0] + 1,) or ptr[0] != 0 or ptr[-1] != len(cols) or np.any(np.diff(ptr) < 0):
            raise ValueError("Invalid CSR row pointers")
        if data.size != cols.size or np.any(cols < 0) or np.any(cols >= a.shape[1]):
            raise ValueError("Invalid CSR column indices")
        rows = np.repeat(np.arange(a.shape[0]), np.diff(ptr))
        if len(cols) > 1 and np.any((rows[1:] == rows[:-1]) & (cols[1:] <= cols[:-1])):
            raise ValueError("CSR rows must have sorted, unique column indices")
        if self.m and np.any(rows // self.m != cols // self.n):
            raise ValueError("CSR must not couple different environments")
        if not np.isfinite(data).all():
            raise ValueError("Matrix values must be finite")
        return csr_matrix((data.astype(np.float64), cols.astype(np.int32), ptr.astype(np.int32)), shape=a.shape)

    def _vector(self, value, name, size, *, boolean=False, bounds=False) -> np.ndarray:
        result = _host(value)
        if result.shape not in ((size,), (self.batch, size // self.batch if self.batch else 0)):
            raise ValueError(f"{name} has the wrong shape")
        if boolean:
            if result.dtype != np.bool_:
                raise ValueError(f"{name} must be boolean")
        else:
            if result.dtype.kind not in "fiu":
                raise ValueError(f"{name} must be real numeric")
            result = result.astype(np.float64)
            invalid = np.isnan(result) if bounds else ~np.isfinite(result)
            if invalid.any():
                raise ValueError(f"{name} contains non-finite values")
        return result.reshape(size).copy()

    def _validated_vectors(self, supplied: dict, full: bool = False) -> dict:
        B, m, n = self.batch, self.m, self.n
        sizes = dict(rhs=B * m, c=B * n, lower=B * n, upper=B * n, tau=B * n, sigma=B * m, inequality_mask=B * m)
        updates = {}
        for name in ("rhs", "c", "lower", "upper", "tau", "sigma", "inequality_mask"):
            if supplied.get(name) is not None:
                updates[name] = self._vector(supplied[name], name, sizes[name], boolean=name == "inequality_mask",
                                             bounds=name in ("lower", "upper"))
        current = {} if full else self._vectors
        lo, hi = updates.get("lower", current.get("lower")), updates.get("upper", current.get("upper"))
        if np.any((lo > hi) | np.isposinf(lo) | np.isneginf(hi)):
            raise ValueError("Bounds require lower <= upper and finite feasible points")
        for name in ("tau", "sigma"):
            if name in updates and np.any(updates[name] <= 0):
                raise ValueError(f"{name} metrics must be strictly positive")
        if "inequality_mask" in updates and m:
            mask = updates["inequality_mask"].reshape(B, m)
            if np.any(mask != mask[0]):
                raise ValueError("inequality_mask must be the same in every environment on this engine")
        return updates

    @staticmethod
    def _theta(theta) -> float:
        theta = float(theta)
        if not np.isfinite(theta) or not 0.0 <= theta <= 1.0:
            raise ValueError("theta must be finite and between zero and one")
        return theta

    # ------------------------------------------------------------------ device state
    def _apply(self, a: Optional[csr_matrix], updates: dict, theta: Optional[float], rebuild: bool) -> None:
        """Copy a validated update into the engine (equality rows first inside it)."""
        B, m, n = self.batch, self.m, self.n
        if rebuild:
            self._vectors = {}
            self._theta_value = theta
        self._vectors.update(updates)
        if theta is not None:
            self._theta_value = theta
        if a is not None:
            self._a = a
        mask = self._vectors["inequality_mask"].reshape(B, m)[0] if m else np.zeros(0, bool)
        perm = np.concatenate([np.flatnonzero(~mask), np.flatnonzero(mask)]).astype(np.int64)
        if rebuild or self.engine is None or not np.array_equal(perm, self._perm):
            self._perm = perm
            self._build_engine()
        vals = self._block_values(self._a)
        V = self._vectors
        rhs = V["rhs"].reshape(B, m)[:, self._perm]
        sigma = V["sigma"].reshape(B, m)[:, self._perm]
        with torch.cuda.stream(self.stream):
            self.engine.update(V["c"].reshape(B, n), rhs, V["lower"].reshape(B, n), V["upper"].reshape(B, n),
                               var_values=vals if vals.shape[1] else None)
            self.engine.set_metric(V["tau"].reshape(B, n), sigma, self._theta_value)
        self._ready = False

    def _build_engine(self) -> None:
        """Union pattern of the environments' blocks (rows permuted equality
        first); every entry carries per-environment values."""
        B, m, n = self.batch, self.m, self.n
        neq = int(m - np.count_nonzero(self._vectors["inequality_mask"].reshape(B, m)[0])) if m else 0
        ptr, cols = self._indptr, self._indices
        rows = np.repeat(np.arange(B * m), np.diff(ptr))
        local_r = np.zeros(rows.size, np.int64)
        if m:
            inv = np.empty(m, np.int64)
            inv[self._perm] = np.arange(m)
            local_r = inv[rows % m]
        pairs = np.unique(local_r * n + (cols % n)) if cols.size else np.zeros(0, np.int64)
        pr, pc = pairs // n, pairs % n
        template = csr_matrix((np.ones(pairs.size), pc.astype(np.int32),
                               np.r_[0, np.cumsum(np.bincount(pr, minlength=m))].astype(np.int32)), shape=(m, n))
        self._entry_env = rows // m if m else np.zeros(0, np.int64)
        self._entry_slot = np.searchsorted(pairs, local_r * n + cols % n) if cols.size else np.zeros(0, np.int64)
        if self.engine is not None:
            self.engine.close()
        with torch.cuda.stream(self.stream):
            self.engine = BatchedLpEngine(template, neq, B, var_positions=np.arange(pairs.size), method="fixed",
                                          theta=self._theta_value, device=self.device)
        self._nnz_union = pairs.size

    def _block_values(self, a: csr_matrix) -> np.ndarray:
        vals = np.zeros((self.batch, self._nnz_union))
        if a.nnz:
            vals[self._entry_env, self._entry_slot] = a.data
        return vals

    # ------------------------------------------------------------------ public API
    def update_problem(self, *, a=None, rhs=None, c=None, lower=None, upper=None, tau=None, sigma=None,
                       inequality_mask=None, theta=None, env_ids=None) -> dict:
        """Validate a complete update before copying (src/gpu_pdhg_workspace.py:209-267)."""
        started = time.perf_counter()
        if env_ids is not None and _ids(env_ids, self.batch) != self.env_ids:
            raise ValueError("env_ids/order changed; explicitly rebuild the workspace")
        matrix = None
        if a is not None:
            matrix = self._matrix_pattern(a)
            if not np.array_equal(matrix.indptr, self._indptr) or not np.array_equal(matrix.indices, self._indices):
                raise ValueError("CSR pattern changed; explicitly rebuild the workspace")
            if tau is None or sigma is None:
                raise ValueError("Matrix updates require explicit tau and sigma metrics")
        updates = self._validated_vectors(dict(rhs=rhs, c=c, lower=lower, upper=upper, tau=tau, sigma=sigma,
                                               inequality_mask=inequality_mask))
        theta = None if theta is None else self._theta(theta)
        validated_at = time.perf_counter()
        # no state has changed before all validation succeeded
        self._apply(matrix, updates, theta, rebuild=False)
        self.synchronize()
        self.update_timing = dict(update_total_seconds=time.perf_counter() - started,
                                  validation_seconds=validated_at - started,
                                  copy_and_synchronize_seconds=time.perf_counter() - validated_at,
                                  matrix_values_updated=a is not None, copied_vector_names=list(updates))
        return dict(self.update_timing)

    def reset(self, x=None, y=None) -> None:
        """Cold or warm start, projected to the box and the dual signs."""
        B, m, n = self.batch, self.m, self.n
        warm_x = None if x is None else self._vector(x, "x", B * n).reshape(B, n)
        warm_y = None if y is None else self._vector(y, "y", B * m).reshape(B, m)[:, self._perm]
        with torch.cuda.stream(self.stream):
            self.engine.reset(warm_x, warm_y)
        self.accepted = np.zeros(B, dtype=bool)
        self.iterations_run = 0
        self._ready = True
        self.synchronize()

    def set_accepted(self, mask) -> None:
        """Freeze certified environments; only reset may unfreeze them."""
        if not self._ready:
            raise RuntimeError("reset is required after a problem update")
        value = self._vector(mask, "accepted mask", self.batch, boolean=True)
        if np.any(self.accepted & ~value):
            raise ValueError("Cannot unfreeze accepted environments without reset")
        with torch.cuda.stream(self.stream):
            self.engine.set_accepted(value)
        self.accepted = value

    def run(self, iterations, *, use_graph=None) -> dict:
        """Enqueue fixed work: full chunks as graph replays, the tail as single launches."""
        iterations = _integer(iterations, "iterations", minimum=0)
        if not self._ready:
            raise RuntimeError("reset is required after a problem update")
        if use_graph is not None and not isinstance(use_graph, (bool, np.bool_)):
            raise ValueError("use_graph must be boolean")
        graph_route = self.use_graph if use_graph is None else bool(use_graph)
        if graph_route and not self.use_graph:
            raise RuntimeError("Workspace has no captured graph; rebuild with use_graph=True")
        started = time.perf_counter()
        chunks, tail = divmod(iterations, self.chunk_size) if graph_route else (0, iterations)
        if iterations:
            with torch.cuda.stream(self.stream):
                self.engine.run(iterations, chunk=self.chunk_size, use_graph=graph_route)
        self.iterations_run += iterations
        self.run_stats = dict(iterations=iterations, total_iterations=self.iterations_run, graph_chunks=chunks,
                              python_tail_iterations=tail, graph_route=graph_route,
                              enqueue_seconds=time.perf_counter() - started,
                              timing_scope="Host enqueue only; caller must synchronize for elapsed wall time")
        return dict(self.run_stats)

    def state(self, *, synchronize=True):
        """x [B, n] and y [B, m] (owned device tensors, the caller's row order)."""
        if not self._ready:
            raise RuntimeError("reset is required after a problem update")
        with torch.cuda.stream(self.stream):
            x, y, _ = self.engine.state()
            if self.m:
                out = torch.empty_like(y)
                out[:, torch.as_tensor(self._perm, device=y.device)] = y
                y = out
        if synchronize:
            self.synchronize()
        return x, y

    def synchronize(self) -> None:
        self.stream.synchronize()

    # current problem data (host copies, flattened as the reference's buffers)
    @property
    def rhs(self):
        return self._vectors["rhs"].copy()

    @property
    def c(self):
        return self._vectors["c"].copy()

    @property
    def lower(self):
        return self._vectors["lower"].copy()

    @property
    def upper(self):
        return self._vectors["upper"].copy()

    @property
    def tau(self):
        return self._vectors["tau"].copy()

    @property
    def sigma(self):
        return self._vectors["sigma"].copy()

    @property
    def inequality_mask(self):
        return self._vectors["inequality_mask"].copy()

    @property
    def theta(self):
        return self._theta_value

    def close(self) -> None:
        if self.engine is not None:
            self.engine.close()
            self.engine = None
