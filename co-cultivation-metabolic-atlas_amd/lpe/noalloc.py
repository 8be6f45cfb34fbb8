"""Handle-free FP64 certificate for callers that cannot allocate (lpe_certify_noalloc).

``DeviceBatch`` uploads a batch's problem once (CSR, the transposed pattern
and the per-LP vectors as device tensors); ``DeviceBatch.certify`` then
writes the certificate of (x, y) into a caller-provided metrics tensor with
no device allocation by the library or by torch, so it may run inside HIP
graph capture or a serving loop that forbids ``hipMalloc``.  The gates and
metrics are lpe_certify's (src/gpu_block_lp.py:52-132).
"""

from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np
import torch
from scipy.sparse import csr_matrix

from . import _lib
from ._lib import check


def capacity() -> int:
    """Doubles of the library's static scratch array (per GPU)."""
    return int(_lib.load().lpe_certify_noalloc_capacity())


def max_batch(m: int, n: int) -> int:
    """Largest batch of m x n LPs one call accepts."""
    need = int(_lib.load().lpe_certify_noalloc_need(int(m), int(n)))
    return capacity() // need if need > 0 else 0


class DeviceBatch:
    """A batch of LPs sharing ``A``'s pattern, resident on the device.

    ``values`` [B, nnz] (CSR value order of the canonicalised ``A``) gives
    per-LP coefficients; None shares ``A.data``.  ``c`` is [n] or [B, n].
    """

    def __init__(self, A, neq: int, rhs, lower, upper, c, *, values=None, device=None):
        self._lib = _lib.load()
        A = csr_matrix(A, dtype=np.float64, copy=True)
        A.sort_indices()
        dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        m, n = A.shape
        B = int(np.asarray(lower).shape[0]) if np.ndim(lower) == 2 else int(lower.shape[0])
        # transposed pattern: rows ascending inside each column, with the CSR
        # position of every entry
        pos = csr_matrix((np.arange(A.nnz, dtype=np.float64) + 1.0, A.indices, A.indptr), shape=A.shape).tocsc()
        pos.sort_indices()
        t_pos = (pos.data - 1.0).astype(np.int32)
        i32 = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.int32), device=dev)  # noqa: E731
        f64 = lambda a: (a.to(device=dev, dtype=torch.float64).contiguous() if isinstance(a, torch.Tensor)  # noqa: E731
                         else torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=dev))
        self.m, self.n, self.neq, self.batch, self.nnz = m, n, int(neq), B, int(A.nnz)
        self.row_ptr, self.col_idx = i32(A.indptr), i32(A.indices)
        self.col_ptr, self.row_idx, self.t_pos = i32(pos.indptr), i32(pos.indices), i32(t_pos)
        self.values = f64(A.data if values is None else values)
        self.c = f64(c)
        self.rhs, self.lower, self.upper = f64(rhs), f64(lower), f64(upper)
        if tuple(self.rhs.shape) != (B, m) or tuple(self.lower.shape) != (B, n) or tuple(self.upper.shape) != (B, n):
            raise ValueError("rhs must be [B, m] and bounds [B, n]")
        if values is not None and tuple(self.values.shape) != (B, self.nnz):
            raise ValueError("values must be [B, nnz]")
        if tuple(self.c.shape) not in ((n,), (B, n)):
            raise ValueError("c must be [n] or [B, n]")
        p = _lib.LpeBatchProblem()
        p.m, p.n, p.neq, p.batch, p.nnz = m, n, self.neq, B, self.nnz
        p.row_ptr, p.col_idx = self.row_ptr.data_ptr(), self.col_idx.data_ptr()
        p.col_ptr, p.row_idx, p.t_pos = self.col_ptr.data_ptr(), self.row_idx.data_ptr(), self.t_pos.data_ptr()
        p.values, p.values_per_lp = self.values.data_ptr(), 0 if values is None else 1
        p.c, p.c_per_lp = self.c.data_ptr(), 1 if self.c.dim() == 2 else 0
        p.rhs, p.lower, p.upper = self.rhs.data_ptr(), self.lower.data_ptr(), self.upper.data_ptr()
        self._p = p
        self.device = dev

    def certify(self, x: torch.Tensor, y: Optional[torch.Tensor], metrics: torch.Tensor,
                eps=(1e-5, 1e-7, 1e-7), batch: Optional[int] = None) -> torch.Tensor:
        """metrics[:batch] <- certificate of (x[:batch], y[:batch]) on torch's
        current stream (the first ``batch`` LPs; default all).  Allocates
        nothing; x, y and metrics must be contiguous FP64 device tensors."""
        B = self.batch if batch is None else int(batch)
        for name, t, width in (("x", x, self.n), ("y", y, self.m), ("metrics", metrics, _lib.NMETRIC)):
            if t is None:
                if name == "y" and self.m == 0:
                    continue
                raise ValueError(f"{name} is required")
            if t.dtype != torch.float64 or not t.is_contiguous() or t.dim() != 2 or t.shape[1] != width \
                    or t.shape[0] < B or t.device != self.device:
                raise ValueError(f"{name} must be a contiguous float64 [>= {B}, {width}] tensor on {self.device}")
        p = _lib.LpeBatchProblem.from_buffer_copy(self._p)
        p.batch = B
        stream = torch.cuda.current_stream(self.device).cuda_stream
        check(self._lib.lpe_certify_noalloc(C.byref(p), x.data_ptr(), None if y is None else y.data_ptr(),
                                            metrics.data_ptr(), float(eps[0]), float(eps[1]), float(eps[2]),
                                            stream))
        return metrics
