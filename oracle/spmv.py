"""CSR products in storage order (test infrastructure; see oracle/__init__.py).

The engine's kernels sum each row's products sequentially in CSR storage
order with separate multiply and add roundings (no FMA); these functions
reproduce that order exactly, vectorised across rows, so integer-valued and
general inputs can be compared bit for bit.

Pinned by: scipy.sparse CSR matvec on integer-valued data (exact in any
order) and to 1e-15 relative otherwise (tests/test_oracle.py).
"""

from __future__ import annotations

import numpy as np
from scipy.sparse import csr_matrix


def csr_matvec_ordered(A: csr_matrix, x: np.ndarray) -> np.ndarray:
    """y = A x for x of shape [n] or [B, n] (returns [m] or [B, m])."""
    A = csr_matrix(A)
    x = np.asarray(x)
    if x.dtype not in (np.float32, np.float64):
        x = x.astype(np.float64)
    batched = x.ndim == 2
    X = x if batched else x[None, :]
    m = A.shape[0]
    counts = np.diff(A.indptr)
    out = np.zeros((X.shape[0], m), dtype=np.result_type(A.data.dtype, X.dtype))
    if A.nnz == 0:
        return out if batched else out[0]
    kmax = int(counts.max())
    for k in range(kmax):
        rows = np.flatnonzero(counts > k)
        pos = A.indptr[rows] + k
        prod = A.data[pos][None, :] * X[:, A.indices[pos]]
        out[:, rows] = out[:, rows] + prod
    return out if batched else out[0]


def csr_rmatvec_ordered(A: csr_matrix, y: np.ndarray) -> np.ndarray:
    """A^T y summed per column in increasing row order (the engine's CSC order)."""
    At = csr_matrix(A).T.tocsr()
    At.sort_indices()
    return csr_matvec_ordered(At, y)
