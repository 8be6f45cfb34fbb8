"""LP request normalisation at the backend boundary.

A request is ``(c, kwargs)`` in ``scipy.optimize.linprog`` form, the shape the
reference's simulator hands to its backends (src/community_solver.py:1200-1205,
src/cpu_repeated_lp.py:783).  ``normalize`` produces the canonical tuple
``(A, rhs, lower, upper, c, neq)`` with rows ``[A_eq; A_ub]`` in sorted,
duplicate-free CSR, exactly as src/cpu_repeated_lp.py:119-147 (``_problem``)
and src/lp_bounds.py:12-51 (``split_variable_bounds``) do.
"""

from __future__ import annotations

import numpy as np
from scipy.sparse import csr_matrix, vstack


def split_variable_bounds(bounds, n: int):
    """Lower/upper arrays: None -> [0, inf); a pair broadcasts; None -> +-inf."""
    if bounds is None:
        return np.zeros(n), np.full(n, np.inf)
    values = np.asarray(bounds)
    if values.shape == (2,):
        values = np.broadcast_to(values, (n, 2))
    if values.shape != (n, 2):
        raise ValueError("Bounds must be a pair or one pair per variable")
    if values.dtype.kind in "biuf":
        return np.array(values[:, 0], dtype=float, copy=True), np.array(values[:, 1], dtype=float, copy=True)
    values = np.asarray(bounds, dtype=object)
    if values.shape == (2,):
        values = np.broadcast_to(values, (n, 2))
    lower = np.array([-np.inf if v is None else v for v in values[:, 0]], dtype=float)
    upper = np.array([np.inf if v is None else v for v in values[:, 1]], dtype=float)
    return lower, upper


def normalize(objective, kwargs: dict):
    """Validate one continuous LP request; return ``(A, rhs, lower, upper, c, neq)``."""
    c = np.asarray(objective, dtype=np.float64)
    if c.ndim != 1 or not len(c) or not np.isfinite(c).all():
        raise ValueError("A finite one-dimensional objective is required")
    n = len(c)
    integrality = kwargs.get("integrality")
    if integrality is not None and np.any(np.asarray(integrality)):
        raise ValueError("The batched LP engine supports continuous LPs only")
    if kwargs.get("method", "highs-ds") not in ("highs", "highs-ds", "highs-ipm", "pdhg"):
        raise ValueError("Unsupported LP method")
    blocks, rhs = [], []
    for matrix_name, rhs_name in (("A_eq", "b_eq"), ("A_ub", "b_ub")):
        supplied = kwargs.get(matrix_name)
        matrix = csr_matrix((0, n), dtype=float) if supplied is None else csr_matrix(supplied, dtype=float)
        if matrix.shape[1] != n or not np.isfinite(matrix.data).all():
            raise ValueError("Invalid LP matrix")
        values = kwargs.get(rhs_name)
        values = np.empty(0) if values is None and not matrix.shape[0] else np.asarray(values, dtype=float)
        if values.shape != (matrix.shape[0],) or not np.isfinite(values).all():
            raise ValueError("Each matrix row requires a finite RHS")
        blocks.append(matrix)
        rhs.append(values)
    a = vstack(blocks, format="csr")
    a.sum_duplicates()
    a.eliminate_zeros()
    a.sort_indices()
    neq = blocks[0].shape[0]
    lower, upper = split_variable_bounds(kwargs.get("bounds"), n)
    if np.isnan(lower).any() or np.isnan(upper).any() or (lower > upper).any():
        raise ValueError("Invalid variable bounds")
    return a, np.concatenate(rhs), lower, upper, c.copy(), neq


def problem_request(problem):
    """Inverse of ``normalize``: a canonical tuple as a ``(c, kwargs)`` request
    (the adapter of src/lp_trace.py:18-40)."""
    a, rhs, lower, upper, c, neq = problem
    a = csr_matrix(a)
    m = a.shape[0]
    kwargs = dict(bounds=np.column_stack((lower, upper)), method="highs-ds")
    if neq:
        kwargs.update(A_eq=a[:neq], b_eq=np.asarray(rhs[:neq]))
    if neq < m:
        kwargs.update(A_ub=a[neq:], b_ub=np.asarray(rhs[neq:]))
    return np.asarray(c, dtype=np.float64).copy(), kwargs


def pattern_key(a, neq: int):
    """Hashable structure of a canonical matrix (shape, equality count, pattern)."""
    a = csr_matrix(a)
    return (a.shape, int(neq), a.indptr.tobytes(), a.indices.tobytes())
