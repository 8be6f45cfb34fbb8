"""Original-LP traces: lossless per-LP files plus a checksummed manifest.

The on-disk format is the reference's (src/lp_trace.py:18-96), so traces it
captured from real dFBA runs (scripts/capture_dfba_lp_trace.py:33-66) replay
here unchanged:

* one ``.npz`` per LP (no pickles) with ``a_data, a_indices, a_indptr,
  a_shape, rhs, lower, upper, c, neq, reference_x, reference_y``;
* a manifest entry per file: ``filename``, ``sha256`` of the file,
  ``problem_sha256`` (``problem_hash`` below), ``rows``, ``columns``,
  ``nonzeros``, ``equalities`` and optionally ``step`` (>= 1), ``stage``
  (``maxmin``, ``aggregate`` or ``exchange``) and ``environment_id`` (>= 0),
  which must agree with the name ``lp_{step:03d}_{stage index}_{env:03d}.npz``.

Loading re-validates everything: the path stays inside the trace directory,
the file checksum, the problem hash after passing through the backend's own
request normalisation, the dimensions and the labels.  ``replay`` solves the
traced LPs through a batched backend and compares each answer with the
recorded reference solution.
"""

from __future__ import annotations

import hashlib
from pathlib import Path

import numpy as np
from scipy.sparse import csr_matrix

from .requests import normalize, problem_request

STAGES = ("maxmin", "aggregate", "exchange")
_INTEGER_FIELDS = {"shape", "indices", "indptr", "neq"}


def problem_hash(problem) -> str:
    """SHA-256 over (name, shape, little-endian values) of every field, in the
    order shape, data, indices, indptr, rhs, lower, upper, c, neq; integer
    fields as int64, the rest as float64 (src/lp_trace.py:18-29)."""
    a, rhs, lower, upper, c, neq = problem
    digest = hashlib.sha256()
    fields = (("shape", a.shape), ("data", a.data), ("indices", a.indices), ("indptr", a.indptr),
              ("rhs", rhs), ("lower", lower), ("upper", upper), ("c", c), ("neq", neq))
    for name, value in fields:
        arr = np.asarray(value, dtype="<i8" if name in _INTEGER_FIELDS else "<f8")
        digest.update(name.encode())
        digest.update(np.asarray(arr.shape, dtype="<i8").tobytes())
        digest.update(arr.tobytes())
    return digest.hexdigest()


def trace_filename(step: int, stage: str, environment_id: int) -> str:
    return f"lp_{step:03d}_{STAGES.index(stage)}_{environment_id:03d}.npz"


def write_trace_lp(path, problem, *, reference_x, reference_y) -> dict:
    """Write one LP and its reference solution; returns the manifest entry."""
    path = Path(path)
    if path.exists():
        raise FileExistsError(path)
    a, rhs, lower, upper, c, neq = problem
    x, y = np.asarray(reference_x), np.asarray(reference_y)
    if x.shape != c.shape or y.shape != rhs.shape or not (np.isfinite(x).all() and np.isfinite(y).all()):
        raise ValueError("Invalid reference solution")
    np.savez_compressed(path, a_data=a.data, a_indices=a.indices, a_indptr=a.indptr, a_shape=a.shape,
                        rhs=rhs, lower=lower, upper=upper, c=c, neq=neq, reference_x=x, reference_y=y)
    return dict(filename=path.name, sha256=hashlib.sha256(path.read_bytes()).hexdigest(),
                problem_sha256=problem_hash(problem), rows=a.shape[0], columns=a.shape[1],
                nonzeros=a.nnz, equalities=int(neq))


def load_trace_lp(directory, entry):
    """(problem, reference_x, reference_y) of one manifest entry, validated."""
    directory = Path(directory).resolve()
    path = (directory / entry["filename"]).resolve()
    if not path.is_relative_to(directory) or not path.is_file():
        raise ValueError("Untrusted LP trace path or checksum")
    if hashlib.sha256(path.read_bytes()).hexdigest() != entry["sha256"]:
        raise ValueError("Untrusted LP trace path or checksum")
    with np.load(path, allow_pickle=False) as data:
        a = csr_matrix((data["a_data"], data["a_indices"], data["a_indptr"]), shape=tuple(data["a_shape"]))
        problem = (a, data["rhs"], data["lower"], data["upper"], data["c"], int(data["neq"]))
        x, y = np.array(data["reference_x"]), np.array(data["reference_y"])
    rebuilt = normalize(*problem_request(problem))
    if problem_hash(rebuilt) != entry["problem_sha256"]:
        raise ValueError("LP trace input changed during restoration")
    a, rhs, _, _, c, neq = rebuilt
    if (entry.get("rows"), entry.get("columns"), entry.get("nonzeros"), entry.get("equalities")) != (
            a.shape[0], a.shape[1], a.nnz, neq):
        raise ValueError("LP trace dimension metadata does not match its payload")
    if any(k in entry for k in ("step", "stage", "environment_id")):
        stage, step, env = entry.get("stage"), entry.get("step"), entry.get("environment_id")
        if stage not in STAGES or not (isinstance(step, int) and step >= 1 and isinstance(env, int) and env >= 0):
            raise ValueError("Invalid LP trace labels")
        if entry["filename"] != trace_filename(step, stage, env):
            raise ValueError("LP trace labels disagree with the recorded filename")
    if x.shape != c.shape or y.shape != rhs.shape or not (np.isfinite(x).all() and np.isfinite(y).all()):
        raise ValueError("Invalid reference solution")
    return rebuilt, x, y


def replay(directory, manifest, backend, *, rel_tol: float = 1e-6) -> dict:
    """Solve every traced LP with ``backend.solve_batch`` (grouped by step and
    stage, environments as IDs) and compare with the recorded references.
    Returns per-entry results: certified, relative objective error."""
    entries = list(manifest["entries"] if isinstance(manifest, dict) else manifest)
    groups: dict = {}
    for i, e in enumerate(entries):
        groups.setdefault((e.get("step"), e.get("stage")), []).append(i)
    rows = [None] * len(entries)
    for (_, stage), idx in sorted(groups.items(), key=lambda kv: (kv[0][0] or 0, str(kv[0][1]))):
        loaded = [load_trace_lp(directory, entries[i]) for i in idx]
        requests = []
        for (problem, _, _) in loaded:
            c, kw = problem_request(problem)
            if stage is not None:
                kw["_stage"] = stage
            requests.append((c, kw))
        env_ids = [entries[i].get("environment_id", i) for i in idx]
        results = backend.solve_batch(requests, environment_ids=env_ids)
        for i, (problem, x_ref, _), r in zip(idx, loaded, results):
            ref_obj = float(problem[4] @ x_ref)
            err = None if not r.success else abs(r.fun - ref_obj) / max(1.0, abs(ref_obj))
            rows[i] = dict(filename=entries[i]["filename"], certified=bool(r.success), reference_objective=ref_obj,
                           objective=r.fun, relative_objective_error=err,
                           within_tolerance=bool(r.success and err <= rel_tol))
    return dict(entries=rows, certified=sum(r["certified"] for r in rows),
                within_tolerance=sum(r["within_tolerance"] for r in rows), total=len(rows))
