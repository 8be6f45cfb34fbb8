"""LP trace files and manifest (src/lp_trace.py:18-96 format): round trip,
tamper detection, path confinement, label checks; GPU replay through the
batched backend against HiGHS-solved references."""

import hashlib
import json

import numpy as np
import pytest
from scipy.sparse import csr_matrix

from lpe import trace, workloads
from oracle import highs


def _problem(seed=0):
    ub = workloads.uptake_batch(batch=2, seed=seed, m=40, n=60)
    return ub.problem(0)


def _reference(problem):
    r = highs.solve(*problem)
    assert r.status == "optimal"
    return r.x, r.y


def test_problem_hash_layout_known_answer():
    """The hash is SHA-256 over name, int64 shape and little-endian values of
    each field in the fixed order; rebuilt here byte by byte for a 1x2 LP."""
    a = csr_matrix(np.array([[2.0, 0.0]]))
    problem = (a, np.array([1.0]), np.array([0.0, -1.0]), np.array([3.0, 4.0]), np.array([1.0, 2.0]), 1)
    d = hashlib.sha256()
    for name, arr in (("shape", np.array([1, 2], "<i8")), ("data", np.array([2.0], "<f8")),
                      ("indices", np.array([0], "<i8")), ("indptr", np.array([0, 1], "<i8")),
                      ("rhs", np.array([1.0], "<f8")), ("lower", np.array([0.0, -1.0], "<f8")),
                      ("upper", np.array([3.0, 4.0], "<f8")), ("c", np.array([1.0, 2.0], "<f8")),
                      ("neq", np.array(1, "<i8"))):
        d.update(name.encode()); d.update(np.asarray(arr.shape, "<i8").tobytes()); d.update(arr.tobytes())
    assert trace.problem_hash(problem) == d.hexdigest()
    changed = (a, np.array([1.0 + 1e-15]), *problem[2:])
    assert trace.problem_hash(changed) != d.hexdigest()


def test_round_trip_and_labels(tmp_path):
    p = _problem()
    x, y = _reference(p)
    name = trace.trace_filename(3, "exchange", 7)
    assert name == "lp_003_2_007.npz"
    entry = trace.write_trace_lp(tmp_path / name, p, reference_x=x, reference_y=y)
    entry.update(step=3, stage="exchange", environment_id=7)
    back, xb, yb = trace.load_trace_lp(tmp_path, entry)
    assert trace.problem_hash(back) == trace.problem_hash(p) == entry["problem_sha256"]
    np.testing.assert_array_equal(xb, x)
    np.testing.assert_array_equal(yb, y)
    with pytest.raises(FileExistsError):
        trace.write_trace_lp(tmp_path / name, p, reference_x=x, reference_y=y)
    bad = dict(entry, environment_id=8)
    with pytest.raises(ValueError, match="labels disagree"):
        trace.load_trace_lp(tmp_path, bad)
    with pytest.raises(ValueError, match="labels"):
        trace.load_trace_lp(tmp_path, dict(entry, stage="other"))
    with pytest.raises(ValueError, match="dimension"):
        trace.load_trace_lp(tmp_path, dict(entry, rows=entry["rows"] + 1))


def test_tampered_file_and_path_escape_rejected(tmp_path):
    p = _problem()
    x, y = _reference(p)
    entry = trace.write_trace_lp(tmp_path / "lp.npz", p, reference_x=x, reference_y=y)
    raw = bytearray((tmp_path / "lp.npz").read_bytes())
    raw[-10] ^= 1
    (tmp_path / "lp.npz").write_bytes(bytes(raw))
    with pytest.raises(ValueError, match="checksum"):
        trace.load_trace_lp(tmp_path, entry)
    outside = tmp_path.parent / "outside.npz"
    outside.write_bytes(b"x")
    with pytest.raises(ValueError):
        trace.load_trace_lp(tmp_path, dict(entry, filename="../outside.npz"))


def test_invalid_reference_solution_rejected(tmp_path):
    p = _problem()
    x, y = _reference(p)
    with pytest.raises(ValueError):
        trace.write_trace_lp(tmp_path / "a.npz", p, reference_x=x[:-1], reference_y=y)
    with pytest.raises(ValueError):
        trace.write_trace_lp(tmp_path / "b.npz", p, reference_x=np.full_like(x, np.nan), reference_y=y)


@pytest.mark.gpu
def test_replay_through_backend(device, tmp_path):
    from lpe.backend import GpuLpBackend
    ub = workloads.uptake_batch(batch=4, seed=5, m=120, n=180)
    entries = []
    for env in range(4):
        p = ub.problem(env)
        x, y = _reference(p)
        name = trace.trace_filename(1, "maxmin", env)
        e = trace.write_trace_lp(tmp_path / name, p, reference_x=x, reference_y=y)
        e.update(step=1, stage="maxmin", environment_id=env)
        entries.append(e)
    manifest = {"entries": entries}
    (tmp_path / "manifest.json").write_text(json.dumps(manifest))
    backend = GpuLpBackend(precision="fp32")
    try:
        report = trace.replay(tmp_path, json.loads((tmp_path / "manifest.json").read_text()), backend)
    finally:
        backend.close()
    assert report["total"] == 4 and report["certified"] == 4 and report["within_tolerance"] == 4, report
