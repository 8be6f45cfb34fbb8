"""CPU tests of the host side: the C-ABI libraries load and export every
declared symbol, and the request normalisation / pattern logic in front of the
engine (src/cpu_repeated_lp.py:119-147, src/gpu_pdhg_corrector.py:35-72)."""

import ctypes
import os
import re

import numpy as np
import pytest
from scipy.sparse import csr_matrix

from conftest import PKG, ROOT

from lpe import requests
from lpe.corrector import _validated_problem, shared_pattern

LIBS = {"lpe.h": os.path.join(PKG, "lpe", "liblpe.so"),
        "lpe_reduction.h": os.path.join(PKG, "lpe", "liblpe.so"),
        "warmstart.h": os.path.join(PKG, "lpe", "libwarmstart.so")}


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:lpe|ws)_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header", sorted(LIBS))
def test_library_exports_every_declared_symbol(header):
    path = LIBS[header]
    if not os.path.exists(path):
        pytest.fail(f"{path} is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    names = declared(header)
    assert len(names) >= 6
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_bindings_cover_the_headers():
    from lpe import _lib, warmstart
    assert set(declared("lpe.h")) | set(declared("lpe_reduction.h")) == set(_lib.EXPORTS)
    assert set(declared("warmstart.h")) == set(warmstart.EXPORTS)


def test_versions_are_callable_without_device():
    from lpe import warmstart
    lib = ctypes.CDLL(LIBS["lpe.h"])
    lib.lpe_version.restype = ctypes.c_char_p
    assert b"lpe" in lib.lpe_version().lower() or lib.lpe_version()
    assert b"gfx950" in warmstart.load().ws_version()


# --------------------------------------------------------------- requests
def test_normalize_stacks_equalities_first_sorted_and_summed():
    c = np.array([1.0, -2.0, 0.5])
    a_eq = np.array([[1.0, 0.0, 1.0]])
    a_ub = csr_matrix((np.array([2.0, 1.0, 1.0, 0.0]), np.array([2, 0, 0, 1]), np.array([0, 4])), shape=(1, 3))
    a, rhs, lo, hi, cc, neq = requests.normalize(c, dict(A_eq=a_eq, b_eq=[1.0], A_ub=a_ub, b_ub=[4.0],
                                                         bounds=[(0, 5), (None, 2), (-1, None)]))
    assert neq == 1 and a.shape == (2, 3)
    np.testing.assert_array_equal(a.toarray(), [[1, 0, 1], [2, 0, 2]])       # duplicates summed, zero dropped
    assert a.has_sorted_indices and a.nnz == 4
    np.testing.assert_array_equal(rhs, [1.0, 4.0])
    np.testing.assert_array_equal(lo, [0, -np.inf, -1])
    np.testing.assert_array_equal(hi, [5, 2, np.inf])
    cc[0] = 99.0
    assert c[0] == 1.0                                                        # owned copy


def test_normalize_default_bounds_and_pair_broadcast():
    _, _, lo, hi, _, _ = requests.normalize([1.0, 1.0], dict(A_ub=[[1.0, 1.0]], b_ub=[1.0]))
    np.testing.assert_array_equal(lo, [0, 0])
    np.testing.assert_array_equal(hi, [np.inf, np.inf])
    _, _, lo, hi, _, _ = requests.normalize([1.0, 1.0], dict(A_ub=[[1.0, 1.0]], b_ub=[1.0], bounds=(-2, 3)))
    np.testing.assert_array_equal(lo, [-2, -2])
    np.testing.assert_array_equal(hi, [3, 3])


@pytest.mark.parametrize("kwargs,objective", [
    (dict(integrality=[1, 0]), [1.0, 1.0]),
    (dict(method="interior-point"), [1.0, 1.0]),
    (dict(A_eq=[[1.0, 1.0]]), [1.0, 1.0]),                      # rows without a rhs
    (dict(A_eq=[[1.0, np.nan]], b_eq=[1.0]), [1.0, 1.0]),
    (dict(bounds=[(1, 0), (0, 1)]), [1.0, 1.0]),
    (dict(), [np.inf, 1.0]),
    (dict(), []),
])
def test_normalize_rejects_invalid_requests(kwargs, objective):
    with pytest.raises(ValueError):
        requests.normalize(objective, kwargs)


def test_problem_request_round_trip():
    rng = np.random.default_rng(0)
    dense = np.where(rng.random((5, 7)) < 0.4, rng.integers(-3, 4, (5, 7)), 0).astype(float)
    a = csr_matrix(dense)
    a.eliminate_zeros()
    problem = (a, rng.normal(size=5), -rng.random(7), 1 + rng.random(7), rng.normal(size=7), 2)
    c, kwargs = requests.problem_request(problem)
    back = requests.normalize(c, kwargs)
    np.testing.assert_array_equal(back[0].toarray(), a.toarray())
    for got, want in zip(back[1:5], problem[1:5]):
        np.testing.assert_array_equal(got, want)
    assert back[5] == 2
    assert requests.pattern_key(back[0], 2) == requests.pattern_key(a, 2)


def test_shared_pattern_union_and_varying_positions():
    a1 = csr_matrix(np.array([[1.0, 0.0, 2.0], [0.0, 3.0, 0.0]]))
    a2 = csr_matrix(np.array([[1.0, 5.0, 2.0], [0.0, 4.0, 0.0]]))
    template, varying = shared_pattern([a1, a2])[:2]
    assert template.nnz == 4                      # union pattern, LP 0's values (0 at (0, 1))
    np.testing.assert_array_equal(template.toarray(), a1.toarray())
    # positions whose value differs between the matrices are the varying ones
    coords = [(r, template.indices[k]) for r in range(2) for k in range(template.indptr[r], template.indptr[r + 1])]
    assert sorted(coords[p] for p in np.asarray(varying).ravel()) == [(0, 1), (1, 1)]


def test_validated_problem_copies_and_rejects():
    a = csr_matrix(np.eye(2))
    p = (a, np.ones(2), np.zeros(2), np.ones(2), np.ones(2), 2)
    out = _validated_problem(p)
    out[1][0] = 7.0
    assert p[1][0] == 1.0
    with pytest.raises(ValueError):
        _validated_problem((a, np.ones(3), np.zeros(2), np.ones(2), np.ones(2), 2))
    with pytest.raises(ValueError):
        _validated_problem((a, np.ones(2), np.ones(2), np.zeros(2), np.ones(2), 2))


def test_ctypes_structs_match_the_c_layout(tmp_path):
    """Size and every field offset of the ctypes mirrors equal the C compiler's
    layout of include/lpe.h and include/lpe_reduction.h (gcc), so the Python
    layer and the library agree."""
    import shutil
    import subprocess
    from lpe import _lib
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    structs = {"lpe_desc": _lib.LpeDesc, "lpe_vectors": _lib.LpeVectors,
               "lpe_solve_opts": _lib.LpeSolveOpts, "lpe_solve_info": _lib.LpeSolveInfo,
               "lpe_reduction_desc": _lib.LpeReductionDesc, "lpe_reduction_dims": _lib.LpeReductionDims,
               "lpe_reduced": _lib.LpeReduced, "lpe_batch_problem": _lib.LpeBatchProblem}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "lpe_reduction.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, py in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)
