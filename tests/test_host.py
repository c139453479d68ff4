"""CPU tests of the host side: the C-ABI library loads and exports every
declared symbol, the generators are deterministic and produce the stated
structures, the MatrixMarket reader follows the reference's reader
(include/common/read.h), and the product fails loudly without a GPU.
No compute call here touches a GPU."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp

import dcu_amd as D
from oracle import binding as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_MATRICES = "/root/reference/matrices"


def declared_c_functions():
    names = set()
    for h in ("spai_mi355x.h",):
        text = open(os.path.join(ROOT, "include", h)).read()
        names |= set(re.findall(r"\b(spai_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    lib = D.load_library()
    declared = declared_c_functions()
    assert declared == set(D.C_SYMBOLS), declared ^ set(D.C_SYMBOLS)
    for name in declared:
        assert hasattr(lib, name), name
    hdr = open(os.path.join(ROOT, "include", "spai_multi_dcu_api.h")).read()
    for name, mangled in D.CXX_SYMBOLS.items():
        assert re.search(r"\b" + name + r"\s*\(", hdr), name
        assert hasattr(lib, mangled), mangled


def test_tile_shape_candidates():
    # each size bucket's candidates gather exactly its T_k, hold at most its
    # |J_k| with a lane per x entry (the last bucket's hold all of it), and a
    # column handed on by a bucket (|I_k| <= T_k of that bucket) fits every
    # candidate of the next; the last bucket never hands on; candidate 0 is
    # the default; the 2D 5-point and 3D 7-point stencils have candidates of
    # their own width
    bucket_t, bucket_nj = (32, 64, 128), (8, 8, 16)
    assert [D.tile_shape_count(t) for t in range(4)][3] == 0
    for t in range(3):
        n = D.tile_shape_count(t)
        assert 1 <= n <= D.TUNE_MAX_CANDIDATES
        shapes = [D.tile_shape(t, c) for c in range(n)]
        assert len({tuple(s.values()) for s in shapes}) == n
        for s in shapes:
            lanes, epl = s["lanes"], s["entries_per_lane"]
            assert lanes * epl == bucket_t[t] and s["max_cols"] <= bucket_nj[t]
            assert t < 2 or s["max_cols"] == bucket_nj[t]
            assert lanes >= s["max_cols"] and 64 % lanes == 0 and epl & (epl - 1) == 0
            rows = lanes * s["rows_per_lane"]
            if t > 0:
                assert rows >= bucket_t[t - 1]
            if t == 2:
                assert rows >= bucket_t[t]
    assert any(D.tile_shape(0, c)["max_cols"] == 5 for c in range(D.tile_shape_count(0)))
    assert any(D.tile_shape(1, c)["max_cols"] == 7 for c in range(D.tile_shape_count(1)))
    assert [D.tile_shape(t, 0) for t in range(3)] == [
        {"lanes": 8, "entries_per_lane": 4, "max_cols": 8, "rows_per_lane": 2},
        {"lanes": 16, "entries_per_lane": 4, "max_cols": 8, "rows_per_lane": 2},
        {"lanes": 64, "entries_per_lane": 2, "max_cols": 16, "rows_per_lane": 2}]
    with pytest.raises(D.SpaiError):
        D.tile_shape(0, D.tile_shape_count(0))
    with pytest.raises(D.SpaiError):
        D.tile_shape(3, 0)


def test_tuner_entry_points_reject_bad_arguments():
    # argument checks that run before any device work
    lib = D.load_library()
    shape = (ctypes.c_int * 3)(0, 0, 0)
    t = D.Tuning()
    assert lib.spai_system_autotune(None, D.PRECISION_MIXED, D.SCHEDULE_BALANCED_NNZ, 3, ctypes.byref(t)) != 0
    assert lib.spai_system_set_tile_shapes(None, shape) != 0
    assert lib.spai_system_get_tile_shapes(None, shape) != 0
    assert lib.spai_tile_shape_get(0, 0, None) != 0
    assert lib.spai_tile_shape_get(-1, 0, ctypes.byref(D.TileShape())) != 0
    assert lib.spai_tile_shape_count(-1) == 0 and lib.spai_tile_shape_count(5) == 0


def test_struct_layout_matches_reference():
    # dataType.h:24-42: int*, double*, int*, 4 ints
    assert D.CSC_Matrix.mIndex.offset == 0
    assert D.CSC_Matrix.mData.offset == 8
    assert D.CSC_Matrix.mPtr.offset == 16
    assert [D.CSC_Matrix.nRow.offset, D.CSC_Matrix.nCol.offset, D.CSC_Matrix.n.offset,
            D.CSC_Matrix.nonzeroes.offset] == [24, 28, 32, 36]
    assert ctypes.sizeof(D.CSC_Matrix) == 40


def test_generators_deterministic_and_structured():
    a1 = D.Problem.generate(1, 1).system(0)
    a2 = D.Problem.generate(1, 1).system(0)
    a3 = D.Problem.generate(1, 2).system(0)
    assert np.array_equal(a1.val, a2.val) and np.array_equal(a1.rowidx, a2.rowidx)
    assert not np.array_equal(a1.val, a3.val)
    # 4096 rows, 4 off-diagonals + diagonal per row, a_ii = 1.5 sum |a_ij|
    A = sp.csc_matrix((a1.val, a1.rowidx, a1.colptr), shape=(4096, 4096)).tocsr()
    assert A.nnz == 5 * 4096
    d = A.diagonal()
    off = np.abs(A).sum(axis=1).A1 - np.abs(d)
    assert np.allclose(d, 1.5 * off, rtol=1e-14)
    b = D.Problem.generate(1, 1).rhs(0)
    assert np.allclose(b, A @ np.ones(4096), rtol=1e-14)
    # rows inside each column ascending
    for c in range(0, 4096, 97):
        seg = a1.rowidx[a1.colptr[c]:a1.colptr[c + 1]]
        assert np.all(np.diff(seg) > 0)


def test_convection_diffusion_structure():
    A = D.Problem.convection_diffusion(8, 10.0).system(0)
    M = sp.csc_matrix((A.val, A.rowidx, A.colptr), shape=(64, 64)).toarray()
    h, eps = 1.0 / 9, 0.1
    assert np.isclose(M[9, 9], 4 * eps + 1.5 * h)
    assert np.isclose(M[9, 8], -eps - h)       # west, upwind
    assert np.isclose(M[9, 10], -eps)          # east
    assert np.isclose(M[9, 1], -eps - 0.5 * h)  # south, upwind
    assert np.isclose(M[9, 17], -eps)          # north
    assert A.nnz == 5 * 64 - 4 * 8


def test_batched_generator_ranges():
    P = D.Problem.batched(64, 4)
    assert P.num_systems == 64
    for s in range(64):
        A = P.system(s)
        if s % 2 == 0:
            m = int(round(np.sqrt(A.n)))
            assert m * m == A.n
        else:
            m = int(round(A.n ** (1 / 3)))
            assert m ** 3 == A.n
        assert 1000 <= A.n <= 210000
        assert P.rhs(s).shape == (A.n,)


def test_grf_diffusion_field_statistics():
    P = D.Problem.generate(5, 5)
    A = P.system(0)
    assert A.n == 64 ** 3
    M = sp.csc_matrix((A.val, A.rowidx, A.colptr), shape=(A.n, A.n))
    assert abs(M - M.T).max() <= 1e-12 * abs(M).max()  # symmetric FV operator
    assert P.target(0) is not None and P.target(0).shape == (A.n,)


def test_mtx_reader_vs_reference_reader():
    if not os.path.isdir(REF_MATRICES):
        pytest.skip("reference fixtures not present")
    R = O.ref_lib()
    for fname in ("circuit_2.mtx", "fill_3elt.mtx"):
        path = os.path.join(REF_MATRICES, fname)
        mine = D.Problem.read_mtx(path).system(0)
        if R is None:
            pytest.skip("oracle/_ref not built")
        n, nnz = ctypes.c_int(), ctypes.c_int()
        h = R.ref_read_mtx(path.encode(), ctypes.byref(n), ctypes.byref(nnz))
        cp = np.empty(n.value + 1, dtype=np.int32)
        ri = np.empty(nnz.value, dtype=np.int32)
        va = np.empty(nnz.value)
        R.ref_mtx_get(h, cp.ctypes.data, ri.ctypes.data, va.ctypes.data)
        R.ref_mtx_free(h)
        assert mine.n == n.value
        assert np.array_equal(mine.colptr, cp), fname
        assert np.array_equal(mine.rowidx, ri), fname  # file order inside columns
        assert np.array_equal(mine.val, va), fname


def test_mtx_reader_semantics(tmp_path):
    # symmetric mirrored in file order, explicit zero dropped
    p = tmp_path / "s.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real symmetric\n% c\n3 3 4\n1 1 2.0\n3 1 -1.0\n2 2 0.0\n"
                 "3 3 4.0\n")
    A = D.Problem.read_mtx(str(p)).system(0)
    assert list(A.colptr) == [0, 2, 2, 4]
    assert list(A.rowidx) == [0, 2, 0, 2]
    assert list(A.val) == [2.0, -1.0, -1.0, 4.0]
    q = tmp_path / "g.mtx"
    q.write_text("%%MatrixMarket matrix coordinate pattern general\n2 2 3\n1 1\n2 1\n2 2\n")
    B = D.Problem.read_mtx(str(q)).system(0)
    assert list(B.colptr) == [0, 2, 3] and list(B.val) == [1.0, 1.0, 1.0]
    with pytest.raises(D.SpaiError):
        D.Problem.read_mtx(str(tmp_path / "missing.mtx"))


def test_no_gpu_fails_loudly():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(D.SpaiError):
        D.Context(0)


def test_cli_help_runs_without_gpu():
    out = subprocess.run([D.CLI_PATH, "--help"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0
    assert "--algo static|dynamic" in out.stdout
