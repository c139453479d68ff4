"""C = A M and the SPAI quality ||A M - I||_F (the reference's
SpMMWithoutSparsity, include/common/SpMM_lib.h:15, and the diagnostic it
serves).

Pinned by scipy's sparse product (a library, independent of the oracle) on
the M the GPU built, and by properties of the SPAI least squares: a
diagonal A gives A M = I; a nested, larger pattern never raises a column's
residual; a column with E_k = -1 (k outside I_k) has M(:, k) = 0 and
residual exactly 1.
"""
import gzip
import os
import shutil

import numpy as np
import pytest
import scipy.sparse as sp

import dcu_amd as D
from test_gpu_patterns import missing_diagonal_matrix

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def scipy_mats(A, S):
    n = A.n
    As = sp.csc_matrix((A.val, A.rowidx, A.colptr), shape=(n, n))
    cp, ri = S.m_pattern()
    Ms = sp.csc_matrix((S.m_values(), ri, cp), shape=(n, n))
    return As, Ms


def check_product(A, S, label):
    As, Ms = scipy_mats(A, S)
    C = (As @ Ms).tocsc()
    C.sort_indices()
    cp, ri, va = S.product()
    G = sp.csc_matrix((va, ri, cp), shape=(A.n, A.n))
    # rows ascending inside every column of C
    for k in range(0, A.n, max(1, A.n // 50)):
        assert np.all(np.diff(ri[cp[k]:cp[k + 1]]) > 0)
    scale = abs(As).max() * abs(Ms).max() * max(1, np.diff(S.m_pattern()[0]).max())
    diff = abs(G - C).max() if C.nnz else 0.0
    # scipy drops exact zeros; the GPU keeps the structural pattern (I_k)
    assert C.nnz <= G.nnz
    R = (C - sp.identity(A.n, format="csc")).tocsc()
    frob_ref = sp.linalg.norm(R)
    col_ref = np.sqrt(np.asarray(R.multiply(R).sum(axis=0)).ravel())
    frob, cols = S.quality()
    print(f"{label}: nnz(C) {G.nnz} (scipy {C.nnz}), max |C - C_scipy| {diff:.2e} (scale {scale:.2e}), "
          f"||AM - I||_F {frob:.6e} (scipy {frob_ref:.6e})", flush=True)
    assert diff <= 1e-14 * scale
    assert abs(frob - frob_ref) <= 1e-12 * max(frob_ref, 1.0)
    assert np.max(np.abs(cols - col_ref)) <= 1e-12 * max(col_ref.max(), 1.0)
    return frob, cols


@pytest.mark.parametrize("precision", [D.PRECISION_FP64, D.PRECISION_MIXED])
def test_product_random4096(ctx, precision):
    P = D.Problem.generate(1, 1)
    A = P.system(0)
    S = D.System(ctx, A)
    S.build(precision, D.SCHEDULE_BALANCED_NNZ)
    check_product(A, S, f"random4096 prec {precision}")


def test_quality_diagonal_is_exact(ctx):
    n = 300
    d = np.linspace(1.0, 4.0, n)
    A = D.CSC(n, np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32), d)
    S = D.System(ctx, A)
    S.build(D.PRECISION_FP64, D.SCHEDULE_BALANCED_NNZ)
    frob, cols = S.quality()
    assert frob <= 1e-14


def test_quality_larger_pattern_never_worse(ctx):
    # J_k of A is inside J_k of E + A + A^T: each column's least-squares
    # residual can only fall
    P = D.Problem.generate(1, 1)
    A = P.system(0)
    S = D.System(ctx, A)
    S.build(D.PRECISION_FP64, D.SCHEDULE_BALANCED_NNZ)
    f_a, c_a = check_product(A, S, "random4096 pattern A")
    S.set_pattern(D.PATTERN_I_PLUS_A_PLUS_AT)
    S.build(D.PRECISION_FP64, D.SCHEDULE_BALANCED_NNZ)
    f_p, c_p = check_product(A, S, "random4096 pattern E+A+A^T")
    assert np.all(c_p <= c_a + 1e-12)
    assert f_p < f_a


def test_quality_missing_diagonal_columns(ctx):
    A = missing_diagonal_matrix()
    S = D.System(ctx, A)
    S.build(D.PRECISION_FP64, D.SCHEDULE_BALANCED_NNZ)
    frob, cols = check_product(A, S, "missing diagonal, pattern A")
    m = S.m_values()
    zero = np.maximum.reduceat(np.abs(m), A.colptr[:-1]) == 0
    assert zero.sum() > 0
    assert np.all(cols[zero] == 1.0)


def test_product_circuit_2_wide_columns(ctx, tmp_path):
    # 1,285-wide columns: up to 17,948 gathered entries per column of C
    path = tmp_path / "circuit_2.mtx"
    with gzip.open(os.path.join(GOLD, "matrices", "circuit_2.mtx.gz"), "rb") as f, open(path, "wb") as g:
        shutil.copyfileobj(f, g)
    A = D.Problem.read_mtx(str(path)).system(0)
    S = D.System(ctx, A)
    S.build(D.PRECISION_FP64, D.SCHEDULE_BALANCED_NNZ)
    check_product(A, S, "circuit_2")
