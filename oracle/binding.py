"""ctypes binding of the CPU oracle (oracle/spai_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the timed CPU baseline.  The
product (dcu_amd) never imports it.

Also binds oracle/_ref/libref_host.so when it was built from the reference
tree (computeN1MAX / computeN2MAX, CSC2CSR, readMatrixToCSC).
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(_HERE, "liboracle.so")
REF_LIB = os.path.join(_HERE, "_ref", "libref_host.so")

_lib = None
_ref = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_LIB):
            raise RuntimeError(f"oracle not built at {ORACLE_LIB}: run __graft_entry__.build()")
        L = ctypes.CDLL(ORACLE_LIB)
        vp, i = ctypes.c_void_p, ctypes.c_int
        L.oracle_column_sets.argtypes = [i, vp, vp, vp, vp, vp, vp]
        L.oracle_spai.argtypes = [i, vp, vp, vp, i, i, i, vp]
        L.oracle_spai_pattern.argtypes = [i, vp, vp, vp, vp, vp, i, i, i, vp]
        L.oracle_csc2csr.argtypes = [i, i, vp, vp, vp, vp, vp, vp, vp]
        L.oracle_spmv.argtypes = [i, vp, vp, vp, vp, vp]
        L.oracle_spmv.restype = None
        L.oracle_bicgstab.argtypes = [i, vp, vp, vp, vp, vp, vp, vp, vp, ctypes.c_double, i, vp]
        L.oracle_bicgstab_history.argtypes = [i, vp, vp, vp, vp, vp, vp, vp, vp, ctypes.c_double, i, vp, vp]
        L.oracle_true_residual.argtypes = [i, vp, vp, vp, vp, vp]
        L.oracle_true_residual.restype = ctypes.c_double
        L.oracle_adjoint.argtypes = [i, vp, vp, vp, vp, vp, vp, ctypes.c_double, i, vp, vp, vp]
        L.oracle_num_threads.restype = i
        L.oracle_set_threads.argtypes = [i]
        L.oracle_set_threads.restype = None
        L.oracle_set_rho_mode.argtypes = [i]
        L.oracle_set_rho_mode.restype = None
        L.oracle_set_shadow_fp32.argtypes = [i]
        L.oracle_set_shadow_fp32.restype = None
        _lib = L
    return _lib


def ref_lib() -> Optional[ctypes.CDLL]:
    """The reference's host helpers, or None when oracle/_ref was not built."""
    global _ref
    if _ref is None and os.path.exists(REF_LIB):
        R = ctypes.CDLL(REF_LIB)
        vp, i = ctypes.c_void_p, ctypes.c_int
        R.ref_compute_n1max.argtypes = [i, vp, vp, vp]
        R.ref_compute_n2max.argtypes = [i, vp]
        R.ref_csc2csr.argtypes = [i, vp, vp, vp, vp, vp, vp]
        R.ref_read_mtx.argtypes = [ctypes.c_char_p, ctypes.POINTER(i), ctypes.POINTER(i)]
        R.ref_read_mtx.restype = vp
        R.ref_mtx_get.argtypes = [vp, vp, vp, vp]
        R.ref_mtx_get.restype = None
        R.ref_mtx_free.argtypes = [vp]
        R.ref_mtx_free.restype = None
        _ref = R
    return _ref


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def num_threads() -> int:
    return lib().oracle_num_threads()


def set_threads(t: int) -> None:
    lib().oracle_set_threads(int(t))


def set_rho_mode(mode: int) -> None:
    """0: rho' = (r*, r) as the reference; 1: (r*, s) - omega (r*, t)."""
    lib().oracle_set_rho_mode(int(mode))


def iteration_spread(a_csr, m_csr, b, tol: float, max_iter: int, threads=(1, 2, 3, 4, 5, 6, 7, 8, 12, 16), x0=None,
                     shadow32: bool = False):
    """Oracle iteration counts under different reduction orders (thread
    counts): BiCGSTAB's count is a chaotic function of rounding on
    non-normal problems, so GPU parity is judged against this spread.
    x0: the initial guess (zeros by default); shadow32 as in bicgstab."""
    before = num_threads()
    its = []
    start = np.zeros(len(b)) if x0 is None else x0
    try:
        for t in threads:
            set_threads(t)
            its.append(bicgstab(a_csr, m_csr, b, start, tol, max_iter, shadow32=shadow32)[1])
    finally:
        set_threads(before)
    return its


def column_sets(n, colptr, rowidx):
    colptr, rowidx = _i32(colptr), _i32(rowidx)
    n1, n2, t, e = (np.empty(n, dtype=np.int32) for _ in range(4))
    st = lib().oracle_column_sets(n, colptr.ctypes.data, rowidx.ctypes.data, n1.ctypes.data, n2.ctypes.data,
                                  t.ctypes.data, e.ctypes.data)
    assert st == 0
    return n1, n2, t, e


def spai(n, colptr, rowidx, val, qr_mode: int = 0, col_begin: int = 0, col_end: int = -1) -> np.ndarray:
    """M's values in A's CSC order; qr_mode 0 = MGS (reference order), 1 = Householder."""
    colptr, rowidx, val = _i32(colptr), _i32(rowidx), _f64(val)
    m = np.zeros(max(len(val), 1))
    st = lib().oracle_spai(n, colptr.ctypes.data, rowidx.ctypes.data, val.ctypes.data, qr_mode, col_begin,
                           n if col_end < 0 else col_end, m.ctypes.data)
    assert st == 0
    return m[: len(val)]


def spai_pattern(n, colptr, rowidx, val, pcolptr, prowidx, qr_mode: int = 0) -> np.ndarray:
    """M's values on the pattern (pcolptr, prowidx), in its CSC order."""
    colptr, rowidx, val = _i32(colptr), _i32(rowidx), _f64(val)
    pcolptr, prowidx = _i32(pcolptr), _i32(prowidx)
    m = np.zeros(max(len(prowidx), 1))
    st = lib().oracle_spai_pattern(n, colptr.ctypes.data, rowidx.ctypes.data, val.ctypes.data, pcolptr.ctypes.data,
                                   prowidx.ctypes.data, qr_mode, 0, n, m.ctypes.data)
    assert st == 0
    return m[: len(prowidx)]


def csc2csr(n, colptr, rowidx, val) -> Tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray]:
    colptr, rowidx, val = _i32(colptr), _i32(rowidx), _f64(val)
    nnz = int(colptr[n])
    rp = np.empty(n + 1, dtype=np.int32)
    ci = np.empty(max(nnz, 1), dtype=np.int32)
    cv = np.empty(max(nnz, 1))
    perm = np.empty(max(nnz, 1), dtype=np.int32)
    st = lib().oracle_csc2csr(n, n, colptr.ctypes.data, rowidx.ctypes.data, val.ctypes.data, rp.ctypes.data,
                              ci.ctypes.data, cv.ctypes.data, perm.ctypes.data)
    assert st == 0
    return rp, ci[:nnz], cv[:nnz], perm[:nnz]


def spmv(rp, ci, v, x) -> np.ndarray:
    rp, ci, v, x = _i32(rp), _i32(ci), _f64(v), _f64(x)
    y = np.empty(len(rp) - 1)
    lib().oracle_spmv(len(rp) - 1, rp.ctypes.data, ci.ctypes.data, v.ctypes.data, x.ctypes.data, y.ctypes.data)
    return y


def bicgstab(a_csr, m_csr, b, x0, tol: float, max_iter: int, history: bool = False, shadow32: bool = False):
    """Returns (x, iterations, info) with info = (||r||, ||r0||, converged);
    with history=True also the per-iteration ||r_k|| as a fourth value.
    shadow32=True takes r* = r0 rounded to FP32 (the mixed-precision solve's
    storage of the shadow residual)."""
    rp, ci, v = (_i32(a_csr[0]), _i32(a_csr[1]), _f64(a_csr[2]))
    b = _f64(b)
    x = np.array(x0, dtype=np.float64, copy=True)
    info = np.zeros(3)
    if m_csr is None:
        mrp = mci = mv = None
        mp = (None, None, None)
    else:
        mrp, mci, mv = _i32(m_csr[0]), _i32(m_csr[1]), _f64(m_csr[2])
        mp = (mrp.ctypes.data, mci.ctypes.data, mv.ctypes.data)
    hist = np.zeros(max(int(max_iter), 1)) if history else None
    L = lib()
    L.oracle_set_shadow_fp32(1 if shadow32 else 0)
    try:
        it = L.oracle_bicgstab_history(len(rp) - 1, rp.ctypes.data, ci.ctypes.data, v.ctypes.data, mp[0], mp[1],
                                       mp[2], b.ctypes.data, x.ctypes.data, float(tol), int(max_iter),
                                       info.ctypes.data, None if hist is None else hist.ctypes.data)
    finally:
        L.oracle_set_shadow_fp32(0)
    if history:
        k = it if info[2] else it - 1
        return x, it, info, hist[:k]
    return x, it, info


def true_residual(a_csr, b, x) -> float:
    rp, ci, v = (_i32(a_csr[0]), _i32(a_csr[1]), _f64(a_csr[2]))
    b, x = _f64(b), _f64(x)
    return lib().oracle_true_residual(len(rp) - 1, rp.ctypes.data, ci.ctypes.data, v.ctypes.data, b.ctypes.data,
                                      x.ctypes.data)


def adjoint(n, colptr, rowidx, val, mval, x, xt, tol: float, max_iter: int):
    """dL/dA (CSC order) and dL/db for L = 0.5||x - xt||^2."""
    colptr, rowidx, val = _i32(colptr), _i32(rowidx), _f64(val)
    mval = None if mval is None else _f64(mval)
    x, xt = _f64(x), _f64(xt)
    dA = np.empty(max(len(val), 1))
    db = np.empty(n)
    info = np.zeros(4)
    it = lib().oracle_adjoint(n, colptr.ctypes.data, rowidx.ctypes.data, val.ctypes.data,
                              None if mval is None else mval.ctypes.data, x.ctypes.data, xt.ctypes.data, float(tol),
                              int(max_iter), dA.ctypes.data, db.ctypes.data, info.ctypes.data)
    return dA[: len(val)], db, it, info
