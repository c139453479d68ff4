"""The reference's C++ entry points, called the way its driver calls them.

src/spai_multi_dcu.cpp:2386-2606 runs: initMultiDCU, Run{Static,Dynamic}SPAI_
MultiDCU_MPI on a host CSC (M comes back as device CSC arrays the caller
frees), cuCSC2CSR of A and M on the device, then cublas2_pbicgstabv2 with
b = 1, x0 = 1, an absolute tolerance of 1e-7 and MAX_ITER = 10000
(runBiCGSTABOnRank0, :1878-1951), and cleanupMultiDCU.  Here the same calls go
through the library's exported C++ symbols.  Checked: M (pattern(A), values
within 1e-10 of the oracle per column), the reference's 0.0 return for an
empty column range, CSR arrays identical to scipy's, the iteration count
against the oracle's spread over reduction orders, ||b - A x|| below the
tolerance, and b left unchanged (the reference overwrites it; this build
does not).
"""
import ctypes

import numpy as np
import pytest
import scipy.sparse as sp

import dcu_amd as D
from oracle import binding as O
from parity import count_ok, oracle_counts
from test_gpu_spai import column_rel_err

pytestmark = pytest.mark.gpu


class Mat(ctypes.Structure):  # CSC_Matrix / CSR_Matrix (include/spai_types.h)
    _fields_ = [("mIndex", ctypes.c_void_p), ("mData", ctypes.c_void_p), ("mPtr", ctypes.c_void_p),
                ("nRow", ctypes.c_int), ("nCol", ctypes.c_int), ("n", ctypes.c_int), ("nonzeroes", ctypes.c_int)]


def hip_runtime():
    path = next(l.split()[-1] for l in open("/proc/self/maps") if "libamdhip64" in l)
    hip = ctypes.CDLL(path)
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    return hip


def to_host(hip, ptr, n, dtype):
    out = np.empty(n, dtype=dtype)
    assert hip.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0  # device to host
    return out


def cxx(lib):
    f = {k: getattr(lib, v) for k, v in D.CXX_SYMBOLS.items()}
    f["initMultiDCU"].restype = ctypes.c_void_p
    f["initMultiDCU"].argtypes = [ctypes.c_int]
    f["cleanupMultiDCU"].argtypes = [ctypes.c_void_p]
    for name in ("RunStaticSPAI_MultiDCU_MPI", "RunDynamicSPAI_MultiDCU_MPI"):
        f[name].restype = ctypes.c_float
        f[name].argtypes = [ctypes.c_void_p, ctypes.POINTER(Mat), ctypes.POINTER(Mat), ctypes.c_int, ctypes.c_int]
    f["cublas2_pbicgstabv2"].restype = ctypes.c_int
    f["cublas2_pbicgstabv2"].argtypes = [ctypes.POINTER(Mat), ctypes.POINTER(Mat), ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_double, ctypes.c_int]
    f["cuCSC2CSR"].argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 6
    return f


@pytest.mark.parametrize("tune", ["0", "1"])
@pytest.mark.parametrize("entry", ["RunStaticSPAI_MultiDCU_MPI", "RunDynamicSPAI_MultiDCU_MPI"])
def test_reference_driver_sequence(ctx, entry, tune, monkeypatch):
    import torch

    monkeypatch.setenv("SPAI_MI355X_AUTOTUNE", tune)  # 1: tile shapes tuned inside Run*

    lib = D.load_library()
    f = cxx(lib)
    hip = hip_runtime()
    A = D.Problem.generate(1, 1).system(0)  # random 4096, 5 nnz per column
    n, nnz = A.n, A.nnz
    hA = A.struct()
    hostA = Mat(ctypes.cast(hA.mIndex, ctypes.c_void_p), ctypes.cast(hA.mData, ctypes.c_void_p),
                ctypes.cast(hA.mPtr, ctypes.c_void_p), n, n, n, nnz)
    mc = f["initMultiDCU"](1)
    assert mc
    M = Mat()
    assert f[entry](mc, ctypes.byref(hostA), ctypes.byref(M), 5, 5) == 0.0  # empty range, as the reference
    ms = f[entry](mc, ctypes.byref(hostA), ctypes.byref(M), 0, n)
    assert ms > 0.0 and M.nonzeroes == nnz and M.n == n
    m_ptr = to_host(hip, M.mPtr, n + 1, np.int32)
    m_idx = to_host(hip, M.mIndex, nnz, np.int32)
    m_val = to_host(hip, M.mData, nnz, np.float64)
    assert np.array_equal(m_ptr, A.colptr) and np.array_equal(m_idx, A.rowidx)
    m_ref = O.spai(n, A.colptr, A.rowidx, A.val, 0)
    err = column_rel_err(A, m_val, m_ref)

    # device CSC of A, then CSR of A and M through cuCSC2CSR
    dev = torch.device("cuda:0")
    a_ptr = torch.from_numpy(A.colptr.astype(np.int32)).to(dev)
    a_idx = torch.from_numpy(A.rowidx.astype(np.int32)).to(dev)
    a_val = torch.from_numpy(A.val.astype(np.float64)).to(dev)
    csr = {k: (torch.empty(n + 1, dtype=torch.int32, device=dev), torch.empty(nnz, dtype=torch.int32, device=dev),
               torch.empty(nnz, dtype=torch.float64, device=dev)) for k in ("A", "M")}
    f["cuCSC2CSR"](n, n, nnz, a_val.data_ptr(), a_idx.data_ptr(), a_ptr.data_ptr(), csr["A"][2].data_ptr(),
                   csr["A"][1].data_ptr(), csr["A"][0].data_ptr())
    f["cuCSC2CSR"](n, n, nnz, M.mData, M.mIndex, M.mPtr, csr["M"][2].data_ptr(), csr["M"][1].data_ptr(),
                   csr["M"][0].data_ptr())
    torch.cuda.synchronize()
    As = sp.csc_matrix((A.val, A.rowidx, A.colptr), shape=(n, n)).tocsr()
    As.sort_indices()
    assert np.array_equal(csr["A"][0].cpu().numpy(), As.indptr)
    assert np.array_equal(csr["A"][1].cpu().numpy(), As.indices)
    assert np.array_equal(csr["A"][2].cpu().numpy(), As.data)
    Ms = sp.csc_matrix((m_val, m_idx, m_ptr), shape=(n, n)).tocsr()
    Ms.sort_indices()
    assert np.array_equal(csr["M"][2].cpu().numpy(), Ms.data)

    def mat(c):
        return Mat(csr[c][1].data_ptr(), csr[c][2].data_ptr(), csr[c][0].data_ptr(), n, n, n, nnz)

    dA, dM = mat("A"), mat("M")
    b = torch.ones(n, dtype=torch.float64, device=dev)
    x = torch.ones(n, dtype=torch.float64, device=dev)
    its = f["cublas2_pbicgstabv2"](ctypes.byref(dA), ctypes.byref(dM), b.data_ptr(), x.data_ptr(), 1e-7, 10000)
    torch.cuda.synchronize()
    xh = x.cpu().numpy()
    r = np.linalg.norm(np.ones(n) - As @ xh)
    Acsr = O.csc2csr(n, A.colptr, A.rowidx, A.val)[:3]
    Mcsr = O.csc2csr(n, A.colptr, A.rowidx, m_val)[:3]
    counts = oracle_counts(Acsr, Mcsr, np.ones(n), 1e-7, 10000, threads=(1, 4), x0=np.ones(n))
    print(f"{entry} (autotune {tune}): build {ms:.3f} ms, M err {err:.2e}, iterations {its} "
          f"(oracle {sorted(counts)}), ||b - A x|| {r:.2e}", flush=True)
    assert err <= 1e-10
    assert count_ok(its, counts)
    assert r < 1.5e-7
    assert torch.equal(b, torch.ones_like(b))
    for p in (M.mPtr, M.mIndex, M.mData):
        assert hip.hipFree(p) == 0
    f["cleanupMultiDCU"](mc)
