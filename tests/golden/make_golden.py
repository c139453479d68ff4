"""Writes the golden fixtures under tests/golden/ from independent libraries
(numpy LAPACK least squares, scipy.sparse.linalg.bicgstab), not from the
oracle, so that the oracle can be checked against them.

  spai_lstsq.npz      small sparse matrices (random diagonally dominant,
                      2D conv-diff, 3D diffusion, a missing-diagonal case) in
                      CSC, with M(:,k) = argmin ||A(:,J_k) m - e_k|| computed
                      by numpy.linalg.lstsq on the dense A(:,J_k) (all n rows,
                      so I_k is never formed), and |I_k| by brute force
                      (nonzero rows of A(:,J_k)).
  bicgstab_scipy.npz  small systems with the scipy BiCGSTAB iterate after
                      k = 1..6 iterations (right preconditioner M from the
                      lstsq fixture; scipy's preconditioned BiCGSTAB uses the
                      same recurrences as bicgstab_solver.h:106-232).

Run: python tests/golden/make_golden.py
"""
import os

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spl

HERE = os.path.dirname(os.path.abspath(__file__))


def random_dd(n, k, rng):
    rows, cols, vals = [], [], []
    for i in range(n):
        cs = rng.choice(np.setdiff1d(np.arange(n), [i]), size=k, replace=False)
        vs = rng.uniform(-1, 1, size=k)
        rows += [i] * (k + 1)
        cols += list(cs) + [i]
        vals += list(vs) + [1.5 * np.abs(vs).sum()]
    return sp.csc_matrix((vals, (rows, cols)), shape=(n, n))


def convdiff(m, pe):
    h = 1.0 / (m + 1)
    eps = 1.0 / pe
    n = m * m
    A = sp.lil_matrix((n, n))
    for j in range(m):
        for i in range(m):
            r = i + m * j
            A[r, r] = 4 * eps + h * 1.5
            if i > 0:
                A[r, r - 1] = -eps - h * 1.0
            if i < m - 1:
                A[r, r + 1] = -eps
            if j > 0:
                A[r, r - m] = -eps - h * 0.5
            if j < m - 1:
                A[r, r + m] = -eps
    return A.tocsc()


def diffusion3d(m, rng):
    kappa = np.exp(rng.standard_normal(m ** 3))
    n = m ** 3
    A = sp.lil_matrix((n, n))

    def idx(i, j, k):
        return i + m * (j + m * k)

    for k in range(m):
        for j in range(m):
            for i in range(m):
                r = idx(i, j, k)
                diag = 0.0
                for di, dj, dk in [(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)]:
                    ii, jj, kk = i + di, j + dj, k + dk
                    if 0 <= ii < m and 0 <= jj < m and 0 <= kk < m:
                        c = 2 * kappa[r] * kappa[idx(ii, jj, kk)] / (kappa[r] + kappa[idx(ii, jj, kk)])
                        A[r, idx(ii, jj, kk)] = -c
                        diag += c
                    else:
                        diag += 2 * kappa[r]
                A[r, r] = diag
    return A.tocsc()


def missing_diag():
    d = np.array([[0.0, 0.0, 0.0, 0.0, 1.0],
                  [2.0, 3.0, 0.0, 0.0, 0.0],
                  [0.0, 1.0, 4.0, 1.0, 0.0],
                  [0.0, 0.0, 1.0, 5.0, 0.0],
                  [1.0, 0.0, 0.0, 2.0, 6.0]])
    return sp.csc_matrix(d)


def lstsq_spai(A):
    A = sp.csc_matrix(A)
    A.sort_indices()
    n = A.shape[0]
    Ad = A.toarray()
    m = np.zeros(A.nnz)
    n1 = np.zeros(n, dtype=np.int32)
    for k in range(n):
        J = A.indices[A.indptr[k]:A.indptr[k + 1]]
        if len(J) == 0:
            continue
        sub = Ad[:, J]
        n1[k] = int(np.count_nonzero(np.any(sub != 0, axis=1)))
        e = np.zeros(n)
        e[k] = 1.0
        if not np.any(sub[k] != 0):
            continue  # k not in I_k: the reference sets M(:,k) = 0
        m[A.indptr[k]:A.indptr[k + 1]] = np.linalg.lstsq(sub, e, rcond=None)[0]
    return A, m, n1


def main():
    rng = np.random.default_rng(20261203)
    mats = {
        "rand40": random_dd(40, 4, rng),
        "rand90": random_dd(90, 3, rng),
        "convdiff7": convdiff(7, 10.0),
        "diff3d4": diffusion3d(4, rng),
        "missing5": missing_diag(),
    }
    out = {}
    for name, A in mats.items():
        A, m, n1 = lstsq_spai(A)
        out[f"{name}_n"] = np.array(A.shape[0])
        out[f"{name}_colptr"] = A.indptr.astype(np.int32)
        out[f"{name}_rowidx"] = A.indices.astype(np.int32)
        out[f"{name}_val"] = A.data
        out[f"{name}_m"] = m
        out[f"{name}_n1"] = n1
    np.savez_compressed(os.path.join(HERE, "spai_lstsq.npz"), names=np.array(list(mats)), **out)

    bout = {}
    for name in ["rand40", "rand90", "convdiff7", "diff3d4"]:
        n = int(out[f"{name}_n"])
        A = sp.csc_matrix((out[f"{name}_val"], out[f"{name}_rowidx"], out[f"{name}_colptr"]), shape=(n, n))
        M = sp.csc_matrix((out[f"{name}_m"], out[f"{name}_rowidx"], out[f"{name}_colptr"]), shape=(n, n))
        b = rng.uniform(size=n)
        bout[f"{name}_b"] = b
        xs = []
        for k in range(1, 7):
            x, _ = spl.bicgstab(A.tocsr(), b, x0=np.zeros(n), rtol=1e-300, atol=0.0, maxiter=k,
                                M=spl.aslinearoperator(M.tocsr()))
            xs.append(x)
        bout[f"{name}_x"] = np.array(xs)
        xs = []
        for k in range(1, 7):
            x, _ = spl.bicgstab(A.tocsr(), b, x0=np.zeros(n), rtol=1e-300, atol=0.0, maxiter=k)
            xs.append(x)
        bout[f"{name}_x_noprec"] = np.array(xs)
    np.savez_compressed(os.path.join(HERE, "bicgstab_scipy.npz"), **bout)
    print("wrote", os.listdir(HERE))


if __name__ == "__main__":
    main()
