"""GPU checks of the hand-written adjoint (no reference counterpart).

L = 0.5 ||x - x_t||^2 with A x = b.  The adjoint solves A^T lambda = x - x_t
with BiCGSTAB preconditioned by M^T and returns dL/db = lambda,
dL/dA_ij = -lambda_i x_j on A's pattern.  Checked against the FP64 CPU
gradient of the oracle (<= 1e-8 relative), against central finite
differences with h = 1e-6 on random entries (solves at 1e-13 so that the
differencing noise tol/h stays near 1e-7), on small grids and on the full
64^3 workload (100 entries), and through the torch.autograd.Function
wrapper.
"""
import numpy as np
import pytest

import dcu_amd as D
from oracle import binding as O

pytestmark = pytest.mark.gpu

TOL = 1e-13


def setup(ctx, m=12, seed=5):
    P = D.Problem.grf(m, seed)
    A = P.system(0)
    S = D.System(ctx, A)
    S.build(D.PRECISION_FP64, D.SCHEDULE_BALANCED_NNZ)
    return P, A, S, P.rhs(0), P.target(0)


def loss_of(S, b, xt):
    x, rep, oc = S.solve(b, rel_tol=TOL, max_iter=5000)
    assert oc[0]["status"] == D.SOLVE_CONVERGED
    return 0.5 * float(np.sum((x - xt) ** 2)), x


def test_adjoint_matches_oracle_gradient(ctx):
    P, A, S, b, xt = setup(ctx)
    L, x = loss_of(S, b, xt)
    dA, db, rep = S.adjoint(x, x - xt, rel_tol=TOL, max_iter=5000)
    assert rep["all_converged"]
    m_ref = O.spai(A.n, A.colptr, A.rowidx, A.val, 0)
    dA_ref, db_ref, it, info = O.adjoint(A.n, A.colptr, A.rowidx, A.val, m_ref, x, xt,
                                         TOL * np.linalg.norm(x - xt), 5000)
    assert np.linalg.norm(db - db_ref) <= 1e-8 * np.linalg.norm(db_ref)
    assert np.linalg.norm(dA - dA_ref) <= 1e-8 * np.linalg.norm(dA_ref)
    # closed form on the pattern: dA = -lambda_i x_j
    cols = np.repeat(np.arange(A.n), np.diff(A.colptr))
    assert np.allclose(dA, -db[A.rowidx] * x[cols], rtol=1e-14, atol=0)


def test_adjoint_central_differences(ctx):
    P, A, S, b, xt = setup(ctx, m=10, seed=7)
    L0, x = loss_of(S, b, xt)
    dA, db, rep = S.adjoint(x, x - xt, rel_tol=TOL, max_iter=5000)
    rng = np.random.default_rng(0)
    h = 1e-6
    vals = A.val.copy()
    worst = 0.0
    for p in rng.choice(A.nnz, 20, replace=False):
        vp = vals.copy()
        vp[p] += h
        S.set_values(vp)
        Lp, _ = loss_of(S, b, xt)
        vm = vals.copy()
        vm[p] -= h
        S.set_values(vm)
        Lm, _ = loss_of(S, b, xt)
        fd = (Lp - Lm) / (2 * h)
        err = abs(fd - dA[p]) / max(abs(dA[p]), 1e-3 * np.max(np.abs(dA)))
        worst = max(worst, err)
    S.set_values(vals)
    for i in rng.choice(A.n, 10, replace=False):
        bp = b.copy()
        bp[i] += h
        Lp, _ = loss_of(S, bp, xt)
        bm = b.copy()
        bm[i] -= h
        Lm, _ = loss_of(S, bm, xt)
        fd = (Lp - Lm) / (2 * h)
        err = abs(fd - db[i]) / max(abs(db[i]), 1e-3 * np.max(np.abs(db)))
        worst = max(worst, err)
    print(f"worst relative FD mismatch {worst:.2e}", flush=True)
    assert worst <= 1e-5


def test_adjoint_full_size_grf64(ctx):
    # the differentiation workload at full size: 3D 7-point diffusion on a
    # 64^3 grid (262,144 unknowns), kappa from a Gaussian random field, b and
    # x_t N(0,1); central differences (h = 1e-6) on 100 random entries of A
    # and the FP64 CPU gradient of the oracle.  M stays the one built from
    # the unperturbed A (it only preconditions).
    P = D.Problem.generate(5, 5)
    A = P.system(0)
    b, xt = P.rhs(0), P.target(0)
    assert A.n == 64 ** 3
    S = D.System(ctx, A)
    S.build(D.PRECISION_FP64, D.SCHEDULE_BALANCED_NNZ)
    L0, x = loss_of(S, b, xt)
    dA, db, rep = S.adjoint(x, x - xt, rel_tol=TOL, max_iter=5000)
    assert rep["all_converged"]
    m_ref = O.spai(A.n, A.colptr, A.rowidx, A.val, 0)
    dA_ref, db_ref, it, info = O.adjoint(A.n, A.colptr, A.rowidx, A.val, m_ref, x, xt,
                                         TOL * np.linalg.norm(x - xt), 5000)
    e_db = np.linalg.norm(db - db_ref) / np.linalg.norm(db_ref)
    e_dA = np.linalg.norm(dA - dA_ref) / np.linalg.norm(dA_ref)
    rng = np.random.default_rng(64)
    h = 1e-6
    vals = A.val.copy()
    scale = np.max(np.abs(dA))
    errs = []

    def perturbed(v):
        # warm start from x: the solve only forms the O(h) correction, to an
        # absolute residual far below h (at 1e-13 of ||b|| from x0 = 0 the
        # two solves' errors would differ by about 1e-3 of the derivative)
        S.set_values(v)
        xs, _, oc = S.solve(b, x0=x, abs_tol=1e-12, max_iter=5000)
        assert oc[0]["status"] in (D.SOLVE_CONVERGED, D.SOLVE_CONVERGED_AT_START)
        return xs

    for p in rng.choice(A.nnz, 100, replace=False):
        vp = vals.copy()
        vp[p] += h
        xp = perturbed(vp)
        vp[p] = vals[p] - h
        xm = perturbed(vp)
        # L(+h) - L(-h) = 0.5 sum (x+ - x-)(x+ + x- - 2 x_t), summed without
        # forming the two losses (about 1e5 each, 1e-8 apart)
        fd = 0.5 * float(np.sum((xp - xm) * (xp + xm - 2.0 * xt))) / (2 * h)
        errs.append(abs(fd - dA[p]) / max(abs(dA[p]), 1e-3 * scale))
    S.set_values(vals)
    errs = np.array(errs)
    print(f"grf64: adjoint {rep['iterations_max']} its; vs oracle gradient: dA {e_dA:.2e} db {e_db:.2e}; "
          f"central differences on 100 entries: median {np.median(errs):.2e} worst {errs.max():.2e}", flush=True)
    assert e_db <= 1e-8 and e_dA <= 1e-8
    # the differences' own floor: x+ - x- is of order h, and the perturbed
    # solves reach an absolute residual of about 1e-12 (the FP64 floor of
    # b - A x at this size), so the difference quotient carries errors of
    # 1e-5 to 1e-4 of the derivative (the small grids above reach 1e-6)
    assert np.median(errs) <= 1e-4 and errs.max() <= 1e-3


def test_autograd_wrapper(ctx):
    import torch

    from dcu_amd.autograd import spai_solve

    P, A, S, b, xt = setup(ctx, m=8, seed=9)
    dev = torch.device("cuda:0")
    vals = torch.tensor(A.val, dtype=torch.float64, device=dev, requires_grad=True)
    bt = torch.tensor(b, dtype=torch.float64, device=dev, requires_grad=True)
    xtt = torch.tensor(xt, dtype=torch.float64, device=dev)
    x = spai_solve(vals, bt, S, 1e-13, 5000)
    loss = 0.5 * ((x - xtt) ** 2).sum()
    loss.backward()
    xh = x.detach().cpu().numpy()
    dA, db, _ = S.adjoint(xh, xh - xt, rel_tol=1e-13, max_iter=5000)
    assert np.allclose(vals.grad.cpu().numpy(), dA, rtol=1e-10, atol=1e-14)
    assert np.allclose(bt.grad.cpu().numpy(), db, rtol=1e-10, atol=1e-14)


def test_adjoint_follows_new_values_and_m(ctx):
    """The adjoint keeps A^T and M^T in its own layout: after set_values and
    a rebuild of M (in either precision) it solves with the new operators.
    With A scaled by 2, lambda halves."""
    P, A, S, b, xt = setup(ctx, m=16, seed=3)
    L, x = loss_of(S, b, xt)
    g = x - xt
    _, db1, rep = S.adjoint(x, g, rel_tol=TOL, max_iter=5000)
    assert rep["all_converged"]
    S.set_values(2.0 * A.val)
    for precision in (D.PRECISION_FP64, D.PRECISION_MIXED):
        S.build(precision, D.SCHEDULE_BALANCED_NNZ)
        dA2, db2, rep = S.adjoint(x, g, rel_tol=TOL, max_iter=5000)
        assert rep["all_converged"]
        err = np.linalg.norm(db2 - 0.5 * db1) / np.linalg.norm(db1)
        print(f"precision {precision}: |db(2A) - db(A)/2| / |db(A)| = {err:.2e}", flush=True)
        assert err <= 1e-10
    S.set_values(A.val)
