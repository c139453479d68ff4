"""Pins the CPU oracle (oracle/spai_oracle.c) before anything is compared
with it.  CPU only.

Each check is chosen so that a plausible mistake in the restatement (a
dropped term, a wrong index or sign, a transposed operand, a wrong iteration
count) fails it:
  - SPAI vs numpy LAPACK least squares on the full dense A(:,J_k) (golden
    fixtures, no I_k involved), MGS vs Householder, closed forms (diagonal A,
    dense pattern -> inverse, block-diagonal with dense blocks -> inverse),
    the normal equations A(:,J)^T (A(:,J) m - e) = 0;
  - |I_k| vs brute force and n1max/n2max vs the reference's own host helpers
    (include/common/init.h) compiled into oracle/_ref;
  - CSC -> CSR vs scipy and the reference's CSC2CSR (include/common/assemble.h);
  - BiCGSTAB iterate-by-iterate vs scipy.sparse.linalg.bicgstab, M = A^-1
    converging in one iteration, a 2x2 case by hand, the MAX_ITER semantics
    of bicgstab_solver.h:132,231 and a planted mistake that must be caught;
  - the FP32 shadow residual (shadow32) by a 2x2 iteration by hand with
    r~ = fl32(r0), bit-identity with the FP64 solve when r0 is exactly
    representable in FP32, and convergence with the recursive residual still
    tracking b - A x otherwise.
"""
import ctypes
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

from oracle import binding as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def gold_mats():
    z = np.load(os.path.join(GOLD, "spai_lstsq.npz"))
    for name in z["names"]:
        name = str(name)
        yield name, int(z[f"{name}_n"]), z[f"{name}_colptr"], z[f"{name}_rowidx"], z[f"{name}_val"], \
            z[f"{name}_m"], z[f"{name}_n1"]


@pytest.mark.parametrize("qr_mode", [0, 1])
def test_spai_matches_lstsq_golden(qr_mode):
    for name, n, cp, ri, va, m_gold, _ in gold_mats():
        m = O.spai(n, cp, ri, va, qr_mode)
        scale = np.max(np.abs(m_gold))
        assert np.max(np.abs(m - m_gold)) <= 1e-12 * scale, name


@pytest.mark.parametrize("flag", [0, 1, 2])
def test_spai_pattern_matches_lstsq(flag):
    # M on an alternative pattern P: column k minimises ||A[:, J] m - e_k||
    # over all n rows, J = P's column k; numpy lstsq on the dense A pins it
    from patterns import pattern_csc
    for name, n, cp, ri, va, _, _ in gold_mats():
        if n > 200:
            continue
        pcp, pri = pattern_csc(n, cp, ri, flag)
        m = O.spai_pattern(n, cp, ri, va, pcp, pri, 0)
        Ad = np.zeros((n, n))
        for c in range(n):
            Ad[ri[cp[c]:cp[c + 1]], c] = va[cp[c]:cp[c + 1]]
        ref = np.zeros(len(pri))
        for k in range(n):
            J = pri[pcp[k]:pcp[k + 1]]
            if len(J) == 0:
                continue
            e = np.zeros(n)
            e[k] = 1.0
            ref[pcp[k]:pcp[k + 1]] = np.linalg.lstsq(Ad[:, J], e, rcond=None)[0]
        scale = np.max(np.abs(ref))
        assert np.max(np.abs(m - ref)) <= 1e-11 * scale, (name, flag)
        # Householder agrees too
        mh = O.spai_pattern(n, cp, ri, va, pcp, pri, 1)
        assert np.max(np.abs(mh - ref)) <= 1e-11 * scale, (name, flag)


def test_row_sets_match_brute_force_golden():
    for name, n, cp, ri, va, _, n1_gold in gold_mats():
        n1, n2, t, e = O.column_sets(n, cp, ri)
        assert np.array_equal(n1, n1_gold), name
        assert np.array_equal(n2, np.diff(cp)), name
        # T_k = sum of the column counts of J_k
        cnt = np.diff(cp)
        t_ref = np.array([cnt[ri[cp[k]:cp[k + 1]]].sum() for k in range(n)])
        assert np.array_equal(t, t_ref), name
        # E_k >= 0 exactly when row k occurs in some column of J_k
        A = sp.csc_matrix((va, ri, cp), shape=(n, n)).toarray()
        for k in range(n):
            J = ri[cp[k]:cp[k + 1]]
            assert (e[k] >= 0) == bool(np.any(A[k, J] != 0)), (name, k)


def test_spai_normal_equations():
    import dcu_amd as D

    P = D.Problem.generate(1, 5)
    A = P.system(0)
    m = O.spai(A.n, A.colptr, A.rowidx, A.val, 0)
    Ad = sp.csc_matrix((A.val, A.rowidx, A.colptr), shape=(A.n, A.n))
    for k in range(0, A.n, 61):
        J = A.rowidx[A.colptr[k]:A.colptr[k + 1]]
        sub = Ad[:, J].toarray()
        e = np.zeros(A.n)
        e[k] = 1.0
        r = sub @ m[A.colptr[k]:A.colptr[k + 1]] - e
        assert np.max(np.abs(sub.T @ r)) <= 1e-13


def test_spai_closed_forms():
    # diagonal
    n = 50
    d = np.linspace(1, 2, n)
    m = O.spai(n, np.arange(n + 1), np.arange(n), d, 0)
    assert np.max(np.abs(m * d - 1)) <= 1e-15
    # dense pattern: every column's least-squares problem is the full solve
    rng = np.random.default_rng(1)
    Ad = rng.standard_normal((12, 12)) + 12 * np.eye(12)
    S = sp.csc_matrix(Ad)
    for qr in (0, 1):
        m = O.spai(12, S.indptr, S.indices, S.data, qr)
        M = sp.csc_matrix((m, S.indices, S.indptr), shape=(12, 12)).toarray()
        assert np.max(np.abs(M - np.linalg.inv(Ad))) <= 1e-13
    # block diagonal with dense blocks: M = A^-1 blockwise
    blocks = [rng.standard_normal((k, k)) + 3 * k * np.eye(k) for k in (3, 5, 4)]
    Bd = sp.block_diag(blocks, format="csc")
    m = O.spai(12, Bd.indptr, Bd.indices, Bd.data, 0)
    M = sp.csc_matrix((m, Bd.indices, Bd.indptr), shape=(12, 12)).toarray()
    assert np.max(np.abs(M - np.linalg.inv(Bd.toarray()))) <= 1e-13


def test_spai_column_range_untouched_outside():
    _, n, cp, ri, va, m_gold, _ = next(gold_mats())
    m = O.spai(n, cp, ri, va, 0, 10, 20)
    assert np.all(m[:cp[10]] == 0) and np.all(m[cp[20]:] == 0)
    assert np.max(np.abs(m[cp[10]:cp[20]] - m_gold[cp[10]:cp[20]])) <= 1e-12 * np.max(np.abs(m_gold))


def test_csc2csr_vs_scipy():
    for name, n, cp, ri, va, _, _ in gold_mats():
        rp, ci, cv, perm = O.csc2csr(n, cp, ri, va)
        R = sp.csc_matrix((va, ri, cp), shape=(n, n)).tocsr()
        R.sort_indices()
        assert np.array_equal(rp, R.indptr) and np.array_equal(ci, R.indices) and np.array_equal(cv, R.data)
        assert np.array_equal(cv[perm], va)


def test_oracle_vs_reference_host_helpers():
    R = O.ref_lib()
    if R is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    import dcu_amd as D

    cases = [(n, cp, ri, va) for _, n, cp, ri, va, _, _ in gold_mats()]
    for cfg in (1,):
        A = D.Problem.generate(cfg, 9).system(0)
        cases.append((A.n, A.colptr, A.rowidx, A.val))
    for n, cp, ri, va in cases:
        cp, ri, va = (np.ascontiguousarray(cp, dtype=np.int32), np.ascontiguousarray(ri, dtype=np.int32),
                      np.ascontiguousarray(va, dtype=np.float64))
        n1, n2, _, _ = O.column_sets(n, cp, ri)
        assert R.ref_compute_n1max(n, cp.ctypes.data, ri.ctypes.data, va.ctypes.data) == n1.max()
        assert R.ref_compute_n2max(n, cp.ctypes.data) == n2.max()
        rp = np.empty(n + 1, dtype=np.int32)
        ci = np.empty(len(ri), dtype=np.int32)
        cv = np.empty(len(ri))
        R.ref_csc2csr(n, cp.ctypes.data, ri.ctypes.data, va.ctypes.data, rp.ctypes.data, ci.ctypes.data,
                      cv.ctypes.data)
        orp, oci, ocv, _ = O.csc2csr(n, cp, ri, va)
        assert np.array_equal(rp, orp) and np.array_equal(ci, oci) and np.array_equal(cv, ocv)


def _golden_systems():
    z = np.load(os.path.join(GOLD, "bicgstab_scipy.npz"))
    for name, n, cp, ri, va, m, _ in gold_mats():
        if f"{name}_b" not in z:
            continue
        yield name, n, cp, ri, va, m, z[f"{name}_b"], z[f"{name}_x"], z[f"{name}_x_noprec"]


def test_bicgstab_iterates_match_scipy_golden():
    for name, n, cp, ri, va, m, b, xs, xs_np in _golden_systems():
        A = O.csc2csr(n, cp, ri, va)[:3]
        M = O.csc2csr(n, cp, ri, m)[:3]
        for k in range(1, 7):
            # max_iter = k + 1 runs exactly k iterations (iter_times starts at 1)
            x, it, info = O.bicgstab(A, M, b, np.zeros(n), 0.0, k + 1)
            assert it == k + 1 and info[2] == 0
            assert np.linalg.norm(x - xs[k - 1]) <= 1e-12 * np.linalg.norm(xs[k - 1]), (name, k)
            # unpreconditioned iterations on rand90 pass close to a breakdown
            # (||r|| jumps by 20x at k = 6) and amplify rounding to ~3e-12
            x, it, info = O.bicgstab(A, None, b, np.zeros(n), 0.0, k + 1)
            assert np.linalg.norm(x - xs_np[k - 1]) <= 1e-10 * np.linalg.norm(xs_np[k - 1]), (name, k)


def test_bicgstab_rho_by_linearity_matches_scipy_golden():
    # rho' = (r*, s) - omega (r*, t) (the GPU's form) is (r*, r) by linearity:
    # the same scipy iterates, to rounding
    try:
        O.set_rho_mode(1)
        for name, n, cp, ri, va, m, b, xs, _ in _golden_systems():
            A = O.csc2csr(n, cp, ri, va)[:3]
            M = O.csc2csr(n, cp, ri, m)[:3]
            for k in range(1, 7):
                x, it, info = O.bicgstab(A, M, b, np.zeros(n), 0.0, k + 1)
                assert np.linalg.norm(x - xs[k - 1]) <= 1e-12 * np.linalg.norm(xs[k - 1]), (name, k)
        # a planted slip in that form (+ omega instead of - omega) is caught
        name, n, cp, ri, va, m, b, xs, _ = next(_golden_systems())
        A = sp.csc_matrix((va, ri, cp), shape=(n, n)).tocsr()
        M = sp.csc_matrix((m, ri, cp), shape=(n, n)).tocsr()
        r = b.copy()
        rs, p = r.copy(), r.copy()
        y = M @ p
        ay = A @ y
        alpha = (rs @ r) / (rs @ ay)
        s = r - alpha * ay
        z = M @ s
        t = A @ z
        omega = (t @ s) / (t @ t)
        x = alpha * y + omega * z
        r = s - omega * t
        rho_wrong = rs @ s + omega * (rs @ t)
        beta = (rho_wrong / (rs @ rs)) * (alpha / omega)
        p = r + beta * (p - omega * ay)
        y = M @ p
        ay = A @ y
        alpha = rho_wrong / (rs @ ay)
        s = r - alpha * ay
        z = M @ s
        t = A @ z
        omega = (t @ s) / (t @ t)
        x = x + alpha * y + omega * z
        assert np.linalg.norm(x - xs[1]) > 1e-6 * np.linalg.norm(xs[1])
    finally:
        O.set_rho_mode(0)


def test_true_residual_matches_numpy():
    # oracle_true_residual judges every "true residual < 1e-8" GPU assertion:
    # it must equal numpy's ||b - A x||_2 on the golden systems (x = 0, the
    # scipy iterates, random x), and see a planted error in x
    rng = np.random.default_rng(11)
    for name, n, cp, ri, va, m, b, xs, _ in _golden_systems():
        A = sp.csc_matrix((va, ri, cp), shape=(n, n))
        Acsr = O.csc2csr(n, cp, ri, va)[:3]
        for x in (np.zeros(n), xs[0], xs[-1], rng.standard_normal(n)):
            ref = np.linalg.norm(b - A @ x)
            got = O.true_residual(Acsr, b, x)
            assert abs(got - ref) <= 1e-13 * max(ref, np.linalg.norm(b)), (name, got, ref)
        assert O.true_residual(Acsr, b, np.zeros(n)) == pytest.approx(np.linalg.norm(b), rel=1e-15)
        # the exact solution has a residual at rounding level; one entry off
        # by d moves it to |d| ||A(:, i)||
        xe = spl.spsolve(A.tocsc(), b)
        assert O.true_residual(Acsr, b, xe) <= 1e-13 * np.linalg.norm(b)
        i = int(rng.integers(n))
        xb = xe.copy()
        xb[i] += 1e-3
        col = np.linalg.norm(A[:, i].toarray())
        assert O.true_residual(Acsr, b, xb) == pytest.approx(1e-3 * col, rel=1e-6), name


def test_residual_history_matches_scipy_iterates():
    # ||r_k|| after iteration k (the recursive residual the stop test reads)
    # equals ||b - A x_k|| of scipy's k-th iterate to rounding; the history
    # has one entry per executed iteration
    for name, n, cp, ri, va, m, b, xs, _ in _golden_systems():
        A = sp.csc_matrix((va, ri, cp), shape=(n, n))
        Acsr = O.csc2csr(n, cp, ri, va)[:3]
        Mcsr = O.csc2csr(n, cp, ri, m)[:3]
        x, it, info, hist = O.bicgstab(Acsr, Mcsr, b, np.zeros(n), 0.0, 7, history=True)
        assert it == 7 and len(hist) == 6
        for k in range(1, 7):
            ref = np.linalg.norm(b - A @ xs[k - 1])
            assert abs(hist[k - 1] - ref) <= 1e-10 * np.linalg.norm(b) + 1e-8 * ref, (name, k, hist[k - 1], ref)
        assert hist[-1] == info[0]
        # on convergence the last entry is the one below the threshold
        tol = 1e-6 * np.linalg.norm(b)
        x, it, info, hist = O.bicgstab(Acsr, Mcsr, b, np.zeros(n), tol, 1000, history=True)
        assert info[2] == 1 and len(hist) == it and hist[-1] < tol and np.all(hist[:-1] >= tol)


def test_bicgstab_exact_inverse_one_iteration():
    rng = np.random.default_rng(2)
    Ad = rng.standard_normal((20, 20)) + 20 * np.eye(20)
    A = sp.csr_matrix(Ad)
    Minv = sp.csr_matrix(np.linalg.inv(Ad))
    b = rng.standard_normal(20)
    x, it, info = O.bicgstab((A.indptr, A.indices, A.data), (Minv.indptr, Minv.indices, Minv.data), b,
                             np.zeros(20), 1e-10, 100)
    assert it == 1 and info[2] == 1
    assert np.linalg.norm(Ad @ x - b) <= 1e-12 * np.linalg.norm(b)


def test_bicgstab_two_by_two_by_hand():
    # A = [[2,1],[0,3]], b = [3,3], x0 = 0, M = I.  By hand: r0 = (3,3),
    # rho = 18, A r0 = (9,9), alpha = 18/54 = 1/3, s = (0,0), z = 0, t = 0,
    # omega = 0/0 -> the reference has no guard; use x0 with r0 not an
    # eigenvector instead: b = [1, 2] gives alpha = 5/((1,2).(4,6)) = 5/16.
    A = sp.csr_matrix(np.array([[2.0, 1.0], [0.0, 3.0]]))
    b = np.array([1.0, 2.0])
    x1, it, info = O.bicgstab((A.indptr, A.indices, A.data), None, b, np.zeros(2), 0.0, 2)
    alpha = 5.0 / 16.0
    s = b - alpha * np.array([4.0, 6.0])
    t = A.toarray() @ s
    omega = (t @ s) / (t @ t)
    x_hand = alpha * b + omega * s
    assert np.allclose(x1, x_hand, rtol=1e-15, atol=1e-15)
    # after one iteration r = (1/16, 1/16) is an eigenvector of A, so the
    # second iteration has s = 0 and omega = 0/0: the reference has no
    # breakdown guard (bicgstab_solver.h:196-199) and neither does the oracle
    x2, it, info = O.bicgstab((A.indptr, A.indices, A.data), None, b, np.zeros(2), 0.0, 3)
    assert np.allclose(A.toarray() @ x1 - b, [1.0 / 16, 1.0 / 16], rtol=0, atol=1e-15)
    assert np.all(np.isnan(x2))


def test_shadow32_two_by_two_by_hand():
    # shadow32: r* = fl32(r0).  One iteration by hand (van der Vorst's
    # BiCGSTAB with a shadow vector r~ other than r0): rho0 = (r~, r0),
    # alpha = rho0 / (r~, A r0), s = r0 - alpha A r0, omega = (As, s)/(As, As),
    # x1 = alpha r0 + omega s.  b = (0.1, 0.3) round to FP32 with different
    # relative errors (1.5e-8, 4.0e-8), so r~ is not a multiple of r0 and
    # alpha moves by ~1e-8 if rho0 or (r~, A r0) used r0 instead of r~.
    A = sp.csr_matrix(np.array([[2.0, 1.0], [0.0, 3.0]]))
    Ad = A.toarray()
    b = np.array([0.1, 0.3])
    rt = b.astype(np.float32).astype(np.float64)
    v = Ad @ b
    alpha = (rt @ b) / (rt @ v)
    s = b - alpha * v
    t = Ad @ s
    omega = (t @ s) / (t @ t)
    x_hand = alpha * b + omega * s
    x1, it, info = O.bicgstab((A.indptr, A.indices, A.data), None, b, np.zeros(2), 0.0, 2, shadow32=True)
    assert np.allclose(x1, x_hand, rtol=1e-15, atol=0)
    # the FP64 shadow differs in alpha at the 1e-8 level
    alpha64 = (b @ b) / (b @ v)
    assert abs(alpha - alpha64) > 1e-10 * abs(alpha64)
    x64, _, _ = O.bicgstab((A.indptr, A.indices, A.data), None, b, np.zeros(2), 0.0, 2)
    assert not np.allclose(x1, x64, rtol=1e-12, atol=0)


def test_shadow32_rounds_only_the_shadow_residual():
    # With r0 = b exactly representable in FP32 (x0 = 0, b of small binary
    # fractions), fl32(r0) = r0 and the shadow32 solve is bit-identical to the
    # FP64 one; were p, y or z rounded too, the iterates would differ.  With
    # a general b both converge, the counts within 1 on these regularly
    # converging systems, and the recursive residual still tracks b - A x.
    # One thread: OpenMP combines the threads' partial sums in no fixed order.
    before = O.num_threads()
    O.set_threads(1)
    try:
        _shadow32_checks()
    finally:
        O.set_threads(before)


def _shadow32_checks():
    for name, n, cp, ri, va, m, b, xs, _ in _golden_systems():
        A = O.csc2csr(n, cp, ri, va)[:3]
        M = O.csc2csr(n, cp, ri, m)[:3]
        bq = np.round(b * 64.0) / 64.0
        assert np.array_equal(bq.astype(np.float32).astype(np.float64), bq)
        xa, ita, _ = O.bicgstab(A, M, bq, np.zeros(n), 0.0, 7)
        xb, itb, _ = O.bicgstab(A, M, bq, np.zeros(n), 0.0, 7, shadow32=True)
        assert ita == itb and np.array_equal(xa, xb), name
        bg = b + 1e-3 * np.sin(np.arange(n))  # entries not representable in FP32
        assert not np.array_equal(bg.astype(np.float32).astype(np.float64), bg)
        x5a, _, _ = O.bicgstab(A, M, bg, np.zeros(n), 0.0, 6)
        x5b, _, _ = O.bicgstab(A, M, bg, np.zeros(n), 0.0, 6, shadow32=True)
        d = np.linalg.norm(x5a - x5b) / np.linalg.norm(x5a)
        assert 0.0 < d < 1e-5, (name, d)
        tol = 1e-10 * np.linalg.norm(bg)
        xa, ita, ia = O.bicgstab(A, M, bg, np.zeros(n), tol, 1000)
        xb, itb, ib = O.bicgstab(A, M, bg, np.zeros(n), tol, 1000, shadow32=True)
        assert ia[2] == 1 and ib[2] == 1 and abs(ita - itb) <= 1, (name, ita, itb)
        assert abs(O.true_residual(A, bg, xb) - ib[0]) <= 1e-12 * np.linalg.norm(bg), name


def test_bicgstab_max_iter_semantics():
    # bicgstab_solver.h: iter_times starts at 1 and the loop runs while
    # iter_times < MAX_ITER, so MAX_ITER - 1 iterations and a return of MAX_ITER
    name, n, cp, ri, va, m, b, xs, _ = next(_golden_systems())
    A = O.csc2csr(n, cp, ri, va)[:3]
    _, _, _, _, _, _, _, _, xs_np = next(_golden_systems())
    x, it, info = O.bicgstab(A, None, b, np.zeros(n), 0.0, 4)
    assert it == 4 and info[2] == 0
    # MAX_ITER = 4 executes 3 iterations: scipy's 3-iteration iterate
    assert np.linalg.norm(x - xs_np[2]) <= 1e-12 * np.linalg.norm(xs_np[2])


def test_planted_mistake_is_caught():
    # Swapping the roles of (t,s) and (t,t) in omega, a plausible slip, must
    # break the scipy iterate comparison; done on the CPU with numpy.
    name, n, cp, ri, va, m, b, xs, _ = next(_golden_systems())
    A = sp.csc_matrix((va, ri, cp), shape=(n, n)).tocsr()
    M = sp.csc_matrix((m, ri, cp), shape=(n, n)).tocsr()
    r = b.copy()
    rs = r.copy()
    p = r.copy()
    y = M @ p
    ay = A @ y
    alpha = (rs @ r) / (rs @ ay)
    s = r - alpha * ay
    z = M @ s
    t = A @ z
    omega_wrong = (t @ t) / (t @ s)
    x_wrong = alpha * y + omega_wrong * z
    assert np.linalg.norm(x_wrong - xs[0]) > 1e-6 * np.linalg.norm(xs[0])


def test_adjoint_oracle_finite_differences():
    # small 2D conv-diff: dL/db and dL/dA against central differences
    name, n, cp, ri, va, m, b, _, _ = list(_golden_systems())[2]
    rng = np.random.default_rng(4)
    xt = rng.standard_normal(n)
    A = O.csc2csr(n, cp, ri, va)[:3]
    x, it, info = O.bicgstab(A, None, b, np.zeros(n), 1e-14 * np.linalg.norm(b), 1000)
    dA, db, it2, info2 = O.adjoint(n, cp, ri, va, m, x, xt, 1e-14, 1000)

    def loss(vals, bb):
        Ad = sp.csc_matrix((vals, ri, cp), shape=(n, n)).toarray()
        xx = np.linalg.solve(Ad, bb)
        return 0.5 * np.sum((xx - xt) ** 2)

    h = 1e-6
    for p in rng.choice(len(va), 10, replace=False):
        vp, vm = va.copy(), va.copy()
        vp[p] += h
        vm[p] -= h
        fd = (loss(vp, b) - loss(vm, b)) / (2 * h)
        assert abs(fd - dA[p]) <= 1e-6 * max(1.0, abs(fd)), (p, fd, dA[p])
    for i in rng.choice(n, 5, replace=False):
        bp, bm = b.copy(), b.copy()
        bp[i] += h
        bm[i] -= h
        fd = (loss(va, bp) - loss(va, bm)) / (2 * h)
        assert abs(fd - db[i]) <= 1e-6 * max(1.0, abs(fd))
