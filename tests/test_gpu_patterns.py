"""Alternative sparsity patterns of M (the reference's GetSparsityPattern,
include/common/cuFormatConversion.h:598-974: 0 = A^T, 1 = E + A,
2 = E + A + A^T), next to pattern(A).

The GPU's pattern must equal the one scipy builds (tests/patterns.py), M
(FP64) must be within 1e-10 of the oracle's Householder M on that pattern
(per column, relative to max|M|; the oracle's pattern mode is pinned against
numpy lstsq in test_oracle.py) and within max(1e-10, 2x the oracle's own
MGS-vs-Householder gap) of its MGS M, the reference's QR: A^T and
E + A + A^T turn the power-law rows into columns with |J_k| up to ~1000,
whose least squares are ill-conditioned enough that MGS and Householder
differ by up to 6e-10 in the oracle itself.  The mixed build must be within
4*2^-23, and the solver must converge with M on each pattern, on both
solver layouts (SELL-64 for the stencil and random rows, row tiles for
power-law rows).
"""
import numpy as np
import pytest
import scipy.sparse as sp

import dcu_amd as D
from oracle import binding as O
from parity import count_ok, oracle_counts
from patterns import pattern_csc

pytestmark = pytest.mark.gpu
FLAGS = [D.PATTERN_AT, D.PATTERN_I_PLUS_A, D.PATTERN_I_PLUS_A_PLUS_AT]


def col_err(pcp, m_gpu, m_ref):
    starts = pcp[:-1][np.diff(pcp) > 0]
    mx = np.maximum.reduceat(np.abs(m_ref), starts)
    dx = np.maximum.reduceat(np.abs(m_gpu - m_ref), starts)
    ok = mx > 0
    return float(np.max(dx[ok] / mx[ok])) if ok.any() else 0.0


def check_pattern(ctx, mats, flag, label):
    S = D.System(ctx, mats if len(mats) > 1 else mats[0])
    S.set_pattern(flag)
    cp, ri = S.m_pattern()
    # the batch is block-diagonal: patterns per system, then offsets
    off_r, off_z = 0, 0
    refs = []
    for A in mats:
        pcp, pri = pattern_csc(A.n, A.colptr, A.rowidx, flag)
        assert np.array_equal(cp[off_r:off_r + A.n + 1] - off_z, pcp), (label, flag)
        assert np.array_equal(ri[off_z:off_z + len(pri)] - off_r, pri), (label, flag)
        mgs = O.spai_pattern(A.n, A.colptr, A.rowidx, A.val, pcp, pri, 0)
        hh = O.spai_pattern(A.n, A.colptr, A.rowidx, A.val, pcp, pri, 1)
        refs.append((pcp, pri, mgs, hh, col_err(pcp, mgs, hh)))
        off_r += A.n
        off_z += len(pri)
    for prec, sched, tol in ((D.PRECISION_FP64, D.SCHEDULE_BALANCED_NNZ, 1e-10),
                             (D.PRECISION_FP64, D.SCHEDULE_BALANCED_COLUMNS, 1e-10),
                             (D.PRECISION_MIXED, D.SCHEDULE_BALANCED_NNZ, 4 * 2.0 ** -23)):
        rep = S.build(prec, sched)
        m = S.m_values()
        off, worst_h, worst_m, gap = 0, 0.0, 0.0, 0.0
        for pcp, pri, mgs, hh, g in refs:
            worst_h = max(worst_h, col_err(pcp, m[off:off + len(pri)], hh))
            worst_m = max(worst_m, col_err(pcp, m[off:off + len(pri)], mgs))
            gap = max(gap, g)
            off += len(pri)
        print(f"{label} pattern {flag} prec {prec} sched {sched}: layout {S.layout()[0]} tiers {rep['tier_columns']} "
              f"M err vs Householder {worst_h:.2e}, vs MGS {worst_m:.2e} (oracle MGS-Householder gap {gap:.2e})",
              flush=True)
        assert off == len(m)
        assert worst_h <= tol, (label, flag, prec, sched, worst_h)
        assert worst_m <= max(tol, 2 * gap), (label, flag, prec, sched, worst_m, gap)
    return S, [(pcp, pri, mgs) for pcp, pri, mgs, _, _ in refs]


@pytest.mark.parametrize("flag", FLAGS)
def test_pattern_random4096(ctx, flag):
    P = D.Problem.generate(1, 1)
    A = P.system(0)
    b = P.rhs(0)
    S, refs = check_pattern(ctx, [A], flag, "random4096")
    S.build(D.PRECISION_FP64, D.SCHEDULE_BALANCED_NNZ)
    x, rep, oc = S.solve(b, rel_tol=1e-8, max_iter=1000)
    pcp, pri, mr = refs[0]
    Acsr = O.csc2csr(A.n, A.colptr, A.rowidx, A.val)[:3]
    Mcsr = O.csc2csr(A.n, pcp, pri, mr)[:3]
    counts = oracle_counts(Acsr, Mcsr, b, 1e-8 * np.linalg.norm(b), 1000, threads=(1, 4, 16))
    true = O.true_residual(Acsr, b, x) / np.linalg.norm(b)
    print(f"random4096 pattern {flag}: gpu {oc[0]['iterations']} oracle {sorted(counts)} true rel {true:.2e}",
          flush=True)
    assert oc[0]["status"] == D.SOLVE_CONVERGED
    assert true < 1.5e-8
    assert count_ok(oc[0]["iterations"], counts)


@pytest.mark.parametrize("flag", FLAGS)
def test_pattern_power_law_tiles(ctx, flag):
    P = D.Problem.power_law(20000, 5)
    S, _ = check_pattern(ctx, [P.system(0)], flag, "powerlaw20k")
    assert S.layout()[0] == "row_tiles"


@pytest.mark.parametrize("flag", FLAGS)
def test_pattern_batched_solve(ctx, flag):
    P = D.Problem.batched(4, 9)
    mats = P.systems()
    S, refs = check_pattern(ctx, mats, flag, "batched4")
    assert S.layout()[0].startswith("sell64")
    b = np.concatenate([P.rhs(s) for s in range(len(mats))])
    S.build(D.PRECISION_MIXED, D.SCHEDULE_BALANCED_NNZ)
    x, rep, oc = S.solve(b, rel_tol=1e-8, use_graph=True, true_residual_restart=True)
    off = S.row_offsets
    for s, A in enumerate(mats):
        bs = b[off[s]:off[s + 1]]
        Acsr = O.csc2csr(A.n, A.colptr, A.rowidx, A.val)[:3]
        true = O.true_residual(Acsr, bs, x[off[s]:off[s + 1]]) / np.linalg.norm(bs)
        assert oc[s]["status"] == D.SOLVE_CONVERGED, (s, oc[s])
        assert true < 1e-8, (s, true)


def missing_diagonal_matrix(n=400, seed=3):
    """Column k holds 3 random rows, row k + 1 and, for 9 columns in 10, the
    diagonal: every row is non-empty, and on pattern(A) some columns have
    k outside I_k (E_k = -1), hence a zero column of M (Sol_SSPAIv10)."""
    rng = np.random.default_rng(seed)
    rows, cols, vals = [], [], []
    for k in range(n):
        rs = set(rng.choice(n, 3, replace=False).tolist())
        rs.discard(k)
        if k % 10:
            rs.add(k)
        rs.add((k + 1) % n)
        for r in sorted(rs):
            rows.append(r)
            cols.append(k)
            vals.append(4.0 if r == k else rng.uniform(-1, 1))
    A = sp.csc_matrix((vals, (rows, cols)), shape=(n, n))
    A.sum_duplicates()
    A.sort_indices()
    return D.CSC(n, A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float64))


def zero_columns(cp, m):
    return int(np.sum(np.maximum.reduceat(np.abs(m), cp[:-1]) == 0))


def test_pattern_fills_missing_diagonal(ctx):
    # E + A + A^T puts row k's neighbours into J_k, so k is in I_k whenever
    # row k is non-empty: no zero column, where pattern(A) has several
    A = missing_diagonal_matrix()
    n = A.n
    zero_a_ref = zero_columns(A.colptr, O.spai(n, A.colptr, A.rowidx, A.val, 0))
    S = D.System(ctx, A)
    S.build(D.PRECISION_FP64, D.SCHEDULE_BALANCED_NNZ)
    m_a = S.m_values()
    zero_a = zero_columns(A.colptr, m_a)
    S.set_pattern(D.PATTERN_I_PLUS_A_PLUS_AT)
    S.build(D.PRECISION_FP64, D.SCHEDULE_BALANCED_NNZ)
    cp, ri = S.m_pattern()
    m_p = S.m_values()
    zero_p = zero_columns(cp, m_p)
    ref = O.spai_pattern(n, A.colptr, A.rowidx, A.val, cp, ri, 0)
    print(f"missing diagonal: zero columns of M on pattern(A) {zero_a} (oracle {zero_a_ref}), on E+A+A^T "
          f"{zero_p}; M err {col_err(cp, m_p, ref):.2e}", flush=True)
    assert zero_a == zero_a_ref and zero_a > 0
    assert zero_p == 0
    assert col_err(cp, m_p, ref) <= 1e-10
    # back to pattern(A): the same M as before, bit for bit
    S.set_pattern(D.PATTERN_A)
    S.build(D.PRECISION_FP64, D.SCHEDULE_BALANCED_NNZ)
    assert np.array_equal(S.m_values(), m_a)


def test_pattern_adjoint_matches_pattern_a(ctx):
    # the adjoint solve's result A^-T g does not depend on the preconditioner
    # beyond the solve tolerance: M^T on E + A + A^T agrees with M^T on A
    P = D.Problem.convection_diffusion(40, 5.0)
    A = P.system(0)
    rng = np.random.default_rng(1)
    b = rng.uniform(size=A.n)
    S = D.System(ctx, A)
    S.build(D.PRECISION_FP64, D.SCHEDULE_BALANCED_NNZ)
    x, _, _ = S.solve(b, rel_tol=1e-12, max_iter=2000)
    g = rng.standard_normal(A.n)
    dA0, db0, _ = S.adjoint(x, g, rel_tol=1e-12, max_iter=2000)
    S.set_pattern(D.PATTERN_I_PLUS_A_PLUS_AT)
    S.build(D.PRECISION_FP64, D.SCHEDULE_BALANCED_NNZ)
    dA1, db1, _ = S.adjoint(x, g, rel_tol=1e-12, max_iter=2000)
    assert np.linalg.norm(db1 - db0) <= 1e-9 * np.linalg.norm(db0)
    assert np.linalg.norm(dA1 - dA0) <= 1e-9 * np.linalg.norm(dA0)
