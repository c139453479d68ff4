"""GPU parity of the fused BiCGSTAB (bicgstab_solver.h:106-232 recurrences)
against the CPU oracle.

Bar: iteration counts within +-1 of the oracle run with the oracle's own M,
true relative residual ||b - A x|| / ||b|| <= 1e-8 (times a 1.5 slack for the
difference between the recursive and the true residual), solutions within
1e-6 relative of the oracle's (both are 1e-8-residual solutions).
"""
import numpy as np
import scipy.sparse as sp
import pytest

import dcu_amd as D
from oracle import binding as O
from parity import count_ok, oracle_counts

pytestmark = pytest.mark.gpu


def oracle_solve(A, m_vals, b, rel_tol=1e-8, max_iter=1000, x0=None):
    rp, ci, cv, _ = O.csc2csr(A.n, A.colptr, A.rowidx, A.val)
    M = None
    if m_vals is not None:
        mrp, mci, mcv, _ = O.csc2csr(A.n, A.colptr, A.rowidx, m_vals)
        M = (mrp, mci, mcv)
    x0 = np.zeros(A.n) if x0 is None else x0
    x, it, info = O.bicgstab((rp, ci, cv), M, b, x0, rel_tol * np.linalg.norm(b), max_iter)
    return x, it, info, (rp, ci, cv)


@pytest.mark.parametrize("use_graph", [False, True])
def test_random4096_solve(ctx, use_graph):
    P = D.Problem.generate(1, 1)
    A = P.system(0)
    b = P.rhs(0)
    S = D.System(ctx, A)
    S.build(D.PRECISION_FP64, D.SCHEDULE_BALANCED_COLUMNS)
    x, rep, oc = S.solve(b, rel_tol=1e-8, max_iter=1000, use_graph=use_graph)
    m_ref = O.spai(A.n, A.colptr, A.rowidx, A.val, 0)
    xr, it, info, Acsr = oracle_solve(A, m_ref, b)
    true_rel = O.true_residual(Acsr, b, x) / np.linalg.norm(b)
    print(f"gpu iters {oc[0]['iterations']} oracle {it} true_rel {true_rel:.3e} ms {rep['ms']:.3f}", flush=True)
    assert oc[0]["status"] == D.SOLVE_CONVERGED
    assert abs(oc[0]["iterations"] - it) <= 1
    assert true_rel <= 1.5e-8
    assert abs(oc[0]["true_residual"] / np.linalg.norm(b) - true_rel) <= 1e-10
    assert np.linalg.norm(x - xr) <= 1e-6 * np.linalg.norm(xr)


def test_graph_and_loop_agree_bitwise(ctx):
    P = D.Problem.generate(1, 2)
    A = P.system(0)
    b = P.rhs(0)
    S = D.System(ctx, A)
    S.build(D.PRECISION_FP64, D.SCHEDULE_BALANCED_NNZ)
    x1, _, o1 = S.solve(b, use_graph=False)
    x2, _, o2 = S.solve(b, use_graph=True, graph_iters=4)
    x3, _, o3 = S.solve(b, use_graph=True, graph_iters=4)  # graph replayed from the cache
    assert o1[0]["iterations"] == o2[0]["iterations"] == o3[0]["iterations"]
    assert np.array_equal(x1, x2) and np.array_equal(x2, x3)


@pytest.mark.parametrize("nsys", [1, 3])
def test_planned_final_graphs(ctx, nsys):
    # A graph-replayed solve sizes its last graphs from the iterations the
    # previous solve on the same operands and stop rule launched.  The plan
    # changes launch counts only: right-hand sides that need fewer or more
    # iterations than the previous one (a wrong plan) give the same x and
    # counts, bit for bit, as a System that has solved nothing before.
    mats = [D.Problem.convection_diffusion(48 + 8 * s, 5.0 + 3.0 * s).system(0) for s in range(nsys)]
    n = sum(A.n for A in mats)
    rng = np.random.default_rng(3)
    rhs = [np.ones(n), rng.standard_normal(n), rng.uniform(0.0, 1.0, n), np.ones(n)]
    # references first: a context's systems share one solver, whose plan is
    # dropped whenever another system's operands come in, so each new System
    # solves unplanned
    refs = []
    for b in rhs:
        F = D.System(ctx, mats if nsys > 1 else mats[0])
        F.build(D.PRECISION_FP64, D.SCHEDULE_BALANCED_NNZ)
        xf, _, of = F.solve(b, rel_tol=1e-8, max_iter=1000, use_graph=True)
        F.close()
        refs.append((xf, of))
    # then one System solves them in a row: each solve after the first is
    # planned from the one before
    S = D.System(ctx, mats if nsys > 1 else mats[0])
    S.build(D.PRECISION_FP64, D.SCHEDULE_BALANCED_NNZ)
    counts = []
    for b, (xf, of) in zip(rhs, refs):
        x, _, oc = S.solve(b, rel_tol=1e-8, max_iter=1000, use_graph=True)
        counts.append([o["iterations"] for o in oc])
        assert [o["iterations"] for o in oc] == [o["iterations"] for o in of]
        assert [o["status"] for o in oc] == [o["status"] for o in of]
        assert np.array_equal(x, xf)
    print(f"nsys={nsys}: iterations per right-hand side {counts}", flush=True)
    S.close()


def test_unpreconditioned_matches_oracle(ctx):
    # Unpreconditioned BiCGSTAB on this conv-diff converges erratically: the
    # oracle alone needs 101..105 iterations depending on its OpenMP reduction
    # order (1, 2, 8 threads).  Parity is therefore checked on the iterate
    # after a fixed 20 iterations, and the full solve within 10% of the count.
    P = D.Problem.convection_diffusion(40, 5.0)
    A = P.system(0)
    b = np.random.default_rng(3).uniform(size=A.n)
    S = D.System(ctx, A)
    x20, _, oc20 = S.solve(b, rel_tol=0.0, abs_tol=1e-300, max_iter=21, precondition=False)
    xr20, it20, _, Acsr = oracle_solve(A, None, b, rel_tol=0.0, max_iter=21)
    assert oc20[0]["iterations"] == it20 == 21
    assert np.linalg.norm(x20 - xr20) <= 1e-9 * np.linalg.norm(xr20)
    x, rep, oc = S.solve(b, rel_tol=1e-8, max_iter=2000, precondition=False)
    xr, it, info, Acsr = oracle_solve(A, None, b, max_iter=2000)
    assert oc[0]["status"] == D.SOLVE_CONVERGED
    assert abs(oc[0]["iterations"] - it) <= 0.1 * it
    assert O.true_residual(Acsr, b, x) / np.linalg.norm(b) <= 1.5e-8


def test_mixed_precision_solve(ctx):
    P = D.Problem.generate(1, 4)
    A = P.system(0)
    b = P.rhs(0)
    S = D.System(ctx, A)
    S.build(D.PRECISION_MIXED, D.SCHEDULE_BALANCED_NNZ)
    x, rep, oc = S.solve(b, rel_tol=1e-8)
    m_ref = O.spai(A.n, A.colptr, A.rowidx, A.val, 0)
    xr, it, info, Acsr = oracle_solve(A, m_ref, b)
    true_rel = O.true_residual(Acsr, b, x) / np.linalg.norm(b)
    print(f"mixed: gpu iters {oc[0]['iterations']} oracle(fp64 M) {it} true_rel {true_rel:.3e}", flush=True)
    assert oc[0]["status"] == D.SOLVE_CONVERGED
    assert abs(oc[0]["iterations"] - it) <= 1
    assert true_rel <= 1.5e-8


def test_batched_solve_per_system(ctx):
    # The conv-diff systems of the batch converge irregularly: the oracle's
    # own count moves by up to +-60 iterations between OpenMP thread counts
    # (a different summation order in the dot products).  Parity: the first
    # 10 iterates agree to 1e-8, and the final count passes tests/parity.py's
    # count_ok against the oracle's counts over reduction orders and both
    # forms of rho'.  Where convergence is regular every sample agrees and
    # the bar is +-1.
    P = D.Problem.batched(6, 5)
    mats = P.systems()
    S = D.System(ctx, mats)
    S.build(D.PRECISION_FP64, D.SCHEDULE_BALANCED_NNZ)
    b = np.concatenate([P.rhs(s) for s in range(len(mats))])
    off = S.row_offsets
    refs = []
    for s, A in enumerate(mats):
        bs = b[off[s]:off[s + 1]]
        m_ref = O.spai(A.n, A.colptr, A.rowidx, A.val, 0)
        Acsr = O.csc2csr(A.n, A.colptr, A.rowidx, A.val)[:3]
        Mcsr = O.csc2csr(A.n, A.colptr, A.rowidx, m_ref)[:3]
        x10 = O.bicgstab(Acsr, Mcsr, bs, np.zeros(A.n), 0.0, 11)[0]
        spread = oracle_counts(Acsr, Mcsr, bs, 1e-8 * np.linalg.norm(bs), 1000)
        refs.append((bs, Acsr, x10, spread))
    x10g, _, oc10 = S.solve(b, rel_tol=0.0, abs_tol=1e-300, max_iter=11)
    for s, (bs, Acsr, x10, spread) in enumerate(refs):
        assert oc10[s]["iterations"] == 11
        xs = x10g[off[s]:off[s + 1]]
        assert np.linalg.norm(xs - x10) <= 1e-8 * np.linalg.norm(x10), s
    for use_graph in (False, True):
        x, rep, oc = S.solve(b, rel_tol=1e-8, use_graph=use_graph)
        for s, (bs, Acsr, x10, spread) in enumerate(refs):
            xs = x[off[s]:off[s + 1]]
            true_rel = O.true_residual(Acsr, bs, xs) / np.linalg.norm(bs)
            print(f"system {s}: gpu {oc[s]['iterations']} oracle counts {sorted(spread)}", flush=True)
            assert oc[s]["status"] == D.SOLVE_CONVERGED, (s, oc[s])
            assert count_ok(oc[s]["iterations"], spread), (s, oc[s]["iterations"], sorted(spread))
            assert true_rel <= 1.5e-8, (s, true_rel)


def test_true_residual_restart(ctx):
    # with the restart option every system ends with ||b - A x|| < TOL itself
    P = D.Problem.batched(8, 4)
    mats = P.systems()
    S = D.System(ctx, mats)
    S.build(D.PRECISION_MIXED, D.SCHEDULE_BALANCED_NNZ)
    b = np.concatenate([P.rhs(s) for s in range(len(mats))])
    x, rep, oc = S.solve(b, rel_tol=1e-8, use_graph=True, true_residual_restart=True)
    off = S.row_offsets
    for s, A in enumerate(mats):
        bs = b[off[s]:off[s + 1]]
        Acsr = O.csc2csr(A.n, A.colptr, A.rowidx, A.val)[:3]
        true = O.true_residual(Acsr, bs, x[off[s]:off[s + 1]])
        assert oc[s]["status"] == D.SOLVE_CONVERGED
        assert true < oc[s]["tol"] * (1 + 1e-9), (s, true, oc[s])
        # two evaluations of ||b - A x|| in different summation orders differ
        # by at most the rounding of b - A x per row, |b_i - (Ax)_i| error <=
        # (k_i + 1) u (|b_i| + (|A||x|)_i) with k_i nonzeros in row i; taken
        # twice (GPU and host), with a factor 2 of margin
        Ad = sp.csr_matrix((Acsr[2], Acsr[1], Acsr[0]), shape=(A.n, A.n))
        xs = x[off[s]:off[s + 1]]
        k = np.diff(Acsr[0]) + 1
        bound = 4 * np.linalg.norm(k * 2.0 ** -53 * (np.abs(bs) + abs(Ad) @ np.abs(xs)))
        assert abs(oc[s]["true_residual"] - true) <= bound, (s, oc[s]["true_residual"], true, bound)


def test_max_iter_semantics(ctx):
    # the reference returns MAX_ITER after MAX_ITER - 1 unconverged iterations
    P = D.Problem.convection_diffusion(60, 10.0)
    A = P.system(0)
    b = np.ones(A.n)
    S = D.System(ctx, A)
    S.build(D.PRECISION_FP64, D.SCHEDULE_BALANCED_NNZ)
    x, rep, oc = S.solve(b, rel_tol=1e-14, max_iter=5)
    m_ref = O.spai(A.n, A.colptr, A.rowidx, A.val, 0)
    xr, it, info, _ = oracle_solve(A, m_ref, b, rel_tol=1e-14, max_iter=5)
    assert it == 5 and info[2] == 0
    assert oc[0]["status"] == D.SOLVE_MAX_ITER and oc[0]["iterations"] == 5
    assert np.linalg.norm(x - xr) <= 1e-10 * np.linalg.norm(xr)


def test_converged_at_start(ctx):
    P = D.Problem.generate(1, 1)
    A = P.system(0)
    S = D.System(ctx, A)
    S.build(D.PRECISION_FP64, D.SCHEDULE_BALANCED_NNZ)
    b = P.rhs(0)  # b = A * ones, so x0 = ones is exact
    x, rep, oc = S.solve(b, x0=np.ones(A.n), rel_tol=1e-8)
    assert oc[0]["status"] == D.SOLVE_CONVERGED_AT_START and oc[0]["iterations"] == 0
    assert np.array_equal(x, np.ones(A.n))


def small_convdiff(m, pe, rng):
    """m x m upwind convection-diffusion (5-point), Dirichlet boundaries."""
    import scipy.sparse as sp
    n = m * m
    h = 1.0 / (m + 1)
    vx, vy = pe * rng.uniform(0.5, 1.0), pe * rng.uniform(0.0, 0.5)
    rows, cols, vals = [], [], []
    for j in range(m):
        for i in range(m):
            r = j * m + i
            diag = 4.0 + h * (vx + vy)
            rows.append(r); cols.append(r); vals.append(diag)
            for di, dj, w in ((-1, 0, -1.0 - h * vx), (1, 0, -1.0), (0, -1, -1.0 - h * vy), (0, 1, -1.0)):
                ii, jj = i + di, j + dj
                if 0 <= ii < m and 0 <= jj < m:
                    rows.append(r); cols.append(jj * m + ii); vals.append(w)
    A = sp.csc_matrix((vals, (rows, cols)), shape=(n, n))
    A.sort_indices()
    return D.CSC(n, A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float64))


def test_batch_of_more_than_1024_systems(ctx):
    # 1100 systems: the active list is rebuilt by the two-kernel compaction
    # (scan, then fill) that batches of more than 1024 systems use
    rng = np.random.default_rng(2)
    mats = [small_convdiff(int(rng.integers(6, 15)), float(rng.uniform(1, 30)), rng) for _ in range(1100)]
    S = D.System(ctx, mats)  # systems of 36-196 rows: row tiles (a 64-row slice would be mostly padding)
    S.build(D.PRECISION_MIXED, D.SCHEDULE_BALANCED_NNZ)
    b = rng.uniform(0, 1, sum(A.n for A in mats))
    x, rep, oc = S.solve(b, rel_tol=1e-8, use_graph=True, true_residual_restart=True)
    off = S.row_offsets
    m_gpu = S.m_values()
    moff = np.concatenate([[0], np.cumsum([A.nnz for A in mats])])
    iters = [o["iterations"] for o in oc]
    print(f"{len(mats)} systems: iterations {min(iters)}..{max(iters)}, graph replays {rep['host_checks']}", flush=True)
    for s, A in enumerate(mats):
        bs = b[off[s]:off[s + 1]]
        Acsr = O.csc2csr(A.n, A.colptr, A.rowidx, A.val)[:3]
        true_rel = O.true_residual(Acsr, bs, x[off[s]:off[s + 1]]) / np.linalg.norm(bs)
        assert oc[s]["status"] == D.SOLVE_CONVERGED, (s, oc[s])
        assert true_rel < 1e-8, (s, true_rel)
        if s % 100 == 0:  # counts on a sample, against the oracle with the GPU's M
            Mcsr = O.csc2csr(A.n, A.colptr, A.rowidx, m_gpu[moff[s]:moff[s + 1]])[:3]
            spread = oracle_counts(Acsr, Mcsr, bs, 1e-8 * np.linalg.norm(bs), 1000, threads=(1, 2), shadow32=True)
            assert count_ok(oc[s]["iterations"], spread), (s, oc[s]["iterations"], sorted(spread))


@pytest.mark.parametrize("cap", [None, 32])
def test_kernel_timers(ctx, cap):
    # the per-kernel timers run every kernel of the iteration on every unit
    # (both K1 forms), on diagonal slices and on 32-bit columns; afterwards a
    # solve still converges (the timers leave no state behind)
    import os
    old = os.environ.get("SPAI_MI355X_SELL_INDEX")
    try:
        if cap is None:
            os.environ.pop("SPAI_MI355X_SELL_INDEX", None)
        else:
            os.environ["SPAI_MI355X_SELL_INDEX"] = str(cap)
        P = D.Problem.batched(4, 3)
        mats = P.systems()
        S = D.System(ctx, mats)
    finally:
        if old is None:
            os.environ.pop("SPAI_MI355X_SELL_INDEX", None)
        else:
            os.environ["SPAI_MI355X_SELL_INDEX"] = old
    S.build(D.PRECISION_MIXED, D.SCHEDULE_BALANCED_NNZ)
    b = np.concatenate([P.rhs(s) for s in range(len(mats))])
    S.solve(b, rel_tol=1e-8, use_graph=True)  # the single-kernel timers run on the last solve's vectors
    for which in range(6):
        assert S.time_spmv(which, 3) > 0.0
    for variant in (0, 1, -1):
        ms = S.time_iteration(3, variant)
        assert ms[0] > 0 and ms[2] > 0 and ms[3] > 0 and ms[4] > 0
        if variant >= 0:
            assert (ms[1] > 0) == (variant == 1)
    x, rep, oc = S.solve(b, rel_tol=1e-8, use_graph=True)
    off = S.row_offsets
    for s, A in enumerate(mats):
        Acsr = O.csc2csr(A.n, A.colptr, A.rowidx, A.val)[:3]
        bs = b[off[s]:off[s + 1]]
        assert oc[s]["status"] == D.SOLVE_CONVERGED
        assert O.true_residual(Acsr, bs, x[off[s]:off[s + 1]]) / np.linalg.norm(bs) < 1.5e-8


def banded_system(n, reach, rng):
    """Nonsymmetric band matrix (offsets -reach..reach where they exist),
    diagonally dominant: a stencil-like system of any order, 1 included."""
    rows, cols, vals = [], [], []
    for i in range(n):
        off = rng.uniform(-1.0, 1.0, 2 * reach + 1)
        for d in range(-reach, reach + 1):
            j = i + d
            if 0 <= j < n and d != 0:
                rows.append(i)
                cols.append(j)
                vals.append(off[d + reach])
        rows.append(i)
        cols.append(i)
        vals.append(2.0 * reach + 1.0)
    A = sp.csc_matrix((vals, (rows, cols)), shape=(n, n))
    A.sort_indices()
    return D.CSC(n, A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float64))


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_of_odd_sizes(ctx, use_graph):
    # systems of 1, 2, 3, 63, 64, 65, 127 and 1031 rows in one batch: partial
    # slices, units of fewer than 8 slices, one-row systems (every diagonal
    # offset but 0 clamps back into the system) and both sweep directions;
    # each system must match its own oracle solve
    rng = np.random.default_rng(2024)
    sizes = [1, 2, 3, 63, 64, 65, 127, 1031]
    mats = [banded_system(n, 2, rng) for n in sizes]
    bs = [rng.uniform(-1.0, 1.0, n) for n in sizes]
    S = D.System(ctx, mats)
    S.build(D.PRECISION_FP64, D.SCHEDULE_BALANCED_NNZ)
    assert S.layout()[0].startswith("sell64")
    x, rep, oc = S.solve(np.concatenate(bs), rel_tol=1e-10, max_iter=500, use_graph=use_graph)
    m = S.m_values()
    off_x = off_m = 0
    for A, b, o in zip(mats, bs, oc):
        xs = x[off_x:off_x + A.n]
        ms = m[off_m:off_m + A.nnz]
        xr, it, info, Acsr = oracle_solve(A, ms, b, rel_tol=1e-10, max_iter=500)
        true = O.true_residual(Acsr, b, xs) / np.linalg.norm(b)
        print(f"n={A.n}: gpu {o['iterations']} ({o['status']}) oracle {it}, true rel {true:.2e}", flush=True)
        if A.n <= 2 * 2 + 1:
            # the band covers the whole matrix, so pattern(A) is full and M is
            # A^-1 up to rounding: s = r - alpha v is zero or at rounding level
            # after the first half step.  Exactly zero gives omega = 0/0, and
            # the reference (no test on ||s||) then runs to MAX_ITER on NaNs,
            # as the oracle does here; at rounding level the next half step
            # converges.  Which one happens is decided by rounding, so only a
            # converged result is checked.
            if o["status"] == D.SOLVE_CONVERGED:
                assert true < 1.5e-10
            else:
                assert o["status"] == D.SOLVE_BREAKDOWN and it == 500
        else:
            assert o["status"] == D.SOLVE_CONVERGED
            assert abs(o["iterations"] - it) <= 1
            assert true < 1.5e-10
            assert np.linalg.norm(xs - xr) <= 1e-8 * np.linalg.norm(xr)
        off_x += A.n
        off_m += A.nnz
    S.close()


@pytest.mark.parametrize("precision", [D.PRECISION_FP64, D.PRECISION_MIXED])
def test_column_blocks_match_plain_tiles(ctx, monkeypatch, precision):
    # One power-law system of more than 2^19 rows: its row-tile layout forms
    # the split iteration's products by column blocks (partial row sums per
    # block, added in block order by the consuming kernel).  The same system
    # created with SPAI_MI355X_COLUMN_BLOCKS=0 runs the plain row tiles.  M
    # is the same; the first 10 iterates and residuals agree with the plain
    # tiles' to 1e-12 (the row sums are added in another order) and with the
    # oracle's (run with the GPU's M) to 1e-8; a longer solve (which stagnates
    # on this matrix, as the oracle does) ends alike, at a comparable
    # residual (rounding differences grow over many unconverged iterations).
    P = D.Problem.power_law(600000, 11)
    A = P.system(0)
    b = P.rhs(0)
    S1 = D.System(ctx, A)
    monkeypatch.setenv("SPAI_MI355X_COLUMN_BLOCKS", "0")
    S2 = D.System(ctx, A)
    monkeypatch.delenv("SPAI_MI355X_COLUMN_BLOCKS")
    assert S1.layout()[0] == "row_tiles_blocked" and S2.layout()[0] == "row_tiles"
    for S in (S1, S2):
        S.set_tile_shapes([0, 0, 0])
        S.build(precision, D.SCHEDULE_BALANCED_NNZ)
    m = S1.m_values()
    assert np.array_equal(m, S2.m_values())
    out = []
    for S in (S1, S2):
        x, _, oc = S.solve(b, rel_tol=0.0, abs_tol=1e-300, max_iter=11, record_history=True, use_graph=True)
        out.append((x, S.residual_history(0), oc[0]))
    (x1, h1, o1), (x2, h2, o2) = out
    assert o1["iterations"] == o2["iterations"] == 11
    d = np.linalg.norm(x1 - x2) / np.linalg.norm(x2)
    dh = np.max(np.abs(h1 - h2) / h2)
    Acsr = O.csc2csr(A.n, A.colptr, A.rowidx, A.val)[:3]
    Mcsr = O.csc2csr(A.n, A.colptr, A.rowidx, m)[:3]
    xo, _, _, ho = O.bicgstab(Acsr, Mcsr, b, np.zeros(A.n), 0.0, 11, history=True,
                              shadow32=precision == D.PRECISION_MIXED)
    eo = np.linalg.norm(x1 - xo) / np.linalg.norm(xo)
    ho_err = np.max(np.abs(h1 - ho) / ho)
    print(f"prec {precision}: blocked vs plain 10 iterates {d:.2e}, history {dh:.2e}; vs oracle {eo:.2e}, "
          f"history {ho_err:.2e}", flush=True)
    assert d <= 1e-12 and dh <= 1e-12
    assert eo <= 1e-8 and ho_err <= 1e-8
    r1 = S1.solve(b, rel_tol=1e-8, max_iter=120, use_graph=True)[2][0]
    r2 = S2.solve(b, rel_tol=1e-8, max_iter=120, use_graph=True)[2][0]
    print(f"120 iterations: blocked {r1['iterations']} res {r1['residual']:.6e}, plain {r2['iterations']} "
          f"res {r2['residual']:.6e}", flush=True)
    assert r1["status"] == r2["status"] and r1["iterations"] == r2["iterations"]
    assert 0.5 < r1["residual"] / r2["residual"] < 2.0
    S1.close()
    S2.close()


def _long_row_system(n, long_rows, long_len, seed):
    # n x n, three random off-diagonal entries per row, and rows `long_rows`
    # with `long_len` entries each; a_ii = 1.2 * sum_j |a_ij| (row dominant)
    rng = np.random.default_rng(seed)
    rows = [np.repeat(np.arange(n), 3)]
    cols = [rng.integers(0, n, 3 * n)]
    for r in long_rows:
        rows.append(np.full(long_len, r))
        cols.append(rng.choice(n, long_len, replace=False))
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    keep = r != c
    r, c = r[keep], c[keep]
    v = rng.uniform(-1.0, 1.0, r.size)
    A = sp.csr_matrix((v, (r, c)), shape=(n, n))
    A.sum_duplicates()
    d = 1.2 * np.asarray(abs(A).sum(axis=1)).ravel() + 1.0
    A = (A + sp.diags(d)).tocsc()
    A.sort_indices()
    return D.CSC(n, A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float64))


def test_column_blocks_rows_beyond_the_staging_buffer(ctx, monkeypatch):
    # Rows with more entries than a tile stages (2048): with column blocks a
    # row of 12000 entries holds about 3000 in each quarter of the columns,
    # so its partial products take the one-row tile path of k_tiles_part;
    # the plain row tiles reduce it with the whole workgroup.  Both agree
    # with each other and with the oracle (run with the GPU's M) over 10
    # iterations.
    A = _long_row_system(600000, [12345, 400000], 12000, 5)
    b = np.random.default_rng(6).uniform(0.0, 1.0, A.n)
    S1 = D.System(ctx, A)
    monkeypatch.setenv("SPAI_MI355X_COLUMN_BLOCKS", "0")
    S2 = D.System(ctx, A)
    monkeypatch.delenv("SPAI_MI355X_COLUMN_BLOCKS")
    assert S1.layout()[0] == "row_tiles_blocked" and S2.layout()[0] == "row_tiles"
    for S in (S1, S2):
        S.build(D.PRECISION_FP64, D.SCHEDULE_BALANCED_NNZ)
    m = S1.m_values()
    assert np.array_equal(m, S2.m_values())
    out = []
    for S in (S1, S2):
        x, _, oc = S.solve(b, rel_tol=0.0, abs_tol=1e-300, max_iter=11, record_history=True, use_graph=True)
        out.append((x, S.residual_history(0), oc[0]))
    (x1, h1, o1), (x2, h2, o2) = out
    assert o1["iterations"] == o2["iterations"] == 11
    Acsr = O.csc2csr(A.n, A.colptr, A.rowidx, A.val)[:3]
    Mcsr = O.csc2csr(A.n, A.colptr, A.rowidx, m)[:3]
    xo, _, _, ho = O.bicgstab(Acsr, Mcsr, b, np.zeros(A.n), 0.0, 11, history=True)
    d = np.linalg.norm(x1 - x2) / np.linalg.norm(x2)
    eo = max(np.linalg.norm(x - xo) / np.linalg.norm(xo) for x in (x1, x2))
    ho_err = max(np.max(np.abs(h - ho) / ho) for h in (h1, h2))
    print(f"long rows: blocked vs plain 10 iterates {d:.2e}; vs oracle {eo:.2e}, history {ho_err:.2e}", flush=True)
    assert d <= 1e-12
    assert eo <= 1e-8 and ho_err <= 1e-8
    S1.close()
    S2.close()
