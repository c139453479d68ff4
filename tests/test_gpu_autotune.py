"""Tile-shape candidates of the SPAI register tiers and the autotuner.

Every candidate of every size bucket is pinned against the oracle on
columns that reach it: M within the usual bars (FP64 1e-10, mixed 4 * 2^-23,
relative per column), |I_k| bit-exact, and the columns counted in the tier
that solved them, which follows from the oracle's (T_k, |J_k|, |I_k|) and
the candidate's rows and columns per group (a column whose |I_k| exceeds the
rows, or whose |J_k| exceeds the columns of a candidate narrower than its
bucket, is solved by the next bucket).  The autotuner must return candidates that exist, time
every candidate of a populated bucket, leave M built with its choice (bit
for bit what a pinned build with that choice gives) and keep the choice for
later builds.
"""
import numpy as np
import pytest

import dcu_amd as D
from oracle import binding as O
from test_gpu_spai import column_rel_err, scattered_columns

pytestmark = pytest.mark.gpu

BUCKET_T = (32, 64, 128)
BUCKET_NJ = (8, 8, 16)
TOL = {D.PRECISION_FP64: 1e-10, D.PRECISION_MIXED: 4 * 2.0 ** -23}


def expected_tiers(n1, n2, t, shape):
    """Columns solved per tier under `shape` (balanced_nnz): the first bucket
    that holds (T_k, |J_k|), then on to the next while |I_k| exceeds the
    candidate's rows or |J_k| its columns; 3 the LDS tier, 4 the global
    tier."""
    rows = [D.tile_shape(b, shape[b])["lanes"] * D.tile_shape(b, shape[b])["rows_per_lane"] for b in range(3)]
    cols = [D.tile_shape(b, shape[b])["max_cols"] for b in range(3)]
    out = np.zeros(5, dtype=np.int64)
    for a, b_, c in zip(n1, n2, t):
        if b_ == 0:
            continue
        tier = next((b for b in range(3) if b_ <= BUCKET_NJ[b] and c <= BUCKET_T[b]), None)
        if tier is None:
            need = c * b_ + b_ * b_ + b_
            tier = 3 if (b_ <= 64 and c <= 1024 and need <= 7168) else 4
        while tier < 2 and (a > rows[tier] or b_ > cols[tier]):
            tier += 1
        out[tier] += 1
    return out.tolist()


def shapes_to_try():
    out = []
    for b in range(3):
        for c in range(D.tile_shape_count(b)):
            s = [0, 0, 0]
            s[b] = c
            if s not in out:
                out.append(s)
    return out


@pytest.fixture(scope="module")
def power_law():
    A = D.Problem.power_law(60000, 7).system(0)
    m_ref = O.spai(A.n, A.colptr, A.rowidx, A.val, 0)
    sets = O.column_sets(A.n, A.colptr, A.rowidx)
    return A, m_ref, sets


@pytest.mark.parametrize("precision", [D.PRECISION_FP64, D.PRECISION_MIXED])
def test_every_shape_matches_oracle(ctx, power_law, precision):
    A, m_ref, (n1_ref, n2_ref, t_ref, _) = power_law
    S = D.System(ctx, A)
    for shape in shapes_to_try():
        S.set_tile_shapes(shape)
        rep = S.build(precision, D.SCHEDULE_BALANCED_NNZ)
        err = column_rel_err(A, S.m_values(), m_ref)
        exp = expected_tiers(n1_ref, n2_ref, t_ref, shape)
        print(f"shape {shape} prec {precision}: tiers {rep['tier_columns']} (expected {exp}) "
              f"build {rep['ms']:.3f} ms err {err:.3e}", flush=True)
        assert rep["tile_shape"] == shape
        assert err <= TOL[precision], (shape, err)
        assert np.array_equal(S.row_set_sizes(), n1_ref)
        assert rep["tier_columns"] == exp
    S.close()


@pytest.mark.parametrize("schedule", [D.SCHEDULE_BALANCED_COLUMNS, D.SCHEDULE_BALANCED_NNZ])
@pytest.mark.parametrize("c0", [0, 1, 2, 3])
def test_overflow_follows_the_rows(ctx, schedule, c0):
    # T_k <= 32 everywhere: the smallest bucket; 16-row candidates hand the
    # columns with |I_k| > 16 to the next bucket, 32-row ones keep them all
    A = scattered_columns()
    n1_ref, n2_ref, t_ref, _ = O.column_sets(A.n, A.colptr, A.rowidx)
    m_ref = O.spai(A.n, A.colptr, A.rowidx, A.val, 0)
    S = D.System(ctx, A)
    S.set_tile_shapes([c0, 0, 0])
    rep = S.build(D.PRECISION_FP64, schedule)
    rows = D.tile_shape(0, c0)["lanes"] * D.tile_shape(0, c0)["rows_per_lane"]
    err = column_rel_err(A, S.m_values(), m_ref)
    print(f"c0 {c0} rows {rows} sched {schedule}: tiers {rep['tier_columns']} err {err:.3e}", flush=True)
    assert err <= 1e-10
    assert rep["tier_columns"][1] == int((n1_ref > rows).sum())
    assert rep["tier_columns"][0] == int((n1_ref <= rows).sum())
    S.close()


def wide_short_columns(n=3000, seed=5):
    """About a third of the columns hold the diagonal and 5-7 random rows,
    the rest the diagonal and 1-2: columns with |J_k| of 6 to 8 whose T_k
    stays within 32 or 64, which the 5- and 7-column candidates hand on."""
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    for k in range(n):
        m = int(rng.integers(5, 8)) if rng.uniform() < 0.33 else int(rng.integers(1, 3))
        rs = set(rng.choice(n, m, replace=False).tolist()) | {k}
        rows += sorted(rs)
        cols += [k] * len(rs)
    A = sp.csc_matrix((rng.uniform(-1, 1, len(rows)), (rows, cols)), shape=(n, n))
    A.setdiag(10.0)
    A.sort_indices()
    return D.CSC(n, A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float64))


@pytest.mark.parametrize("schedule", [D.SCHEDULE_BALANCED_COLUMNS, D.SCHEDULE_BALANCED_NNZ])
@pytest.mark.parametrize("shape", [[4, 4, 0], [4, 5, 0], [0, 4, 0]])
def test_narrow_shapes_hand_wide_columns_on(ctx, schedule, shape):
    # candidates narrower than their bucket (5 or 7 columns): a column with
    # more columns moves to the next bucket (twice for |J_k| = 8 under a
    # 5-column then a 7-column candidate), under both schedules
    A = wide_short_columns()
    n1_ref, n2_ref, t_ref, _ = O.column_sets(A.n, A.colptr, A.rowidx)
    assert ((t_ref <= 32) & (n2_ref > 5)).sum() > 50
    assert ((t_ref > 32) & (t_ref <= 64) & (n2_ref == 8)).sum() > 20
    m_ref = O.spai(A.n, A.colptr, A.rowidx, A.val, 0)
    S = D.System(ctx, A)
    S.set_tile_shapes(shape)
    for precision in (D.PRECISION_FP64, D.PRECISION_MIXED):
        rep = S.build(precision, schedule)
        err = column_rel_err(A, S.m_values(), m_ref)
        print(f"shape {shape} sched {schedule} prec {precision}: tiers {rep['tier_columns']} err {err:.3e}",
              flush=True)
        assert err <= TOL[precision], (shape, err)
        assert np.array_equal(S.row_set_sizes(), n1_ref)
        if schedule == D.SCHEDULE_BALANCED_NNZ:
            assert rep["tier_columns"] == expected_tiers(n1_ref, n2_ref, t_ref, shape)
    S.close()


def test_autotune_batched(ctx):
    P = D.Problem.batched(8, 3)
    mats = P.systems()
    S = D.System(ctx, mats)
    tune = S.autotune(D.PRECISION_MIXED, D.SCHEDULE_BALANCED_NNZ, reps=2)
    print(f"autotune: {tune}", flush=True)
    for b in range(3):
        assert 0 <= tune["shape"][b] < D.tile_shape_count(b)
        if tune["columns"][b] > 0:
            assert all(v is not None and v > 0 for v in tune["ms"][b])
            assert tune["ms"][b][tune["shape"][b]] == min(tune["ms"][b])
        else:
            assert all(v is None for v in tune["ms"][b])
            assert tune["shape"][b] == 0
    assert tune["columns"][0] > 0 and tune["columns"][1] > 0  # 2D 5-point and 3D 7-point columns
    assert S.tile_shapes() == tune["shape"]
    m_tuned = S.m_values()
    rep = S.build(D.PRECISION_MIXED, D.SCHEDULE_BALANCED_NNZ)
    assert rep["tile_shape"] == tune["shape"]
    assert np.array_equal(S.m_values(), m_tuned)
    b = np.concatenate([P.rhs(s) for s in range(len(mats))])
    x, rep, oc = S.solve(b, rel_tol=1e-8, use_graph=True)
    assert all(o["status"] == D.SOLVE_CONVERGED for o in oc)
    S.close()


def test_autotune_static_schedule_with_overflow(ctx):
    # balanced_columns: every column in the static tier (bucket 0 here), the
    # default 16-row shape handing the wider row sets to bucket 1; whatever
    # the tuner picks, M stays within the FP64 bar and the columns are counted
    # where the chosen rows put them
    A = scattered_columns()
    n1_ref, _, _, _ = O.column_sets(A.n, A.colptr, A.rowidx)
    m_ref = O.spai(A.n, A.colptr, A.rowidx, A.val, 0)
    S = D.System(ctx, A)
    tune = S.autotune(D.PRECISION_FP64, D.SCHEDULE_BALANCED_COLUMNS, reps=2)
    print(f"static autotune: {tune}", flush=True)
    assert tune["columns"][0] > 0 and tune["columns"][1] == int((n1_ref > 16).sum())
    assert all(v is not None for v in tune["ms"][0]) and all(v is not None for v in tune["ms"][1])
    assert all(v is None for v in tune["ms"][2])
    err = column_rel_err(A, S.m_values(), m_ref)
    assert err <= 1e-10
    rep = S.build(D.PRECISION_FP64, D.SCHEDULE_BALANCED_COLUMNS)
    rows = D.tile_shape(0, tune["shape"][0])["lanes"] * D.tile_shape(0, tune["shape"][0])["rows_per_lane"]
    assert rep["static_tier"] == 0 and rep["tile_shape"] == tune["shape"]
    assert rep["tier_columns"][1] == int((n1_ref > rows).sum())
    S.close()


@pytest.mark.parametrize("precision", [D.PRECISION_FP64, D.PRECISION_MIXED])
def test_sort_key_widths_agree(ctx, power_law, precision):
    # the register tiers sort (row, entry) keys in 32 bits when the rows fit
    # (every system here), in 64 bits otherwise; the order is the same, so M
    # must be bit-identical, for every tile shape
    import os
    A, m_ref, (n1_ref, _, _, _) = power_law
    for shape in shapes_to_try():
        out = {}
        for width in ("32", "64"):
            old = os.environ.get("SPAI_MI355X_SORT_KEYS")
            os.environ["SPAI_MI355X_SORT_KEYS"] = width
            try:
                S = D.System(ctx, A)
                S.set_tile_shapes(shape)
                rep = S.build(precision, D.SCHEDULE_BALANCED_NNZ)
                out[width] = (S.m_values(), S.row_set_sizes(), rep["tier_columns"])
                S.close()
            finally:
                if old is None:
                    os.environ.pop("SPAI_MI355X_SORT_KEYS", None)
                else:
                    os.environ["SPAI_MI355X_SORT_KEYS"] = old
        assert np.array_equal(out["32"][0], out["64"][0]), shape
        assert np.array_equal(out["32"][1], n1_ref) and np.array_equal(out["64"][1], n1_ref)
        assert out["32"][2] == out["64"][2]


def test_autotune_without_register_tier_columns(ctx):
    # every column gathers ~324 entries from 18 columns: all in the LDS tier,
    # so no bucket is timed and the shapes stay as they were
    import scipy.sparse as sp
    n, per = 300, 18
    rng = np.random.default_rng(11)
    rows, cols = [], []
    for k in range(n):
        rs = set(rng.choice(n, per - 1, replace=False).tolist()) | {k}
        rows += sorted(rs)
        cols += [k] * len(rs)
    A = sp.csc_matrix((rng.uniform(-1, 1, len(rows)), (rows, cols)), shape=(n, n))
    A.setdiag(30.0)
    A.sort_indices()
    A = D.CSC(n, A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float64))
    S = D.System(ctx, A)
    S.set_tile_shapes([1, 2, 1])
    tune = S.autotune(D.PRECISION_FP64, D.SCHEDULE_BALANCED_NNZ, reps=1)
    print(f"no register-tier columns: {tune}", flush=True)
    assert tune["columns"] == [0, 0, 0] and tune["shape"] == [1, 2, 1]
    assert all(v is None for row in tune["ms"] for v in row)
    m_ref = O.spai(A.n, A.colptr, A.rowidx, A.val, 0)
    assert column_rel_err(A, S.m_values(), m_ref) <= 1e-10
    S.close()


def test_shape_arguments_checked(ctx):
    S = D.System(ctx, D.Problem.generate(1, 1).system(0))
    before = S.tile_shapes()
    for bad in ([D.tile_shape_count(0), 0, 0], [0, -1, 0], [0, 0, D.tile_shape_count(2)]):
        with pytest.raises(D.SpaiError):
            S.set_tile_shapes(bad)
    assert S.tile_shapes() == before
    with pytest.raises(D.SpaiError):
        S.autotune(D.PRECISION_FP64, D.SCHEDULE_BALANCED_NNZ, reps=0)
    S.close()
