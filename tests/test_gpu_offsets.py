"""Column storage of the SELL-64 layout.

Three forms, the most compact that applies: diagonal slices (one column -
row offset per column block of a slice, shared by its 64 rows, when a
slice's distinct offsets add at most 10% to the row-length padding), 16-bit
column - row offsets per entry (when every entry has |column - row| < 2^15)
or 32-bit columns.  All gather the same values in the same order (diagonal
slices add exact zeros), so a solve must give bit-identical x, iteration
counts and M with each; SPAI_MI355X_SELL_INDEX = 16 or 32 caps the form.
The edge of the 16-bit range is pinned with a matrix whose longest reach is
2^15 - 1 (fits) or 2^15 (does not), and diagonal slices with a reach far past
it, each checked against the oracle's iteration count and the true residual.
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp

import dcu_amd as D
from oracle import binding as O
from parity import count_ok, oracle_counts

pytestmark = pytest.mark.gpu


def build_system(ctx, mats, cap=None):
    old = os.environ.get("SPAI_MI355X_SELL_INDEX")
    try:
        if cap is None:
            os.environ.pop("SPAI_MI355X_SELL_INDEX", None)
        else:
            os.environ["SPAI_MI355X_SELL_INDEX"] = str(cap)
        return D.System(ctx, mats)
    finally:
        if old is None:
            os.environ.pop("SPAI_MI355X_SELL_INDEX", None)
        else:
            os.environ["SPAI_MI355X_SELL_INDEX"] = old


@pytest.mark.parametrize("precision", [D.PRECISION_FP64, D.PRECISION_MIXED])
def test_forms_bit_identical_batched(ctx, precision):
    P = D.Problem.batched(6, 11)
    mats = P.systems()
    b = np.concatenate([P.rhs(s) for s in range(len(mats))])
    out = {}
    for cap in (None, 16, 32):
        S = build_system(ctx, mats, cap)
        S.build(precision, D.SCHEDULE_BALANCED_NNZ)
        x, rep, oc = S.solve(b, rel_tol=1e-8, use_graph=True, true_residual_restart=True)
        out[cap] = (S.layout(), S.m_values(), x, [o["iterations"] for o in oc], [o["status"] for o in oc])
        S.close()
    print(f"layouts {[out[c][0] for c in out]}, iterations {out[None][3]}", flush=True)
    assert [out[c][0][0] for c in (None, 16, 32)] == ["sell64_diag", "sell64_off16", "sell64"]
    # stencil slices: the distinct offsets barely exceed the row-length
    # padding (a slice whose rows miss different neighbours adds a block)
    assert out[16][0][1] <= out[None][0][1] <= 1.1 * out[16][0][1]
    for cap in (16, 32):
        assert np.array_equal(out[None][1], out[cap][1])
        assert np.array_equal(out[None][2], out[cap][2])
        assert out[None][3] == out[cap][3]
    assert all(st == D.SOLVE_CONVERGED for st in out[None][4])


def test_diagonal_slices_rejected_for_scattered_rows(ctx):
    # random columns: a slice's distinct offsets far exceed its row length
    A = D.Problem.generate(1, 1).system(0)
    S = build_system(ctx, A)
    assert S.layout()[0] == "sell64_off16"
    S.close()


def reach_matrix(n, reach):
    """Tridiagonal (4 on the diagonal, -1 beside it) plus -1 at column
    i + reach and i - reach where these exist: the longest |column - row| is
    `reach`."""
    i = np.arange(n)
    rows = [i, i[1:], i[:-1], i[: n - reach], i[reach:]]
    cols = [i, i[1:] - 1, i[:-1] + 1, i[: n - reach] + reach, i[reach:] - reach]
    vals = [np.full(n, 4.0)] + [np.full(len(r), -1.0) for r in rows[1:]]
    A = sp.csc_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    A.sort_indices()
    return D.CSC(n, A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float64))


@pytest.mark.parametrize("n,reach,cap,layout", [(40000, 32767, 16, "sell64_off16"), (40000, 32768, 16, "sell64"),
                                                 (150000, 100000, None, "sell64_diag")])
def test_offsets_range_edge(ctx, n, reach, cap, layout):
    A = reach_matrix(n, reach)
    b = np.random.default_rng(reach).uniform(-1, 1, n)
    S = build_system(ctx, A, cap)
    S.build(D.PRECISION_FP64, D.SCHEDULE_BALANCED_NNZ)
    assert S.layout()[0] == layout
    x, rep, oc = S.solve(b, rel_tol=1e-10, max_iter=1000)
    Acsr = O.csc2csr(n, A.colptr, A.rowidx, A.val)[:3]
    Mcsr = O.csc2csr(n, A.colptr, A.rowidx, S.m_values())[:3]
    counts = oracle_counts(Acsr, Mcsr, b, 1e-10 * np.linalg.norm(b), 1000, threads=(1, 4))
    true = O.true_residual(Acsr, b, x) / np.linalg.norm(b)
    print(f"reach {reach}: layout {S.layout()[0]}, gpu {oc[0]['iterations']} oracle {sorted(counts)} "
          f"true rel {true:.2e}", flush=True)
    assert oc[0]["status"] == D.SOLVE_CONVERGED
    assert true < 1.5e-10
    assert count_ok(oc[0]["iterations"], counts)
    S.close()
