"""16-bit column offsets in the SELL-64 layout.

The layout stores each entry's column as column - row in 16 bits when every
entry of the pattern fits (|column - row| <= 2^15 - 1, or -2^15), otherwise
32-bit columns.  Both forms gather the same values in the same order, so a
solve must give bit-identical x, iteration counts and M with either
(SPAI_MI355X_SELL_OFFSETS=0 forces 32-bit columns).  The edge of the 16-bit
range is pinned with a matrix whose longest reach is 2^15 - 1 (fits) or 2^15
(does not), each checked against the oracle's iteration count and the true
residual.
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp

import dcu_amd as D
from oracle import binding as O
from parity import count_ok, oracle_counts

pytestmark = pytest.mark.gpu


def build_system(ctx, mats, offsets):
    old = os.environ.get("SPAI_MI355X_SELL_OFFSETS")
    try:
        if offsets:
            os.environ.pop("SPAI_MI355X_SELL_OFFSETS", None)
        else:
            os.environ["SPAI_MI355X_SELL_OFFSETS"] = "0"
        return D.System(ctx, mats)
    finally:
        if old is None:
            os.environ.pop("SPAI_MI355X_SELL_OFFSETS", None)
        else:
            os.environ["SPAI_MI355X_SELL_OFFSETS"] = old


@pytest.mark.parametrize("precision", [D.PRECISION_FP64, D.PRECISION_MIXED])
def test_offsets_bit_identical_batched(ctx, precision):
    P = D.Problem.batched(6, 11)
    mats = P.systems()
    b = np.concatenate([P.rhs(s) for s in range(len(mats))])
    out = {}
    for offsets in (True, False):
        S = build_system(ctx, mats, offsets)
        S.build(precision, D.SCHEDULE_BALANCED_NNZ)
        x, rep, oc = S.solve(b, rel_tol=1e-8, use_graph=True, true_residual_restart=True)
        out[offsets] = (S.layout()[0], S.m_values(), x, [o["iterations"] for o in oc], [o["status"] for o in oc])
        S.close()
    print(f"layouts {out[True][0]} / {out[False][0]}, iterations {out[True][3]}", flush=True)
    assert out[True][0] == "sell64_off16" and out[False][0] == "sell64"
    assert np.array_equal(out[True][1], out[False][1])
    assert np.array_equal(out[True][2], out[False][2])
    assert out[True][3] == out[False][3]
    assert all(st == D.SOLVE_CONVERGED for st in out[True][4])


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


@pytest.mark.parametrize("reach,layout", [(32767, "sell64_off16"), (32768, "sell64")])
def test_offsets_range_edge(ctx, reach, layout):
    n = 40000
    A = reach_matrix(n, reach)
    b = np.random.default_rng(reach).uniform(-1, 1, n)
    S = build_system(ctx, A, True)
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
