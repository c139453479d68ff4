"""The allocation-free solve, spai_system_solve_noalloc, against
spai_system_solve (the entry point it varies) through the C ABI.

Its work vectors, states, counters and reduction partials live in the
library's static device scratch, shared by every call on the device, so two
calls run in a row, the second on a smaller input (a batch of 6 systems,
333,378 rows, then one system of 4096 rows): each must give x, iteration
counts and statuses bit-identical to spai_system_solve with the same options
(no graph), whatever the scratch held before.  Both are also checked against
the oracle's true residual.  A system beyond the scratch's 64 MiB is refused
with SPAI_ERR_UNSUPPORTED, and graphs or histories with SPAI_ERR_INVALID.
"""
import numpy as np
import pytest

import dcu_amd as D
from oracle import binding as O

pytestmark = pytest.mark.gpu


def _solve_ref(S, b, **kw):
    x, rep, oc = S.solve(b, **kw)
    return x, rep, oc


def _true_rel(mats, off, b, x):
    worst = 0.0
    for s, A in enumerate(mats):
        bs = b[off[s]:off[s + 1]]
        Acsr = O.csc2csr(A.n, A.colptr, A.rowidx, A.val)[:3]
        worst = max(worst, O.true_residual(Acsr, bs, x[off[s]:off[s + 1]]) / np.linalg.norm(bs))
    return worst


def test_noalloc_matches_solve_two_calls(ctx):
    import torch

    dev = torch.device("cuda:0")
    P1 = D.Problem.batched(6, 1)  # 333,378 rows in 6 systems
    mats1 = P1.systems()
    b1 = np.concatenate([P1.rhs(s) for s in range(len(mats1))])
    S1 = D.System(ctx, mats1)
    S1.build(D.PRECISION_MIXED, D.SCHEDULE_BALANCED_NNZ)
    P2 = D.Problem.generate(1, 1)  # 4096 rows
    A2 = P2.system(0)
    b2 = P2.rhs(0)
    S2 = D.System(ctx, A2)
    S2.build(D.PRECISION_FP64, D.SCHEDULE_BALANCED_NNZ)
    kw = dict(rel_tol=1e-8, max_iter=1000, use_graph=False, true_residual_restart=True)
    refs = []
    for S, b in ((S1, b1), (S2, b2)):
        x, rep, oc = _solve_ref(S, b, **kw)
        refs.append((x, rep, oc))
    # the two allocation-free calls in a row, the second on the smaller input
    outs = []
    for S, b in ((S1, b1), (S2, b2)):
        db = torch.from_numpy(b).to(dev)
        dx = torch.zeros(S.n, dtype=torch.float64, device=dev)
        print(f"solve_noalloc: {S.nsys} system(s), {S.n} rows", flush=True)
        rep = S.solve_noalloc(db.data_ptr(), dx.data_ptr(), **kw)
        outs.append((dx.cpu().numpy(), rep, S.outcomes()))
    for (S, b, mats), (xr, rr, ocr), (x, r, oc) in zip(((S1, b1, mats1), (S2, b2, [A2])), refs, outs):
        assert np.array_equal(x, xr)
        assert [o["iterations"] for o in oc] == [o["iterations"] for o in ocr]
        assert [o["status"] for o in oc] == [o["status"] for o in ocr]
        assert [o["true_residual"] for o in oc] == [o["true_residual"] for o in ocr]
        assert r["iterations_max"] == rr["iterations_max"] and r["iterations_sum"] == rr["iterations_sum"]
        assert r["all_converged"] and rr["all_converged"]
        worst = _true_rel(mats, S.row_offsets, b, x)
        print(f"{S.nsys} system(s): iterations max {r['iterations_max']}, worst true rel {worst:.2e}", flush=True)
        assert worst < 1e-8
        assert len(S.residual_history(0)) == 0  # no history without allocation
    # again in the same order: the scratch now holds the smaller call's data
    db = torch.from_numpy(b1).to(dev)
    dx = torch.zeros(S1.n, dtype=torch.float64, device=dev)
    S1.solve_noalloc(db.data_ptr(), dx.data_ptr(), **kw)
    assert np.array_equal(dx.cpu().numpy(), refs[0][0])


def test_noalloc_refusals(ctx):
    import torch

    dev = torch.device("cuda:0")
    P = D.Problem.generate(1, 2)
    S = D.System(ctx, P.system(0))
    db = torch.ones(S.n, dtype=torch.float64, device=dev)
    dx = torch.zeros(S.n, dtype=torch.float64, device=dev)
    with pytest.raises(D.SpaiError) as e:  # M not built
        S.solve_noalloc(db.data_ptr(), dx.data_ptr())
    assert e.value.status == D.ERR_NOT_BUILT
    S.build(D.PRECISION_FP64, D.SCHEDULE_BALANCED_NNZ)
    for kw in (dict(use_graph=True), dict(record_history=True)):
        with pytest.raises(D.SpaiError) as e:
            S.solve_noalloc(db.data_ptr(), dx.data_ptr(), **kw)
        assert e.value.status == D.ERR_INVALID
    # a diagonal system of 1,000,000 rows needs more than the 64 MiB scratch
    n = 1_000_000
    big = D.CSC(n, np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32), np.full(n, 2.0))
    B = D.System(ctx, big)
    bb = torch.ones(n, dtype=torch.float64, device=dev)
    bx = torch.zeros(n, dtype=torch.float64, device=dev)
    with pytest.raises(D.SpaiError) as e:
        B.solve_noalloc(bb.data_ptr(), bx.data_ptr(), precondition=False)
    assert e.value.status == D.ERR_UNSUPPORTED
    torch.cuda.synchronize()
    assert torch.count_nonzero(bx).item() == 0  # refused before any work
    B.close()
    # the entry point still works after the refusals
    rep = S.solve_noalloc(db.data_ptr(), dx.data_ptr(), rel_tol=1e-8)
    assert rep["all_converged"]
    x, _, _ = S.solve(np.ones(S.n), rel_tol=1e-8)
    assert np.array_equal(dx.cpu().numpy(), x)
