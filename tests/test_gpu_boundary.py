"""Entry points of the C ABI and the C++ drop-ins beyond the main build and
solve path.

  * cuCSC2CSR / spai_csc_to_csr_device on rectangular matrices (the
    reference converts an m x n CSC, include/common/cuFormatConversion.h:453-480):
    CSR arrays identical to scipy's.
  * spai_system_create_device: M and x bit-identical to a system created
    from the same matrix on the host.
  * record_history + spai_system_residual_history: one value per iteration,
    equal to the oracle's ||r_k|| (run with the GPU's own M) over the first
    iterations.
  * spai_system_time_spmv right after a build (no solve has run): it works
    on the solver's own vectors and leaves the next solve bit-identical.
  * MAX_ITER = 1: the reference's loop runs no iteration (x = x0 returned
    with MAX_ITER), as the oracle.
  * graphs captured on a destroyed system are never replayed on a new one
    with the same sizes: graph and loop solves of the new system agree.
"""
import ctypes

import numpy as np
import pytest
import scipy.sparse as sp

import dcu_amd as D
from oracle import binding as O

pytestmark = pytest.mark.gpu


def rect_csc(m, n, density, seed):
    rng = np.random.default_rng(seed)
    A = sp.random(m, n, density=density, format="csc", random_state=rng, dtype=np.float64)
    A.sort_indices()
    # the converter takes rows inside a column in any order: shuffle them
    ri, va = A.indices.copy(), A.data.copy()
    for c in range(n):
        p0, p1 = A.indptr[c], A.indptr[c + 1]
        q = rng.permutation(p1 - p0)
        ri[p0:p1] = ri[p0:p1][q]
        va[p0:p1] = va[p0:p1][q]
    return A, A.indptr.astype(np.int32), ri.astype(np.int32), va


@pytest.mark.parametrize("shape", [(300, 700), (700, 300), (1, 50), (50, 1)])
@pytest.mark.parametrize("entry", ["c_abi", "cxx"])
def test_csc_to_csr_rectangular(ctx, shape, entry):
    import torch

    m, n = shape
    A, cp, ri, va = rect_csc(m, n, 0.05 if min(m, n) > 1 else 0.5, seed=m * 7 + n)
    nnz = int(cp[-1])
    dev = torch.device("cuda:0")
    d_cp = torch.from_numpy(cp).to(dev)
    d_ri = torch.from_numpy(ri).to(dev)
    d_va = torch.from_numpy(va).to(dev)
    rp = torch.full((m + 1,), -7, dtype=torch.int32, device=dev)
    ci = torch.empty(max(nnz, 1), dtype=torch.int32, device=dev)
    cv = torch.empty(max(nnz, 1), dtype=torch.float64, device=dev)
    lib = D.load_library()
    args = (m, n, nnz, ctypes.c_void_p(d_va.data_ptr()), ctypes.c_void_p(d_ri.data_ptr()),
            ctypes.c_void_p(d_cp.data_ptr()), ctypes.c_void_p(cv.data_ptr()), ctypes.c_void_p(ci.data_ptr()),
            ctypes.c_void_p(rp.data_ptr()))
    if entry == "c_abi":
        assert lib.spai_csc_to_csr_device(*args) == 0
    else:
        f = getattr(lib, D.CXX_SYMBOLS["cuCSC2CSR"])
        f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 6
        f(*args)
    torch.cuda.synchronize()
    R = A.tocsr()
    R.sort_indices()
    assert np.array_equal(rp.cpu().numpy(), R.indptr)
    assert np.array_equal(ci[:nnz].cpu().numpy(), R.indices)
    assert np.array_equal(cv[:nnz].cpu().numpy(), R.data)


def test_create_device_matches_host(ctx):
    import torch

    P = D.Problem.generate(1, 3)
    A = P.system(0)
    b = P.rhs(0)
    dev = torch.device("cuda:0")
    d_cp = torch.from_numpy(A.colptr.astype(np.int32)).to(dev)
    d_ri = torch.from_numpy(A.rowidx.astype(np.int32)).to(dev)
    d_va = torch.from_numpy(A.val.astype(np.float64)).to(dev)
    Sd = D.System.from_device(ctx, A.n, d_cp.data_ptr(), d_ri.data_ptr(), d_va.data_ptr(), A.nnz)
    Sh = D.System(ctx, A)
    outs = []
    for S in (Sh, Sd):
        S.build(D.PRECISION_FP64, D.SCHEDULE_BALANCED_NNZ)
        x, rep, oc = S.solve(b, rel_tol=1e-8, use_graph=True)
        outs.append((S.m_values(), S.row_set_sizes(), x, oc[0]["iterations"], S.layout()))
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1])
    assert np.array_equal(outs[0][2], outs[1][2])
    assert outs[0][3] == outs[1][3] and outs[0][4] == outs[1][4]
    m_ref = O.spai(A.n, A.colptr, A.rowidx, A.val, 0)
    assert np.max(np.abs(outs[1][0] - m_ref)) <= 1e-10 * np.max(np.abs(m_ref))
    # the device arrays were copied: changing them does not reach the system
    d_va.mul_(2.0)
    torch.cuda.synchronize()
    Sd.build(D.PRECISION_FP64, D.SCHEDULE_BALANCED_NNZ)
    assert np.array_equal(Sd.m_values(), outs[0][0])
    Sd.close()
    Sh.close()


@pytest.mark.parametrize("use_graph", [False, True])
def test_residual_history_single(ctx, use_graph):
    P = D.Problem.generate(1, 1)
    A = P.system(0)
    b = P.rhs(0)
    S = D.System(ctx, A)
    S.build(D.PRECISION_FP64, D.SCHEDULE_BALANCED_NNZ)
    x, rep, oc = S.solve(b, rel_tol=1e-8, max_iter=1000, use_graph=use_graph, record_history=True)
    h = S.residual_history(0)
    it = oc[0]["iterations"]
    assert oc[0]["status"] == D.SOLVE_CONVERGED and len(h) == it
    assert h[-1] == pytest.approx(oc[0]["residual"], rel=1e-15)
    tol = oc[0]["tol"]
    assert h[-1] < tol and np.all(h[:-1] >= tol)
    Acsr = O.csc2csr(A.n, A.colptr, A.rowidx, A.val)[:3]
    Mcsr = O.csc2csr(A.n, A.colptr, A.rowidx, S.m_values())[:3]
    _, it_o, _, ho = O.bicgstab(Acsr, Mcsr, b, np.zeros(A.n), tol, 1000, history=True)
    k = min(len(h), len(ho))
    rel = np.abs(h[:k] - ho[:k]) / ho[:k]
    print(f"history: gpu {it} values, oracle {it_o}; max rel diff {rel.max():.2e}", flush=True)
    assert abs(it - it_o) <= 1
    assert rel.max() <= 1e-8
    # a solve without the history leaves none to read
    S.solve(b, rel_tol=1e-8)
    assert len(S.residual_history(0)) == 0


def test_residual_history_batch(ctx):
    P = D.Problem.batched(6, 5)
    mats = P.systems()
    b = np.concatenate([P.rhs(s) for s in range(len(mats))])
    S = D.System(ctx, mats)
    S.build(D.PRECISION_MIXED, D.SCHEDULE_BALANCED_NNZ)
    x, rep, oc = S.solve(b, rel_tol=1e-8, max_iter=1000, use_graph=True, record_history=True)
    m_gpu = S.m_values()
    off = S.row_offsets
    moff = np.concatenate([[0], np.cumsum([A.nnz for A in mats])])
    for s, A in enumerate(mats):
        h = S.residual_history(s)
        assert oc[s]["status"] == D.SOLVE_CONVERGED and len(h) == oc[s]["iterations"], s
        bs = b[off[s]:off[s + 1]]
        Acsr = O.csc2csr(A.n, A.colptr, A.rowidx, A.val)[:3]
        Mcsr = O.csc2csr(A.n, A.colptr, A.rowidx, m_gpu[moff[s]:moff[s + 1]])[:3]
        # the mixed solve stores r* in FP32: the oracle's shadow32 mode
        _, _, _, ho = O.bicgstab(Acsr, Mcsr, bs, np.zeros(A.n), 0.0, 11, history=True, shadow32=True)
        rel = np.abs(h[:10] - ho[:10]) / ho[:10]
        assert rel.max() <= 1e-8, (s, rel.max())
    with pytest.raises(D.SpaiError):
        S.residual_history(len(mats))


def test_time_spmv_right_after_build(ctx):
    # the single-kernel timers work on the solver's own vectors: usable
    # before any solve, and they change neither a caller's x nor the next
    # solve's result
    import torch

    P = D.Problem.generate(1, 4)
    A = P.system(0)
    b = P.rhs(0)
    S1 = D.System(ctx, A)
    S1.build(D.PRECISION_FP64, D.SCHEDULE_BALANCED_NNZ)
    S2 = D.System(ctx, A)
    S2.build(D.PRECISION_FP64, D.SCHEDULE_BALANCED_NNZ)
    for which in range(6):
        assert S2.time_spmv(which, 2) > 0.0
    dev = torch.device("cuda:0")
    db = torch.from_numpy(b).to(dev)
    dx = torch.zeros(A.n, dtype=torch.float64, device=dev)
    S1.solve_device(db.data_ptr(), dx.data_ptr(), rel_tol=1e-8)
    torch.cuda.synchronize()
    x_keep = dx.clone()
    for which in range(6):  # S2's timers after S1's solve on the same context
        assert S2.time_spmv(which, 2) > 0.0
    torch.cuda.synchronize()
    assert torch.equal(dx, x_keep)
    x1, _, o1 = S1.solve(b, rel_tol=1e-8)
    x2, _, o2 = S2.solve(b, rel_tol=1e-8)
    assert np.array_equal(x1, x2) and np.array_equal(x1, x_keep.cpu().numpy())
    assert o1[0]["iterations"] == o2[0]["iterations"]


@pytest.mark.parametrize("use_graph", [False, True])
def test_max_iter_one_runs_no_iteration(ctx, use_graph):
    P = D.Problem.batched(3, 2)
    mats = P.systems()
    b = np.concatenate([P.rhs(s) for s in range(len(mats))])
    for systems in (mats[:1], mats):
        S = D.System(ctx, systems if len(systems) > 1 else systems[0])
        S.build(D.PRECISION_FP64, D.SCHEDULE_BALANCED_NNZ)
        bb = b[:S.n]
        x0 = np.full(S.n, 0.25)
        x, rep, oc = S.solve(bb, x0=x0, rel_tol=1e-8, max_iter=1, use_graph=use_graph)
        assert np.array_equal(x, x0)
        for o in oc:
            assert o["status"] == D.SOLVE_MAX_ITER and o["iterations"] == 1
        A = systems[0]
        Acsr = O.csc2csr(A.n, A.colptr, A.rowidx, A.val)[:3]
        xo, it, info = O.bicgstab(Acsr, None, bb[:A.n], x0[:A.n], 1e-8, 1)
        assert it == 1 and info[2] == 0 and np.array_equal(xo, x0[:A.n])
        S.close()


def test_graphs_not_replayed_on_a_new_system(ctx):
    # systems of identical sizes created after one is destroyed may get the
    # same device addresses; their solves capture their own graphs
    P = D.Problem.generate(1, 6)
    A = P.system(0)
    b = P.rhs(0)
    rng = np.random.default_rng(6)
    for k in range(3):
        vals = A.val * rng.uniform(0.9, 1.1, size=A.nnz)
        Ak = D.CSC(A.n, A.colptr, A.rowidx, vals)
        S = D.System(ctx, Ak)
        S.build(D.PRECISION_FP64, D.SCHEDULE_BALANCED_NNZ)
        xg, _, og = S.solve(b, rel_tol=1e-8, use_graph=True)
        xl, _, ol = S.solve(b, rel_tol=1e-8, use_graph=False)
        assert np.array_equal(xg, xl) and og[0]["iterations"] == ol[0]["iterations"], k
        Acsr = O.csc2csr(A.n, A.colptr, A.rowidx, vals)[:3]
        assert O.true_residual(Acsr, b, xg) / np.linalg.norm(b) <= 1.5e-8
        S.close()


@pytest.mark.parametrize("batched", [False, True])
def test_systems_of_more_than_512_units(ctx, batched):
    """A 650 x 650 conv-diff grid (422,500 rows: 826 SELL-64 units of 512
    rows) has more units than partial-sum groups per system, so its units
    publish through groups of two (the group level with arrival counters)
    while a small system in the same batch publishes one partial per unit.
    The first 10 iterates and residuals agree with the oracle's to 1e-8 and
    the solve converges."""
    big = D.Problem.convection_diffusion(650, 5.0).system(0)
    mats = [big, D.Problem.convection_diffusion(40, 5.0).system(0)] if batched else [big]
    rng = np.random.default_rng(11)
    bs = [rng.uniform(size=A.n) for A in mats]
    S = D.System(ctx, mats if batched else big)
    layout, _ = S.layout()
    assert layout.startswith("sell64") and big.n > 512 * 512
    S.build(D.PRECISION_FP64, D.SCHEDULE_BALANCED_NNZ)
    m_gpu = S.m_values()
    b = np.concatenate(bs)
    x10, _, oc10 = S.solve(b, rel_tol=0.0, abs_tol=1e-300, max_iter=11, record_history=True)
    off = S.row_offsets
    moff = np.concatenate([[0], np.cumsum([A.nnz for A in mats])])
    for s, A in enumerate(mats):
        Acsr = O.csc2csr(A.n, A.colptr, A.rowidx, A.val)[:3]
        Mcsr = O.csc2csr(A.n, A.colptr, A.rowidx, m_gpu[moff[s]:moff[s + 1]])[:3]
        xo, _, _, ho = O.bicgstab(Acsr, Mcsr, bs[s], np.zeros(A.n), 0.0, 11, history=True)
        xs = x10[off[s]:off[s + 1]]
        h = S.residual_history(s)
        err = np.linalg.norm(xs - xo) / np.linalg.norm(xo)
        rel = np.abs(h[:10] - ho[:10]) / ho[:10]
        print(f"system {s} (n={A.n}): 10-iterate rel err {err:.2e}, history {rel.max():.2e}", flush=True)
        assert oc10[s]["iterations"] == 11
        assert err <= 1e-8 and rel.max() <= 1e-8, s
    x, rep, oc = S.solve(b, rel_tol=1e-8, max_iter=3000, use_graph=True)
    for s, A in enumerate(mats):
        Acsr = O.csc2csr(A.n, A.colptr, A.rowidx, A.val)[:3]
        true_rel = O.true_residual(Acsr, bs[s], x[off[s]:off[s + 1]]) / np.linalg.norm(bs[s])
        print(f"system {s}: {oc[s]['iterations']} iterations, true rel {true_rel:.2e}", flush=True)
        assert oc[s]["status"] == D.SOLVE_CONVERGED and true_rel <= 1.5e-8, s
    S.close()
