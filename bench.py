"""Benchmark: SPAI build + SPAI-preconditioned BiCGSTAB on one MI355X.

Default workload (one "step"): the batch of 64 independent systems (sizes
log-uniform in [2e3, 2e5]; even: 2D conv-diff, odd: 3D lognormal diffusion),
per-system relative tolerance 1e-8, x0 = 0:
  - batched mixed path: FP32 SPAI QR + FP64 sweep, FP32 M, balanced_nnz
    queues, all 64 systems in each launch, BiCGSTAB replayed from a hipGraph;
  - plain FP64 path (timed beside it): the same 64 systems one at a time,
    FP64 SPAI and M, one launch set per iteration with a host convergence
    check per iteration.
value = end-to-end systems solved per second of the batched path
(SPAI build + solve, inputs resident in HBM).

Beside it in the same JSON line (skip with --no-legs):
  - batched_fp64_graph: the same batched hipGraph path with FP64 SPAI and M
    (the mixed-precision gain alone);
  - spai_roofline: the SPAI build's algorithmic bytes and flops per build;
  - work: rows x iterations per second of solve (independent of which
    system converges last);
  - legs: the four single-system workloads (random 4096 FP64 static,
    conv-diff 512^2 FP64, power-law 1e6 mixed, 64^3 GRF FP64 with its
    adjoint), each with its kernel roofline, SPAI roofline and a capped CPU
    baseline of the oracle.

Usage: python bench.py [--steps K] [--warmup W] [--workload batched|convdiff|powerlaw|grf|random]
                       [--precision mixed|fp64] [--no-cpu] [--no-plain]
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_GFLOPS = 157300.0  # MI355X vector FP32
FP64_PEAK_GFLOPS = 78600.0   # MI355X vector FP64

WORKLOADS = {
    # name: (generator config, description)
    "random": (1, "random4096: n=4096, 4 off-diagonals/row, static columns"),
    "convdiff": (2, "convdiff512: 512^2 upwind conv-diff, n=262144"),
    "powerlaw": (3, "powerlaw1e6: Zipf(2)-capped rows, n=1e6"),
    "batched": (4, "batched64: 64 systems, n in [2e3,2e5] log-uniform, 2D conv-diff / 3D diffusion"),
    "grf": (5, "grf64: 64^3 lognormal-GRF diffusion, n=262144"),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def kernel_bytes(kind, n, nnz, m_bytes, batched=True):
    """Algorithmic HBM bytes of one launch of an iteration kernel with every
    system active, in the CSR terms of the reference's operators: per
    nonzero a 4 B column index and the value (A: 8 B, M: m_bytes), 4 B per
    row pointer, and each FP64 vector read or written once (the shadow
    residual r*, stored in FP32 next to an FP32 M, at m_bytes).  The solver's
    SELL-64 layouts store less index data (16-bit offsets, or one offset per
    column block of a slice) and some padding; the measured traffic is the
    `traffic` figure."""
    rp = 4 * (n + 1)
    rs = m_bytes * n  # r*
    # x: a batch's K1 updates x every other iteration from two iterations' y
    # and z (one launch reads x, y, z, y', z' and writes x, the next touches
    # none of them: 3 vectors per launch on average); a single system's K1
    # reads x, y, z and writes x every iteration (4)
    xv = 3 if batched else 4
    if kind == "update_mapply_p":  # fused K1: r update + y = M p'; reads M, s, t, p, a; writes r, p', y; + x
        return (m_bytes + 4) * nnz + rp + 8 * n * (7 + xv)
    if kind == "update":          # split K1, streaming half: reads s, t, p, a; writes r, p'; + x
        return 8 * n * (6 + xv)
    if kind == "mapply_p":        # split K1, y = M p'; reads M, p' (gather, once); writes y
        return (m_bytes + 4) * nnz + rp + 8 * n * 2
    if kind == "spmv_alpha":      # a = A y; reads A, y (gather, once), r*; writes a
        return 12 * nnz + rp + 8 * n * 2 + rs
    if kind == "mapply_s":        # z = M s; reads M, r, a; writes s, z
        return (m_bytes + 4) * nnz + rp + 8 * n * 4
    if kind == "spmv_omega":      # t = A z; reads A, z, s, r*; writes t
        return 12 * nnz + rp + 8 * n * 3 + rs
    raise ValueError(kind)


def spai_roofline(S, mats, rep, precision_mixed, ms):
    """The SPAI build against its roofline: algorithmic bytes = 12 B per
    gathered entry of A(:, J_k) (row index + value), A's pattern (4 B per
    nonzero and per column pointer) and M written twice (CSC and the solver's
    layout); flops = sum_k 2 |I_k| |J_k|^2 - 2/3 |J_k|^3 (Householder QR of
    each column's least squares)."""
    import numpy as np

    n1 = S.row_set_sizes().astype(np.float64)
    n2 = np.concatenate([np.diff(A.colptr) for A in mats]).astype(np.float64)
    n = float(sum(A.n for A in mats))
    nnz = float(sum(A.nnz for A in mats))
    s_m = 4.0 if precision_mixed else 8.0
    nbytes = 12.0 * rep["sum_gathered"] + 4.0 * nnz + 4.0 * (n + 1) + 2.0 * s_m * nnz
    flops = float(np.sum(np.maximum(2.0 * n1 * n2 ** 2 - 2.0 / 3.0 * n2 ** 3, 0.0)))
    gbs = nbytes / (ms * 1e-3) / 1e9
    gflops = flops / (ms * 1e-3) / 1e9
    peak = FP32_PEAK_GFLOPS if precision_mixed else FP64_PEAK_GFLOPS
    return {"bound": "hbm", "ms": round(ms, 4), "bytes": int(nbytes), "achieved": round(gbs, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4), "flops": int(flops),
            "gflops": round(gflops, 1), "flop_peak": peak, "flop_frac": round(gflops / peak, 5),
            "sum_gathered": int(rep["sum_gathered"])}


def row_iterations(mats, outcomes, max_status):
    """Sum over systems of rows x iterations actually run (the work a solve
    does, independent of which system is slowest)."""
    tot = 0
    for A, o in zip(mats, outcomes):
        it = o["iterations"] if o["status"] != max_status else o["iterations"] - 1
        tot += A.n * it
    return tot


# spai_system_time_spmv's kernel index for the iteration kernels it can
# launch alone (the split K1's halves have none)
TIME_SPMV_WHICH = {"update_mapply_p": 1, "spmv_alpha": 2, "mapply_s": 3, "spmv_omega": 4}


def launch_ms(S, name, fallback, reps=50):
    """Average device ms of one launch of kernel `name`, from HIP events
    around `reps` back-to-back launches of it alone (every system active), so
    that the events add no gap between launches; the time_iteration figure
    (events between every kernel of the iteration sequence) when the kernel
    cannot be launched alone."""
    which = TIME_SPMV_WHICH.get(name)
    return S.time_spmv(which, reps) if which is not None else fallback


def iteration_kernels(ms5):
    """Names the five slots of System.time_iteration for the K1 form used."""
    if ms5[1] > 0.0:  # split K1
        names = ("update", "mapply_p", "spmv_alpha", "mapply_s", "spmv_omega")
        return dict(zip(names, ms5))
    return dict(zip(("update_mapply_p", "spmv_alpha", "mapply_s", "spmv_omega"), [ms5[0]] + ms5[2:]))


def cpu_solves(O, mats, rhs, budget, iter_cap=None):
    """The oracle (MGS SPAI + BiCGSTAB, FP64, OpenMP over columns and rows)
    on the host cores: cycles over the systems until `budget` seconds are
    spent (at least one solve).  iter_cap: run only that many iterations per
    solve and extrapolate the solve to the iterations the GPU ran (stated in
    the sample)."""
    import numpy as np

    def one(A, r, its):
        t0 = time.perf_counter()
        m = O.spai(A.n, A.colptr, A.rowidx, A.val, 0)
        t1 = time.perf_counter()
        Acsr = O.csc2csr(A.n, A.colptr, A.rowidx, A.val)[:3]
        Mcsr = O.csc2csr(A.n, A.colptr, A.rowidx, m)[:3]
        t2 = time.perf_counter()
        _, it, _ = O.bicgstab(Acsr, Mcsr, r, np.zeros(A.n), 1e-8 * np.linalg.norm(r), its)
        t3 = time.perf_counter()
        return t1 - t0, t3 - t2, it, t2 - t1

    small = min(range(len(mats)), key=lambda s_: mats[s_].n)
    one(mats[small], rhs[small], iter_cap or 1000)  # starts the thread pool
    done, t_cpu, unknowns, note = 0, 0.0, 0, ""
    while t_cpu < budget or done == 0:
        A, r = mats[done % len(mats)], rhs[done % len(mats)]
        if iter_cap:
            ts, tsolve, it, tconv = one(A, r, iter_cap + 1)
            per_it = tsolve / max(it - 1, 1)
            t = ts + tconv + per_it * 1000
            note = (f"oracle SPAI once + {iter_cap} BiCGSTAB iterations, the solve extrapolated to 1000 "
                    f"iterations at {per_it * 1e3:.2f} ms per iteration")
            t_cpu += t
            done += 1
            unknowns += A.n
            break
        ts, tsolve, it, tconv = one(A, r, 1000)
        t_cpu += ts + tsolve + tconv
        done += 1
        unknowns += A.n
    return {"value": round(done / t_cpu, 4), "unit": "solves/s", "cores": O.num_threads(), "kind": "port",
            "sample": (f"{done} solve(s) over the {len(mats)} system(s) of the same workload ({unknowns} unknowns "
                       f"in all), oracle MGS SPAI + BiCGSTAB (FP64) on {cpu_model()}, {t_cpu:.1f}s"
                       + (f"; {note}" if note else ""))}


def run_leg(D, O, ctx, torch, name, precision, schedule, steps, cpu_budget, with_cpu, adjoint=False, cpu_cap=None):
    """One single-system workload at its configured precision and schedule:
    SPAI build + graph-replayed BiCGSTAB (rel. 1e-8, x0 = 0, true-residual
    restart), with the tile shapes tuned once first; per-kernel roofline,
    SPAI-build roofline, work rate and a capped CPU baseline."""
    cfg, desc = WORKLOADS[name]
    t0 = time.time()
    P = D.Problem.generate(cfg, cfg)
    mats = P.systems()
    A = mats[0]
    b = P.rhs(0)
    gen_s = time.time() - t0
    S = D.System(ctx, A)
    mixed = precision == D.PRECISION_MIXED
    tune = S.autotune(precision, schedule, reps=3)
    dev = torch.device("cuda:0")
    b_d = torch.from_numpy(b).to(dev)
    x_d = torch.zeros(A.n, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()

    def step():
        x_d.zero_()
        torch.cuda.synchronize()
        rb = S.build(precision, schedule)
        rs = S.solve_device(b_d.data_ptr(), x_d.data_ptr(), rel_tol=1e-8, max_iter=1000, use_graph=True,
                            true_residual_restart=True)
        return rb, rs

    step()
    torch.cuda.synchronize()
    spai, solve = [], []
    t1 = time.perf_counter()
    for _ in range(steps):
        rb, rs = step()
        spai.append(rb["ms"])
        solve.append(rs["ms"])
    torch.cuda.synchronize()
    el = time.perf_counter() - t1
    oc = S.outcomes()[0]
    m_bytes = 4 if mixed else 8
    kms = iteration_kernels(S.time_iteration(10))
    dom = max(kms, key=kms.get)
    kb = kernel_bytes(dom, A.n, A.nnz, m_bytes, batched=False)
    dom_ms = launch_ms(S, dom, kms[dom])
    ach = kb / (dom_ms * 1e-3) / 1e9
    layout, _ = S.layout()
    solve_ms = float(np.median(solve))
    rows_it = row_iterations([A], [oc], D.SOLVE_MAX_ITER)
    # the whole solve against the roofline: the iterations run x one
    # iteration's algorithmic bytes (fused K1 count), over the solve's time
    # (graph-replayed, so without the launch gaps of the per-kernel events)
    its_run = rows_it // A.n
    iter_bytes = sum(kernel_bytes(k, A.n, A.nnz, m_bytes, False)
                     for k in ("update_mapply_p", "spmv_alpha", "mapply_s", "spmv_omega"))
    solve_gbs = its_run * iter_bytes / (solve_ms * 1e-3) / 1e9
    leg = {"workload": desc, "precision": "mixed" if mixed else "fp64",
           "schedule": "balanced_nnz" if schedule == D.SCHEDULE_BALANCED_NNZ else "balanced_columns",
           "layout": layout, "n": A.n, "nnz": A.nnz, "steps": steps,
           "value": round(steps / el, 3), "unit": "solves/s", "ms_per_step": round(1e3 * el / steps, 3),
           "spai_ms": round(float(np.median(spai)), 4), "solve_ms": round(solve_ms, 4),
           "iterations": oc["iterations"], "converged": oc["status"] == D.SOLVE_CONVERGED,
           "true_rel_residual": oc["true_residual"] / float(np.linalg.norm(b)),
           "row_iterations_per_s": round(rows_it / (solve_ms * 1e-3), 1),
           "tile_shapes": tune["shape"],
           "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "bytes_per_launch": int(kb),
                        "launch_ms": round(dom_ms, 5),
                        "all_kernels_frac": {k: round(kernel_bytes(k, A.n, A.nnz, m_bytes, False) / (v * 1e-3) /
                                                      1e9 / HBM_PEAK_GBS, 4) for k, v in kms.items()},
                        "all_kernels_ms": {k: round(v, 5) for k, v in kms.items()},
                        "iteration_us": round(1e3 * solve_ms / max(its_run, 1), 2),
                        "iteration_bytes": int(iter_bytes),
                        "solve_achieved": round(solve_gbs, 1), "solve_frac": round(solve_gbs / HBM_PEAK_GBS, 4)},
           "spai_roofline": spai_roofline(S, [A], rb, mixed, float(np.median(spai))),
           "generate_s": round(gen_s, 2)}
    if adjoint:
        # the adjoint of the converged solve for L = 0.5 ||x - x_t||^2:
        # A^T lambda = x - x_t, preconditioned by M^T, then dL/dA and dL/db
        xt = P.target(0)
        g_d = x_d - torch.from_numpy(xt if xt is not None else np.zeros(A.n)).to(dev)
        dA_d = torch.empty(A.nnz, dtype=torch.float64, device=dev)
        db_d = torch.empty(A.n, dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        S.adjoint_device(x_d.data_ptr(), g_d.data_ptr(), dA_d.data_ptr(), db_d.data_ptr(), rel_tol=1e-8,
                         max_iter=1000, use_graph=True)
        ams = []
        for _ in range(steps):
            ams.append(S.adjoint_device(x_d.data_ptr(), g_d.data_ptr(), dA_d.data_ptr(), db_d.data_ptr(),
                                        rel_tol=1e-8, max_iter=1000, use_graph=True)["ms"])
        leg["adjoint_ms"] = round(float(np.median(ams)), 4)
        leg["adjoint_iterations"] = S.outcomes()[0]["iterations"]
    if with_cpu:
        leg["cpu_baseline"] = cpu_solves(O, [A], [b], cpu_budget, cpu_cap)
    S.close()
    log(f"leg {name}: {json.dumps(leg)}")
    return leg


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="batched", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of CPU-baseline work")
    ap.add_argument("--graph-iters", type=int, default=8)
    ap.add_argument("--no-autotune", action="store_true",
                    help="build with the default tile shapes instead of tuning them once before the warm-up")
    ap.add_argument("--tune-reps", type=int, default=5)
    ap.add_argument("--profile-kernels", action="store_true",
                    help="after one solve, launch each fused kernel --profile-reps times and a device copy of "
                         "known size (256 MiB read + 256 MiB written) as many times, for counter collection")
    ap.add_argument("--profile-reps", type=int, default=20)
    ap.add_argument("--no-legs", action="store_true",
                    help="skip the batched FP64 leg and the single-workload legs (random, convdiff, powerlaw, grf)")
    ap.add_argument("--leg-steps", type=int, default=3)
    ap.add_argument("--leg-cpu-budget", type=float, default=3.0, help="seconds of CPU-baseline work per leg")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the SPAI build and the batched solve run on a single MI355X "
              "(multi-GPU sharding is not part of this build); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    import dcu_amd as D

    cfg, desc = WORKLOADS[args.workload]
    t0 = time.time()
    P = D.Problem.generate(cfg, args.seed if args.workload == "batched" else cfg)
    mats = P.systems()
    rhs = [P.rhs(s) for s in range(len(mats))]
    log(f"generated {P.name} ({len(mats)} systems) in {time.time() - t0:.1f}s")
    nsys = len(mats)
    n_tot = sum(A.n for A in mats)
    nnz_tot = sum(A.nnz for A in mats)

    ctx = D.Context(0)
    dev_name, cus = ctx.device_info()
    precision = D.PRECISION_MIXED if args.precision == "mixed" else D.PRECISION_FP64
    batched = nsys > 1
    S = D.System(ctx, mats if batched else mats[0])
    dev = torch.device("cuda:0")
    b_d = torch.from_numpy(np.concatenate(rhs)).to(dev)
    x_d = torch.zeros(n_tot, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    bnorm = np.array([np.linalg.norm(r) for r in rhs])

    def step():
        x_d.zero_()
        torch.cuda.synchronize()
        rep_b = S.build(precision, D.SCHEDULE_BALANCED_NNZ)
        rep_s = S.solve_device(b_d.data_ptr(), x_d.data_ptr(), rel_tol=1e-8, max_iter=1000, use_graph=True,
                               graph_iters=args.graph_iters, true_residual_restart=True)
        return rep_b, rep_s

    layout, layout_entries = S.layout()
    if args.profile_kernels:
        step()
        kms = iteration_kernels(S.time_iteration(args.profile_reps))
        copy_bytes = 256 * 1024 * 1024
        ctx.calibration_copy(copy_bytes, args.profile_reps)
        m_bytes = 4 if precision == D.PRECISION_MIXED else 8
        print(json.dumps({"profile_kernels": True, "n": n_tot, "nnz": nnz_tot, "layout": layout,
                          "iteration_kernel_ms": kms,
                          "algorithmic_bytes": {k: kernel_bytes(k, n_tot, nnz_tot, m_bytes, batched) for k in kms},
                          "copy_bytes_read": copy_bytes, "copy_bytes_written": copy_bytes}))
        S.close()
        ctx.close()
        return

    tune = None
    if not args.no_autotune:
        # one-time set-up, like the graph capture: the tile shape of each
        # size bucket is chosen on this batch's own columns, then kept
        tune = S.autotune(precision, D.SCHEDULE_BALANCED_NNZ, reps=args.tune_reps)
        log(f"autotune: shapes {tune['shape']} in {tune['tune_ms']:.0f} ms, tier ms per candidate {tune['ms']}")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    spai_ms, solve_ms, iters = [], [], []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        rb, rs = step()
        spai_ms.append(rb["ms"])
        solve_ms.append(rs["ms"])
        iters.append(rs["iterations_max"])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    oc = S.outcomes()
    true_rel = [o["true_residual"] / bn for o, bn in zip(oc, bnorm)]
    converged = all(o["status"] in (D.SOLVE_CONVERGED, D.SOLVE_CONVERGED_AT_START) for o in oc)
    ms_per_step = 1e3 * elapsed / args.steps
    value = nsys * args.steps / elapsed
    log(f"batched {args.precision}: {ms_per_step:.2f} ms/step, spai {np.median(spai_ms):.3f} ms, solve "
        f"{np.median(solve_ms):.3f} ms, iters(max) {iters[-1]}, worst true rel {max(true_rel):.2e}")
    log("per-system iterations (n): " + ", ".join(f"{o['iterations']}({A.n})" for o, A in zip(oc, mats)))

    # ------------------------------------------------------- roofline ----
    # per-kernel times of the real iteration sequence (in order, every system
    # active, HIP events between the kernels on the library's stream)
    m_bytes = 4 if precision == D.PRECISION_MIXED else 8
    kernel_ms = iteration_kernels(S.time_iteration(20))
    dominant = max(kernel_ms, key=kernel_ms.get)
    kb = kernel_bytes(dominant, n_tot, nnz_tot, m_bytes, batched)
    dom_ms = launch_ms(S, dominant, kernel_ms[dominant])
    achieved = kb / (dom_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": dominant, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "bytes_per_launch": int(kb), "launch_ms": round(dom_ms, 5),
                "layout": layout, "stored_entries": layout_entries,
                "all_kernels_ms": {k: round(v, 5) for k, v in kernel_ms.items()},
                "all_kernels_frac": {k: round(kernel_bytes(k, n_tot, nnz_tot, m_bytes, batched) / (v * 1e-3) /
                                              1e9 / HBM_PEAK_GBS, 4) for k, v in kernel_ms.items()}}
    # HBM traffic per launch from the committed counter summary of this
    # workload (two rocprofv3 --pmc passes, calibrated against a copy kernel)
    traffic_file = os.path.join(ROOT, "profiles", "batched64_traffic.md")
    if args.workload == "batched" and precision == D.PRECISION_MIXED and os.path.exists(traffic_file):
        txt = open(traffic_file).read()
        blob = txt[txt.index("```json") + 7: txt.rindex("```")]
        kern = json.loads(blob)["kernels"].get(dominant)
        if kern:
            roofline["traffic"] = kern["traffic_bytes"]
            roofline["traffic_source"] = "profiles/batched64_traffic.md"
            # the same launch time against the bytes the counters saw (the
            # layout's own index data instead of CSR's)
            tgbs = kern["traffic_bytes"] / (dom_ms * 1e-3) / 1e9
            roofline["traffic_achieved"] = round(tgbs, 1)
            roofline["traffic_frac"] = round(tgbs / HBM_PEAK_GBS, 4)
    copy_ms = ctx.calibration_copy(256 * 1024 * 1024, 10)
    roofline["copy_kernel_GBs"] = round(2 * 256 * 1024 * 1024 / (copy_ms * 1e-3) / 1e9, 1)
    # whole solve: every system's iterations x its per-iteration algorithmic
    # bytes (the fused K1's count: the split form re-reads p')
    per_iter = 0.0
    for A, o in zip(mats, oc):
        it = o["iterations"] if o["status"] != D.SOLVE_MAX_ITER else o["iterations"] - 1
        per_iter += it * sum(kernel_bytes(k, A.n, A.nnz, m_bytes, batched)
                             for k in ("update_mapply_p", "spmv_alpha", "mapply_s", "spmv_omega"))
    solve_gbs = per_iter / (float(np.median(solve_ms)) * 1e-3) / 1e9
    roofline["solve_achieved"] = round(solve_gbs, 1)
    roofline["solve_frac"] = round(solve_gbs / HBM_PEAK_GBS, 4)
    log(f"roofline: {json.dumps(roofline)}")
    # the SPAI build of the last timed step against its own roofline
    spai_rf = spai_roofline(S, mats, rb, precision == D.PRECISION_MIXED, float(np.median(spai_ms)))
    log(f"spai roofline: {json.dumps(spai_rf)}")
    # work-normalised rate: rows x iterations per second of solve, which does
    # not hinge on the slowest system's iteration count
    rows_it = row_iterations(mats, oc, D.SOLVE_MAX_ITER)
    work = {"row_iterations": int(rows_it), "row_iterations_per_s": round(rows_it / (np.median(solve_ms) * 1e-3), 1),
            "iterations_sum": int(sum(o["iterations"] for o in oc))}

    # -------------------------------------- batched FP64 leg (same path) ----
    # the same batch, hipGraph and autotuned shapes, FP64 SPAI and M: the
    # mixed-precision gain alone
    fp64_leg = None
    if batched and precision == D.PRECISION_MIXED and not args.no_legs:
        S64 = D.System(ctx, mats)
        tune64 = S64.autotune(D.PRECISION_FP64, D.SCHEDULE_BALANCED_NNZ, reps=3)

        def step64():
            x_d.zero_()
            torch.cuda.synchronize()
            rb_ = S64.build(D.PRECISION_FP64, D.SCHEDULE_BALANCED_NNZ)
            rs_ = S64.solve_device(b_d.data_ptr(), x_d.data_ptr(), rel_tol=1e-8, max_iter=1000, use_graph=True,
                                   graph_iters=args.graph_iters, true_residual_restart=True)
            return rb_, rs_

        step64()
        torch.cuda.synchronize()
        sp64, so64 = [], []
        t64 = time.perf_counter()
        for _ in range(max(1, min(args.steps, 5))):
            rb64, rs64 = step64()
            sp64.append(rb64["ms"])
            so64.append(rs64["ms"])
        torch.cuda.synchronize()
        el64 = time.perf_counter() - t64
        oc64 = S64.outcomes()
        r64 = row_iterations(mats, oc64, D.SOLVE_MAX_ITER)
        k64 = iteration_kernels(S64.time_iteration(20))
        fp64_leg = {"value": round(nsys * len(sp64) / el64, 3), "unit": "solves/s",
                    "ms_per_step": round(1e3 * el64 / len(sp64), 3), "spai_ms": round(float(np.median(sp64)), 4),
                    "solve_ms": round(float(np.median(so64)), 4),
                    "iterations_max": max(o["iterations"] for o in oc64),
                    "iterations_sum": int(sum(o["iterations"] for o in oc64)),
                    "row_iterations_per_s": round(r64 / (np.median(so64) * 1e-3), 1),
                    "all_converged": all(o["status"] == D.SOLVE_CONVERGED for o in oc64),
                    "worst_true_rel_residual": float(max(o["true_residual"] / bn for o, bn in zip(oc64, bnorm))),
                    "all_kernels_ms": {k: round(v, 5) for k, v in k64.items()},
                    "tile_shapes": tune64["shape"],
                    "spai_roofline": spai_roofline(S64, mats, rb64, False, float(np.median(sp64)))}
        log(f"batched fp64 + graph: {json.dumps(fp64_leg)}")
        S64.close()

    # ------------------------------------------------- plain FP64 path ----
    plain = None
    if not args.no_plain:
        systems = [D.System(ctx, A) for A in mats]
        bs = [torch.from_numpy(r).to(dev) for r in rhs]
        xs = [torch.zeros(A.n, dtype=torch.float64, device=dev) for A in mats]
        torch.cuda.synchronize()

        def plain_step():
            tot_b = tot_s = 0.0
            for Sk, bk, xk in zip(systems, bs, xs):
                xk.zero_()
                torch.cuda.synchronize()
                rb = Sk.build(D.PRECISION_FP64, D.SCHEDULE_BALANCED_NNZ)
                rs = Sk.solve_device(bk.data_ptr(), xk.data_ptr(), rel_tol=1e-8, max_iter=1000, use_graph=False,
                                     check_every=1, true_residual_restart=True)
                tot_b += rb["ms"]
                tot_s += rs["ms"]
            return tot_b, tot_s

        plain_step()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        pb, ps = [], []
        for _ in range(args.plain_steps):
            a, b_ = plain_step()
            pb.append(a)
            ps.append(b_)
        torch.cuda.synchronize()
        pel = time.perf_counter() - t1
        ptrue = max(Sk.outcomes()[0]["true_residual"] / bn for Sk, bn in zip(systems, bnorm))
        plain = {"value": round(nsys * args.plain_steps / pel, 3), "unit": "solves/s",
                 "ms_per_step": round(1e3 * pel / args.plain_steps, 3), "spai_ms": round(float(np.median(pb)), 3),
                 "solve_ms": round(float(np.median(ps)), 3), "worst_true_rel_residual": float(ptrue),
                 "steps": args.plain_steps}
        prow = row_iterations(mats, [Sk.outcomes()[0] for Sk in systems], D.SOLVE_MAX_ITER)
        plain["row_iterations_per_s"] = round(prow / (np.median(ps) * 1e-3), 1)
        log(f"plain fp64: {json.dumps(plain)}")
        for Sk in systems:
            Sk.close()

    # --------------------------------------------------- CPU baseline ----
    cpu = None
    if not args.no_cpu:
        from oracle import binding as O

        def cpu_solve(A, r):
            m = O.spai(A.n, A.colptr, A.rowidx, A.val, 0)
            Acsr = O.csc2csr(A.n, A.colptr, A.rowidx, A.val)[:3]
            Mcsr = O.csc2csr(A.n, A.colptr, A.rowidx, m)[:3]
            O.bicgstab(Acsr, Mcsr, r, np.zeros(A.n), 1e-8 * np.linalg.norm(r), 1000)

        # one untimed solve starts the thread pool; then the systems in turn,
        # cycling, until the budget is spent
        small = min(range(nsys), key=lambda s_: mats[s_].n)
        cpu_solve(mats[small], rhs[small])
        done, t_cpu = 0, 0.0
        while t_cpu < args.cpu_budget:
            A, r = mats[done % nsys], rhs[done % nsys]
            t2 = time.perf_counter()
            cpu_solve(A, r)
            t_cpu += time.perf_counter() - t2
            done += 1
        unknowns = sum(mats[k % nsys].n for k in range(done))
        cpu = {"value": round(done / t_cpu, 4), "unit": "solves/s", "cores": O.num_threads(), "kind": "port",
               "sample": f"{done} solves cycling over the {nsys} system(s) of the same workload ({unknowns} "
                         f"unknowns in all), oracle MGS SPAI + BiCGSTAB (FP64) on {cpu_model()}, "
                         f"{t_cpu:.1f}s"}
        log(f"cpu baseline: {json.dumps(cpu)}")

    # ------------------------------------------- single-workload legs ----
    legs = None
    if batched and not args.no_legs:
        from oracle import binding as O_

        with_cpu = not args.no_cpu
        legs = {
            "random4096": run_leg(D, O_, ctx, torch, "random", D.PRECISION_FP64, D.SCHEDULE_BALANCED_COLUMNS,
                                  args.leg_steps, args.leg_cpu_budget, with_cpu),
            "convdiff512": run_leg(D, O_, ctx, torch, "convdiff", D.PRECISION_FP64, D.SCHEDULE_BALANCED_NNZ,
                                   args.leg_steps, args.leg_cpu_budget, with_cpu),
            "powerlaw1e6": run_leg(D, O_, ctx, torch, "powerlaw", D.PRECISION_MIXED, D.SCHEDULE_BALANCED_NNZ,
                                   max(1, args.leg_steps - 1), args.leg_cpu_budget, with_cpu, cpu_cap=50),
            "grf64": run_leg(D, O_, ctx, torch, "grf", D.PRECISION_FP64, D.SCHEDULE_BALANCED_NNZ, args.leg_steps,
                             args.leg_cpu_budget, with_cpu, adjoint=True),
        }

    line = {
        "metric": "SPAI build ms + BiCGSTAB solves/sec, mixed/batched vs plain FP64 at 1e-8 resid",
        "value": round(value, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if precision == D.PRECISION_MIXED else "f64",
        "data": "synthetic (seeded generators)",
        "config": {"workload": desc, "systems": nsys, "unknowns": n_tot, "nnz": nnz_tot,
                   "precision": args.precision, "schedule": "balanced_nnz", "graph_iters": args.graph_iters,
                   "rel_tol": 1e-8, "seed": args.seed, "device": dev_name, "cus": cus, "layout": layout},
        "spai_ms": round(float(np.median(spai_ms)), 4),
        "spai_phase_ms": {k: round(v, 4) for k, v in rb["phase_ms"].items()},
        "solve_ms": round(float(np.median(solve_ms)), 4),
        "iterations_max": int(iters[-1]),
        "all_converged": bool(converged),
        "worst_true_rel_residual": float(max(true_rel)),
        "work": work,
        "plain_fp64": plain,
        "speedup_vs_plain_fp64": (round(value / plain["value"], 3) if plain else None),
        "batched_fp64_graph": fp64_leg,
        "speedup_vs_batched_fp64": (round(value / fp64_leg["value"], 3) if fp64_leg else None),
        "roofline": roofline,
        "spai_roofline": spai_rf,
        "cpu_baseline": cpu,
        "autotune": ({"tile_shapes": [D.tile_shape(t, c) for t, c in enumerate(tune["shape"])],
                      "tier_ms_per_candidate": [[None if v is None else round(v, 4) for v in r] for r in tune["ms"]],
                      "tune_ms": round(tune["tune_ms"], 1)} if tune else None),
        "legs": legs,
    }
    S.close()
    ctx.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
