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
    row pointer, and each FP64 vector read or written once.  The solver's
    SELL-64 layouts store less index data (16-bit offsets, or one offset per
    column block of a slice) and some padding; the measured traffic is the
    `traffic` figure."""
    rp = 4 * (n + 1)
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
        return 12 * nnz + rp + 8 * n * 3
    if kind == "mapply_s":        # z = M s; reads M, r, a; writes s, z
        return (m_bytes + 4) * nnz + rp + 8 * n * 4
    if kind == "spmv_omega":      # t = A z; reads A, z, s, r*; writes t
        return 12 * nnz + rp + 8 * n * 4
    raise ValueError(kind)


def iteration_kernels(ms5):
    """Names the five slots of System.time_iteration for the K1 form used."""
    if ms5[1] > 0.0:  # split K1
        names = ("update", "mapply_p", "spmv_alpha", "mapply_s", "spmv_omega")
        return dict(zip(names, ms5))
    return dict(zip(("update_mapply_p", "spmv_alpha", "mapply_s", "spmv_omega"), [ms5[0]] + ms5[2:]))


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
    achieved = kb / (kernel_ms[dominant] * 1e-3) / 1e9
    roofline = {"bound": "hbm", "kernel": dominant, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "bytes_per_launch": int(kb), "launch_ms": round(kernel_ms[dominant], 5),
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
        "solve_ms": round(float(np.median(solve_ms)), 4),
        "iterations_max": int(iters[-1]),
        "all_converged": bool(converged),
        "worst_true_rel_residual": float(max(true_rel)),
        "plain_fp64": plain,
        "speedup_vs_plain_fp64": (round(value / plain["value"], 3) if plain else None),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "autotune": ({"tile_shapes": [D.tile_shape(t, c) for t, c in enumerate(tune["shape"])],
                      "tier_ms_per_candidate": [[None if v is None else round(v, 4) for v in r] for r in tune["ms"]],
                      "tune_ms": round(tune["tune_ms"], 1)} if tune else None),
    }
    S.close()
    ctx.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
