#!/usr/bin/env python3
"""Benchmark of the supernodal factor + solve on one MI355X.

Workload (one "step" = one numeric factorization of the HBM-resident matrix
followed by one solve with a resident right-hand side):
  3D 7-point variable-coefficient diffusion on a 48^3 grid (n = 110592),
  face coefficients 10**U(-2,2) from seed 0, diagonal = sum of incident
  coefficients + 1e-3, nested-dissection ordering (NDBOX), FP64 LL^T,
  symPACKOptions defaults (relaxation 8/32/64, supernode width <= 150), nrhs=1.
value = factor GFLOP/s with the reference's flop count sum_j cc_j^2 over the
unrelaxed column counts.  The symbolic analysis and the upload of A and b to
HBM happen before the timed region.

Printed JSON line adds: solves/s and p50/p99 solve latency, the Schur-update
kernel's roofline (HIP-event timing of every launch inside the timed region)
and a CPU baseline: the oracle (single core C restatement of the same
algorithm, same ordering and supernodes) timed on this host.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_MFMA_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0


def progress(msg):
    print("bench: " + msg, file=sys.stderr, flush=True)


def build_problem(k):
    import sympack_amd as sp
    n, cp, ri, v = sp.matrices.diffusion_3d_7pt(k, seed=0)
    return sp.DistSparseMatrix(n, cp, ri, v)


def precision_leg(k, nrhs, reps):
    """FP32 factor + FP64 iterative refinement against the FP64 factor on the
    27-point 60^3 stencil: time to solution (factor + solve of nrhs Gaussian
    right-hand sides, inputs resident in HBM), backward errors, refinement
    steps and the forward difference between the two solutions."""
    import numpy as np
    import torch

    import sympack_amd as sp
    from oracle import numeric as onum

    n, cp, ri, v = sp.matrices.stencil_3d_27pt(k, seed=0)
    A = sp.DistSparseMatrix(n, cp, ri, v)
    dev = torch.device("cuda:0")
    vals = torch.from_numpy(A.nzval).to(dev)
    rng = np.random.default_rng(1)
    B = rng.standard_normal((n, nrhs))
    b = torch.from_numpy(np.ascontiguousarray(B.ravel(order="F"))).to(dev)
    out = {}
    X = {}
    for prec in ("FP64", "FP32_IR"):
        rp = sp.RelaxationParameters()
        rp.SetMaxSize(150)
        M = sp.symPACKMatrix2D()
        M.Init(sp.symPACKOptions(ordering="NDBOX", relax=rp, precision=prec, ir_tol=1e-12, autotune=True))
        M.SymbolicFactorization(A)
        torch.cuda.synchronize()
        M.DistributeMatrixDevice(vals.data_ptr())
        M.Factorize()
        M.SolveDevice(b.data_ptr(), nrhs)
        torch.cuda.synchronize()
        tf, ts = [], []
        for _ in range(reps):
            a = time.perf_counter()
            M.Factorize()
            c = time.perf_counter()
            M.SolveDevice(b.data_ptr(), nrhs)
            torch.cuda.synchronize()
            d = time.perf_counter()
            tf.append((c - a) * 1e3)
            ts.append((d - c) * 1e3)
        x = torch.empty(n * nrhs, dtype=torch.float64, device=dev)
        M.GetSolutionDevice(x.data_ptr(), nrhs)
        X[prec] = x.cpu().numpy().reshape(nrhs, n).T
        st = M.stats()
        key = "fp64" if prec == "FP64" else "fp32_ir"
        out[key + "_ms"] = round(float(np.median(np.add(tf, ts))), 3)
        out[key + "_factor_ms"] = round(float(np.median(tf)), 3)
        out[key + "_solve_ms"] = round(float(np.median(ts)), 3)
        out[key + "_backward_error"] = onum.backward_error(A, X[prec], B)
        if prec == "FP32_IR":
            out["ir_iterations"] = st["ir_iterations"]
            out["ir_converged"] = bool(st["ir_converged"])
        out["flops_sum_cc2"] = st["flops"]
        del M
    out["speedup_fp32_ir_vs_fp64"] = round(out["fp64_ms"] / out["fp32_ir_ms"], 3)
    out["forward_diff_fp32_ir_vs_fp64"] = float(np.linalg.norm(X["FP32_IR"] - X["FP64"]) /
                                                np.linalg.norm(X["FP64"]))
    out["workload"] = ("3D 27-point SPD stencil %d^3 (n=%d), NDBOX, nrhs=%d Gaussian, median of %d "
                       "factor+solve runs" % (k, n, nrhs, reps))
    return out


def ldl_leg(k, nrhs, reps):
    """Static LDL^T of the shifted 7-point Laplacian L_h - 0.3 I on a k^3 grid
    (symmetric indefinite): FP64 LDL^T against FP32 LDL^T + FP64 iterative
    refinement, nrhs uniform[-1,1] right-hand sides resident in HBM.  Forward
    difference against the FP64 factorization's solution and backward errors
    of both."""
    import numpy as np
    import torch

    import sympack_amd as sp
    from oracle import numeric as onum

    n, cp, ri, v = sp.matrices.shifted_laplacian_3d_7pt(k, 0.3)
    A = sp.DistSparseMatrix(n, cp, ri, v)
    dev = torch.device("cuda:0")
    vals = torch.from_numpy(A.nzval).to(dev)
    rng = np.random.default_rng(2)
    B = rng.uniform(-1.0, 1.0, (n, nrhs))
    b = torch.from_numpy(np.ascontiguousarray(B.ravel(order="F"))).to(dev)
    out, X = {}, {}
    for prec in ("FP64", "FP32_IR"):
        rp = sp.RelaxationParameters()
        rp.SetMaxSize(150)
        M = sp.symPACKMatrix2D()
        M.Init(sp.symPACKOptions(ordering="NDBOX", decomposition="LDL", relax=rp, precision=prec,
                                 ir_tol=1e-12, autotune=True))
        M.SymbolicFactorization(A)
        torch.cuda.synchronize()
        M.DistributeMatrixDevice(vals.data_ptr())
        M.Factorize()
        M.SolveDevice(b.data_ptr(), nrhs)
        torch.cuda.synchronize()
        tf, ts = [], []
        for _ in range(reps):
            a = time.perf_counter()
            M.Factorize()
            c = time.perf_counter()
            M.SolveDevice(b.data_ptr(), nrhs)
            torch.cuda.synchronize()
            d = time.perf_counter()
            tf.append((c - a) * 1e3)
            ts.append((d - c) * 1e3)
        x = torch.empty(n * nrhs, dtype=torch.float64, device=dev)
        M.GetSolutionDevice(x.data_ptr(), nrhs)
        X[prec] = x.cpu().numpy().reshape(nrhs, n).T
        st = M.stats()
        key = "fp64" if prec == "FP64" else "fp32_ir"
        out[key + "_factor_ms"] = round(float(np.median(tf)), 3)
        out[key + "_solve_ms"] = round(float(np.median(ts)), 3)
        out[key + "_factor_gflops"] = round(st["flops"] / (np.median(tf) * 1e-3) / 1e9, 1)
        out[key + "_backward_error"] = onum.backward_error(A, X[prec], B)
        if prec == "FP32_IR":
            out["ir_iterations"] = st["ir_iterations"]
            out["ir_converged"] = bool(st["ir_converged"])
        out["flops_sum_cc2"] = st["flops"]
        del M
    out["forward_diff_fp32_ir_vs_fp64"] = float(np.linalg.norm(X["FP32_IR"] - X["FP64"]) /
                                                np.linalg.norm(X["FP64"]))
    out["workload"] = ("shifted 3D 7-point Laplacian L_h - 0.3 I, %d^3 (n=%d), NDBOX, static LDL^T, "
                       "nrhs=%d uniform[-1,1], median of %d factor+solve runs" % (k, n, nrhs, reps))
    return out


def serving_leg(n_requests, rate, precision, slots, seed=0):
    """Request mix through the serving layer: new factors of 2D/3D
    random-coefficient Laplacians (n log-uniform in [2e3, 2e5]) and follow-up
    solves (nrhs 1..32) on resident factors, Poisson arrivals; returns the
    latency percentiles and throughput."""
    import numpy as np

    import sympack_amd as sp
    from sympack_amd.serving import SolverService, make_request_stream

    stream = make_request_stream(n_requests, rate, seed=seed)
    mats, rhs = {}, []
    for t, kind, key, p in stream:
        if kind == "factor":
            dim, k, s = p
            n, cp, ri, v = sp.matrices.random_laplacian("2d" if dim == 2 else "3d", k, s)
            mats[key] = (dim, sp.DistSparseMatrix(n, cp, ri, v))
    for t, kind, key, p in stream:
        if kind == "solve":
            rhs.append(np.random.default_rng(p[1]).standard_normal((mats[key][1].size, p[0])))

    def opts(dim):
        rp = sp.RelaxationParameters()
        rp.SetMaxSize(150)
        return sp.symPACKOptions(ordering="NDGRID" if dim == 2 else "NDBOX", relax=rp, precision=precision)

    # HBM budget of the resident factors: 70% of the device memory free now
    import torch
    budget = int(0.7 * torch.cuda.mem_get_info()[0])
    futs = []
    with SolverService(hbm_budget_bytes=budget, slots=slots) as svc:
        t0 = time.perf_counter()
        j = 0
        for t, kind, key, p in stream:
            dt = t0 + t - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            if kind == "factor":
                futs.append(svc.submit_factor(key, mats[key][1], opts(mats[key][0])))
            else:
                futs.append(svc.submit_solve(key, rhs[j]))
                j += 1
        for f in futs:
            f.result()
        rep = svc.report()
    nf = sum(1 for r in stream if r[1] == "factor")
    rep["requests"] = n_requests
    rep["factor_requests"] = nf
    rep["arrival_rate_per_s"] = rate
    rep["slots"] = slots
    rep["solve_slots"] = 1 if slots > 1 else 0
    rep["hbm_budget_gb"] = round(budget / 1e9, 1)
    rep["precision"] = precision
    return rep


def level_table(M, n, reps, title, cmd):
    """Per-etree-level breakdown of the factorization and the solve from the
    per-launch event trace (host-launched, every kernel kind timed): number
    of supernodes, Schur-update tiles, algorithmic flops, block factor +
    TRSM time, Schur time and rate, forward/backward solve time.  Returns
    markdown text."""
    import numpy as np

    xs = np.asarray(M.GetSupernodalPartition())
    xl, _ = M.GetStructure()
    w = np.diff(xs).astype(np.float64)
    r = np.diff(np.asarray(xl)).astype(np.float64)
    m = r - w
    lv = M.supernode_levels()
    nl = int(lv.max()) + 1
    schur_fl = np.bincount(lv, weights=w * m * (m + 1), minlength=nl)
    diag_fl = np.bincount(lv, weights=w ** 3 / 3 + w * w * m, minlength=nl)
    nsn = np.bincount(lv, minlength=nl)
    M.reset_kernel_stats()
    M.set_profile(True)
    for _ in range(reps):
        M.Factorize()
        M.Solve(np.ones(n), 1)
    M.set_profile(False)
    kinds, grids, ms = M.launch_trace()
    levels = M.launch_levels()
    t = {k: np.zeros(nl) for k in ("blk", "schur", "fwd", "bwd")}
    tiles = np.zeros(nl)
    for k, g, x, L in zip(kinds, grids, ms, levels):
        if L < 0:
            continue
        if k in (0, 1, 7, 9):
            t["blk"][L] += x
        elif k == 2:
            t["schur"][L] += x
            tiles[L] += g
        elif k in (3, 4, 10):
            t["fwd"][L] += x
        elif k in (5, 6, 11):
            t["bwd"][L] += x
    for k in t:
        t[k] /= reps
    tiles /= reps
    lines = ["# %s" % title, "", "Command (one MI355X, `R` = repository root):", "", "```", cmd, "```", "",
             "Host-launched factorizations and nrhs = 1 solves with HIP events around every launch "
             "(%d repetitions, averaged); times in microseconds per factorization / solve.  "
             "Flops are algorithmic: Schur w m (m+1), block factor + TRSM w^3/3 + w^2 m per supernode "
             "(w columns, m rows below the diagonal block)." % reps, "",
             "| level | supernodes | Schur tiles | Schur GFLOP | block us | Schur us | Schur TFLOP/s | "
             "block+Schur us | fwd solve us | bwd solve us |",
             "|---:|---:|---:|---:|---:|---:|---:|---:|---:|---:|"]
    tot = {k: 0.0 for k in ("blk", "schur", "fwd", "bwd")}
    for L in range(nl):
        rate = schur_fl[L] / (t["schur"][L] * 1e-3) / 1e12 if t["schur"][L] > 0 else 0.0
        lines.append("| %d | %d | %d | %.3f | %.1f | %.1f | %.1f | %.1f | %.1f | %.1f |" % (
            L, nsn[L], tiles[L], schur_fl[L] / 1e9, t["blk"][L] * 1e3, t["schur"][L] * 1e3, rate,
            (t["blk"][L] + t["schur"][L]) * 1e3, t["fwd"][L] * 1e3, t["bwd"][L] * 1e3))
        for k in tot:
            tot[k] += t[k][L]
    lines.append("| all | %d | %d | %.3f | %.1f | %.1f | %.1f | %.1f | %.1f | %.1f |" % (
        nsn.sum(), tiles.sum(), schur_fl.sum() / 1e9, tot["blk"] * 1e3, tot["schur"] * 1e3,
        schur_fl.sum() / (tot["schur"] * 1e-3) / 1e12 if tot["schur"] > 0 else 0.0,
        (tot["blk"] + tot["schur"]) * 1e3, tot["fwd"] * 1e3, tot["bwd"] * 1e3))
    lines.append("")
    lines.append("Block factor + TRSM flops in all: %.3f GFLOP." % (diag_fl.sum() / 1e9))
    return "\n".join(lines) + "\n"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--k", type=int, default=48, help="grid edge (48 -> n=110592)")
    ap.add_argument("--solves", type=int, default=50, help="extra solve-only calls for latency")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-precision", action="store_true",
                    help="skip the FP32+IR vs FP64 comparison on the 60^3 27-point stencil")
    ap.add_argument("--precision-k", type=int, default=60)
    ap.add_argument("--no-ldl", action="store_true",
                    help="skip the static LDL^T leg on the 80^3 shifted Laplacian")
    ap.add_argument("--ldl-k", type=int, default=80)
    ap.add_argument("--serving", type=int, default=2000, metavar="N",
                    help="requests of the serving stream run after the headline (FP64 and FP32+IR); "
                         "0 skips it")
    ap.add_argument("--serving-rate", type=float, default=200.0)
    ap.add_argument("--serving-slots", type=int, default=4)
    ap.add_argument("--level-table", metavar="PATH",
                    help="write the per-etree-level factor/solve breakdown (markdown) of the workload "
                         "selected by --level-matrix and exit")
    ap.add_argument("--level-matrix", default="diffusion48",
                    choices=["diffusion48", "stencil27_60", "shifted80"],
                    help="workload of --level-table: the 48^3 diffusion headline, the 60^3 27-point "
                         "stencil or the 80^3 shifted Laplacian (LDL^T)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs the factorization on one GPU only; "
              "multi-GPU runs (--gpus %d) are not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import sympack_amd as sp

    if args.level_table:
        gens = {"diffusion48": ("3D 7-point diffusion 48^3 (n=110592), NDBOX, FP64 LL^T",
                                lambda: sp.matrices.diffusion_3d_7pt(48, seed=0), "LL"),
                "stencil27_60": ("3D 27-point SPD stencil 60^3 (n=216000), NDBOX, FP64 LL^T",
                                 lambda: sp.matrices.stencil_3d_27pt(60, seed=0), "LL"),
                "shifted80": ("shifted 3D 7-point Laplacian 80^3 (n=512000), NDBOX, FP64 static LDL^T",
                              lambda: sp.matrices.shifted_laplacian_3d_7pt(80, 0.3), "LDL")}
        title, gen, fact = gens[args.level_matrix]
        n, cp, ri, v = gen()
        A = sp.DistSparseMatrix(n, cp, ri, v)
        rp = sp.RelaxationParameters()
        rp.SetMaxSize(150)
        M = sp.symPACKMatrix2D()
        M.Init(sp.symPACKOptions(ordering="NDBOX", decomposition=fact, relax=rp, autotune=True))
        M.SymbolicFactorization(A)
        M.DistributeMatrix(A)
        M.Factorize()
        progress("level table: " + title)
        cmd = "python3 $R/bench.py --level-table %s --level-matrix %s" % (args.level_table, args.level_matrix)
        text = level_table(M, n, 5, "Per-level factor and solve times: " + title, cmd)
        with open(args.level_table, "w") as f:
            f.write(text)
        print(text)
        return

    progress("building %d^3 diffusion matrix" % args.k)
    A = build_problem(args.k)
    n = A.size
    rp = sp.RelaxationParameters()
    rp.SetMaxSize(150)
    opts = sp.symPACKOptions(ordering="NDBOX", decomposition="LL", relax=rp, precision="FP64", autotune=True)
    M = sp.symPACKMatrix2D()
    M.Init(opts)
    t0 = time.perf_counter()
    M.SymbolicFactorization(A)
    t_symb = time.perf_counter() - t0

    dev = torch.device("cuda:0")
    vals = torch.from_numpy(A.nzval).to(dev)
    rng = np.random.default_rng(0)
    x_true = rng.uniform(-1.0, 1.0, n)
    b_host = A.matvec(x_true)
    b = torch.from_numpy(b_host).to(dev)
    x = torch.empty(n, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    M.DistributeMatrixDevice(vals.data_ptr())
    st = M.stats()
    flops = st["flops"]

    progress("warmup")
    for _ in range(args.warmup):
        M.Factorize()
        M.SolveDevice(b.data_ptr(), 1)

    # the autotuner ran on the first (warmup) factorization: per-level Schur
    # launch configurations timed on the device
    tl = M.autotune_log()
    autotune = {"levels": len(tl), "changed": sum(1 for r in tl if r["ms_best"] < r["ms_default"]),
                "schur_ms_default": round(sum(r["ms_default"] for r in tl), 4),
                "schur_ms_tuned": round(sum(r["ms_best"] for r in tl), 4),
                "choices": [[r["level"], r["split"], r["order"]] for r in tl]} if tl else None

    # timed region 1: K steps replayed from the captured HIP graphs
    M.reset_kernel_stats()
    torch.cuda.synchronize()
    fac_ms, sol_ms = [], []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        a = time.perf_counter()
        M.Factorize()
        c = time.perf_counter()
        M.SolveDevice(b.data_ptr(), 1)
        d = time.perf_counter()
        fac_ms.append((c - a) * 1e3)
        sol_ms.append((d - c) * 1e3)
    torch.cuda.synchronize()
    t_total = time.perf_counter() - t_start
    launches = {k: M.kernel_stats(k)[0] for k in range(13)}

    # timed region 2: K factorizations launched from the host with HIP events
    # around every Schur-update launch (kind 2) on the library's stream, for
    # the kernel's roofline (events cannot be timed inside a graph replay)
    progress("profiled factorizations")
    M.reset_kernel_stats()
    M.set_profile(True, kinds=[2])
    prof_ms = []
    for _ in range(args.steps):
        a = time.perf_counter()
        M.Factorize()
        prof_ms.append((time.perf_counter() - a) * 1e3)
    M.set_profile(False)
    kinds = {k: M.kernel_stats(k) for k in range(13)}

    # solve latency distribution (factor resident, repeated solves)
    lat = []
    for _ in range(args.solves):
        a = time.perf_counter()
        M.SolveDevice(b.data_ptr(), 1)
        lat.append((time.perf_counter() - a) * 1e3)
    M.GetSolutionDevice(x.data_ptr(), 1)
    torch.cuda.synchronize()
    xh = x.cpu().numpy()
    fwd_err = float(np.linalg.norm(xh - x_true) / np.linalg.norm(x_true))
    from oracle import numeric as onum
    berr = onum.backward_error(A, xh, b_host)

    ms_per_step = t_total * 1e3 / args.steps
    fac_mean = float(np.mean(fac_ms))
    value = flops / (fac_mean * 1e-3) / 1e9

    sl, sms = kinds[2]
    schur_launches_per_step = sl / args.steps
    schur_avg_ms = sms / max(1, sl)
    schur_flops_per_launch = st["schur_flops"] / max(1, schur_launches_per_step)
    achieved_tflops = schur_flops_per_launch / (schur_avg_ms * 1e-3) / 1e12 if sms > 0 else None

    # HBM bytes per Schur launch from the rocprofv3 FETCH_SIZE / WRITE_SIZE
    # passes (scaled by a plain copy of known size) committed under profiles/
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "schur_update_traffic_48cube.json")
    if args.k == 48 and os.path.exists(tpath):
        traffic = json.load(open(tpath))["traffic_bytes_per_launch"]

    cpu = None
    if not args.no_cpu_baseline:
        # the same supernodal LL^T on the host: oracle/supernodal.c's
        # BLAS-backed restatement (per supernode dpotrf, dtrsm, dsyrk of
        # scipy's OpenBLAS, then the relative-index scatter), same ordering,
        # supernodes and structure; assembly not timed, as DistributeMatrix
        # is not part of the reference's factorization time
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
        threads = min(threads, os.cpu_count() or 1)
        S = onum.Structure.of(M)
        progress("cpu baseline: BLAS-backed factor, 1 thread")
        info, F, t1 = onum.factor_blas(S, A, 1)
        progress("cpu baseline: BLAS-backed factor, %d threads x3" % threads)
        tn = []
        for _ in range(3):
            info, F, t = onum.factor_blas(S, A, threads)
            tn.append(t)
        tn_med = float(np.median(tn))
        Xc = onum.solve(S, F, b_host)
        cpu = {"value": round(flops / tn_med / 1e9, 3), "unit": "GFLOP/s", "cores": threads, "kind": "port",
               "value_1core": round(flops / t1 / 1e9, 3),
               "factor_s": round(tn_med, 3), "factor_s_1core": round(t1, 3),
               "forward_diff_vs_gpu": float(np.linalg.norm(Xc - xh) / np.linalg.norm(xh)),
               "sample": "whole factorization of the same %d^3 matrix (same ordering, supernodes, "
                         "structure) by oracle/supernodal.c's BLAS-backed LL^T (per-supernode dpotrf / "
                         "dtrsm / dsyrk of scipy's OpenBLAS + relative-index scatter): %d threads, "
                         "median of 3 = %.2f s; 1 thread, one run = %.2f s" %
                         (args.k, threads, tn_med, t1)}

    progress("precision leg")
    prec = None if args.no_precision else precision_leg(args.precision_k, 16, 3)
    progress("ldl leg")
    ldl = None if args.no_ldl else ldl_leg(args.ldl_k, 64, 3)
    serving = None
    if args.serving:
        progress("serving leg")
        serving = {}
        for p in ("FP64", "FP32_IR"):
            progress("serving leg %s: %d requests at %.0f/s" % (p, args.serving, args.serving_rate))
            serving[p.lower()] = serving_leg(args.serving, args.serving_rate, p, args.serving_slots)
        serving["workload"] = ("%d requests: 2D 5-point / 3D 7-point random-coefficient Laplacians, n "
                               "log-uniform in [2e3, 2e5], 70%% follow-up solves with 1-32 right-hand sides, "
                               "Poisson arrivals at %.0f/s, %d slots" %
                               (args.serving, args.serving_rate, args.serving_slots))

    out = {
        "metric": "factor GFLOP/s; solves/s + p99 latency, FP32+IR vs FP64 at bwd err <=1e-12",
        "value": round(value, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "3D 7-pt variable-coefficient diffusion %d^3 (n=%d), NDBOX ordering, "
                               "FP64 LL^T factor + solve nrhs=1" % (args.k, n),
                   "n": n, "nnz_lower": A.nnz, "ordering": "NDBOX", "decomposition": "LL",
                   "relax": "8/32/64, max width 150", "autotune": True, "nsuper": st["nsuper"],
                   "etree_levels": st["nlevels"], "flops_sum_cc2": flops, "nnz_L": st["nnz_l"]},
        "factor_ms": round(fac_mean, 3),
        "factor_ms_min": round(float(np.min(fac_ms)), 3),
        "solve_ms_in_step": round(float(np.mean(sol_ms)), 3),
        "solves_per_s": round(1e3 / float(np.mean(lat)), 1),
        "solve_p50_ms": round(float(np.percentile(lat, 50)), 3),
        "solve_p99_ms": round(float(np.percentile(lat, 99)), 3),
        "symbolic_s": round(t_symb, 3),
        "forward_error": fwd_err,
        "backward_error": berr,
        "factor_ms_eager_profiled": round(float(np.mean(prof_ms)), 3),
        "launches_per_step": {name: launches[k] / args.steps
                              for k, name in [(0, "potrf"), (7, "trsm"), (9, "potrf_trsm_fused"),
                                              (1, "inpanel_update"),
                                              (2, "schur_update"), (3, "fwd_diag"), (4, "fwd_offdiag"),
                                              (5, "bwd_offdiag"), (6, "bwd_diag"),
                                              (10, "fwd_partitioned_inverse"), (11, "bwd_partitioned_inverse"),
                                              (12, "partitioned_inverse_prep")]},
        "roofline": {"kernel": "Schur-update launches (FP64 MFMA + atomic scatter; k_update<double,1>, "
                               "k_schur_w1 or k_schur_wide per etree level as autotuned)",
                     "timing": "HIP events around each Schur-update launch, %d host-launched "
                               "factorizations after the graph-replayed throughput loop" % args.steps,
                     "bound": "mfma",
                     "achieved": round(achieved_tflops, 3) if achieved_tflops else None,
                     "peak": FP64_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved_tflops / FP64_MFMA_PEAK_TFLOPS, 4) if achieved_tflops else None,
                     "traffic": traffic,
                     "traffic_source": "profiles/schur_update_hbm_traffic_48cube.md" if traffic else None,
                     "flops_per_launch": schur_flops_per_launch,
                     "avg_launch_ms": schur_avg_ms},
        "autotune": autotune,
        "cpu_baseline": cpu,
        "precision_fp32_ir_vs_fp64": prec,
        "ldl_indefinite": ldl,
        "serving": serving,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
