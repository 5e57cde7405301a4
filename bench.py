"""Benchmark: Jacobian+Schur observations/s and LM time-to-tolerance on one MI355X.

Workload (one "step"): one Jacobian + Schur pass of the bundle-adjustment
engine over the 1,000-camera / 500,000-point / ~4.0M-observation scene with
5% 20-px outliers under a 4-px Huber loss, mixed precision (FP32 residuals,
Jacobians and block products; FP64 point-block inverses):
residuals, analytic Jacobian (Huber-corrected, Jacobi-scaled), V_p, its FP64
inverse, g_p, U_c, g_c and the reduced right-hand side.  Inputs are resident
in HBM when the timed region starts.

Also reported:
  * the same pass in FP64;
  * the LM solve (implicit-Schur PCG) in mixed precision and in FP64: time
    until the RMS reprojection error (from the FP64-evaluated cost) is within
    1e-6 relative of the FP64 path's final RMS after its iteration cap, and
    the stricter time until the cost itself is within 1e-6; the mixed run may
    take 1.5x the iterations; the same with SCHUR_JACOBI preconditioning, and
    both preconditioners solved to convergence on the 697k-observation scene;
  * the explicit-Schur LM variant on the same scene (dense 9,000 x 9,000
    reduced camera matrix, FP32 MFMA Cholesky, FP64 refinement): per-phase
    times of one solve and the factorisation's MFMA rate;
  * batched local BA: 512 independent windows (5-40 cameras, 200-5,000
    points) solved by one mba_solve_batched call, windows/s;
  * the 10,000-camera / 4M-point / ~25M-observation scene: the pass and the
    LM solve with block-Jacobi PCG to relative residual 1e-6;
  * the dominant kernel's roofline (HIP events on the engine's stream);
  * the grouped-intrinsics scene (the incremental engine's residual blocks);
  * the CPU baseline: the reference's Ceres Jacobian + Schur pass and the FP64
    oracle's on the host cores (OpenMP), on the same scene.

    python bench.py [--steps K] [--warmup W] [--gpus 1]

Each timed pass loop runs exactly K passes after at least W untimed passes and
at least 0.5 s of them (WARM_S), so that it measures the GPU at its running
clocks rather than its ramp from idle.
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
HUBER_A = 4.0
RADIUS = 1e4
# Warm-up before a timed pass loop: the requested warm-up steps and at least
# this much wall time of passes (the GPU's clocks ramp up over the first tens
# of milliseconds of back-to-back passes: ~6 % slower passes at the start).
WARM_S = 0.5


def warm_up(ctx, steps, torch):
    """Untimed passes: at least `steps`, and for at least WARM_S seconds."""
    t_end = time.perf_counter() + WARM_S
    k = 0
    while k < steps or time.perf_counter() < t_end:
        ctx.jacobian_schur(RADIUS)
        k += 1
        if k % 16 == 0:
            torch.cuda.synchronize()
    torch.cuda.synchronize()
# LM leg: CG forcing and function tolerance tightened from the Ceres defaults
# (0.1, 1e-6) so that both precisions run on toward the minimum instead of
# stopping on slow progress (5% outliers under Huber converge slowly).
LM_ETA = 1e-2
LM_FTOL = 1e-12
RMS_TOL = 1e-6


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cam_pass_bytes(nrows, ne, nf, nchunks, elem):
    """Algorithmic bytes of one camera-pass launch: per camera-major row the
    point index (4 B), observation (2 elem) and loss weight (1 elem); per
    point block its position, Jacobi scale and V^-1 g_p (3 elem each) read
    once; per camera its table row and scale; per chunk 63 partial sums
    written."""
    return (nrows * (4 + 3 * elem) + ne * 9 * elem + nf * (48 + 9) * elem + nchunks * 63 * elem)


def point_pass_bytes(nrows, ne, nc, elem):
    """Per row camera id (4 B), observation (2 elem) and loss weight (1 elem);
    per point start (4 B), FP64 position (24 B) and scale (3 elem) read, and
    V and V^-1 (FP64, 48 B each), g_p (3 elem) and the packed point record
    (8 elem) written; camera projection records (16 elem) once."""
    return nrows * (4 + 3 * elem) + ne * (4 + 24 + 3 * elem + 96 + 3 * elem + 8 * elem) + nc * 16 * elem


# Whole solves, the reference's Ceres on the host cores against the library on
# the GPU, same scenes and options: (name, preset, Huber threshold, reference
# linear solver (0 SPARSE_SCHUR, 1 ITERATIVE_SCHUR + JACOBI), LM iterations).
REF_SOLVE_CASES = [("C1", 1, 0.0, 0, 20), ("C2", 2, 0.0, 0, 20), ("C3", 3, 4.0, 1, 5)]


def reference_solves(mba, O, threads, cpu_model):
    """Optional CPU-baseline leg (--ref-solves, ~1 min): per scene, the
    reference Ceres solve's wall time and final cost beside the library's
    (SPARSE_SCHUR -> explicit Schur FP64; ITERATIVE_SCHUR -> implicit-Schur PCG
    in FP64 and mixed), each library solve the second of two calls."""
    out = []
    for name, preset, huber, ls, iters in REF_SOLVE_CASES:
        d = mba.generate_scene(preset)
        nres = int(d["obs"].shape[0])
        log(f"bench: reference solve {name} ({nres} observations, {threads} threads)")
        sc = O.Scene(d)
        t0 = time.perf_counter()
        r = O.ref_solve(sc, huber_a=huber, linear_solver=ls, max_iters=iters, num_threads=threads)
        ref_wall = time.perf_counter() - t0
        row = {"scene": name, "observations": nres, "lm_iterations": iters,
               "reference": {"solver": "SPARSE_SCHUR" if ls == 0 else "ITERATIVE_SCHUR + JACOBI",
                             "wall_s": ref_wall, "ceres_total_s": r["total_time"],
                             "iterations": len(r["iterations"]) - 1, "final_cost": r["final_cost"],
                             "cores": threads, "cpu": cpu_model},
               "gpu": []}
        runs = ([("explicit Schur FP64", mba.EXPLICIT_SCHUR, mba.PRECISION_FP64)] if ls == 0 else
                [("implicit-Schur PCG FP64", mba.IMPLICIT_SCHUR_PCG, mba.PRECISION_FP64),
                 ("implicit-Schur PCG mixed", mba.IMPLICIT_SCHUR_PCG, mba.PRECISION_MIXED)])
        for label, solver, prec in runs:
            o = mba.default_options(max_num_iterations=iters, linear_solver=solver, precision=prec, huber_a=huber)
            for _ in range(2):  # the first call pays the device set-up
                p = mba.BAProblem(d["cam_init"].copy(), d["pt_init"].copy(), d["obs"], d["cam_idx"], d["pt_idx"])
                t0 = time.perf_counter()
                s = mba.solve(p, o)
                wall = time.perf_counter() - t0
            row["gpu"].append({"solver": label, "wall_s": wall, "preprocess_s": s.preprocessor_time_s,
                               "iterations": s.num_iterations - 1, "final_cost": s.final_cost})
        log(f"  reference {ref_wall:.2f} s (cost {r['final_cost']:.10g}); GPU " +
            ", ".join(f"{g['solver']} {g['wall_s']:.3f} s (cost {g['final_cost']:.10g})" for g in row["gpu"]))
        out.append(row)
        del d, sc
    return out


def run(args):
    import torch

    mba = importlib.import_module("3dreconstruction_amd")
    mba.load_library()
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    log("bench: generating the 4M-observation scene")
    d = mba.generate_scene(3)
    nobs = d["obs"].shape[0]
    nc, npt = d["cam_init"].shape[0], d["pt_init"].shape[0]

    def problem():
        return mba.BAProblem(d["cam_init"].copy(), d["pt_init"].copy(), d["obs"], d["cam_idx"], d["pt_idx"])

    def time_pass(precision):
        ctx = mba.Context(problem(), mba.default_options(precision=precision, huber_a=HUBER_A,
                                                         linear_solver=mba.IMPLICIT_SCHUR_PCG))
        stream = torch.cuda.ExternalStream(ctx.stream, device=dev)
        warm_up(ctx, args.warmup, torch)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record(stream)
        for _ in range(args.steps):
            ctx.jacobian_schur(RADIUS)
        e1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        ms = e0.elapsed_time(e1)
        # Per-kernel durations: HIP events recorded by the engine around
        # each kernel on its stream.
        ctx.set_timing(True)
        kern = np.zeros(3)
        for _ in range(args.steps):
            ctx.jacobian_schur(RADIUS)
            kern += ctx.last_timing()[:3]
        ctx.set_timing(False)
        kern /= args.steps
        ctx.close()
        return ms / args.steps, wall / args.steps, kern

    log("bench: mixed-precision Jacobian+Schur pass")
    ms_mixed, wall_mixed, kern_mixed = time_pass(mba.PRECISION_MIXED)
    log(f"  {ms_mixed:.4f} ms/step (events), kernels point {kern_mixed[0]:.4f} camera {kern_mixed[1]:.4f} ms")
    log("bench: FP64 Jacobian+Schur pass")
    ms_fp64, _, kern_fp64 = time_pass(mba.PRECISION_FP64)
    log(f"  {ms_fp64:.4f} ms/step")

    # Structure sizes for the roofline (from the library's own ordering).
    ne, order, _, _ = mba.schur_structure(problem(), with_pairs=False)
    nf = order.shape[0] - ne
    rows_per_chunk = max(64, min(2048, nobs // 2048))
    rows_per_chunk = (rows_per_chunk + 63) // 64 * 64
    nchunks = int(np.sum(np.ceil(np.bincount(d["cam_idx"], minlength=nc) / rows_per_chunk)))
    # The Jacobian+Schur unit's algorithmic bytes per observation (read 16 B;
    # W = Jc'Jp 108 B written; 60 B of point blocks per point), times the
    # observations one pass covers, over the pass's kernel time (the three
    # launches of one pass, HIP events on the engine's stream).  This engine
    # recomputes W instead of storing it, so it moves fewer bytes than the
    # unit counts: the fraction is an effective one; `traffic` is what the
    # pass's kernels really move (counters), and `dominant_kernel` gives the
    # slower of the point and camera passes against its own bytes.
    unit_bytes = 16 + 108 + 60.0 * npt / nobs
    pass_bytes = unit_bytes * nobs
    achieved = pass_bytes / (kern_mixed[2] * 1e-3) / 1e9
    # Counter figures of the pass kernels (HBM bytes per launch from
    # FETCH_SIZE / WRITE_SIZE, VALU issue share), committed with the
    # rocprofv3 commands that produced them: not measured in this run.
    tdata = {}
    tpath = os.path.join(ROOT, "profiles", "counters_jacobian_schur_4M.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tdata = json.load(f)
    traffic_of = tdata.get("hbm_bytes_per_launch", {})
    valu_of = tdata.get("valu_issue_frac", {})
    # rocprofv3 names of the mixed pass's kernels at this size (1,000 cameras:
    # the point pass runs in its LDS-staged persistent form).
    pass_kernels = ["k_point_pass_lds", "k_cam_pass<float, 64>", "k_cam_finalize<float>"]
    traffic = (sum(traffic_of[k] for k in pass_kernels) if all(k in traffic_of for k in pass_kernels) else None)
    if kern_mixed[1] >= kern_mixed[0]:
        name, kms, nbytes = "k_cam_pass<float, 64>", kern_mixed[1], cam_pass_bytes(nobs, ne, nf, nchunks, 4)
    else:
        name, kms, nbytes = "k_point_pass_lds", kern_mixed[0], point_pass_bytes(nobs, ne, nc, 4)
    dom_achieved = nbytes / (kms * 1e-3) / 1e9
    # The Jacobian+Schur unit of the metric (131.5 B/obs incl. 108 B of
    # W = Jc'Jp this engine never writes), over the whole pass: an effective
    # figure, reported beside the headline.
    unit = {"kernels": " + ".join(pass_kernels), "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "algorithmic_bytes": pass_bytes, "bytes_per_obs": unit_bytes,
            "kernel_ms": kern_mixed[2], "traffic": traffic,
            "note": "effective: the unit counts 108 B/obs of W = Jc'Jp that the engine recomputes, not writes"}

    lm = None
    if not args.no_lm:
        log("bench: LM solves to tolerance (implicit-Schur PCG)")
        res = {}
        for label, prec, cap in (("fp64", mba.PRECISION_FP64, args.lm_iters),
                                 ("mixed", mba.PRECISION_MIXED, (3 * args.lm_iters) // 2)):
            p = problem()
            opts = mba.default_options(precision=prec, huber_a=HUBER_A, linear_solver=mba.IMPLICIT_SCHUR_PCG,
                                       max_num_iterations=cap, eta=LM_ETA, function_tolerance=LM_FTOL)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            s = mba.solve(p, opts)
            torch.cuda.synchronize()
            res[label] = (s, time.perf_counter() - t0)
            log(f"  {label}: final cost {s.final_cost:.10e} in {res[label][1]:.3f} s, "
                f"{s.num_successful_steps}+{s.num_unsuccessful_steps} steps, {s.num_linear_solver_iterations} CG")
        ref_cost = res["fp64"][0].final_cost

        def time_to(s, target):
            for it in s.iterations:
                if it.cost <= target:
                    return s.preprocessor_time_s + it.cumulative_time_s
            return None

        cost_target = ref_cost * (1.0 + RMS_TOL)
        t_mixed, t_fp64 = time_to(res["mixed"][0], cost_target), time_to(res["fp64"][0], cost_target)
        # RMS = sqrt(2 cost / N): RMS within RMS_TOL <=> cost within (1 + RMS_TOL)^2.
        rms_target = ref_cost * (1.0 + RMS_TOL) ** 2
        r_mixed, r_fp64 = time_to(res["mixed"][0], rms_target), time_to(res["fp64"][0], rms_target)
        lm = {"solver": f"implicit-Schur PCG (JACOBI), eta {LM_ETA}, function tolerance {LM_FTOL}",
              "max_iterations": {"fp64": args.lm_iters, "mixed": (3 * args.lm_iters) // 2},
              "target": f"RMS reprojection error (FP64-evaluated) within {RMS_TOL} relative of the FP64 "
                        f"path's final RMS",
              "fp64_final_rms_px": float(np.sqrt(2 * res["fp64"][0].final_cost / nobs)),
              "fp64_final_cost": res["fp64"][0].final_cost, "mixed_final_cost": res["mixed"][0].final_cost,
              "final_cost_rel_diff": abs(res["mixed"][0].final_cost - res["fp64"][0].final_cost)
              / res["fp64"][0].final_cost,
              "time_to_tol_s": {"mixed": r_mixed, "fp64": r_fp64},
              "solve_wall_s": {"mixed": res["mixed"][1], "fp64": res["fp64"][1]},
              "speedup_mixed_vs_fp64": (r_fp64 / r_mixed) if (r_mixed and r_fp64) else None,
              "cost_target": {"target": f"FP64-evaluated cost within {RMS_TOL} relative of the FP64 path's "
                                        f"final cost",
                              "time_to_tol_s": {"mixed": t_mixed, "fp64": t_fp64},
                              "speedup_mixed_vs_fp64": (t_fp64 / t_mixed) if (t_mixed and t_fp64) else None}}

        # The same LM with SCHUR_JACOBI (both precisions, same caps), against
        # the JACOBI FP64 final cost; and both preconditioners to convergence
        # on the 697k-observation scene.
        def lm_record(s, wall, ref, target):
            return {"final_cost": s.final_cost, "final_cost_rel_to_jacobi_fp64": (s.final_cost - ref) / ref,
                    "termination": s.termination_type, "lm_iterations": len(s.iterations) - 1,
                    "steps": [s.num_successful_steps, s.num_unsuccessful_steps],
                    "invalid_steps": sum(1 for it in s.iterations[1:] if not it.step_is_valid),
                    "cg_iterations": s.num_linear_solver_iterations, "solve_wall_s": wall,
                    "time_to_tol_s": time_to(s, target)}

        log("bench: LM with SCHUR_JACOBI")
        sj = {}
        for label, prec, cap in (("fp64", mba.PRECISION_FP64, args.lm_iters),
                                 ("mixed", mba.PRECISION_MIXED, (3 * args.lm_iters) // 2)):
            opts = mba.default_options(precision=prec, huber_a=HUBER_A, linear_solver=mba.IMPLICIT_SCHUR_PCG,
                                       max_num_iterations=cap, eta=LM_ETA, function_tolerance=LM_FTOL,
                                       preconditioner=mba.PRECOND_SCHUR_JACOBI)
            t0 = time.perf_counter()
            s = mba.solve(problem(), opts)
            sj[label] = lm_record(s, time.perf_counter() - t0, ref_cost, rms_target)
            log(f"  {label}: final cost {s.final_cost:.10e} ({sj[label]['final_cost_rel_to_jacobi_fp64']:+.2e}), "
                f"{s.num_successful_steps}+{s.num_unsuccessful_steps} steps, {s.num_linear_solver_iterations} CG, "
                f"{sj[label]['solve_wall_s']:.3f} s")
        lm["schur_jacobi"] = sj
        d2 = mba.generate_scene(2)
        c2 = {}
        for pre_name, pre in (("jacobi", mba.PRECOND_JACOBI), ("schur_jacobi", mba.PRECOND_SCHUR_JACOBI)):
            for label, prec, cap in (("fp64", mba.PRECISION_FP64, 300), ("mixed", mba.PRECISION_MIXED, 450)):
                opts = mba.default_options(precision=prec, linear_solver=mba.IMPLICIT_SCHUR_PCG, preconditioner=pre,
                                           max_num_iterations=cap, eta=LM_ETA, function_tolerance=LM_FTOL)
                best = None
                for _ in range(2):  # the first call pays the device set-up
                    p2 = mba.BAProblem(d2["cam_init"].copy(), d2["pt_init"].copy(), d2["obs"], d2["cam_idx"],
                                       d2["pt_idx"])
                    t0 = time.perf_counter()
                    s = mba.solve(p2, opts)
                    best = (s, time.perf_counter() - t0)
                c2[f"{pre_name}_{label}"] = best
        ref2 = c2["jacobi_fp64"][0].final_cost
        lm["c2_to_convergence"] = {
            "scene": "200 cameras / 100,000 points / 696,955 observations, no loss",
            "target": f"cost within {RMS_TOL} relative of the JACOBI FP64 final cost",
            "runs": {k: lm_record(s, wall, ref2, ref2 * (1.0 + RMS_TOL)) for k, (s, wall) in c2.items()}}
        for k, r in lm["c2_to_convergence"]["runs"].items():
            log(f"  C2 {k}: {r['lm_iterations']} iterations, {r['cg_iterations']} CG, termination "
                f"{r['termination']}, {r['solve_wall_s']:.3f} s, to tolerance {r['time_to_tol_s']}")

    explicit = None
    if not args.no_explicit:
        log("bench: explicit-Schur LM (dense S, MFMA Cholesky)")
        opts = mba.default_options(precision=mba.PRECISION_MIXED, huber_a=HUBER_A,
                                   linear_solver=mba.EXPLICIT_SCHUR, max_num_iterations=args.explicit_iters)
        t0 = time.perf_counter()
        ctx = mba.Context(problem(), opts)
        t_ctx = time.perf_counter() - t0
        ctx.set_timing(True)
        s = ctx.solve()
        ph = ctx.last_timing()[3:]
        ctx.close()
        n = 9 * nf
        solves = max(1, s.num_iterations - 1)
        explicit = {"n": n, "iterations": s.num_iterations - 1, "final_cost": s.final_cost,
                    "context_s": t_ctx, "linear_solver_ms_per_iteration": 1e3 * s.linear_solver_time_s / solves,
                    "last_solve_ms": {"assemble_S": ph[0], "cholesky": ph[1], "triangular_solves": ph[2]},
                    "cholesky_tflops": (n ** 3 / 3.0) / (ph[1] * 1e-3) / 1e12 if ph[1] > 0 else None,
                    "cholesky_mfma_peak_tflops": FP32_PEAK_TFLOPS}
        log(f"  {explicit['linear_solver_ms_per_iteration']:.2f} ms/solve; last: assemble {ph[0]:.2f} "
            f"cholesky {ph[1]:.2f} solves {ph[2]:.2f} ms")

    batched = None
    if not args.no_batched:
        log("bench: 512 batched local-BA windows (mixed, implicit-Schur PCG)")
        ws = [mba.generate_scene(4, w) for w in range(512)]
        probs = [mba.BAProblem(w["cam_init"].copy(), w["pt_init"].copy(), w["obs"], w["cam_idx"], w["pt_idx"])
                 for w in ws]
        opts = mba.default_options(precision=mba.PRECISION_MIXED, linear_solver=mba.IMPLICIT_SCHUR_PCG)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sums = mba.solve_batched(probs, opts)
        wall = time.perf_counter() - t0
        tot, pre = sums[0].total_time_s, sums[0].preprocessor_time_s
        its = [x.num_iterations - 1 for x in sums]
        batched = {"windows": len(ws), "observations": int(sum(w["obs"].shape[0] for w in ws)),
                   "windows_per_s": len(ws) / tot, "windows_per_s_solve_only": len(ws) / (tot - pre),
                   "windows_per_s_wall_incl_python": len(ws) / wall, "preprocess_s": pre, "total_s": tot,
                   "lm_iterations": {"min": min(its), "mean": float(np.mean(its)), "max": max(its)},
                   "all_usable": all(x.IsSolutionUsable() for x in sums)}
        log(f"  {batched['windows_per_s']:.0f} windows/s ({tot:.3f} s, preprocessing {pre:.3f} s)")
        # The same windows with the explicit solver (the DENSE_SCHUR role): every
        # window's dense S factored by its own workgroup in shared launches, FP64.
        probs = [mba.BAProblem(w["cam_init"].copy(), w["pt_init"].copy(), w["obs"], w["cam_idx"], w["pt_idx"])
                 for w in ws]
        torch.cuda.synchronize()
        sums = mba.solve_batched(probs, mba.default_options(linear_solver=mba.EXPLICIT_SCHUR))
        tot, pre = sums[0].total_time_s, sums[0].preprocessor_time_s
        its = [x.num_iterations - 1 for x in sums]
        batched["explicit_fp64"] = {"windows_per_s": len(ws) / tot, "windows_per_s_solve_only": len(ws) / (tot - pre),
                                    "preprocess_s": pre, "total_s": tot,
                                    "lm_iterations": {"min": min(its), "mean": float(np.mean(its)), "max": max(its)},
                                    "all_usable": all(x.IsSolutionUsable() for x in sums)}
        log(f"  explicit FP64: {len(ws) / tot:.0f} windows/s ({tot:.3f} s, preprocessing {pre:.3f} s)")
        del ws, probs

    large = None
    if not args.no_large:
        log("bench: 25M-observation scene")
        dl = mba.generate_scene(5)
        nol = dl["obs"].shape[0]

        def lproblem():
            return mba.BAProblem(dl["cam_init"].copy(), dl["pt_init"].copy(), dl["obs"], dl["cam_idx"],
                                 dl["pt_idx"])
        ctx = mba.Context(lproblem(), mba.default_options(precision=mba.PRECISION_MIXED,
                                                          linear_solver=mba.IMPLICIT_SCHUR_PCG))
        stream = torch.cuda.ExternalStream(ctx.stream, device=dev)
        warm_up(ctx, args.warmup, torch)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(args.steps):
            ctx.jacobian_schur(RADIUS)
        e1.record(stream)
        torch.cuda.synchronize()
        ms_l = e0.elapsed_time(e1) / args.steps
        ctx.close()
        s = mba.solve(lproblem(), mba.default_options(precision=mba.PRECISION_MIXED,
                                                      linear_solver=mba.IMPLICIT_SCHUR_PCG, max_num_iterations=30,
                                                      pcg_r_tolerance=1e-6, eta=1e-10))
        large = {"cameras": dl["cam_init"].shape[0], "points": dl["pt_init"].shape[0], "observations": nol,
                 "pass_ms": ms_l, "pass_obs_per_s": nol / (ms_l * 1e-3),
                 "lm": {"solver": "implicit-Schur PCG (JACOBI), relative residual 1e-6, at most 30 iterations",
                        "iterations": s.num_iterations - 1, "cg_iterations": s.num_linear_solver_iterations,
                        "total_s": s.total_time_s, "preprocess_s": s.preprocessor_time_s,
                        "termination": s.termination_type,
                        "initial_rms_px": float(np.sqrt(2 * s.initial_cost / nol)),
                        "final_rms_px": float(np.sqrt(2 * s.final_cost / nol))}}
        log(f"  pass {ms_l:.3f} ms ({nol / (ms_l * 1e-3):.3e} obs/s); LM {s.total_time_s:.2f} s")
        del dl

    grouped = None
    if not args.no_grouped:
        log("bench: grouped-intrinsics scene (the incremental engine's residual blocks)")
        dg = mba.generate_grouped_scene(6)
        ngo = dg["obs"].shape[0]
        gconst = np.zeros(dg["cam_init"].shape[0], np.uint8)
        gconst[0] = 1  # the incremental engine holds its first camera

        def gproblem(subset):
            return mba.BAProblem(dg["cam_init"].copy(), dg["pt_init"].copy(), dg["obs"], dg["cam_idx"],
                                 dg["pt_idx"], gconst, intrinsics=dg["intr_init"].copy(),
                                 intrinsic_index=dg["intr_idx"],
                                 intrinsic_subset=np.full(dg["intr_init"].shape[0], subset, np.uint8))
        gpass = {}
        for label, prec in (("mixed", mba.PRECISION_MIXED), ("fp64", mba.PRECISION_FP64)):
            ctx = mba.Context(gproblem(0), mba.default_options(precision=prec, huber_a=HUBER_A,
                                                               linear_solver=mba.IMPLICIT_SCHUR_PCG))
            stream = torch.cuda.ExternalStream(ctx.stream, device=dev)
            warm_up(ctx, args.warmup, torch)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(args.steps):
                ctx.jacobian_schur(RADIUS)
            e1.record(stream)
            torch.cuda.synchronize()
            ms_g = e0.elapsed_time(e1) / args.steps
            ctx.close()
            gpass[label] = {"ms_per_step": ms_g, "obs_per_s": ngo / (ms_g * 1e-3)}
        gsolves = {}
        for label, ls, prec, subset in (("sparse_schur_role_fp64_focal_only", mba.EXPLICIT_SCHUR, mba.PRECISION_FP64,
                                         0x3E),
                                        ("sparse_schur_role_fp64_all_intrinsics", mba.EXPLICIT_SCHUR,
                                         mba.PRECISION_FP64, 0),
                                        ("iterative_schur_mixed_all_intrinsics", mba.IMPLICIT_SCHUR_PCG,
                                         mba.PRECISION_MIXED, 0)):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            s = mba.solve(gproblem(subset), mba.default_options(linear_solver=ls, precision=prec, huber_a=HUBER_A))
            wall = time.perf_counter() - t0
            gsolves[label] = {"iterations": s.num_iterations - 1, "successful": s.num_successful_steps,
                              "termination": s.termination_type, "total_s": s.total_time_s,
                              "preprocess_s": s.preprocessor_time_s, "linear_solver_s": s.linear_solver_time_s,
                              "wall_s": wall, "initial_rms_px": float(np.sqrt(2 * s.initial_cost / ngo)),
                              "final_rms_px": float(np.sqrt(2 * s.final_cost / ngo))}
            log(f"  {label}: {s.num_iterations - 1} iterations in {s.total_time_s:.3f} s, RMS "
                f"{gsolves[label]['initial_rms_px']:.3f} -> {gsolves[label]['final_rms_px']:.4f} px")
        grouped = {"cameras": dg["cam_init"].shape[0], "intrinsic_groups": dg["intr_init"].shape[0],
                   "points": dg["pt_init"].shape[0], "observations": ngo,
                   "options": "Huber 4 px, first camera constant, Ceres defaults (50 iterations)",
                   "pass": gpass, "lm": gsolves}
        log(f"  pass mixed {gpass['mixed']['ms_per_step']:.3f} ms, fp64 {gpass['fp64']['ms_per_step']:.3f} ms")
        del dg

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O  # CPU baseline leg only

        threads = max(1, min(16, os.cpu_count() or 1))
        cpu_model = "unknown CPU"
        try:
            with open("/proc/cpuinfo") as f:
                for line in f:
                    if line.startswith("model name"):
                        cpu_model = line.split(":", 1)[1].strip()
                        break
        except OSError:
            pass
        sc = O.Scene(dict(cam_init=d["cam_init"], pt_init=d["pt_init"], obs=d["obs"], cam_idx=d["cam_idx"],
                          pt_idx=d["pt_idx"]))
        log(f"bench: CPU baseline (FP64 oracle, {threads} threads)")
        O.jacobian_schur(sc, huber_a=HUBER_A, num_threads=threads)
        reps = args.cpu_reps
        t0 = time.perf_counter()
        for _ in range(reps):
            O.jacobian_schur(sc, huber_a=HUBER_A, num_threads=threads)
        dt = (time.perf_counter() - t0) / reps
        port = {"value": nobs / dt, "unit": "obs/s", "cores": threads, "kind": "port",
                "sample": f"FP64 oracle Jacobian+Schur pass over the full {nobs}-observation scene, "
                          f"{reps} passes after one warm-up, OpenMP {threads} threads on {cpu_model} "
                          f"({os.cpu_count()} logical CPUs visible)"}
        cpu = port
        if O.ref_available() and not args.no_cpu_ref:
            # The reference's own Ceres (built from its sources), the same scene and loss.
            log(f"bench: CPU baseline (reference Ceres SPARSE_SCHUR pass, {threads} threads)")
            sec, setup, nres = O.ref_jacobian_schur_pass(sc, huber_a=HUBER_A, num_threads=threads, reps=1)
            cpu = {"value": nres / sec, "unit": "obs/s", "cores": threads, "kind": "reference",
                   "sample": f"the reference's Ceres 1.10 SPARSE_SCHUR iteration pass (ProgramEvaluator::Evaluate "
                             f"with the Huber loss, ScaleColumns, SchurEliminator<2,3,9>::Eliminate into the "
                             f"block-sparse S) over the full {nres}-observation scene: one timed pass after one "
                             f"warm-up ({sec:.2f} s; problem building and ordering {setup:.1f} s untimed), "
                             f"OpenMP {threads} threads on {cpu_model}",
                   "port": port}
        log(f"  CPU {cpu['kind']}: {cpu['value']:.3g} obs/s")
        if args.ref_solves and O.ref_available():
            cpu["solves"] = reference_solves(mba, O, threads, cpu_model)

    value = nobs / (ms_mixed * 1e-3)
    out = {
        "metric": "Jacobian+Schur obs/s; LM time-to-RMS-tol, mixed vs FP64 on GPU, 4M obs",
        "value": value, "unit": "obs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_mixed, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 (FP64 point-block inverses)", "data": "synthetic",
        "config": {"workload": "jacobian_schur_pass", "cameras": nc, "points": npt, "observations": nobs,
                   "huber_a_px": HUBER_A, "outliers": "5% x 20 px", "precision": "mixed",
                   "parallelism": "single GPU"},
        "fp64_pass": {"ms_per_step": ms_fp64, "value": nobs / (ms_fp64 * 1e-3), "unit": "obs/s",
                      "mixed_speedup": ms_fp64 / ms_mixed},
        "kernels_ms": {"point_pass": kern_mixed[0], "camera_pass": kern_mixed[1], "pass_total": kern_mixed[2]},
        "roofline": {"bound": "hbm", "kernel": name, "achieved": dom_achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": dom_achieved / HBM_PEAK_GBS, "traffic": traffic_of.get(name),
                     "algorithmic_bytes": nbytes, "kernel_ms": kms,
                     "limiter": ("valu" if valu_of.get(name, 0.0) > 0.6 else
                                 "hbm" if dom_achieved / HBM_PEAK_GBS > 0.6 else
                                 "latency (neither VALU issue nor HBM bandwidth saturated)"),
                     "valu_issue_frac": valu_of.get(name),
                     "counters_source": "profiles/counters_jacobian_schur_4M.json (rocprofv3 --pmc runs; "
                                        "traffic and VALU share are committed profile values, not measured "
                                        "by this run)",
                     "jacobian_schur_unit": unit},
        "lm": lm,
        "explicit_schur": explicit,
        "batched_windows": batched,
        "large_scene": large,
        "grouped_intrinsics": grouped,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def pmc_target(kind):
    import torch

    mba = importlib.import_module("3dreconstruction_amd")
    a = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
    a.fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    d = mba.generate_scene(3)
    p = mba.BAProblem(d["cam_init"], d["pt_init"], d["obs"], d["cam_idx"], d["pt_idx"])
    if kind == "lm_mixed":
        # Two whole mixed LM solves with the LM leg's options, 0.3 s apart: the
        # second (no device set-up) is the one to read in a kernel trace.
        for _ in range(2):
            q = mba.BAProblem(d["cam_init"].copy(), d["pt_init"].copy(), d["obs"], d["cam_idx"], d["pt_idx"])
            mba.solve(q, mba.default_options(precision=mba.PRECISION_MIXED, huber_a=HUBER_A,
                                             linear_solver=mba.IMPLICIT_SCHUR_PCG, max_num_iterations=450,
                                             eta=LM_ETA, function_tolerance=LM_FTOL))
            time.sleep(0.3)
        return
    if kind == "explicit":
        ctx = mba.Context(p, mba.default_options(precision=mba.PRECISION_MIXED, huber_a=HUBER_A,
                                                 linear_solver=mba.EXPLICIT_SCHUR, max_num_iterations=2))
        ctx.solve()
    else:
        prec = mba.PRECISION_FP64 if kind == "pass_fp64" else mba.PRECISION_MIXED
        ctx = mba.Context(p, mba.default_options(precision=prec, huber_a=HUBER_A,
                                                 linear_solver=mba.IMPLICIT_SCHUR_PCG))
        for _ in range(5):
            ctx.jacobian_schur(RADIUS)
    ctx.synchronize()
    ctx.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--lm-iters", type=int, default=300)
    ap.add_argument("--cpu-reps", type=int, default=2)
    ap.add_argument("--no-lm", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-cpu-ref", action="store_true", help="CPU baseline from the oracle port only")
    ap.add_argument("--ref-solves", action="store_true",
                    help="also time whole solves of the reference's Ceres (host cores) against the library "
                         "on C1-C3 (cpu_baseline.solves; about a minute more)")
    ap.add_argument("--no-explicit", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-large", action="store_true")
    ap.add_argument("--no-grouped", action="store_true")
    ap.add_argument("--explicit-iters", type=int, default=5)
    ap.add_argument("--pmc-target", nargs="?", const="pass", choices=["pass", "pass_fp64", "explicit", "lm_mixed"],
                    help="target run for rocprofv3 (--pmc or --kernel-trace): a 256 MiB device copy (known "
                         "bytes, the counter calibration) followed by five mixed-precision Jacobian+Schur "
                         "passes ('pass'), five FP64 ones ('pass_fp64'), a two-iteration explicit-Schur "
                         "solve ('explicit') or two whole mixed LM solves with the LM leg's options "
                         "('lm_mixed'); prints nothing else")
    args = ap.parse_args()
    if args.pmc_target:
        pmc_target(args.pmc_target)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the engine runs a problem on a single MI355X; "
              "multi-GPU runs are not part of this benchmark.", file=sys.stderr)
        sys.exit(2)
    run(args)


if __name__ == "__main__":
    main()
