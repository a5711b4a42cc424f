#!/usr/bin/env python3
"""Benchmark: walk-on-boundary paths/s and time-to-RMSE on one MI355X.

Headline workload (geodesic_mixed_3d): the learned-control-variate configuration.  A
frequency-100 geodesic sphere (200,000 triangles) radially displaced to
r = 1 + 0.2 sin(3 theta) cos(2 phi), Dirichlet where the face centroid has z > 0 and Neumann
elsewhere, data from u = exp(0.6 x) cos(0.6 y) + x z; single-layer BIE, interior, backward
estimator, RIS with 64 candidates, first-kind constant 2, path length 4
(data/config_general_3d_interior_Mixed.json); 1,000,000 interior points.  A step is one
wob_estimate_domain call of --paths paths at every point (walk kernel + FP64 per-point
reduction), inputs resident in HBM, fresh Philox streams per step (path_base advances).
`value` is plain WoB paths/s.  Prints ONE JSON line.

Also reported: the walk kernel's roofline position from its own HIP-event duration (the
library brackets the kernel on its launch stream) and the work model of
wobtoolbox_amd/roofline.py; a CPU baseline (the OpenMP restatement of the reference in
oracle/, reference hit pick and per-point mt19937 streams) timed on a bounded sample of the
same workload on this host's cores; the time to a common RMSE of plain WoB and of WoB with
the learned control variate (training data and training on the control variate's clock);
and legs for the other configurations (level-5 icosphere Neumann, the 64-scene exterior
batch, the Poisson problem, the virtual-point-light app, scene construction).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "WoB paths/s and time-to-RMSE 1e-3 (learned CV vs plain WoB), 1 MI355X"
HCV = 16   # WOB_EST_HARMONIC_CV


def cpu_baseline(workload, cores, paths, target_s=10.0):
    """Reference-mode OpenMP oracle on a deterministic subset (every k-th point, `paths` paths
    per point as in a GPU step).

    k is chosen from a short calibration run so that the timed sample takes about target_s."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    sc = oracle_lib.OracleScene(workload.elements)
    n_all = len(workload.points)
    cal = np.arange(0, n_all, max(1, n_all // max(cores * 4, 64)), dtype=np.uint32)[: max(cores * 4, 64)]
    t0 = time.perf_counter()
    sc.estimate_domain(workload.solver, workload.points[cal], paths, seed=0, threads=cores, pick_mode=0, rng_mode=1,
                       point_index=cal)
    dt = time.perf_counter() - t0
    rate = len(cal) * paths / dt
    want = max(cores, int(target_s * rate / paths))
    stride = max(1, n_all // want)
    sel = np.arange(0, n_all, stride, dtype=np.uint32)[:want]
    n = len(sel)
    t0 = time.perf_counter()
    sc.estimate_domain(workload.solver, workload.points[sel], paths, seed=0, threads=cores, pick_mode=0, rng_mode=1,
                       point_index=sel)
    dt = time.perf_counter() - t0
    return {"value": n * paths / dt, "unit": "paths/s", "cores": cores, "kind": "port",
            "sample": f"{n} points (every {stride}th evaluation point) x {paths} paths = {n * paths} paths in "
                      f"{dt:.1f} s; reference hit pick (64-entry buffer), per-point mt19937 seeded as "
                      f"runtime_utils.cuh:89-103, OpenMP static schedule over points"}


def mlp_train_roofline(steps, batch, n_outputs, seconds):
    """BF16 MFMA training throughput: forward 2 x (4 x 128 x 128 + 128 x n_out) FLOP per sample
    (64 Fourier frequencies -> 128 features -> 4 ReLU layers of 128 -> outputs), a training step
    about 3 x forward (forward, activation and weight gradients)."""
    from wobtoolbox_amd.roofline import MI355X_PEAKS
    fwd = 2 * (4 * 128 * 128 + 128 * 16)
    flop = 3 * fwd * batch * steps
    tf = flop / seconds / 1e12
    return {"bound": "mfma", "achieved": tf, "peak": MI355X_PEAKS["bf16_TFLOPs"], "unit": "TFLOP/s",
            "frac": tf / MI355X_PEAKS["bf16_TFLOPs"], "seconds": seconds, "steps": steps, "batch": batch,
            "flop_per_sample_step": 3 * fwd}


def walk_timing(est, pts, paths, steps, warmup, step0=0):
    """Warm-up and timed steps of one domain estimate of `paths` paths at every point, then the
    walk kernel's own duration (HIP events on its launch stream) over a few more steps.
    Returns (paths/s, ms per step, mean walk-kernel ms, sums, sumsq, paths per point done)."""
    import torch
    n = pts.shape[0]
    V = est.scene.dim + 1
    s = torch.zeros((n, V), dtype=torch.float64, device=pts.device)
    q = torch.zeros_like(s)
    step = step0
    for _ in range(warmup):
        est.compute_domain(pts, paths, seed=0, path_base=step * paths, out=(s, q))
        step += 1
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        est.compute_domain(pts, paths, seed=0, path_base=step * paths, out=(s, q))
        step += 1
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    est.time_walk = True
    walk_ms = []
    for _ in range(max(3, min(steps, 10))):
        est.compute_domain(pts, paths, seed=0, path_base=step * paths, out=(s, q))
        walk_ms.append(est.last_launch.walk_ms)
        step += 1
    est.time_walk = False
    return steps * n * paths / elapsed, elapsed / steps * 1e3, float(np.mean(walk_ms)), s, q, (step - step0) * paths


def roofline_block(name, n_elements, path_length, launch_paths, walk_ms, candidates=64):
    """The walk kernel against MI355X's peaks, priced with the work model of roofline.py.
    `bound` is what rocprof's counters show limits it: "hbm" (record bytes against the 8 TB/s
    HBM peak) or the compute roof, "mfma" (model FP32 operations against the 157.3 TFLOP/s
    FP32 peak; the walk has no matrix instructions, it runs on the vector ALUs).  The other roof
    and the counter-measured HBM bytes are given beside it."""
    from wobtoolbox_amd.roofline import BOUND, MEASURED_TRAFFIC, MI355X_PEAKS, ris_per_path, workload_model
    flop, byts = workload_model(name, n_elements, path_length, candidates)
    s = walk_ms * 1e-3
    tflops = flop * launch_paths / s / 1e12
    gbps = byts * launch_paths / s / 1e9
    traffic = MEASURED_TRAFFIC.get(name)
    mem = {"achieved": gbps, "peak": MI355X_PEAKS["hbm_GBps"], "unit": "GB/s", "frac": gbps / MI355X_PEAKS["hbm_GBps"]}
    comp = {"achieved": tflops, "peak": MI355X_PEAKS["fp32_TFLOPs"], "unit": "TFLOP/s",
            "frac": tflops / MI355X_PEAKS["fp32_TFLOPs"], "unit_note": "FP32 vector ALU"}
    bound = BOUND[name]
    main, other = (mem, comp) if bound == "hbm" else (comp, mem)
    extra = {}
    if name == "geodesic_mixed_3d":
        # the fixed figure with a binary CDF descent per candidate instead of the guide record
        _, desc = ris_per_path(3, n_elements, path_length, candidates, descent=True)
        extra = {"binary_descent_model_bytes_per_path": desc,
                 "binary_descent_model_GBps": desc * launch_paths / s / 1e9}
    return {"bound": bound, **{k: main[k] for k in ("achieved", "peak", "unit", "frac")}, "traffic": traffic,
            "kernel": "walk_kernel", "avg_kernel_ms": walk_ms,
            "algorithmic_flop_per_path": flop, "algorithmic_bytes_per_path": byts, **extra,
            "fp32_valu" if bound == "hbm" else "model_bytes": other,
            "measured_hbm": None if traffic is None else
            {"bytes_per_launch": traffic, "achieved": traffic / s / 1e9, "unit": "GB/s",
             "frac": traffic / s / 1e9 / MI355X_PEAKS["hbm_GBps"]}}


def time_to_rmse_leg(args, torch, wob, G, w, scene, pts):
    """Plain WoB vs the learned control variates on the headline workload, all progressive.

    Each arm runs passes for the same walk-time budget; a control-variate arm's clock starts
    before its training data are generated and includes training (the MLP) or the fit (the
    harmonic control variate, whose arm stops at RMSE 1e-3).  The common target is the larger
    of the plain and MLP arms' final RMSE, which both reach inside the budget; the times are
    the first checkpoints (pass counts growing by 1.25) at or below it.  RMSE is against the
    exact solution, with the rms standard error beside it so that a bias would show."""
    from wobtoolbox_amd import cv as CV
    from wobtoolbox_amd import progressive
    exact_u = torch.as_tensor(w.exact(w.points)[:, 0]).to(pts.device)
    plain = wob.Estimator.from_config(scene, w.solver)
    L = w.solver["path_length"]
    n = pts.shape[0]
    print("[bench] time-to-RMSE: plain arm", file=sys.stderr, flush=True)
    curve_plain = progressive.run(plain, pts, exact_u, args.cv_paths_per_pass, args.cv_budget, seed=1)
    print("[bench] time-to-RMSE: control-variate arm", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pos, idx, tgt = CV.training_data(plain, args.cv_train_points, seed=2, paths_per_sample=args.cv_train_paths)
    net = CV.ControlVariate(3, L, sigma=args.cv_sigma, seed=3)
    torch.cuda.synchronize()
    t_train0 = time.perf_counter()
    net.train(pos, idx, tgt, batch=16384, steps=args.cv_train_steps, lr=2e-3)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t_train0
    net.bind(scene, L)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    cvest = wob.Estimator.from_config(scene, w.solver, flags=2, cv_samples=args.cv_samples)
    curve_cv = progressive.run(cvest, pts, exact_u, args.cv_paths_per_pass, t_setup + args.cv_budget, t0=t0, seed=1)
    # the harmonic control variate: a fundamental-solution expansion fitted to the data at
    # boundary samples (fit on its clock), then walks of the residual, until RMSE 1e-3
    print("[bench] time-to-RMSE: harmonic control-variate arm", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t0h = time.perf_counter()
    hfit = scene.fit_harmonic_cv(n_sources=args.hcv_sources, radius_scale=args.hcv_radius, n_train=args.hcv_train,
                                 volume_samples=0, seed=4)
    torch.cuda.synchronize()
    t_hfit = time.perf_counter() - t0h
    hest = wob.Estimator.from_config(scene, w.solver, flags=HCV)
    curve_h = progressive.run(hest, pts, exact_u, args.hcv_paths_per_pass, t_hfit + args.cv_budget, t0=t0h, seed=1,
                              stop_rmse=1e-3)
    scene.set_harmonic_cv(None)
    target = max(curve_plain[-1][2], curve_cv[-1][2])
    tp = progressive.time_to_rmse(curve_plain, target)
    tc = progressive.time_to_rmse(curve_cv, target)
    th = progressive.time_to_rmse(curve_h, target)
    th3 = progressive.time_to_rmse(curve_h, 1e-3)
    # walk throughput of each arm over its last checkpoint interval
    def rate(curve, t_off=0.0):
        t, k = curve[-1][0] - t_off, curve[-1][1]
        return n * k / t
    # variance x time per arm: RMSE^2 x walk seconds; at the path counts reached the RMSE is the
    # noise (RMSE ~ mean SE), so the time to 1e-3 is an extrapolation of this product
    eff_p = curve_plain[-1][2] ** 2 * curve_plain[-1][0]
    eff_c = curve_cv[-1][2] ** 2 * (curve_cv[-1][0] - t_setup)
    tp_last, kp, rp, sep = curve_plain[-1]
    # truncation bias of the path-length series: RMSE^2 = bias^2 + noise^2 over the points
    bias2 = rp * rp - sep * sep
    bias_bound = float(np.sqrt(max(bias2, 0.0) + 2.0 * 1.4 * sep * sep / np.sqrt(n)))
    return {"workload": w.name, "elements": w.elements.n, "points": n,
            "solver": {k: w.solver[k] for k in ("bie_model", "problem", "path_length", "num_resampling_candidates",
                                                 "first_kind_equation_scaling_constant")},
            "target_rmse": target, "target_rule": "the larger final RMSE of the two arms (both reach it)",
            "plain_s": tp, "cv_s": tc, "speedup": None if (tp is None or tc is None) else tp / tc,
            "cv_setup_s": t_setup, "plain_paths_per_s": rate(curve_plain), "cv_paths_per_s": rate(curve_cv, t_setup),
            "cv_training_walks": args.cv_train_points * L * args.cv_train_paths,
            "cv_config": {"train_points_per_value": args.cv_train_points, "paths_per_target": args.cv_train_paths,
                          "train_steps": args.cv_train_steps, "batch": 16384, "sigma": args.cv_sigma,
                          "cv_samples": args.cv_samples},
            "mlp_train": mlp_train_roofline(args.cv_train_steps, 16384, L, t_train),
            "variance_time_ratio": eff_p / eff_c,
            "harmonic_cv": {"s_to_target": th, "speedup_vs_plain": None if (tp is None or th is None) else tp / th,
                            "fit_s": t_hfit, "fit_rms_residual": hfit, "n_sources": args.hcv_sources,
                            "radius_scale": args.hcv_radius, "n_train": args.hcv_train,
                            "paths_per_pass": args.hcv_paths_per_pass,
                            "paths_per_s": rate(curve_h, t_hfit),
                            "kind": "fundamental solutions outside the domain, FP64 least squares on the walk's "
                                    "boundary data at uniform boundary samples (value rows at Dirichlet points, "
                                    "normal-derivative rows at Neumann points)"},
            "rmse_1e-3": {"harmonic_cv_s": th3, "harmonic_cv_reached": th3 is not None,
                          "harmonic_cv_paths_per_point": None if th3 is None else
                          next(k for t, k, r, _ in curve_h if r <= 1e-3),
                          "speedup_vs_plain_extrapolated": None if th3 is None else (eff_p / 1e-6) / th3,
                          "speedup_vs_mlp_cv_extrapolated": None if th3 is None else (t_setup + eff_c / 1e-6) / th3,
                          "plain_reached": bool(rp <= 1e-3), "mlp_cv_reached": bool(curve_cv[-1][2] <= 1e-3),
                          "extrapolated_plain_s": eff_p / 1e-6, "extrapolated_mlp_cv_s": t_setup + eff_c / 1e-6,
                          "truncation_bias_rms_bound": bias_bound,
                          "note": "plain WoB and the MLP control variate do not reach 1e-3 at these budgets (per-path "
                                  "standard deviation about 200 x |u|): their times are RMSE^2 x walk time / 1e-6, noise "
                                  "only; the harmonic control variate's time is measured, fit included"},
            "plain_curve": [[round(t, 3), k, r, se] for t, k, r, se in curve_plain],
            "cv_curve": [[round(t, 3), k, r, se] for t, k, r, se in curve_cv],
            "harmonic_cv_curve": [[round(t, 3), k, r, se] for t, k, r, se in curve_h]}


def icosphere_leg(args, torch, wob, G):
    """Level-5 icosphere (20,480 triangles), displaced; Neumann data of x^2 - y^2 + 0.5 z; direct
    BIE, path length 6 (line-intersection and boundary-value sampling); 100,000 points x 256
    paths per step."""
    print("[bench] icosphere leg", file=sys.stderr, flush=True)
    w = G.icosphere_neumann_3d()
    paths = w.solver["num_sample_paths"]
    pts = torch.as_tensor(w.points).to("cuda:0")
    # the BVH width from the autotuner (a 16-path walk at every point on each width)
    t0 = time.perf_counter()
    width, width_ms = wob.tune_bvh_width(w.elements, w.solver, pts, 16)
    tune_s = time.perf_counter() - t0
    print(f"[bench] icosphere BVH width {width} (walk ms by width {width_ms})", file=sys.stderr, flush=True)
    scene = wob.SceneHost(w.elements, bvh_width=width)
    est = wob.Estimator.from_config(scene, w.solver)
    value, ms, walk_ms, s, q, k = walk_timing(est, pts, paths, args.leg_steps, 2)
    m = (s[:, 0] / k).cpu().numpy()
    var = np.maximum((q[:, 0] / k).cpu().numpy() - m * m, 0)
    d = m - w.exact(w.points)[:, 0]
    d -= d.mean()
    out = {"workload": w.name, "elements": w.elements.n, "points": len(w.points), "paths_per_step": len(w.points) * paths,
           "paths_per_s": value, "ms_per_step": ms,
           "roofline": roofline_block(w.name, w.elements.n, w.solver["path_length"], len(w.points) * paths, walk_ms),
           "accuracy": {"rmse_u": float(np.sqrt((d ** 2).mean())), "mean_se_u": float(np.sqrt(var / k).mean()),
                        "paths_per_point": k, "note": "Neumann problem: u compared after removing the mean offset"},
           "bvh_width": {"tuned": width, "walk_ms_by_width": width_ms, "tune_s": tune_s},
           "solver": "Direct, InteriorDomain, BackwardEstimator, NeumannProblem, path_length 6"}
    # the harmonic control variate on the same walk (fit on the clock, one step of paths)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rms = scene.fit_harmonic_cv(n_sources=64, n_train=16384, seed=7)
    torch.cuda.synchronize()
    fit_s = time.perf_counter() - t0
    hest = wob.Estimator.from_config(scene, w.solver, flags=HCV)
    hv, hms, _, hs_, hq_, hk = walk_timing(hest, pts, paths, 1, 1)
    hm = (hs_[:, 0] / hk).cpu().numpy()
    hvar = np.maximum((hq_[:, 0] / hk).cpu().numpy() - hm * hm, 0)
    hd = hm - w.exact(w.points)[:, 0]
    hd -= hd.mean()
    out["control_variate"] = {"kind": "harmonic (64 fundamental solutions outside the scene)", "fit_s": fit_s,
                              "fit_rms_residual": rms, "paths_per_s": hv, "rmse_u": float(np.sqrt((hd ** 2).mean())),
                              "mean_se_u": float(np.sqrt(hvar / hk).mean()), "paths_per_point": hk,
                              "per_path_variance_ratio": float(np.mean(var) / max(np.mean(hvar), 1e-300))}
    return out


def batch_leg(args, torch, wob, G):
    """64 exterior star polygons (Neumann / Robin / mixed, single layer, RIS 64) in one launch per pass.

    The autotuner picks paths per lane, block size and persistent mode from timed passes;
    the timed passes replay the captured hipGraph.  Reports paths/s and the RMSE of u against
    the exact log-source solution per problem kind."""
    print("[bench] batch leg: building 64 star scenes", file=sys.stderr, flush=True)
    ws = G.star_exterior_batch_2d()
    dev = torch.device("cuda:0")
    scenes = [wob.SceneHost(w.elements) for w in ws]
    pts = [torch.as_tensor(w.points).to(dev) for w in ws]
    paths = args.batch_paths
    entries = [(wob.Estimator.from_config(sc, w.solver), p, paths) for sc, w, p in zip(scenes, ws, pts)]
    b = wob.Batch(entries, persistent=True, use_graph=True)
    t0 = time.perf_counter()
    best, best_ms = b.tune(seed=1)
    tune_s = time.perf_counter() - t0
    print(f"[bench] batch tuned in {tune_s:.1f} s: {best} ({best_ms:.1f} ms/pass)", file=sys.stderr, flush=True)
    b.run(1, seed=0)   # warm-up pass (graph capture)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.run(args.batch_passes, seed=0)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    n_points = sum(len(w.points) for w in ws)
    total = n_points * paths * args.batch_passes
    per_kind = {}
    for w, (s, q) in zip(ws, b.outputs):
        k = b.passes * paths
        m = (s[:, 0] / k).cpu().numpy()
        var = np.maximum((q[:, 0] / k).cpu().numpy() - m * m, 0) / k
        err = m - w.exact(w.points.astype(np.float64))[:, 0]
        e, v = per_kind.setdefault(w.solver["problem"], ([], []))
        e.append(err)
        v.append(var)
    # RMSE against the exact solution beside the mean SE; RMSE^2 - mean SE^2 estimates the
    # squared bias (the truncation of the path-length series), which the noise would hide
    rmse, mean_se, bias = {}, {}, {}
    for kind, (e, v) in per_kind.items():
        e, v = np.concatenate(e), np.concatenate(v)
        rmse[kind] = float(np.sqrt(np.mean(e ** 2)))
        mean_se[kind] = float(np.mean(np.sqrt(v)))
        bias[kind] = float(np.sqrt(max(np.mean(e ** 2) - np.mean(v), 0.0)))
    out = {"workload": "star_exterior_batch_2d", "scenes": len(ws), "segments": int(sum(w.elements.n for w in ws)),
           "points": n_points, "paths_per_point_per_pass": paths, "passes": args.batch_passes,
           "paths_per_s": total / el, "ms_per_pass": el / args.batch_passes * 1e3, "tuned": best,
           "tune_s": tune_s, "rmse_u_by_problem": rmse, "mean_se_u_by_problem": mean_se,
           "bias_rms_estimate_by_problem": bias, "paths_per_point_total": b.passes * paths,
           "solver": "IndirectSingleLayer, ExteriorDomain, BackwardEstimator, RIS 64, eps 1e-5; path_length 7/4/6 "
                     "and fkesc 1/1/2 for Neumann/Robin/mixed scenes"}
    if not args.no_batch_cv:
        errs = [e for e, _ in per_kind.values()]
        plain_rmse_all = float(np.sqrt(np.mean(np.concatenate([np.concatenate(e) for e in errs]) ** 2)))
        plain_var = {kind: float(np.mean(np.concatenate(v))) * b.passes * paths for kind, (_, v) in per_kind.items()}
        out["control_variate"] = batch_cv(args, torch, wob, ws, scenes, pts, best, plain_rmse_all,
                                          b.passes * el / args.batch_passes, b.passes * paths, plain_var)
    return out


def batch_cv(args, torch, wob, ws, scenes, pts, best, plain_rmse_all, plain_s, plain_paths, plain_var):
    """The harmonic control variate on the batched exterior scenes: one exterior fit per scene
    (fundamental solutions inside the body, no constant; fits on the clock), then the same
    batched launch with WOB_EST_HARMONIC_CV and the tuned configuration.  Reports paths/s, and
    per problem kind the RMSE against the exact solution beside the standard error; the
    standard errors come from the spread of the per-pass means (the control variate's
    variances are far below the resolution of FP32 sums of squared values near u), and the
    time to plain WoB's final RMSE over all points."""
    print("[bench] batch control variate: 64 exterior fits", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fit_res = [sc.fit_harmonic_cv(n_sources=args.batch_cv_sources, n_train=8192, seed=7, exterior=True)
               for sc in scenes]
    torch.cuda.synchronize()
    fit_s = time.perf_counter() - t0
    paths = args.batch_paths
    entries = [(wob.Estimator.from_config(sc, w.solver, flags=HCV), p, paths) for sc, w, p in zip(scenes, ws, pts)]
    cfg = {k: best[k] for k in ("paths_per_lane", "block_size", "schedule", "regen_threshold")}
    b = wob.Batch(entries, persistent=bool(best["persistent"]), use_graph=True, **cfg)
    b.run(1, seed=1)   # warm-up (graph capture), then reset the sums: the timed passes start at 0
    for s, q in b.outputs:
        s.zero_()
        q.zero_()
    b.passes = 0
    reps = max(args.batch_passes, 4)
    snaps, el = [], 0.0
    prev = [s[:, 0].clone() for s, _ in b.outputs]
    for r in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b.run(1, seed=0)
        torch.cuda.synchronize()
        el += time.perf_counter() - t0
        cur = [s[:, 0].clone() for s, _ in b.outputs]
        snaps.append([((c - p_) / paths).cpu().numpy() for c, p_ in zip(cur, prev)])
        prev = cur
    first_pass_s = el / reps
    n_points = sum(len(w.points) for w in ws)
    per_kind = {}
    err1_all = []
    for i, w in enumerate(ws):
        reps_i = np.stack([snaps[r][i] for r in range(reps)])          # (reps, points)
        ex = w.exact(w.points.astype(np.float64))[:, 0]
        m = reps_i.mean(0)
        se = reps_i.std(0, ddof=1) / np.sqrt(reps)
        e, v, pv = per_kind.setdefault(w.solver["problem"], ([], [], []))
        e.append(m - ex)
        v.append(se ** 2)
        pv.append(reps_i.var(0, ddof=1) * paths)
        err1_all.append(reps_i[0] - ex)
    rmse, mean_se, bias, var_path = {}, {}, {}, {}
    for kind, (e, v, pv) in per_kind.items():
        e, v = np.concatenate(e), np.concatenate(v)
        rmse[kind] = float(np.sqrt(np.mean(e ** 2)))
        mean_se[kind] = float(np.mean(np.sqrt(v)))
        bias[kind] = float(np.sqrt(max(np.mean(e ** 2) - np.mean(v), 0.0)))
        var_path[kind] = float(np.mean(np.concatenate(pv)))
    rmse1_all = float(np.sqrt(np.mean(np.concatenate(err1_all) ** 2)))
    t_cv = fit_s + first_pass_s
    return {"kind": "harmonic, exterior (fundamental solutions inside each body, no constant term; FP64 least "
                    "squares on the walk's boundary data)",
            "n_sources": args.batch_cv_sources, "fits": len(scenes), "fit_s": fit_s,
            "fit_rms_residual_max": float(max(fit_res)),
            "paths_per_s": n_points * paths * reps / el, "ms_per_pass": el / reps * 1e3,
            "paths_per_point_total": reps * paths, "rmse_u_by_problem": rmse, "mean_se_u_by_problem": mean_se,
            "se_from": f"spread of {reps} per-pass means",
            "bias_rms_estimate_by_problem": bias, "per_path_variance_by_problem": var_path,
            "plain_per_path_variance_by_problem": plain_var,
            "variance_ratio_by_problem": {k: plain_var[k] / var_path[k] for k in var_path if var_path[k] > 0},
            "time_to_plain_rmse": {"plain_rmse_all": plain_rmse_all, "plain_s": plain_s,
                                   "plain_paths_per_point": plain_paths,
                                   "cv_rmse_all_after_one_pass": rmse1_all,
                                   "cv_s": t_cv if rmse1_all <= plain_rmse_all else None,
                                   "cv_paths_per_point": paths,
                                   "speedup": plain_s / t_cv if rmse1_all <= plain_rmse_all else None,
                                   "plain_s_to_cv_rmse_extrapolated": plain_s * (plain_rmse_all / rmse1_all) ** 2,
                                   "note": "control variate: all fits plus its first pass; the plain arm's time to "
                                           "the control variate's one-pass RMSE is noise-only (RMSE^2 scaling), "
                                           "its truncation bias (Robin, mixed) would keep it above"}}


def poisson_leg(args, torch, wob, G):
    """Poisson problem (16 Gaussians, interior Dirichlet, double layer, path length 4) on the
    200k-triangle geodesic sphere: 8 point sets of 16,384 .. 524,288 points as 8 entries of one
    batched launch per pass (autotuned, hipGraph replays).  Reports paths/s and the RMSE of u
    against u = N_f + H."""
    print("[bench] poisson leg: building the 200k-triangle scene", file=sys.stderr, flush=True)
    w = G.poisson_gaussian_3d()
    dev = torch.device("cuda:0")
    scene = wob.SceneHost(w.elements)
    scene.set_volume_gaussians(w.extra["centers"], w.extra["widths"], w.extra["amplitudes"])
    est = wob.Estimator.from_config(scene, w.solver)
    sets = w.extra["point_sets"]
    pts = [torch.as_tensor(p).to(dev) for p in sets]
    paths = args.poisson_paths
    b = wob.Batch([(est, p, paths, j) for j, p in enumerate(pts)], persistent=True, use_graph=True)
    t0 = time.perf_counter()
    best, best_ms = b.tune(seed=1)
    tune_s = time.perf_counter() - t0
    print(f"[bench] poisson tuned in {tune_s:.1f} s: {best} ({best_ms:.1f} ms/pass)", file=sys.stderr, flush=True)
    b.run(1, seed=0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.run(args.poisson_passes, seed=0)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    n_points = sum(len(p) for p in sets)
    k = b.passes * paths
    rmse, se_mean = [], []
    for p, (s, q) in zip(sets, b.outputs):
        m = (s[:, 0] / k).cpu().numpy()
        var = np.maximum((q[:, 0] / k).cpu().numpy() - m * m, 0)
        rmse.append(float(np.sqrt(np.mean((m - w.exact(p)[:, 0]) ** 2))))
        se_mean.append(float(np.mean(np.sqrt(var / k))))
    out = {"workload": w.name, "elements": w.elements.n, "point_sets": [len(p) for p in sets], "points": n_points,
           "paths_per_point_per_pass": paths, "passes": args.poisson_passes,
           "paths_per_s": n_points * paths * args.poisson_passes / el,
           "ms_per_pass": el / args.poisson_passes * 1e3, "tuned": best, "tune_s": tune_s,
           "rmse_u_by_set": rmse, "mean_se_u_by_set": se_mean, "paths_per_point_total": k,
           "solver": "IndirectDoubleLayer, InteriorDomain, BackwardEstimator, DirichletProblem, path_length 4, "
                     "eps 1e-4, num_volume_samples 1 (Gaussian-mixture source, inside test per sample)"}
    if not args.no_poisson_cv:
        out["control_variate"] = poisson_cv(args, torch, wob, w, scene, pts, sets, b, rmse, se_mean)
    return out


def poisson_cv(args, torch, wob, w, scene, pts, sets, plain_batch, plain_rmse, plain_se):
    """The learned harmonic control variate on the Poisson workload: a fundamental-solution
    expansion fitted to the walk's own known-term data at boundary samples (its fit on the
    clock), then the same batched launch with WOB_EST_CONTROL_VARIATE.  Reports its paths/s,
    RMSE and mean SE per point set beside plain WoB's at the same path count, the per-path
    variance ratio, and the time of each arm to a common RMSE (progressive passes over all
    points in one call, the fit on the control variate's clock)."""
    from wobtoolbox_amd import progressive
    print("[bench] poisson control variate: fit", file=sys.stderr, flush=True)
    cvkw = dict(n_sources=args.poisson_cv_sources, radius_scale=1.5, n_train=16384, volume_samples=16,
                eps=w.solver["eps"], seed=7)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rms_fit = scene.fit_harmonic_cv(**cvkw)
    torch.cuda.synchronize()
    fit_s = time.perf_counter() - t0
    est = wob.Estimator.from_config(scene, w.solver, flags=HCV)
    paths = args.poisson_paths
    b = wob.Batch([(est, p, paths, j) for j, p in enumerate(pts)], persistent=True, use_graph=True)
    best, _ = b.tune(seed=1)
    b.run(1, seed=0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.run(args.poisson_passes, seed=0)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    n_points = sum(len(p) for p in sets)
    k = b.passes * paths
    rmse, se_mean, var_cv, var_plain = [], [], [], []
    for p, (s, q), (ps, pq) in zip(sets, b.outputs, plain_batch.outputs):
        m = (s[:, 0] / k).cpu().numpy()
        var = np.maximum((q[:, 0] / k).cpu().numpy() - m * m, 0)
        rmse.append(float(np.sqrt(np.mean((m - w.exact(p)[:, 0]) ** 2))))
        se_mean.append(float(np.mean(np.sqrt(var / k))))
        var_cv.append(var)
        mp = (ps[:, 0] / k).cpu().numpy()
        var_plain.append(np.maximum((pq[:, 0] / k).cpu().numpy() - mp * mp, 0))
    ratio = float(np.mean(np.concatenate(var_plain)) / np.mean(np.concatenate(var_cv)))
    # time to a common RMSE, both arms progressive over all points in one call per pass
    allp = torch.cat(pts)
    exact_u = torch.as_tensor(w.exact(np.concatenate(sets))[:, 0]).to(allp.device)
    plain = wob.Estimator.from_config(scene, w.solver)
    print("[bench] poisson time-to-RMSE: plain arm", file=sys.stderr, flush=True)
    curve_p = progressive.run(plain, allp, exact_u, args.poisson_ttr_paths, args.poisson_ttr_budget, seed=3)
    print("[bench] poisson time-to-RMSE: control-variate arm", file=sys.stderr, flush=True)
    scene.set_harmonic_cv(None)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    scene.fit_harmonic_cv(**cvkw)
    cvest = wob.Estimator.from_config(scene, w.solver, flags=HCV)
    curve_c = progressive.run(cvest, allp, exact_u, args.poisson_ttr_paths, args.poisson_ttr_budget, t0=t0, seed=3)
    target = max(curve_p[-1][2], curve_c[-1][2])
    tp = progressive.time_to_rmse(curve_p, target)
    tc = progressive.time_to_rmse(curve_c, target)
    return {"kind": "harmonic (fundamental solutions outside the domain, FP64 least squares on the walk's "
                    "known-term data at boundary samples)",
            "n_sources": args.poisson_cv_sources, "fit_s": fit_s, "fit_rms_residual": rms_fit,
            "paths_per_s": n_points * paths * args.poisson_passes / el, "ms_per_pass": el / args.poisson_passes * 1e3,
            "tuned": best, "rmse_u_by_set": rmse, "mean_se_u_by_set": se_mean,
            "plain_rmse_u_by_set": plain_rmse, "plain_mean_se_u_by_set": plain_se, "paths_per_point_total": k,
            "variance_ratio": ratio,
            "time_to_rmse": {"target_rmse": target, "target_rule": "the larger final RMSE of the two arms",
                             "plain_s": tp, "cv_s": tc, "speedup": None if (tp is None or tc is None) else tp / tc,
                             "paths_per_pass": args.poisson_ttr_paths, "points": int(allp.shape[0]),
                             "plain_curve": [[round(t, 3), kk, r, se_] for t, kk, r, se_ in curve_p],
                             "cv_curve": [[round(t, 3), kk, r, se_] for t, kk, r, se_ in curve_c]}}


# FP32 operations per (point, VPL) pair of the 2D single-layer aggregation (add, mul, cmp,
# max, log2, rcp = 1; FMA = 2): r (2), |r|^2 (3), clamp (1), log2 (1), value FMA (2), rcp (1),
# r = 0 test (1), gradient scale (1), two gradient FMAs (4).
VPL_FLOP_PER_PAIR_2D_SL = 16


def vpl_leg(args, torch, wob, G):
    """The virtual-point-light app (apps/potential_flow_with_vpl.cu, data/config_potential_flow_
    with_vpl.json): 100,000 single-layer Neumann walks (RIS 1024, path length 4) per iteration,
    aggregated at the 39,280 grid points (256^2) outside the four obstacles.  Reports the
    iteration time split into the path phase and the aggregation, and the aggregation's FP32
    VALU rate."""
    print("[bench] vpl leg", file=sys.stderr, flush=True)
    w = G.potential_flow_vpl_2d(grid_res=256)
    dev = torch.device("cuda:0")
    scene = wob.SceneHost(w.elements)
    est = wob.Estimator.from_config(scene, w.solver)
    pts = torch.as_tensor(w.points[w.extra["mask"]]).to(dev)
    n_vpl = int(w.solver["num_sample_paths"])
    acc = torch.zeros((pts.shape[0], 3), dtype=torch.float64, device=dev)
    rec, c = est.compute_vpl_paths(n_vpl, seed=0, path_base=0)
    est.aggregate_vpl(rec, c, pts, out=acc)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3 * args.vpl_iterations)]
    t0 = time.perf_counter()
    for it in range(args.vpl_iterations):
        ev[3 * it].record()
        rec, c = est.compute_vpl_paths(n_vpl, seed=0, path_base=1 + it)
        ev[3 * it + 1].record()
        est.aggregate_vpl(rec, c, pts, out=acc)
        ev[3 * it + 2].record()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    path_ms = float(np.mean([ev[3 * i].elapsed_time(ev[3 * i + 1]) for i in range(args.vpl_iterations)]))
    agg_ms = float(np.mean([ev[3 * i + 1].elapsed_time(ev[3 * i + 2]) for i in range(args.vpl_iterations)]))
    pairs = pts.shape[0] * n_vpl
    g = (acc[:, 1:] / ((1 + args.vpl_iterations) * n_vpl)).cpu().numpy()
    p = w.points[w.extra["mask"]]
    free = (p[:, 0] < -0.4) & (np.abs(p[:, 1]) < 0.45)
    tflops = VPL_FLOP_PER_PAIR_2D_SL * pairs / (agg_ms * 1e-3) / 1e12
    return {"workload": w.name, "elements": w.elements.n, "points": int(pts.shape[0]), "vpls_per_iteration": n_vpl,
            "iterations": args.vpl_iterations, "ms_per_iteration": el / args.vpl_iterations * 1e3,
            "path_phase_ms": path_ms, "aggregate_ms": agg_ms, "pairs_per_s": pairs / (agg_ms * 1e-3),
            "vpl_paths_per_s": n_vpl / (path_ms * 1e-3),
            "aggregate_fp32_valu": {"achieved": tflops, "peak": 157.3, "unit": "TFLOP/s", "frac": tflops / 157.3,
                                    "flop_per_pair": VPL_FLOP_PER_PAIR_2D_SL},
            "free_stream_velocity_x": float(g[free, 0].mean()),
            "solver": "IndirectSingleLayer, InteriorDomain, BackwardEstimator, NeumannProblem, path_length 4, "
                      "RIS 1024, eps 1e-4"}


def scene_build_leg(args, torch, wob, G, headline):
    """Scene construction with the BVH built on the host (binned SAH) and on the device (LBVH
    with SAH treelet restructuring, collapsed to the same 4-wide layouts), on the 200k-triangle
    geodesic sphere, and the headline walk on a device-built scene (same results; the tree
    differs)."""
    print("[bench] scene build leg", file=sys.stderr, flush=True)
    geo = G.geodesic_mixed_3d(n_points=1000)
    out = {"workload": "geodesic_mixed_3d", "elements": int(geo.elements.n)}
    for name, dev_bvh in (("host_bvh", False), ("device_bvh", True)):
        ts = []
        for _ in range(2):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            sc = wob.SceneHost(geo.elements, device_bvh=dev_bvh)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
            info = sc.info()
            sc.close()
        out[name] = {"create_s": min(ts), "nodes": int(info["n_nodes"]), "depth": int(info["depth"])}
    # the headline walk on a device-built tree
    w, pts, paths = headline
    sc = wob.SceneHost(w.elements, device_bvh=True)
    est = wob.Estimator.from_config(sc, w.solver)
    est.time_walk = True
    ms = []
    for i in range(3):
        est.compute_domain(pts, paths, seed=0, path_base=i * paths)
        ms.append(est.last_launch.walk_ms)
    out["headline_walk_ms_device_bvh"] = float(np.mean(ms[1:]))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--points", type=int, default=1_000_000, help="headline evaluation points")
    ap.add_argument("--paths", type=int, default=16, help="headline paths per point per step")
    ap.add_argument("--headline-only", action="store_true", help="skip every leg (profiling runs)")
    ap.add_argument("--only", choices=["icosphere", "batch", "poisson", "vpl"],
                    help="run only this leg and print its JSON object (profiling runs)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--leg-steps", type=int, default=10)
    ap.add_argument("--no-icosphere", action="store_true")
    ap.add_argument("--no-time-to-rmse", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-paths", type=int, default=16, help="paths per point per batch pass")
    ap.add_argument("--batch-passes", type=int, default=4)
    ap.add_argument("--no-batch-cv", action="store_true", help="skip the batch leg's harmonic control variate")
    ap.add_argument("--batch-cv-sources", type=int, default=64)
    ap.add_argument("--no-poisson", action="store_true")
    ap.add_argument("--poisson-paths", type=int, default=16, help="paths per point per Poisson pass")
    ap.add_argument("--poisson-passes", type=int, default=4)
    ap.add_argument("--no-poisson-cv", action="store_true")
    ap.add_argument("--poisson-cv-sources", type=int, default=64)
    ap.add_argument("--poisson-ttr-paths", type=int, default=2, help="paths per point per progressive pass")
    ap.add_argument("--poisson-ttr-budget", type=float, default=4.0, help="walk seconds per time-to-RMSE arm")
    ap.add_argument("--no-vpl", action="store_true")
    ap.add_argument("--vpl-iterations", type=int, default=5)
    ap.add_argument("--no-scene-build", action="store_true")
    ap.add_argument("--cv-budget", type=float, default=15.0, help="walk seconds per arm")
    ap.add_argument("--cv-paths-per-pass", type=int, default=4)
    ap.add_argument("--cv-train-points", type=int, default=1048576, help="boundary points per value function")
    ap.add_argument("--cv-train-paths", type=int, default=16, help="walks averaged per training point")
    ap.add_argument("--cv-train-steps", type=int, default=1500)
    ap.add_argument("--cv-sigma", type=float, default=0.5, help="scale of the Gaussian Fourier frequencies")
    ap.add_argument("--cv-samples", type=int, default=16, help="samples of the domain-level control-variate integral")
    ap.add_argument("--hcv-sources", type=int, default=192, help="harmonic control variate: fundamental solutions")
    ap.add_argument("--hcv-radius", type=float, default=2.0, help="harmonic control variate: source radius / scene radius")
    ap.add_argument("--hcv-train", type=int, default=32768, help="harmonic control variate: boundary samples of the fit")
    ap.add_argument("--hcv-paths-per-pass", type=int, default=1, help="harmonic control variate: paths per point per pass")
    ap.add_argument("--profile-copy-mib", type=int, default=0,
                    help="after the timed steps, copy this many MiB once (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (--gpus {args.gpus} requested). "
              "Points would shard across GPUs with no exchange step; only one GPU is measured.", file=sys.stderr)
        sys.exit(2)
    if args.headline_only:
        args.no_icosphere = args.no_time_to_rmse = args.no_batch = args.no_poisson = True
        args.no_vpl = args.no_scene_build = args.no_cpu_baseline = True

    import torch
    import wobtoolbox_amd as wob
    from wobtoolbox_amd import geometry as G

    if args.only:
        leg = {"icosphere": icosphere_leg, "batch": batch_leg, "poisson": poisson_leg, "vpl": vpl_leg}[args.only]
        out = leg(args, torch, wob, G)
        if args.profile_copy_mib:
            src = torch.ones(args.profile_copy_mib * (1 << 18), dtype=torch.float32, device="cuda:0")
            torch.empty_like(src).copy_(src)
            torch.cuda.synchronize()
        print(json.dumps(out), flush=True)
        return

    print("[bench] building the headline workload geodesic_mixed_3d", file=sys.stderr, flush=True)
    w = G.geodesic_mixed_3d(n_points=args.points)
    dev = torch.device("cuda:0")
    scene = wob.SceneHost(w.elements)
    est = wob.Estimator.from_config(scene, w.solver)
    pts = torch.as_tensor(w.points).to(dev)
    n = pts.shape[0]
    paths = args.paths
    print(f"[bench] headline: {args.warmup} warm-up and {args.steps} timed steps", file=sys.stderr, flush=True)
    value, ms_per_step, walk_ms, s, q, k = walk_timing(est, pts, paths, args.steps, args.warmup)
    if args.profile_copy_mib:
        src = torch.ones(args.profile_copy_mib * (1 << 18), dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"[bench] copied {args.profile_copy_mib} MiB", file=sys.stderr, flush=True)
    m = (s[:, 0] / k).cpu().numpy()
    var = np.maximum((q[:, 0] / k).cpu().numpy() - m * m, 0)
    err = m - w.exact(w.points)[:, 0]

    legs = {}
    if not args.no_icosphere:
        legs["icosphere"] = icosphere_leg(args, torch, wob, G)
    if not args.no_batch:
        legs["batch"] = batch_leg(args, torch, wob, G)
    if not args.no_poisson:
        legs["poisson"] = poisson_leg(args, torch, wob, G)
    if not args.no_vpl:
        legs["vpl"] = vpl_leg(args, torch, wob, G)
    if not args.no_scene_build:
        legs["scene_build"] = scene_build_leg(args, torch, wob, G, (w, pts, paths))
    ttr = None
    if not args.no_time_to_rmse:
        ttr = time_to_rmse_leg(args, torch, wob, G, w, scene, pts)

    cores = min(16, os.cpu_count() or 1)
    cpu = None
    if not args.no_cpu_baseline:
        print(f"[bench] cpu baseline on {cores} threads", file=sys.stderr, flush=True)
        cpu = cpu_baseline(w, cores, paths)

    line = {
        "metric": METRIC, "value": value, "unit": "paths/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": w.name, "elements": w.elements.n, "points": n, "paths_per_point": paths,
                   "paths_per_step": n * paths, "path_length": w.solver["path_length"],
                   "bie_model": w.solver["bie_model"], "problem": w.solver["problem"],
                   "estimator": w.solver["estimator"],
                   "num_resampling_candidates": w.solver["num_resampling_candidates"],
                   "paths_per_lane": est.last_launch.used_paths_per_lane,
                   "work_items": est.last_launch.used_work_items},
        "roofline": roofline_block(w.name, w.elements.n, w.solver["path_length"], n * paths, walk_ms,
                                   w.solver["num_resampling_candidates"]),
        "cpu_baseline": cpu,
        "time_to_rmse": ttr,
        "accuracy": {"rmse_u": float(np.sqrt((err ** 2).mean())), "mean_se_u": float(np.sqrt(var / k).mean()),
                     "paths_per_point": k, "note": "u against the exact solution; RMSE ~ mean SE: noise-limited"},
        **legs,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
