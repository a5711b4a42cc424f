"""Progressive estimation with RMSE tracking (the reference apps' main loop).

The reference apps run passes of num_sample_paths per point until max_time, normalise, and
log "samples, seconds, rmse_u, rmse_grad..." to rmse_list.csv (apps/general_2d.cu:318-367,
237, 358-359).  Here every pass adds fresh Philox paths (path_base advances) into FP64
accumulators; the RMSE of u against an exact solution is computed on the GPU at geometric
checkpoints, and the seconds are the wall time of the estimation (walk launches included).
"""
from __future__ import annotations

import math
import time


def run(estimator, points, exact_u, paths_per_pass, time_budget_s, t0=None, seed=0, neumann_shift=False,
        csv_path=None, max_passes=1_000_000, growth=1.25, stop_rmse=None):
    """Returns a list of (seconds since t0, paths per point, rmse_u, rms standard error of u)
    checkpoints, taken after pass counts growing by `growth`, until the time budget is spent
    or (stop_rmse) the first checkpoint with rmse_u <= stop_rmse.  With no bias, rmse_u and the
    rms standard error agree; rmse^2 - se^2 estimates the squared bias.

    points: device tensor (n, dim); exact_u: device FP64 tensor (n,) of exact values.
    t0: time origin (perf_counter); setup work done before the call (training) counts.
    """
    import torch
    n = points.shape[0]
    dim = points.shape[1]
    s = torch.zeros((n, dim + 1), dtype=torch.float64, device=points.device)
    q = torch.zeros_like(s)
    if t0 is None:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
    curve = []
    passes = 0
    next_check = 1
    while passes < max_passes:
        estimator.compute_domain(points, paths_per_pass, seed=seed, path_base=passes * paths_per_pass, out=(s, q))
        passes += 1
        if passes >= next_check:
            torch.cuda.synchronize()
            t = time.perf_counter() - t0
            k = passes * paths_per_pass
            m = s[:, 0] / k
            err = m - exact_u
            if neumann_shift:
                err = err - err.mean()
            rmse = float(torch.sqrt((err * err).mean()))
            se = float(torch.sqrt((torch.clamp(q[:, 0] / k - m * m, min=0) / k).mean()))
            curve.append((t, k, rmse, se))
            if t >= time_budget_s or (stop_rmse is not None and rmse <= stop_rmse):
                break
            next_check = max(passes + 1, int(math.ceil(passes * growth)))
    if csv_path:
        with open(csv_path, "w") as f:
            for t, k, r, _ in curve:
                f.write(f"{k},{t},{r}\n")
    return curve


def time_to_rmse(curve, target):
    """Seconds (since the curve's origin) of the first checkpoint with rmse <= target, or None."""
    for t, k, r, _ in curve:
        if r <= target:
            return t
    return None
