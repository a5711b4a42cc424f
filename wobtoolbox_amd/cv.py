"""Learned control variate: a Fourier-feature MLP trained on the walk's own boundary values.

The network (BF16 MFMA kernels, FP32 accumulation and Adam masters) models V_r(y), the
expected value of a single-layer walk with r bounces left from boundary point y.  Training
data are walks of r bounces from uniform boundary points (wob_estimate_boundary with
WOB_EST_BOUNDARY_DENSITY).  Bound to a scene as a per-vertex value table, it drives
WOB_EST_CONTROL_VARIATE: zero-mean corrections that keep plain WoB's expectation.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi
from ._capi import check, lib


def _torch():
    import torch
    return torch


def _stream(stream):
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


class ControlVariate:
    def __init__(self, dim, n_outputs, sigma=1.0, seed=0):
        h = C.c_void_p()
        check(lib().wob_cv_create(dim, n_outputs, float(sigma), int(seed), C.byref(h)))
        self._h = h
        self.dim = dim
        self.n_outputs = n_outputs
        npar = C.c_uint32()
        check(lib().wob_cv_get_info(h, C.byref(npar), None))
        self.n_params = npar.value

    def params(self):
        p = np.zeros(self.n_params, np.float32)
        f = np.zeros(64 * 3, np.float32)
        check(lib().wob_cv_get_params(self._h, p.ctypes.data, f.ctypes.data))
        return p, f

    def set_params(self, params=None, freq=None):
        p = None if params is None else np.ascontiguousarray(params, np.float32)
        f = None if freq is None else np.ascontiguousarray(freq, np.float32)
        check(lib().wob_cv_set_params(self._h, None if p is None else p.ctypes.data,
                                      None if f is None else f.ctypes.data))

    def train(self, pos, out_index, target, batch=16384, steps=1000, lr=2e-3, record_loss=False, stream=None):
        n = pos.shape[0]
        losses = np.zeros(steps, np.float32) if record_loss else None
        check(lib().wob_cv_train(self._h, pos.data_ptr(), out_index.data_ptr(), target.data_ptr(), n, batch, steps,
                                 float(lr), None if losses is None else losses.ctypes.data, _stream(stream)))
        return losses

    def gradient(self, pos, out_index, target, stream=None):
        g = np.zeros(self.n_params, np.float32)
        loss = C.c_float()
        check(lib().wob_cv_gradient(self._h, pos.data_ptr(), out_index.data_ptr(), target.data_ptr(), pos.shape[0],
                                    g.ctypes.data, C.byref(loss), _stream(stream)))
        return g, loss.value

    def infer(self, pos, stream=None):
        torch = _torch()
        out = torch.zeros((pos.shape[0], self.n_outputs), dtype=torch.float32, device=pos.device)
        check(lib().wob_cv_infer(self._h, pos.contiguous().data_ptr(), pos.shape[0], out.data_ptr(), _stream(stream)))
        return out

    def bind(self, scene, n_values, stream=None):
        check(lib().wob_scene_set_control_variate(scene.handle, self._h, None, n_values, _stream(stream)))

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().wob_cv_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def bind_table(scene, table, n_values, stream=None):
    """Bind a caller table [n_values][n_elements][dim] (device tensor) as the control variate."""
    check(lib().wob_scene_set_control_variate(scene.handle, None, table.contiguous().data_ptr(), n_values,
                                              _stream(stream)))


def sample_boundary(scene, n, seed=0, stream_id=0x40000000, stream=None):
    torch = _torch()
    rec = torch.zeros((n, 2 * scene.dim + 4), dtype=torch.float32, device="cuda")
    check(lib().wob_sample_boundary(scene.handle, n, int(seed), stream_id, rec.data_ptr(), _stream(stream)))
    return rec


def training_data(estimator, n_per_value, seed=0, paths_per_sample=1, stream=None):
    """Boundary walks of r = 1..path_length bounces from uniform boundary points.

    Each sample's target is the mean of `paths_per_sample` walks from its point.
    Returns (positions, output index r - 1, target) as device tensors, shuffled."""
    torch = _torch()
    from . import Estimator
    L = estimator.path_length
    pos, idx, val = [], [], []
    for r in range(1, L + 1):
        rec = sample_boundary(estimator.scene, n_per_value, seed=seed + 7919 * r, stream=stream)
        e = Estimator(**{**estimator.__dict__, "path_length": r,
                         "flags": (estimator.flags | _capi.EST_BOUNDARY_DENSITY) & ~_capi.EST_CONTROL_VARIATE})
        e.last_launch = None
        s, q = e.compute_boundary(rec, paths_per_sample, seed=seed + 104729 * r, stream_id=0x20000000 + r,
                                  stream=stream)
        pos.append(rec[:, :estimator.scene.dim])
        idx.append(torch.full((n_per_value,), r - 1, dtype=torch.int32, device=rec.device))
        val.append((s / paths_per_sample).float())
    pos, idx, val = torch.cat(pos), torch.cat(idx), torch.cat(val)
    g = torch.Generator(device="cpu").manual_seed(seed)
    perm = torch.randperm(pos.shape[0], generator=g).to(pos.device)
    return pos[perm].contiguous(), idx[perm].contiguous(), val[perm].contiguous()


def _timing_main():
    """`python -m wobtoolbox_amd.cv [--steps N] [--batch B] [--samples M]`: times the MLP's
    training steps and inference on random data (the network of the time-to-RMSE workload:
    3D inputs, 4 outputs) and prints one JSON line."""
    import argparse
    import json
    import time
    torch = _torch()
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--samples", type=int, default=1 << 20)
    args = ap.parse_args()
    g = torch.Generator(device="cuda").manual_seed(0)
    pos = (torch.rand((args.samples, 3), device="cuda", generator=g) * 2 - 1).contiguous()
    idx = torch.randint(0, 4, (args.samples,), dtype=torch.int32, device="cuda", generator=g)
    tgt = (torch.sin(pos[:, 0] * 3) + pos[:, 1] * pos[:, 2]).contiguous()
    net = ControlVariate(3, 4, sigma=1.0, seed=1)
    net.train(pos, idx, tgt, batch=args.batch, steps=20)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    net.train(pos, idx, tgt, batch=args.batch, steps=args.steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    net.infer(pos)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    for _ in range(5):
        net.infer(pos)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    flop_step = 405504 * args.batch          # forward + backward per sample, 128-wide layers
    step_ms = (t1 - t0) / args.steps * 1e3
    print(json.dumps({"train_ms_per_step": step_ms, "batch": args.batch,
                      "train_tflops": flop_step / (step_ms * 1e-3) / 1e12,
                      "infer_ms_per_million": (t3 - t2) / 5 * 1e3 * (1 << 20) / args.samples}), flush=True)


if __name__ == "__main__":
    _timing_main()
