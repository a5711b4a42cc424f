"""MI355X walk-on-boundary engine: Python mirror of the reference's Scene / Estimator API.

The reference (rsugimoto/WoBToolbox) exposes, on its hot path,
  SceneHost(elements)                                        wob/wob/scene.cuh:264-299
  Estimator<..., BIE, Domain, Estimator, Problem>(...)       wob/wob/estimator.cuh:118-152
  compute_sample_path_contribution_domain / _boundary        wob/wob/estimator.cuh:154-222
  utils::run<DeviceFunctor>(args, n, bie, domain, est, prob) wob/wob/runtime_utils.cuh:105-153
  string_to_*_type / *_type_to_string                        bie_models.cuh:24-51, estimator.cuh:25-54
This package keeps those names and meanings; the work runs in libwob_hip.so (hand-written
gfx950 HIP kernels behind the C ABI of include/wob_c.h).  PyTorch only provides device
memory and streams.  There is no CPU fallback: a missing library or a missing GPU raises.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _capi
from ._capi import WobError, check, lib
from .geometry import Elements

__all__ = ["SceneHost", "Estimator", "Batch", "WobError", "Elements", "tune_bvh_width", "string_to_bie_model_type",
           "string_to_domain_type",
           "string_to_estimator_type", "string_to_problem_type", "string_to_equation_kernel_type",
           "bie_model_type_to_string", "domain_type_to_string", "estimator_type_to_string",
           "problem_type_to_string", "Dirichlet", "Neumann", "Robin"]

Dirichlet, Neumann, Robin = 0, 1, 2
PoissonKernel = 0
IndirectSingleLayer, IndirectDoubleLayer, Direct = 0, 1, 2
InteriorDomain, ExteriorDomain = 0, 1
BackwardEstimator, ForwardEstimator = 0, 1
DirichletProblem, NeumannProblem, RobinProblem, MixedProblem = 0, 1, 2, 3

_BIE = ["IndirectSingleLayer", "IndirectDoubleLayer", "Direct"]
_DOM = ["InteriorDomain", "ExteriorDomain"]
_EST = ["BackwardEstimator", "ForwardEstimator"]
_PROB = ["DirichletProblem", "NeumannProblem", "RobinProblem", "MixedProblem"]


# string <-> enum (bie_models.cuh:24-51, estimator.cuh:25-54, equation_kernel.cuh:25-34);
# unknown strings map to the *TypeCount value, as in the reference.
def string_to_equation_kernel_type(s): return 0 if s == "PoissonKernel" else 1
def string_to_bie_model_type(s): return _BIE.index(s) if s in _BIE else len(_BIE)
def string_to_domain_type(s): return _DOM.index(s) if s in _DOM else len(_DOM)
def string_to_estimator_type(s): return _EST.index(s) if s in _EST else len(_EST)
def string_to_problem_type(s): return _PROB.index(s) if s in _PROB else len(_PROB)
def bie_model_type_to_string(v): return _BIE[v] if 0 <= v < 3 else "InvalidBIEModelType"
def domain_type_to_string(v): return _DOM[v] if 0 <= v < 2 else "InvalidDomainType"
def estimator_type_to_string(v): return _EST[v] if 0 <= v < 2 else "InvalidEstimatorType"
def problem_type_to_string(v): return _PROB[v] if 0 <= v < 4 else "InvalidProblemType"


def _enum(v, conv):
    return conv(v) if isinstance(v, str) else int(v)


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("the walk-on-boundary engine needs a GPU (torch.cuda.is_available() is False)")
    return torch


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _stream_ptr(stream):
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def _scene_desc(el):
    return _capi.SceneDesc(el.dim, el.n, el.vertices.ctypes.data, el.normals.ctypes.data,
                           el.boundary_values.ctypes.data, el.boundary_types.ctypes.data, el.robin_alphas.ctypes.data)


def tune_bvh_width(elements: Elements, solver: dict, points, paths_per_point, seed=0, device_bvh=False, stream=None):
    """BVH width autotuner (wob_tune_bvh_width): times a walk of the solver (a reference JSON
    config dict) at the device points on scenes with nodes of at most 2, 3 and 4 children.
    Returns (best width, {width: walk ms or nan})."""
    _torch()
    el = elements.astype_contiguous()
    d = _scene_desc(el)
    ed = Estimator.from_config(None, solver).desc()
    pts = points.contiguous()
    best = C.c_uint32()
    ms = (C.c_float * 3)()
    check(lib().wob_tune_bvh_width(C.byref(d), _capi.SCENE_DEVICE_BVH if device_bvh else 0, C.byref(ed), _ptr(pts),
                                   pts.shape[0], int(paths_per_point), int(seed), C.byref(best), ms,
                                   _stream_ptr(stream)))
    return best.value, {w: float(ms[w - 2]) for w in (2, 3, 4)}


class SceneHost:
    """Device scene: wide BVH, CDFs and element data (SceneHost, scene.cuh:264-326)."""

    def __init__(self, elements: Elements, device_bvh: bool = False, bvh_width: int = 4):
        """device_bvh: build the BVH on the GPU (wob_scene_create_ex, WOB_SCENE_DEVICE_BVH)
        instead of the host's SAH build; bvh_width: the most children a node takes (2..4; see
        tune_bvh_width).  Walks give the same results either way."""
        _torch()
        el = elements.astype_contiguous()
        self.elements = el
        self.dim = el.dim
        d = _scene_desc(el)
        h = C.c_void_p()
        flags = (_capi.SCENE_DEVICE_BVH if device_bvh else 0) | _capi.scene_bvh_width(bvh_width)
        check(lib().wob_scene_create_ex(C.byref(d), flags, C.byref(h)))
        self._h = h
        self._has_volume = False

    @property
    def handle(self):
        return self._h

    def info(self) -> dict:
        i = _capi.SceneInfo()
        check(lib().wob_scene_get_info(self._h, C.byref(i)))
        return {f: getattr(i, f) for f, _ in i._fields_}

    def closest_points(self, points, stream=None):
        """Nearest element (int32, caller order), distance and boundary value at the closest
        point for device points (n, dim) (apps/simple_wob_and_wos.cu:48-79)."""
        torch = _torch()
        pts = points.contiguous()
        n = pts.shape[0]
        e = torch.empty(n, dtype=torch.int32, device=pts.device)
        d = torch.empty(n, dtype=torch.float32, device=pts.device)
        v = torch.empty(n, dtype=torch.float32, device=pts.device)
        check(lib().wob_closest_points(self._h, _ptr(pts), n, _ptr(e), _ptr(d), _ptr(v), _stream_ptr(stream)))
        return e, d, v

    def walk_on_spheres(self, points, paths_per_point, eps=1e-4, seed=0, path_base=0, point_base=0, stream_id=0,
                        max_steps=0, out=None, stream=None):
        """Walk-on-spheres path sums of an interior Dirichlet problem (WoSFunc,
        apps/simple_wob_and_wos.cu:116-141): (sum, sumsq) FP64 tensors of shape (n,),
        accumulated into `out` if given."""
        torch = _torch()
        pts = points.contiguous()
        n = pts.shape[0]
        if out is None:
            s = torch.zeros(n, dtype=torch.float64, device=pts.device)
            q = torch.zeros(n, dtype=torch.float64, device=pts.device)
        else:
            s, q = out
        check(lib().wob_wos_estimate(self._h, _ptr(pts), n, point_base, paths_per_point, seed, path_base, stream_id,
                                     float(eps), max_steps, _ptr(s), _ptr(q), _stream_ptr(stream)))
        return s, q

    def set_volume_cache(self, grid_res, domain_size, cache):
        """Grid source over (-domain_size/2, domain_size/2)^dim (SceneHost::set_volume_cache,
        scene.cuh:301-306); cache: grid_res^dim values, x fastest.  None removes the source."""
        if cache is None:
            check(lib().wob_scene_set_volume_grid(self._h, 0, 0.0, None))
            self._has_volume = False
            return
        c = np.ascontiguousarray(cache, np.float32).reshape(-1)
        if c.size != grid_res ** self.dim:
            raise ValueError(f"cache has {c.size} values, expected grid_res^dim = {grid_res ** self.dim}")
        check(lib().wob_scene_set_volume_grid(self._h, int(grid_res), float(domain_size), c.ctypes.data))
        self._has_volume = True

    def set_volume_gaussians(self, centers, widths, amplitudes, mask=True):
        """Gaussian-mixture source sum_j A_j exp(-|x - c_j|^2 / (2 w_j^2)); with mask, samples
        outside the domain contribute 0."""
        c = np.ascontiguousarray(centers, np.float32).reshape(-1, self.dim)
        w = np.ascontiguousarray(widths, np.float32).reshape(-1)
        a = np.ascontiguousarray(amplitudes, np.float32).reshape(-1)
        if not (len(c) == len(w) == len(a)):
            raise ValueError("centers, widths and amplitudes must have one entry per Gaussian")
        check(lib().wob_scene_set_volume_gaussians(self._h, len(w), c.ctypes.data, w.ctypes.data, a.ctypes.data,
                                                   0 if mask else _capi.VOLUME_NO_MASK))
        self._has_volume = len(w) > 0

    @staticmethod
    def _hcv_desc(n_sources, radius_scale, n_train, volume_samples, eps, ridge, seed, stream_id, exterior=False,
                  zero_charge=False):
        flags = (_capi.HCV_EXTERIOR if exterior else 0) | (_capi.HCV_ZERO_CHARGE if zero_charge else 0)
        return _capi.HcvDesc(int(n_sources), float(radius_scale), int(n_train), int(volume_samples), float(eps),
                             float(ridge), int(seed), int(stream_id), flags)

    def fit_harmonic_cv(self, n_sources=64, radius_scale=None, n_train=16384, volume_samples=16, eps=1e-4,
                        ridge=1e-10, seed=0, stream_id=0, exterior=False, zero_charge=False, stream=None):
        """Fit and bind the harmonic control variate u_h = c0 + sum_j w_j phi(|x - s_j|^2) of
        WOB_EST_HARMONIC_CV walks (wob_scene_fit_harmonic_cv) to the walk's known-term data
        at n_train uniform boundary samples; volume_samples is ignored without a volume
        source.  exterior=True fits the exterior domain's (sources inside the body at
        radius_scale x the smallest centre-to-element distance, default 0.75; c0 = 0), and
        zero_charge=True with it keeps sum_j w_j = 0 (exterior double-layer walks need it); the
        interior fit's sources lie at radius_scale x the bounding radius (default 1.5).
        Returns the RMS residual of the fit."""
        _torch()
        if not self._has_volume:
            volume_samples = 0
        if radius_scale is None:
            radius_scale = 0.75 if exterior else 1.5
        d = self._hcv_desc(n_sources, radius_scale, n_train, volume_samples, eps, ridge, seed, stream_id, exterior,
                           zero_charge)
        rms = C.c_float()
        check(lib().wob_scene_fit_harmonic_cv(self._h, C.byref(d), C.byref(rms), _stream_ptr(stream)))
        return rms.value

    def harmonic_cv_targets(self, n_train=16384, volume_samples=16, eps=1e-4, seed=0, stream_id=0, n_sources=1,
                            exterior=False, stream=None):
        """The fit's training data: BoundaryPoint records (n, 2 dim + 4) whose boundary value
        is the target."""
        torch = _torch()
        if not self._has_volume:
            volume_samples = 0
        d = self._hcv_desc(n_sources, 0.75 if exterior else 1.5, n_train, volume_samples, eps, 0.0, seed, stream_id,
                           exterior)
        rec = torch.zeros((n_train, 2 * self.dim + 4), dtype=torch.float32, device="cuda")
        check(lib().wob_harmonic_cv_targets(self._h, C.byref(d), _ptr(rec), _stream_ptr(stream)))
        return rec

    def set_harmonic_cv(self, sources=None, c0=0.0):
        """Bind explicit sources (n, dim + 1): positions then weights; None unbinds."""
        if sources is None:
            check(lib().wob_scene_set_harmonic_cv(self._h, 0, None, 0.0))
            return
        s = np.ascontiguousarray(sources, np.float32).reshape(-1, self.dim + 1)
        check(lib().wob_scene_set_harmonic_cv(self._h, len(s), s.ctypes.data, float(c0)))

    def harmonic_cv(self):
        """(sources (n, dim + 1), c0) of the bound harmonic control variate (n = 0: none)."""
        m = C.c_uint32()
        check(lib().wob_scene_get_harmonic_cv(self._h, C.byref(m), None, None))
        s = np.zeros((m.value, self.dim + 1), np.float32)
        c0 = C.c_float()
        check(lib().wob_scene_get_harmonic_cv(self._h, C.byref(m), s.ctypes.data if m.value else None, C.byref(c0)))
        return s, c0.value

    def eval_harmonic_cv(self, points, stream=None):
        """u_h and grad u_h at device points (n, dim) with the walk's arithmetic: (n, dim + 1)."""
        torch = _torch()
        pts = points.contiguous()
        out = torch.zeros((pts.shape[0], self.dim + 1), dtype=torch.float32, device=pts.device)
        check(lib().wob_eval_harmonic_cv(self._h, _ptr(pts), pts.shape[0], _ptr(out), _stream_ptr(stream)))
        return out

    def intersect_lines(self, origins, dirs, max_hits=64, stream=None):
        """All-hit line query: (counts, ids, t) for device tensors of lines."""
        torch = _torch()
        n = origins.shape[0]
        counts = torch.zeros(n, dtype=torch.int32, device=origins.device)
        ids = torch.zeros((n, max_hits), dtype=torch.int32, device=origins.device)
        ts = torch.zeros((n, max_hits), dtype=torch.float32, device=origins.device)
        check(lib().wob_intersect_lines(self._h, _ptr(origins.contiguous()), _ptr(dirs.contiguous()), n, max_hits,
                                        _ptr(counts), _ptr(ids), _ptr(ts), _stream_ptr(stream)))
        return counts, ids, ts

    def sample_lines(self, points, seed=0, path=0, stream_id=0, stream=None):
        """One line-intersection sample per point with the walk's sampler: (records, inv_pdf);
        point i uses the Philox stream (seed, path, i, stream_id)."""
        torch = _torch()
        n = points.shape[0]
        rec = torch.zeros((n, 2 * self.dim + 4), dtype=torch.float32, device=points.device)
        inv = torch.zeros(n, dtype=torch.float32, device=points.device)
        check(lib().wob_sample_lines(self._h, _ptr(points.contiguous()), n, seed, path, stream_id, _ptr(rec),
                                     _ptr(inv), _stream_ptr(stream)))
        return rec, inv

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().wob_scene_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class Estimator:
    """The reference's Estimator template arguments and constructor parameters."""

    scene: SceneHost
    bie_model: object = "IndirectDoubleLayer"
    domain: object = "InteriorDomain"
    estimator: object = "BackwardEstimator"
    problem: object = "DirichletProblem"
    path_length: int = 4
    num_resampling_candidates: int = 1
    eps: float = float(np.finfo(np.float32).eps)
    first_kind_equation_scaling_constant: float = 1.0
    num_volume_samples: int = 0
    flags: int = 0
    cv_samples: int = 0
    paths_per_lane: int = 0
    block_size: int = 0
    schedule: int = 0            # _capi.SCHEDULE_AUTO / SCHEDULE_LANE / SCHEDULE_WAVEFRONT
    regen_threshold: int = 0
    launch_flags: int = 0        # _capi.LAUNCH_ONE_LANE_PER_ITEM

    @classmethod
    def from_config(cls, scene, cfg: dict, **overrides):
        keys = ["bie_model", "domain", "estimator", "problem", "path_length", "num_resampling_candidates", "eps",
                "first_kind_equation_scaling_constant", "num_volume_samples", "flags", "cv_samples"]
        kw = {k: cfg[k] for k in keys if k in cfg}
        kw.update(overrides)
        return cls(scene, **kw)

    def desc(self) -> _capi.EstimatorDesc:
        return _capi.EstimatorDesc(
            PoissonKernel, _enum(self.bie_model, string_to_bie_model_type), _enum(self.domain, string_to_domain_type),
            _enum(self.estimator, string_to_estimator_type), _enum(self.problem, string_to_problem_type),
            int(self.path_length), int(self.num_resampling_candidates), int(self.num_volume_samples), float(self.eps),
            float(self.first_kind_equation_scaling_constant), int(self.flags), int(self.cv_samples))

    time_walk: bool = False
    last_launch: object = None

    def _launch(self):
        lc = _capi.LaunchConfig(int(self.paths_per_lane), int(self.block_size),
                                (_capi.LAUNCH_TIME_WALK if self.time_walk else 0) | int(self.launch_flags), 0.0, 0, 0,
                                int(self.schedule), int(self.regen_threshold), 0)
        self.last_launch = lc
        return lc

    def compute_domain(self, points, paths_per_point, seed=0, path_base=0, point_base=0, stream_id=0, out=None,
                       stream=None):
        """Accumulate [u, grad u] path sums and squared sums at device points (n, dim).

        Returns (sum, sumsq) FP64 tensors of shape (n, dim + 1), added into `out` if given.
        """
        torch = _torch()
        pts = points.contiguous()
        n = pts.shape[0]
        V = self.scene.dim + 1
        if out is None:
            s = torch.zeros((n, V), dtype=torch.float64, device=pts.device)
            q = torch.zeros((n, V), dtype=torch.float64, device=pts.device)
        else:
            s, q = out
        d, lc = self.desc(), self._launch()
        check(lib().wob_estimate_domain(self.scene.handle, C.byref(d), _ptr(pts), n, point_base, paths_per_point,
                                        seed, path_base, stream_id, _ptr(s), _ptr(q), C.byref(lc),
                                        _stream_ptr(stream)))
        return s, q

    def compute_boundary(self, records, paths_per_point, seed=0, path_base=0, point_base=0, stream_id=0, out=None,
                         stream=None):
        """Accumulate on-boundary path sums at BoundaryPoint records (n, 2*dim + 4)."""
        torch = _torch()
        rec = records.contiguous()
        n = rec.shape[0]
        if out is None:
            s = torch.zeros(n, dtype=torch.float64, device=rec.device)
            q = torch.zeros(n, dtype=torch.float64, device=rec.device)
        else:
            s, q = out
        d, lc = self.desc(), self._launch()
        check(lib().wob_estimate_boundary(self.scene.handle, C.byref(d), _ptr(rec), n, point_base, paths_per_point,
                                          seed, path_base, stream_id, _ptr(s), _ptr(q), C.byref(lc),
                                          _stream_ptr(stream)))
        return s, q

    def compute_vpl_paths(self, n_paths, seed=0, path_base=0, stream_id=0, device="cuda", out=None, stream=None):
        """Path phase of the VPL method (apps/potential_flow_with_vpl.cu:73-88): n_paths uniform
        boundary start points and the boundary-density walk value at each.

        Returns (records (n, 2*dim + 4) float32, contributions (n,) float32) device tensors,
        written into `out` = (records, contributions) if given."""
        torch = _torch()
        W = 2 * self.scene.dim + 4
        if out is not None:
            rec, c = out
            if rec.shape != (n_paths, W) or c.shape != (n_paths,) or not (rec.is_contiguous() and c.is_contiguous()):
                raise ValueError("out must be contiguous ((n_paths, 2 dim + 4), (n_paths,)) float32 tensors")
        else:
            rec = torch.empty((n_paths, W), dtype=torch.float32, device=device)
            c = torch.empty(n_paths, dtype=torch.float32, device=device)
        d, lc = self.desc(), self._launch()
        check(lib().wob_vpl_paths(self.scene.handle, C.byref(d), int(n_paths), seed, path_base, stream_id, _ptr(rec),
                                  _ptr(c), C.byref(lc), _stream_ptr(stream)))
        return rec, c

    def aggregate_vpl(self, records, contrib, points, out=None, stream=None):
        """Aggregation of the VPL method (apps/potential_flow_with_vpl.cu:44-71): adds
        sum_j c_j [solution_domain_unknown, gradient_domain_unknown](x_i, y_j, |boundary|) into an
        FP64 (n, dim + 1) tensor (created when `out` is None) and returns it."""
        torch = _torch()
        rec, c, pts = records.contiguous(), contrib.contiguous(), points.contiguous()
        n = pts.shape[0]
        s = torch.zeros((n, self.scene.dim + 1), dtype=torch.float64, device=pts.device) if out is None else out
        d = self.desc()
        check(lib().wob_vpl_aggregate(self.scene.handle, C.byref(d), _ptr(rec), _ptr(c), rec.shape[0], _ptr(pts), n,
                                      _ptr(s), _stream_ptr(stream)))
        return s


class Batch:
    """Many (scene, estimator, point set) entries in one launch per pass (wob_batch_*)."""

    def __init__(self, entries, boundary=False, paths_per_lane=0, block_size=0, persistent=False, use_graph=False,
                 schedule=0, regen_threshold=0):
        """entries: list of (Estimator, device points, paths_per_point[, stream_id[, path_base]]).

        Each entry gets FP64 (sum, sumsq) device buffers, accumulated over passes."""
        torch = _torch()
        self.outputs = []
        arr = (_capi.BatchEntry * len(entries))()
        self._keep = []
        for i, ent in enumerate(entries):
            est, pts, paths = ent[0], ent[1].contiguous(), int(ent[2])
            sid = int(ent[3]) if len(ent) > 3 else i
            pbase = int(ent[4]) if len(ent) > 4 else 0
            V = 1 if boundary else est.scene.dim + 1
            shape = (pts.shape[0],) if boundary else (pts.shape[0], V)
            s = torch.zeros(shape, dtype=torch.float64, device=pts.device)
            q = torch.zeros_like(s)
            self.outputs.append((s, q))
            self._keep += [pts, est]
            arr[i] = _capi.BatchEntry(est.scene.handle, est.desc(), pts.data_ptr(), pts.shape[0], paths, 0, sid, pbase,
                                      s.data_ptr(), q.data_ptr())
        cfg = _capi.BatchConfig(int(boundary), paths_per_lane, block_size, int(persistent), int(use_graph),
                                int(schedule), int(regen_threshold))
        h = C.c_void_p()
        check(lib().wob_batch_create(arr, len(entries), C.byref(cfg), C.byref(h)))
        self._h = h
        self.paths = [int(e[2]) for e in entries]
        self.passes = 0

    def run(self, n_passes=1, seed=0, stream=None):
        check(lib().wob_batch_run(self._h, seed, n_passes, _stream_ptr(stream)))
        self.passes += n_passes

    def tune(self, seed=0, stream=None):
        best = _capi.BatchConfig()
        ms = C.c_float()
        check(lib().wob_batch_tune(self._h, seed, C.byref(best), C.byref(ms), _stream_ptr(stream)))
        return {f: getattr(best, f) for f, _ in best._fields_}, ms.value

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().wob_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def mean_and_se(s, q, n_paths):
    """Per-point mean and standard error from path sums (FP64)."""
    m = s / n_paths
    var = (q / n_paths - m * m).clamp_min(0) if hasattr(s, "clamp_min") else np.maximum(q / n_paths - m * m, 0)
    return m, (var / n_paths) ** 0.5


def eval_kernels(dim, y, x, ny, nx, eps, stream=None):
    torch = _torch()
    n = y.shape[0]
    out = torch.zeros((n, 3 + 2 * dim), dtype=torch.float32, device=y.device)
    check(lib().wob_eval_kernels(dim, n, _ptr(y), _ptr(x), _ptr(ny), _ptr(nx), float(eps), _ptr(out),
                                 _stream_ptr(stream)))
    return out


def eval_weights(dim, bie_model, domain, eps, fkesc, x, y, inv_pdf, xdomain, u_split, stream=None):
    torch = _torch()
    n = x.shape[0]
    out = torch.zeros((n, 5 + dim), dtype=torch.float32, device=x.device)
    check(lib().wob_eval_weights(dim, _enum(bie_model, string_to_bie_model_type), _enum(domain, string_to_domain_type),
                                 float(eps), float(fkesc), n, _ptr(x), _ptr(y), _ptr(inv_pdf), _ptr(xdomain),
                                 _ptr(u_split), _ptr(out), _stream_ptr(stream)))
    return out


def philox4x32_10(ctr, key, stream=None):
    torch = _torch()
    n = ctr.shape[0]
    out = torch.zeros((n, 4), dtype=torch.int32, device=ctr.device)
    check(lib().wob_philox4x32_10(n, _ptr(ctr), _ptr(key), _ptr(out), _stream_ptr(stream)))
    return out
