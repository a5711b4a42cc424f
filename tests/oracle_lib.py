"""ctypes binding of the CPU oracle (oracle/build/libwob_oracle.so).

Test infrastructure: imported only by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, always as the checker or the CPU baseline, never as the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "build", "libwob_oracle.so")

ENUMS = {
    "IndirectSingleLayer": 0, "IndirectDoubleLayer": 1, "Direct": 2,
    "InteriorDomain": 0, "ExteriorDomain": 1,
    "BackwardEstimator": 0, "ForwardEstimator": 1,
    "DirichletProblem": 0, "NeumannProblem": 1, "RobinProblem": 2, "MixedProblem": 3,
}


class EstDesc(C.Structure):
    _fields_ = [("bie_model", C.c_int32), ("domain", C.c_int32), ("estimator", C.c_int32), ("problem", C.c_int32),
                ("path_length", C.c_uint32), ("num_resampling_candidates", C.c_uint32),
                ("num_volume_samples", C.c_uint32), ("eps", C.c_float),
                ("first_kind_equation_scaling_constant", C.c_float), ("pick_mode", C.c_int32),
                ("rng_mode", C.c_int32), ("use_bvh", C.c_int32), ("exclude_origin", C.c_int32),
                ("independent_volume", C.c_int32), ("boundary_density", C.c_int32), ("harmonic_cv", C.c_int32)]


def build():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.orc_scene_create.restype = P
        L.orc_scene_create.argtypes = [C.c_int, C.c_int, P, P, P, P, P]
        L.orc_scene_destroy.argtypes = [P]
        L.orc_scene_totals.argtypes = [P, P]
        L.orc_scene_cdfs.argtypes = [P, P, P]
        L.orc_scene_set_volume_grid.argtypes = [P, C.c_uint, C.c_float, P]
        L.orc_scene_set_volume_gaussians.argtypes = [P, C.c_uint, P, P, P, C.c_int]
        L.orc_inside.argtypes = [P, P]
        L.orc_inside.restype = C.c_int
        L.orc_inside_grid.argtypes = [P, P, P]
        L.orc_inside_grid.restype = C.c_int
        L.orc_intersect_line.argtypes = [P, P, P, C.c_int, C.c_int, P, P]
        L.orc_intersect_line.restype = C.c_int
        L.orc_estimate_domain.argtypes = [P, C.POINTER(EstDesc), P, P, C.c_int, C.c_uint, C.c_uint64, C.c_uint64,
                                          C.c_uint, C.c_int, P, P, P]
        L.orc_estimate_boundary.argtypes = [P, C.POINTER(EstDesc), P, C.c_int, C.c_uint, C.c_uint64, C.c_uint64,
                                            C.c_uint, C.c_int, P, P, P]
        L.orc_eval_kernels_f32.argtypes = [C.c_int, C.c_int, P, P, P, P, C.c_float, P]
        L.orc_eval_kernels_f64.argtypes = [C.c_int, C.c_int, P, P, P, P, C.c_double, P]
        L.orc_eval_weights.argtypes = [C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_int, P, P, P, P, P, P]
        L.orc_eval_domain_terms.argtypes = [C.c_int, C.c_int, C.c_int, C.c_float, C.c_int, P, P, P, P]
        L.orc_philox4x32_10.argtypes = [P, P, P]
        L.orc_triple32.argtypes = [C.c_uint32]
        L.orc_triple32.restype = C.c_uint32
        L.orc_sample_line.argtypes = [P, P, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, P]
        L.orc_sample_line.restype = C.c_float
        L.orc_intersect_lines.argtypes = [P, C.c_int, P, P, C.c_int, C.c_int, C.c_int, P, P, P]
        L.orc_sample_lines.argtypes = [P, C.c_int, P, C.c_uint64, C.c_uint32, C.c_uint32, C.c_int, P, P]
        L.orc_sample_uniform.argtypes = [P, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, P]
        L.orc_sample_uniform.restype = C.c_float
        L.orc_closest_points.argtypes = [P, C.c_int, P, P, P, P]
        L.orc_simple_wob.argtypes = [P, P, C.c_int, C.c_uint, C.c_uint, C.c_uint64, C.c_uint64, C.c_uint, C.c_int, P,
                                     P]
        L.orc_wos.argtypes = [P, P, C.c_int, C.c_uint, C.c_uint64, C.c_uint64, C.c_uint, C.c_float, C.c_uint, C.c_int,
                              P, P, P]
        L.orc_vpl_aggregate.argtypes = [C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_int, P, P, C.c_int, P,
                                        C.c_int, P, P]
        L.orc_scene_set_harmonic_cv.argtypes = [P, C.c_uint, P, C.c_float]
        L.orc_eval_harmonic_cv.argtypes = [P, C.c_int, P, P]
        L.orc_hcv_targets.argtypes = [P, C.c_int, C.c_uint64, C.c_uint32, C.c_uint, C.c_float, P]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def est_desc(solver: dict, pick_mode=1, rng_mode=0, use_bvh=1, num_volume_samples=0) -> EstDesc:
    """Solver dict (reference JSON keys + the device estimator flags) -> oracle descriptor.
    flags bit 0: exclude the origin element of boundary line samples; bit 2: boundary density
    (the walk's value at boundary records); bit 3: independent volume samples; bit 4: the
    scene's harmonic control variate (single and double layer)."""
    fl = int(solver.get("flags", 0))
    nvs = num_volume_samples or int(solver.get("num_volume_samples", 0))
    return EstDesc(ENUMS[solver["bie_model"]], ENUMS[solver["domain"]], ENUMS[solver["estimator"]],
                   ENUMS[solver["problem"]], solver["path_length"], solver.get("num_resampling_candidates", 1),
                   nvs, solver.get("eps", float(np.finfo(np.float32).eps)),
                   solver.get("first_kind_equation_scaling_constant", 1.0), pick_mode, rng_mode, use_bvh, fl & 1,
                   (fl >> 3) & 1, (fl >> 2) & 1, (fl >> 4) & 1)


class OracleScene:
    def __init__(self, el):
        self.el = el
        self.dim = el.dim
        self.h = lib().orc_scene_create(el.dim, el.n, _p(el.vertices), _p(el.normals), _p(el.boundary_values),
                                        _p(el.boundary_types), _p(el.robin_alphas))
        assert self.h

    def set_harmonic_cv(self, sources, c0):
        """Bind the harmonic control variate: sources (m, dim + 1) = positions then weights."""
        s = np.ascontiguousarray(sources, np.float32).reshape(-1, self.dim + 1)
        lib().orc_scene_set_harmonic_cv(self.h, len(s), _p(s), float(c0))

    def eval_harmonic_cv(self, points):
        pts = np.ascontiguousarray(points, np.float32).reshape(-1, self.dim)
        out = np.zeros((len(pts), self.dim + 1), np.float32)
        lib().orc_eval_harmonic_cv(self.h, len(pts), _p(pts), _p(out))
        return out

    def hcv_targets(self, n, seed=0, stream=0, volume_samples=0, eps=1e-4):
        """BoundaryPoint records (n, 2 dim + 4) of the harmonic control variate's training
        samples, the boundary value replaced by the target."""
        rec = np.zeros((n, 2 * self.dim + 4), np.float32)
        lib().orc_hcv_targets(self.h, n, seed, stream, volume_samples, eps, _p(rec))
        return rec

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_scene_destroy(self.h)
            self.h = None

    def totals(self):
        t = np.zeros(2, np.float32)
        lib().orc_scene_totals(self.h, _p(t))
        return t

    def cdfs(self):
        a = np.zeros(self.el.n, np.float32)
        b = np.zeros(self.el.n, np.float32)
        lib().orc_scene_cdfs(self.h, _p(a), _p(b))
        return a, b

    def set_volume_grid(self, grid_res, domain_size, cache):
        cache = np.ascontiguousarray(cache, np.float32)
        lib().orc_scene_set_volume_grid(self.h, grid_res, domain_size, _p(cache))

    def set_volume_gaussians(self, centers, widths, amplitudes, mask=True):
        c = np.ascontiguousarray(centers, np.float32)
        w = np.ascontiguousarray(widths, np.float32)
        a = np.ascontiguousarray(amplitudes, np.float32)
        lib().orc_scene_set_volume_gaussians(self.h, len(w), _p(c), _p(w), _p(a), int(mask))

    def inside_grid(self, points):
        """(verdict, cell) per point from the masked Gaussian source's occupancy grid (cell 2:
        the point's own ray decided)."""
        pts = np.ascontiguousarray(points, np.float32)
        cell = C.c_int(0)
        out = np.zeros(len(pts), np.int32)
        cells = np.zeros(len(pts), np.int32)
        for i in range(len(pts)):
            out[i] = lib().orc_inside_grid(self.h, _p(pts[i]), C.byref(cell))
            cells[i] = cell.value
        return out, cells

    def inside(self, points):
        pts = np.ascontiguousarray(points, np.float32)
        return np.array([lib().orc_inside(self.h, _p(pts[i])) for i in range(len(pts))], bool)

    def intersect(self, origin, direction, use_bvh=True, cap=256):
        o = np.ascontiguousarray(origin, np.float32)
        d = np.ascontiguousarray(direction, np.float32)
        ids = np.zeros(cap, np.int32)
        ts = np.zeros(cap, np.float32)
        c = lib().orc_intersect_line(self.h, _p(o), _p(d), int(use_bvh), cap, _p(ids), _p(ts))
        m = min(c, cap)
        return c, ids[:m], ts[:m]

    def intersect_many(self, origins, directions, use_bvh=False, cap=64, threads=0):
        """All hits of n lines (brute force by default): counts (n,), ids and t (n, cap)."""
        o = np.ascontiguousarray(origins, np.float32)
        d = np.ascontiguousarray(directions, np.float32)
        n = o.shape[0]
        counts = np.zeros(n, np.int32)
        ids = np.full((n, cap), -1, np.int32)
        ts = np.zeros((n, cap), np.float32)
        lib().orc_intersect_lines(self.h, n, _p(o), _p(d), int(use_bvh), cap, threads, _p(counts), _p(ids), _p(ts))
        return counts, ids, ts

    def sample_lines(self, points, seed=0, path=0, stream=0, pick_mode=1):
        """One line sample from each point x_i on the Philox stream (seed, path, i, stream):
        boundary records (n, 2 dim + 4) and inverse pdfs (n,)."""
        x = np.ascontiguousarray(points, np.float32)
        n = x.shape[0]
        rec = np.zeros((n, 2 * self.dim + 4), np.float32)
        inv = np.zeros(n, np.float32)
        lib().orc_sample_lines(self.h, n, _p(x), seed, path, stream, pick_mode, _p(rec), _p(inv))
        return rec, inv

    def estimate_domain(self, solver, points, paths, seed=0, path_base=0, stream=0, threads=0, pick_mode=1,
                        rng_mode=0, point_index=None, return_paths=False, num_volume_samples=0):
        pts = np.ascontiguousarray(points, np.float32)
        n = pts.shape[0]
        D = self.dim
        s = np.zeros((n, D + 1), np.float64)
        q = np.zeros((n, D + 1), np.float64)
        pi = None if point_index is None else np.ascontiguousarray(point_index, np.uint32)
        po = np.zeros((n, paths, D + 1), np.float32) if return_paths else None
        d = est_desc(solver, pick_mode, rng_mode, 1, num_volume_samples)
        lib().orc_estimate_domain(self.h, C.byref(d), _p(pts), _p(pi), n, paths, seed, path_base, stream, threads,
                                  _p(s), _p(q), _p(po))
        return (s, q, po) if return_paths else (s, q)

    def estimate_boundary(self, solver, records, paths, seed=0, path_base=0, stream=0, threads=0, pick_mode=1,
                          return_paths=False):
        r = np.ascontiguousarray(records, np.float32)
        n = r.shape[0]
        s = np.zeros(n, np.float64)
        q = np.zeros(n, np.float64)
        po = np.zeros((n, paths), np.float32) if return_paths else None
        d = est_desc(solver, pick_mode, 0, 1, 0)
        lib().orc_estimate_boundary(self.h, C.byref(d), _p(r), n, paths, seed, path_base, stream, threads, _p(s),
                                    _p(q), _p(po))
        return (s, q, po) if return_paths else (s, q)

    def sample_uniform(self, seed, path, point, stream):
        """One uniform boundary sample from the Philox stream (seed, path, point, stream)."""
        rec = np.zeros(2 * self.dim + 4, np.float32)
        inv = lib().orc_sample_uniform(self.h, seed, path, point, stream, _p(rec))
        return rec, inv

    def closest_points(self, points):
        """Brute-force nearest element, distance and interpolated boundary value per point."""
        x = np.ascontiguousarray(points, np.float32)
        n = x.shape[0]
        e = np.zeros(n, np.int32)
        d = np.zeros(n, np.float32)
        v = np.zeros(n, np.float32)
        lib().orc_closest_points(self.h, n, _p(x), _p(e), _p(d), _p(v))
        return e, d, v

    def wos(self, points, paths, eps, seed=0, path_base=0, stream=0, max_steps=1 << 16, threads=0,
            return_paths=False):
        """Walk on spheres with the device's Philox streams: (sum, sumsq[, per-path values])."""
        x = np.ascontiguousarray(points, np.float32)
        n = x.shape[0]
        s = np.zeros(n, np.float64)
        q = np.zeros(n, np.float64)
        po = np.zeros((n, paths), np.float32) if return_paths else None
        lib().orc_wos(self.h, _p(x), n, paths, seed, path_base, stream, float(eps), max_steps, threads, _p(s), _p(q),
                      _p(po))
        return (s, q, po) if return_paths else (s, q)

    def simple_wob(self, points, paths, path_length, seed=0, path_base=0, stream=0, pick_mode=1):
        """The app's double-layer walk (WoBFunc): (sum, per-path values)."""
        x = np.ascontiguousarray(points, np.float32)
        n = x.shape[0]
        s = np.zeros(n, np.float64)
        po = np.zeros((n, paths), np.float32)
        lib().orc_simple_wob(self.h, _p(x), n, paths, path_length, seed, path_base, stream, pick_mode, _p(s), _p(po))
        return s, po

    def vpl_paths(self, solver, n, seed=0, path_base=0, stream=0, threads=0):
        """The VPL path phase with the device's streams: start point i from (seed, path_base, i,
        stream | 2^30), then one boundary-density walk from (seed, path_base, i, stream)."""
        rec = np.stack([self.sample_uniform(seed, path_base, i, stream | (1 << 30))[0] for i in range(n)])
        sol = dict(solver, flags=int(solver.get("flags", 0)) | 4)
        s, _ = self.estimate_boundary(sol, rec, 1, seed=seed, path_base=path_base, stream=stream, threads=threads)
        return rec, s.astype(np.float32)


def vpl_aggregate(dim, bie, domain, eps, inv_pdf, records, contrib, points, threads=0):
    """FP64 sums of c_j * [solution_domain_unknown, gradient_domain_unknown](x_i, y_j, inv_pdf)
    and of their absolute values, each (n_points, dim + 1)."""
    r = np.ascontiguousarray(records, np.float32)
    c = np.ascontiguousarray(contrib, np.float32)
    x = np.ascontiguousarray(points, np.float32)
    s = np.zeros((x.shape[0], dim + 1), np.float64)
    a = np.zeros_like(s)
    lib().orc_vpl_aggregate(dim, bie, domain, float(eps), float(inv_pdf), r.shape[0], _p(r), _p(c), x.shape[0],
                            _p(x), threads, _p(s), _p(a))
    return s, a


def eval_kernels(dim, y, x, ny, nx, eps, f64=False):
    n = y.shape[0]
    W = 3 + 2 * dim
    if f64:
        a = [np.ascontiguousarray(v, np.float64) for v in (y, x, ny, nx)]
        out = np.zeros((n, W), np.float64)
        lib().orc_eval_kernels_f64(dim, n, *[_p(v) for v in a], float(eps), _p(out))
    else:
        a = [np.ascontiguousarray(v, np.float32) for v in (y, x, ny, nx)]
        out = np.zeros((n, W), np.float32)
        lib().orc_eval_kernels_f32(dim, n, *[_p(v) for v in a], float(eps), _p(out))
    return out


def eval_weights(dim, bie, domain, eps, fkesc, xb, yb, inv_pdf, xd, u_split):
    n = xb.shape[0]
    out = np.zeros((n, 5 + dim), np.float32)
    a = [np.ascontiguousarray(v, np.float32) for v in (xb, yb, inv_pdf, xd, u_split)]
    lib().orc_eval_weights(dim, bie, domain, float(eps), float(fkesc), n, *[_p(v) for v in a], _p(out))
    return out


def eval_domain_terms(dim, bie, domain, eps, xd, yb, inv_pdf):
    """Per (domain point, boundary record) pair: [known, unknown, grad known (dim), grad unknown
    (dim)] of the model's domain representation (no volume source)."""
    n = xd.shape[0]
    out = np.zeros((n, 2 + 2 * dim), np.float32)
    a = [np.ascontiguousarray(v, np.float32) for v in (xd, yb, inv_pdf)]
    lib().orc_eval_domain_terms(dim, bie, domain, float(eps), n, *[_p(v) for v in a], _p(out))
    return out


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    o = np.zeros(4, np.uint32)
    lib().orc_philox4x32_10(_p(c), _p(k), _p(o))
    return o
