"""Synthetic scenes and evaluation sets for the walk-on-boundary workloads.

Geometry and boundary data are generated in FP64 with numpy and rounded to FP32 once,
in the element layout of the reference's ``ElementXs`` (wob/wob/scene.cuh:378-382 for
segments, :646-650 for triangles): per element the vertices, one normal, per-vertex
boundary values, a boundary type and per-vertex Robin coefficients.

2D segments follow the reference harness convention (apps/general_2d.cu:246-249): the
normal of segment (v1, v2) is (t_y, -t_x) with t = normalize(v1 - v2), so polygons are
emitted clockwise to get outward normals.  3D normals are face normals
normalize((b - a) x (c - a)) of consistently wound, outward-facing triangles
(apps/general_3d.cu:251-252 uses igl::per_face_normals, the same formula).

Every workload is named by what it is:
  * ``circle_dirichlet_2d``   unit-circle 1024-gon, g = exp(x) cos(y), 256^2 grid in the disk
  * ``icosphere_neumann_3d``  level-5 icosphere (20,480 triangles), radially displaced,
                              Neumann data of u = x^2 - y^2 + 0.5 z, 100k interior points
  * ``star_exterior_batch_2d`` 64 star polygons, exterior Neumann / Robin / mixed data
  * ``geodesic_mixed_3d``     frequency-100 geodesic sphere (200,000 triangles), displaced,
                              Dirichlet for z > 0, Neumann elsewhere
  * ``potential_flow_vpl_2d`` Neumann square with four circular obstacles, 256^2 grid
                              (the virtual-point-light app's scene)
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

DIRICHLET, NEUMANN, ROBIN = 0, 1, 2


@dataclass
class Elements:
    """Host copy of a boundary: per-element arrays, FP32 (int32 for the type)."""

    dim: int
    vertices: np.ndarray        # (n, dim, dim)
    normals: np.ndarray         # (n, dim)
    boundary_values: np.ndarray  # (n, dim)
    boundary_types: np.ndarray  # (n,)
    robin_alphas: np.ndarray    # (n, dim)

    @property
    def n(self) -> int:
        return int(self.vertices.shape[0])

    def astype_contiguous(self) -> "Elements":
        return Elements(
            self.dim,
            np.ascontiguousarray(self.vertices, np.float32),
            np.ascontiguousarray(self.normals, np.float32),
            np.ascontiguousarray(self.boundary_values, np.float32),
            np.ascontiguousarray(self.boundary_types, np.int32),
            np.ascontiguousarray(self.robin_alphas, np.float32),
        )


@dataclass
class Workload:
    name: str
    elements: Elements
    points: np.ndarray                 # (m, dim) FP32 evaluation points
    exact: callable                    # FP64 exact solution: pts (m, dim) -> (m, dim + 1) [u, grad u]
    solver: dict = field(default_factory=dict)
    extra: dict = field(default_factory=dict)


# ----------------------------------------------------------------------------------------
# analytic harmonic functions (FP64)
# ----------------------------------------------------------------------------------------
def exact_exp_cos_2d(p):
    p = np.asarray(p, np.float64)
    ex = np.exp(p[:, 0])
    return np.stack([ex * np.cos(p[:, 1]), ex * np.cos(p[:, 1]), -ex * np.sin(p[:, 1])], 1)


def exact_quadratic_3d(p):
    p = np.asarray(p, np.float64)
    x, y, z = p[:, 0], p[:, 1], p[:, 2]
    return np.stack([x * x - y * y + 0.5 * z, 2 * x, -2 * y, np.full_like(x, 0.5)], 1)


def exact_expcos_xz_3d(p):
    p = np.asarray(p, np.float64)
    x, y, z = p[:, 0], p[:, 1], p[:, 2]
    e = np.exp(0.6 * x)
    return np.stack([e * np.cos(0.6 * y) + x * z, 0.6 * e * np.cos(0.6 * y) + z, -0.6 * e * np.sin(0.6 * y), x], 1)


def make_log_sources_2d(sources, charges):
    s = np.asarray(sources, np.float64)
    q = np.asarray(charges, np.float64)

    def exact(p):
        p = np.asarray(p, np.float64)
        d = p[:, None, :] - s[None, :, :]
        r2 = (d ** 2).sum(-1)
        u = (q[None, :] * 0.5 * np.log(r2)).sum(1)
        g = (q[None, :, None] * d / r2[:, :, None]).sum(1)
        return np.concatenate([u[:, None], g], 1)

    return exact


# ----------------------------------------------------------------------------------------
# 2D polygons
# ----------------------------------------------------------------------------------------
def polygon_elements(verts_ccw: np.ndarray, types, value_fn, robin_alpha=None) -> Elements:
    """Segments of a closed polygon given counter-clockwise, emitted clockwise.

    value_fn(points (k,2), normals (k,2), types (k,)) -> per-vertex boundary value; for
    Robin elements it must return u + du/dn (the reference's harness, general_2d.cu:265-268).
    """
    v = np.asarray(verts_ccw, np.float64)
    n = v.shape[0]
    v1 = np.roll(v, -1, axis=0)   # v_{i+1}
    v2 = v                        # v_i
    t = v1 - v2
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    nrm = np.stack([t[:, 1], -t[:, 0]], 1)
    types = np.broadcast_to(np.asarray(types, np.int32), (n,)).copy()
    verts = np.stack([v1, v2], 1)                      # (n, 2 vertices, 2)
    pts = verts.reshape(-1, 2)
    nr = np.repeat(nrm, 2, axis=0)
    ty = np.repeat(types, 2)
    bv = value_fn(pts, nr, ty).reshape(n, 2)
    if robin_alpha is None:
        al = np.where(types[:, None] == ROBIN, 1.0, 0.0) * np.ones((n, 2))
    else:
        al = np.repeat(np.asarray(robin_alpha, np.float64)[:, None], 2, 1)
    return Elements(2, verts, nrm, bv, types, al).astype_contiguous()


def inside_convex_polygon(points, elements: Elements):
    p = np.asarray(points, np.float64)
    a = elements.vertices[:, 0, :].astype(np.float64)
    nrm = elements.normals.astype(np.float64)
    side = np.einsum("pnd,nd->pn", p[:, None, :] - a[None, :, :], nrm)
    return np.all(side < 0, axis=1)


def inside_polygon_winding(points, verts_ccw):
    """Point-in-polygon by the crossing rule (FP64)."""
    p = np.asarray(points, np.float64)
    v = np.asarray(verts_ccw, np.float64)
    w = np.roll(v, -1, axis=0)
    inside = np.zeros(len(p), bool)
    for a, b in zip(v, w):
        cond = (a[1] > p[:, 1]) != (b[1] > p[:, 1])
        xint = a[0] + (p[:, 1] - a[1]) * (b[0] - a[0]) / np.where(b[1] != a[1], b[1] - a[1], 1.0)
        inside ^= cond & (p[:, 0] < xint)
    return inside


def circle_dirichlet_2d(n_segments: int = 1024, grid_res: int = 256, grid_stride: int = 1) -> Workload:
    """Unit circle polyline, Dirichlet data g = exp(x) cos(y) (also the exact solution).

    Evaluation: the cell-centred grid_res^2 grid over [-1, 1]^2 (x = (i + 1/2) / (grid_res/2) - 1)
    clipped to the polygon.  grid_stride > 1 keeps every grid_stride-th kept point.
    """
    th = 2 * np.pi * np.arange(n_segments) / n_segments
    v = np.stack([np.cos(th), np.sin(th)], 1)
    el = polygon_elements(v, DIRICHLET, lambda p, nr, ty: exact_exp_cos_2d(p)[:, 0])
    c = (np.arange(grid_res) + 0.5) / (grid_res / 2) - 1.0
    X, Y = np.meshgrid(c, c, indexing="xy")
    g = np.stack([X.ravel(), Y.ravel()], 1)
    keep = inside_convex_polygon(g, el)
    pts = g[keep][::grid_stride]
    solver = dict(bie_model="IndirectDoubleLayer", domain="InteriorDomain", estimator="BackwardEstimator",
                  problem="DirichletProblem", path_length=8, num_resampling_candidates=1,
                  eps=float(np.finfo(np.float32).eps), first_kind_equation_scaling_constant=1.0,
                  num_sample_paths=1024)
    return Workload("circle_dirichlet_2d", el, pts.astype(np.float32), exact_exp_cos_2d, solver)


# ----------------------------------------------------------------------------------------
# 3D spheres
# ----------------------------------------------------------------------------------------
def _icosahedron():
    t = (1 + 5 ** 0.5) / 2
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    # orient outward
    a, b, c = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
    nrm = np.cross(b - a, c - a)
    flip = (nrm * (a + b + c)).sum(1) < 0
    f[flip] = f[flip][:, [0, 2, 1]]
    return v, f


def icosphere(level: int):
    """Midpoint-subdivided icosahedron projected to the unit sphere (vertices, faces)."""
    v, f = _icosahedron()
    vlist = list(v)
    for _ in range(level):
        cache = {}

        def mid(i, j):
            key = (min(i, j), max(i, j))
            if key not in cache:
                m = vlist[i] + vlist[j]
                m = m / np.linalg.norm(m)
                cache[key] = len(vlist)
                vlist.append(m)
            return cache[key]

        nf = []
        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        f = np.array(nf, np.int64)
    return np.array(vlist), f


def geodesic_sphere(freq: int):
    """Frequency-`freq` geodesic sphere: each icosahedron face split into freq^2 triangles.

    Shared vertices are computed from the same integer barycentric weights summed in a
    canonical corner order, so adjacent faces get bit-identical vertices.
    """
    v, f = _icosahedron()
    key_to_idx = {}
    verts = []
    faces = []
    for a, b, c in f:
        corners = (a, b, c)
        local = {}
        for i in range(freq + 1):
            for j in range(freq + 1 - i):
                w = {a: freq - i - j, b: i, c: j}
                key = tuple(sorted((k, w[k]) for k in corners if w[k] > 0))
                idx = key_to_idx.get(key)
                if idx is None:
                    p = np.zeros(3)
                    for k, wk in key:
                        p = p + wk * v[k]
                    p = p / np.linalg.norm(p)
                    idx = len(verts)
                    verts.append(p)
                    key_to_idx[key] = idx
                local[(i, j)] = idx
        for i in range(freq):
            for j in range(freq - i):
                faces.append([local[(i, j)], local[(i + 1, j)], local[(i, j + 1)]])
                if i + j + 1 < freq:
                    faces.append([local[(i + 1, j)], local[(i + 1, j + 1)], local[(i, j + 1)]])
    return np.array(verts), np.array(faces, np.int64)


def displace_radial(v):
    """r = 1 + 0.2 sin(3 theta) cos(2 phi), theta = acos z, phi = atan2(y, x)."""
    th = np.arccos(np.clip(v[:, 2], -1, 1))
    ph = np.arctan2(v[:, 1], v[:, 0])
    r = 1.0 + 0.2 * np.sin(3 * th) * np.cos(2 * ph)
    return v * r[:, None]


def smooth_radius(dirs):
    d = dirs / np.linalg.norm(dirs, axis=1, keepdims=True)
    th = np.arccos(np.clip(d[:, 2], -1, 1))
    ph = np.arctan2(d[:, 1], d[:, 0])
    return 1.0 + 0.2 * np.sin(3 * th) * np.cos(2 * ph)


def mesh_elements(V, F, types, value_fn, robin_alpha=None) -> Elements:
    V = np.asarray(V, np.float64)
    a, b, c = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    nrm = np.cross(b - a, c - a)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    n = F.shape[0]
    types = np.broadcast_to(np.asarray(types, np.int32), (n,)).copy()
    verts = np.stack([a, b, c], 1)
    pts = verts.reshape(-1, 3)
    bv = value_fn(pts, np.repeat(nrm, 3, 0), np.repeat(types, 3)).reshape(n, 3)
    if robin_alpha is None:
        al = np.zeros((n, 3))
    else:
        al = np.repeat(np.asarray(robin_alpha, np.float64)[:, None], 3, 1)
    return Elements(3, verts, nrm, bv, types, al).astype_contiguous()


def interior_points_star(n_points: int, rng: np.random.Generator, margin: float = 0.01, box: float = 1.25):
    """Uniform points inside the displaced sphere, at least `margin` inside the smooth surface."""
    out = []
    need = n_points
    while need > 0:
        q = rng.uniform(-box, box, size=(max(4 * need, 1024), 3))
        r = np.linalg.norm(q, axis=1)
        keep = r < smooth_radius(q) - margin
        q = q[keep][:need]
        out.append(q)
        need -= len(q)
    return np.concatenate(out, 0)


def icosphere_neumann_3d(n_points: int = 100_000, level: int = 5, seed: int = 1001) -> Workload:
    V, F = icosphere(level)
    V = displace_radial(V)

    def neumann(p, nr, ty):
        g = exact_quadratic_3d(p)[:, 1:]
        return (g * nr).sum(1)

    el = mesh_elements(V, F, NEUMANN, neumann)
    rng = np.random.default_rng(seed)
    pts = interior_points_star(n_points, rng)
    solver = dict(bie_model="Direct", domain="InteriorDomain", estimator="BackwardEstimator",
                  problem="NeumannProblem", path_length=6, num_resampling_candidates=1,
                  eps=float(np.finfo(np.float32).eps), first_kind_equation_scaling_constant=1.0,
                  num_sample_paths=256)
    return Workload("icosphere_neumann_3d", el, pts.astype(np.float32), exact_quadratic_3d, solver,
                    extra=dict(vertices=V, faces=F))


def geodesic_mixed_3d(n_points: int = 1_000_000, freq: int = 100, seed: int = 3001) -> Workload:
    V, F = geodesic_sphere(freq)
    V = displace_radial(V)
    cz = V[F].mean(1)[:, 2]
    types = np.where(cz > 0, DIRICHLET, NEUMANN).astype(np.int32)

    def data(p, nr, ty):
        e = exact_expcos_xz_3d(p)
        return np.where(ty == DIRICHLET, e[:, 0], (e[:, 1:] * nr).sum(1))

    el = mesh_elements(V, F, types, data)
    rng = np.random.default_rng(seed)
    pts = interior_points_star(n_points, rng)
    solver = dict(bie_model="IndirectSingleLayer", domain="InteriorDomain", estimator="BackwardEstimator",
                  problem="MixedProblem", path_length=4, num_resampling_candidates=64, eps=1e-5,
                  first_kind_equation_scaling_constant=2.0, num_sample_paths=256)
    return Workload("geodesic_mixed_3d", el, pts.astype(np.float32), exact_expcos_xz_3d, solver,
                    extra=dict(vertices=V, faces=F))


# ----------------------------------------------------------------------------------------
# reference app grids; the virtual-point-light app's scene
# ----------------------------------------------------------------------------------------
def grid_points(grid_res: int, dim: int = 2, domain_size: float = 1.0) -> np.ndarray:
    """Cell-centred grid of the reference apps (utils.cuh:92-99), FP32 arithmetic as there:
    point idx has coordinates ((idx / grid_res^k) % grid_res + 0.5) * (size / grid_res) - size / 2."""
    n = grid_res ** dim
    idx = np.arange(n, dtype=np.int64)
    dx = np.float32(domain_size) / np.float32(grid_res)
    out = np.empty((n, dim), np.float32)
    for k in range(dim):
        u = (idx // grid_res ** k) % grid_res if k < dim - 1 else idx // grid_res ** k
        out[:, k] = (u.astype(np.float32) + np.float32(0.5)) * dx - np.float32(domain_size / 2)
    return out


def potential_flow_vpl_2d(grid_res: int = 256, circles: bool = True) -> Workload:
    """The scene of apps/potential_flow_with_vpl.cu:141-184: the square [-0.5, 0.5]^2 as four
    Neumann segments with flux -1 (left), +1 (right) and 0 (top, bottom), and four circular
    obstacles (360, 180, 180 and 90 segments) with zero flux, normals pointing into the
    obstacles.  Evaluation: the grid_res^2 grid over the square; `extra["mask"]` marks the
    points outside every obstacle (the reference writes NaN at the others).  Without the
    obstacles u = x + const, so grad u = (1, 0) is exact."""
    segs = [((-0.5, -0.5), (-0.5, 0.5), (-1.0, 0.0), -1.0), ((-0.5, 0.5), (0.5, 0.5), (0.0, 1.0), 0.0),
            ((0.5, 0.5), (0.5, -0.5), (1.0, 0.0), 1.0), ((0.5, -0.5), (-0.5, -0.5), (0.0, -1.0), 0.0)]
    V, N, B = [], [], []
    for a, b, nr, g in segs:
        V.append([a, b]); N.append(nr); B.append([g, g])
    obstacles = [((0.2, 0.1), 0.25, 360), ((-0.25, 0.3), 0.15, 180), ((-0.20, -0.25), 0.2, 180),
                 ((0.25, -0.3), 0.05, 90)] if circles else []
    for (cx, cy), r, m in obstacles:
        for i in range(m):
            t0, t1 = 2 * np.pi * i / m, 2 * np.pi * (i + 1) / m
            v1 = np.array([cx + r * np.cos(t0), cy + r * np.sin(t0)])
            v2 = np.array([cx + r * np.cos(t1), cy + r * np.sin(t1)])
            t = (v1 - v2) / np.linalg.norm(v1 - v2)
            V.append([v1, v2]); N.append([t[1], -t[0]]); B.append([0.0, 0.0])
    n = len(V)
    el = Elements(2, np.asarray(V, np.float64).astype(np.float32), np.asarray(N, np.float64).astype(np.float32),
                  np.asarray(B, np.float64).astype(np.float32), np.full(n, NEUMANN, np.int32),
                  np.zeros((n, 2), np.float32))
    pts = grid_points(grid_res, 2, 1.0)
    mask = np.ones(len(pts), bool)
    for (cx, cy), r, _ in obstacles:
        mask &= ((pts[:, 0] - np.float32(cx)) ** 2 + (pts[:, 1] - np.float32(cy)) ** 2) >= np.float32(r) ** 2

    def exact(p):
        p = np.asarray(p, np.float64)
        return np.stack([p[:, 0], np.ones(len(p)), np.zeros(len(p))], 1)
    solver = dict(bie_model="IndirectSingleLayer", domain="InteriorDomain", estimator="BackwardEstimator",
                  problem="NeumannProblem", path_length=4, num_resampling_candidates=1024, eps=1e-4,
                  first_kind_equation_scaling_constant=1.0, num_sample_paths=100000)
    return Workload("potential_flow_vpl_2d", el, pts, None if circles else exact, solver,
                    dict(mask=mask, grid_res=grid_res, obstacles=obstacles))


# ----------------------------------------------------------------------------------------
# Poisson: the reference app's square with a constant grid source
# ----------------------------------------------------------------------------------------
def square_poisson_2d(source: float = 5.0, per_side: int = 128, grid_res: int = 64, n_points: int = 1024,
                      seed: int = 5001) -> Workload:
    """Square (-1, 1)^2 with a constant source on a grid cache covering it, as in
    apps/poisson_equation.cu:105-135 (there: grid 256, source 5, data 1).  Exact solution
    u = source r^2 / 4 + exp(x) cos(y) (Laplacian(u) = source); each side is split into
    `per_side` segments so the linearly interpolated data follow u closely."""
    t = np.linspace(-1.0, 1.0, per_side + 1)[:-1]
    one = np.ones(per_side)
    v = np.concatenate([np.stack([t, -one], 1), np.stack([one, t], 1), np.stack([-t, one], 1), np.stack([-one, -t], 1)])

    def exact(p):
        p = np.asarray(p, np.float64)
        x, y = p[:, 0], p[:, 1]
        e = np.exp(x)
        return np.stack([source * (x * x + y * y) / 4 + e * np.cos(y), source * x / 2 + e * np.cos(y),
                         source * y / 2 - e * np.sin(y)], 1)

    el = polygon_elements(v, DIRICHLET, lambda p, nr, ty: exact(p)[:, 0])
    pts = np.random.default_rng(seed).uniform(-0.9, 0.9, (n_points, 2)).astype(np.float32)
    solver = dict(bie_model="IndirectDoubleLayer", domain="InteriorDomain", estimator="BackwardEstimator",
                  problem="DirichletProblem", path_length=4, eps=1e-4, num_volume_samples=16, num_sample_paths=1000)
    return Workload("square_poisson_2d", el, pts, exact, solver,
                    extra=dict(grid_res=grid_res, domain_size=2.0,
                               cache=np.full(grid_res * grid_res, source, np.float32)))


# ----------------------------------------------------------------------------------------
# Poisson: Gaussian-mixture source on the geodesic sphere
# ----------------------------------------------------------------------------------------
def gaussian_sources_3d(n: int = 16, seed: int = 4001):
    """16 isotropic Gaussians: centres uniform in the ball r < 0.7, widths U(0.05, 0.2),
    amplitudes U(-5, 5)."""
    rng = np.random.default_rng(seed)
    c = []
    while len(c) < n:
        q = rng.uniform(-0.7, 0.7, 3)
        if np.linalg.norm(q) < 0.7:
            c.append(q)
    return np.array(c), rng.uniform(0.05, 0.2, n), rng.uniform(-5.0, 5.0, n)


def newtonian_gaussians_3d(centers, widths, amps):
    """N with Laplacian(N) = sum_j A_j exp(-|x - c_j|^2 / (2 w_j^2)) in R^3 (the potential of a
    Gaussian charge): N_j(r) = -A_j w_j^3 sqrt(pi/2) erf(r / (sqrt(2) w_j)) / r.  Returns a
    function p -> [N, grad N] (FP64)."""
    from scipy.special import erf
    c = np.asarray(centers, np.float64)
    w = np.asarray(widths, np.float64)
    a = np.asarray(amps, np.float64)

    def fn(p):
        p = np.asarray(p, np.float64)
        d = p[:, None, :] - c[None, :, :]
        r = np.sqrt((d ** 2).sum(-1))
        s = np.sqrt(2.0) * w[None, :]
        K = a[None, :] * w[None, :] ** 3 * math.sqrt(math.pi / 2.0)
        small = r < 1e-6
        rs = np.where(small, 1.0, r)
        er = erf(rs / s)
        N = np.where(small, -K * 2.0 / (math.sqrt(math.pi) * s), -K * er / rs)
        dN = -K * (2.0 / math.sqrt(math.pi) * np.exp(-(rs / s) ** 2) / (s * rs) - er / rs ** 2)
        dN = np.where(small, 0.0, dN)
        g = (dN / rs)[:, :, None] * d
        return np.concatenate([N.sum(1)[:, None], g.sum(1)], 1)

    return fn


def poisson_exact_3d(centers, widths, amps):
    """u = N_f + H with H the harmonic exp(0.6x)cos(0.6y) + xz, so Laplacian(u) = f."""
    nf = newtonian_gaussians_3d(centers, widths, amps)
    return lambda p: nf(p) + exact_expcos_xz_3d(p)


def poisson_gaussian_3d(freq: int = 100, seed: int = 4001, point_sets=None, V=None, F=None) -> Workload:
    """Interior Dirichlet Poisson problem on the displaced geodesic sphere: f = 16 Gaussians,
    g = u restricted to the boundary.  Double layer, backward estimator, path length 4 with
    the volume term (data/config_poisson_equation.json); evaluation point sets of 16,384 x
    32^(j/7) points, j = 0..7, drawn inside the mesh."""
    if V is None:
        V, F = geodesic_sphere(freq)
        V = displace_radial(V)
    cen, wid, amp = gaussian_sources_3d(seed=seed)
    exact = poisson_exact_3d(cen, wid, amp)
    el = mesh_elements(V, F, DIRICHLET, lambda p, nr, ty: exact(p)[:, 0])
    rng = np.random.default_rng(seed + 1)
    sizes = point_sets if point_sets is not None else [int(round(16384 * 32 ** (j / 7))) for j in range(8)]
    sets = [interior_points_star(n, rng).astype(np.float32) for n in sizes]
    solver = dict(bie_model="IndirectDoubleLayer", domain="InteriorDomain", estimator="BackwardEstimator",
                  problem="DirichletProblem", path_length=4, eps=1e-4, num_volume_samples=1, num_sample_paths=256)
    return Workload("poisson_gaussian_3d", el, np.concatenate(sets), exact, solver,
                    extra=dict(vertices=V, faces=F, point_sets=sets, centers=cen, widths=wid, amplitudes=amp))


# ----------------------------------------------------------------------------------------
# batched exterior star polygons
# ----------------------------------------------------------------------------------------
def star_exterior_batch_2d(n_scenes: int = 64, seed: int = 2001, min_segments: int = 256,
                           max_segments: int = 16384, min_points: int = 1000, max_points: int = 64000,
                           point_scale: float = 1.0, kinds=None):
    """64 star polygons r = 1 + sum_k a_k cos(k th + phi_k); exterior data from log sources.

    Returns a list of Workloads (one per scene).  Each scene's problem type is drawn from
    {Neumann, Robin, Mixed}; mixed scenes draw a type per segment (general_2d.cu:242-253
    draws them from std::random_device; here they are seeded).  kinds (optional, one per
    scene: 0 Neumann, 1 Robin, 2 mixed) overrides the drawn kind.
    """
    rng = np.random.default_rng(seed)
    scenes = []
    for s in range(n_scenes):
        ak = rng.uniform(0, 0.08 / np.arange(1, 7))
        phk = rng.uniform(0, 2 * np.pi, 6)
        ns = int(round(math.exp(rng.uniform(math.log(min_segments), math.log(max_segments)))))
        th = 2 * np.pi * np.arange(ns) / ns
        r = 1 + (ak[None, :] * np.cos(np.arange(1, 7)[None, :] * th[:, None] + phk[None, :])).sum(1)
        v = np.stack([r * np.cos(th), r * np.sin(th)], 1)
        src = []
        while len(src) < 4:
            c = rng.uniform(-0.5, 0.5, 2)
            if np.linalg.norm(c) < 0.5:
                src.append(c)
        q = rng.uniform(-1, 1, 4)
        q -= q.mean()
        exact = make_log_sources_2d(np.array(src), q)
        kind = int(rng.integers(0, 3))  # 0 Neumann, 1 Robin, 2 Mixed
        if kinds is not None:
            kind = int(kinds[s])
        if kind == 0:
            types = np.full(ns, NEUMANN, np.int32)
        elif kind == 1:
            types = np.full(ns, ROBIN, np.int32)
        else:
            types = rng.integers(0, 3, ns).astype(np.int32)
        alpha = np.where(types == ROBIN, rng.uniform(0.5, 2.0, ns), 0.0)
        # per-vertex data: Dirichlet u, Neumann du/dn, Robin du/dn + alpha u (bie_models.cuh:295-298),
        # with n the element normal as emitted (out of the polygon, for interior and exterior
        # problems alike, as in general_2d.cu:248-249).
        tmp = polygon_elements(v, types, lambda p, nr, ty: np.zeros(len(p)))
        pts2 = tmp.vertices.reshape(-1, 2).astype(np.float64)
        nr2 = np.repeat(tmp.normals.astype(np.float64), 2, 0)
        ty2 = np.repeat(types, 2)
        al2 = np.repeat(alpha, 2)
        e = exact(pts2)
        dn = (e[:, 1:] * nr2).sum(1)
        bv = np.where(ty2 == DIRICHLET, e[:, 0], np.where(ty2 == NEUMANN, dn, dn + al2 * e[:, 0]))
        el = Elements(2, tmp.vertices, tmp.normals, bv.reshape(ns, 2), types,
                      np.repeat(alpha[:, None], 2, 1)).astype_contiguous()
        npnt = int(round(math.exp(rng.uniform(math.log(min_points), math.log(max_points))) * point_scale))
        rr = np.sqrt(rng.uniform(1.5 ** 2, 4.0 ** 2, npnt))
        tt = rng.uniform(0, 2 * np.pi, npnt)
        pts = np.stack([rr * np.cos(tt), rr * np.sin(tt)], 1)
        pl = {0: "NeumannProblem", 1: "RobinProblem", 2: "MixedProblem"}[kind]
        solver = dict(bie_model="IndirectSingleLayer", domain="ExteriorDomain", estimator="BackwardEstimator",
                      problem=pl, path_length={0: 7, 1: 4, 2: 6}[kind], num_resampling_candidates=64, eps=1e-5,
                      first_kind_equation_scaling_constant=2.0 if kind == 2 else 1.0, num_sample_paths=256)
        scenes.append(Workload(f"star_{s}", el, pts.astype(np.float32), exact, solver))
    return scenes
