"""CPU: the oracle's closest-point query and walk on spheres (apps/simple_wob_and_wos.cu)
pinned against closed forms.

* Distances to a square and to a cube are min over the faces; the boundary value at the foot
  point of linear data is that data's value there.
* Point-triangle distance: an independent FP64 computation (projection onto the plane when
  it falls inside, otherwise the nearest of the three edges) for points in all seven Voronoi
  regions of a triangle.
* Walk on spheres: harmonic Dirichlet data in a disk and on a sphere, estimate within 4 SE
  of the exact solution plus the eps / polygon-approximation bias.
"""
import numpy as np

from wobtoolbox_amd import geometry as G


def square_elements(n_per_side=8):
    """Clockwise square [-0.5, 0.5]^2 (outward normals), Dirichlet data g = x + 2y."""
    c = [(-0.5, -0.5), (-0.5, 0.5), (0.5, 0.5), (0.5, -0.5)]
    V = []
    for i in range(4):
        a, b = np.array(c[i]), np.array(c[(i + 1) % 4])
        for k in range(n_per_side):
            V.append(a + (b - a) * k / n_per_side)
    V = np.array(V)
    return G.polygon_elements(V[::-1], G.DIRICHLET, lambda p, nr, ty: p[:, 0] + 2 * p[:, 1])


def test_closest_points_square(oracle):
    el = square_elements()
    os_ = oracle.OracleScene(el)
    rng = np.random.default_rng(0)
    x = rng.uniform(-0.45, 0.45, (500, 2)).astype(np.float32)
    e, d, v = os_.closest_points(x)
    xd = x.astype(np.float64)
    exact = np.min(np.stack([0.5 - xd[:, 0], 0.5 + xd[:, 0], 0.5 - xd[:, 1], 0.5 + xd[:, 1]], 1), 1)
    np.testing.assert_allclose(d, exact, atol=2e-7)
    # the chosen element is a nearest one (FP32 near-ties go to the smaller id) and the value
    # is the data interpolated at the clamped foot point on it
    a, b = el.vertices[e, 0].astype(np.float64), el.vertices[e, 1].astype(np.float64)
    t = np.clip(((xd - a) * (b - a)).sum(1) / ((b - a) ** 2).sum(1), 0, 1)
    foot = a + t[:, None] * (b - a)
    np.testing.assert_allclose(np.linalg.norm(foot - xd, axis=1), exact, atol=2e-7)
    np.testing.assert_allclose(v, foot[:, 0] + 2 * foot[:, 1], atol=2e-6)
    # points outside the square: distance to the box
    y = rng.uniform(-1.5, 1.5, (300, 2)).astype(np.float32)
    y = y[np.max(np.abs(y), 1) > 0.55]
    _, d2, _ = os_.closest_points(y)
    ex = np.linalg.norm(np.maximum(np.abs(y.astype(np.float64)) - 0.5, 0), axis=1)
    np.testing.assert_allclose(d2, ex, atol=2e-7)


def _tri_dist_fp64(p, a, b, c):
    n = np.cross(b - a, c - a)
    n /= np.linalg.norm(n)
    q = p - np.dot(p - a, n) * n
    # inside test by barycentric signs of the projection
    def side(u, v, w):
        return np.dot(np.cross(v - u, q - u), n)
    if side(a, b, c) >= 0 and side(b, c, a) >= 0 and side(c, a, b) >= 0:
        return np.linalg.norm(p - q)
    def seg(u, v):
        t = np.clip(np.dot(p - u, v - u) / np.dot(v - u, v - u), 0, 1)
        return np.linalg.norm(u + t * (v - u) - p)
    return min(seg(a, b), seg(b, c), seg(c, a))


def test_closest_points_triangle_regions(oracle):
    a, b, c = np.array([0.1, -0.2, 0.05]), np.array([0.9, 0.1, -0.1]), np.array([0.3, 0.8, 0.2])
    V = np.array([[a, b, c]], np.float32)
    nr = np.cross(b - a, c - a)
    nr /= np.linalg.norm(nr)
    el = G.Elements(3, V, nr[None].astype(np.float32), np.array([[1.0, 2.0, 4.0]], np.float32),
                    np.zeros(1, np.int32), np.zeros((1, 3), np.float32))
    os_ = oracle.OracleScene(el)
    rng = np.random.default_rng(1)
    # points around the triangle: every vertex, edge and face region is visited
    cen = (a + b + c) / 3
    x = (cen + rng.normal(scale=0.8, size=(2000, 3))).astype(np.float32)
    _, d, v = os_.closest_points(x)
    exact = np.array([_tri_dist_fp64(p.astype(np.float64), a.astype(np.float32).astype(np.float64),
                                     b.astype(np.float32).astype(np.float64), c.astype(np.float32).astype(np.float64))
                      for p in x])
    np.testing.assert_allclose(d, exact, atol=3e-6)
    assert v.min() >= 1.0 - 1e-5 and v.max() <= 4.0 + 1e-5


def test_walk_on_spheres_disk(oracle):
    el = G.circle_dirichlet_2d(n_segments=512, grid_res=8).elements    # data exp(x) cos(y)
    os_ = oracle.OracleScene(el)
    rng = np.random.default_rng(2)
    r = 0.7 * np.sqrt(rng.uniform(size=24))
    th = rng.uniform(0, 2 * np.pi, 24)
    x = np.stack([r * np.cos(th), r * np.sin(th)], 1).astype(np.float32)
    n = 512
    s, q = os_.wos(x, n, eps=1e-4, seed=3)
    m = s / n
    se = np.sqrt(np.maximum(q / n - m * m, 0) / n)
    exact = np.exp(x[:, 0].astype(np.float64)) * np.cos(x[:, 1].astype(np.float64))
    z = np.abs(m - exact) / (se + 1e-12)
    assert np.all(np.abs(m - exact) < 4 * se + 2e-3), z.max()
    assert np.mean(z) < 1.3


def test_walk_on_spheres_sphere(oracle):
    V, F = G.icosphere(3)
    el = G.mesh_elements(V, F, G.DIRICHLET, lambda p, nr, ty: p[:, 0] ** 2 - p[:, 1] ** 2 + 0.5 * p[:, 2])
    os_ = oracle.OracleScene(el)
    rng = np.random.default_rng(4)
    d = rng.normal(size=(12, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    x = (d * rng.uniform(0, 0.6, (12, 1))).astype(np.float32)
    n = 128
    s, q = os_.wos(x, n, eps=1e-3, seed=5)
    m = s / n
    se = np.sqrt(np.maximum(q / n - m * m, 0) / n)
    xd = x.astype(np.float64)
    exact = xd[:, 0] ** 2 - xd[:, 1] ** 2 + 0.5 * xd[:, 2]
    # linear data on a level-3 icosphere: interpolation bias up to ~1e-2
    assert np.all(np.abs(m - exact) < 4 * se + 2e-2)


def test_simple_wob_equals_double_layer_estimator(oracle):
    """WoBFunc (simple_wob_and_wos.cu:81-113) and the general estimator's interior Dirichlet
    double-layer walk with the same path length consume the same draws and give the same path
    values up to FP32 rounding (the app runner's walk-on-boundary mode relies on this).  With
    path length 1 WoBFunc never reaches its halving step, so its paths are twice the
    estimator's (whose single term is the halved last one).  In 3D a ray from a boundary point
    can re-hit that point's own triangle at a grazing angle, where the estimator's weight uses
    the clamped cosine (scene.cuh:836-838) and WoBFunc's stays m sign(.): those few paths
    differ (the reference app is 2D, where this does not happen)."""
    for el, x in [(G.circle_dirichlet_2d(n_segments=256, grid_res=8).elements,
                   np.array([[0.1, 0.2], [-0.5, 0.3], [0.0, -0.7]], np.float32)),
                  (G.mesh_elements(*G.icosphere(2), G.DIRICHLET, lambda p, nr, ty: p[:, 0] + p[:, 1] * p[:, 2]),
                   np.array([[0.1, 0.2, 0.0], [-0.3, 0.3, 0.4]], np.float32))]:
        os_ = oracle.OracleScene(el)
        for L in (1, 2, 4):
            solver = dict(bie_model="IndirectDoubleLayer", domain="InteriorDomain", estimator="BackwardEstimator",
                          problem="DirichletProblem", path_length=L, num_resampling_candidates=1,
                          eps=float(np.finfo(np.float32).eps), first_kind_equation_scaling_constant=1.0)
            _, _, po = os_.estimate_domain(solver, x, 256, seed=9, return_paths=True)
            _, ps = os_.simple_wob(x, 256, L, seed=9)
            ref = po[:, :, 0] * (2.0 if L == 1 else 1.0)
            same = np.abs(ps - ref) <= 2e-5 * (np.abs(ref) + np.abs(ref).max())
            if el.dim == 2:
                assert same.all()
            else:
                assert same.mean() > 0.9
