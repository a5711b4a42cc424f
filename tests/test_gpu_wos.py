"""GPU: closest points and walk on spheres (apps/simple_wob_and_wos.cu) through the C ABI.

* wob_closest_points equals the oracle's brute-force scan bit for bit (element, distance,
  value) for points inside, outside and on the boundary of 2D and 3D scenes.
* wob_wos_estimate follows the oracle's walks on the same Philox streams (paths diverge only
  where the device's sincospi and the oracle's double-precision cos/sin round differently),
  and matches harmonic solutions within standard errors.
* Walk on spheres and the double-layer walk on boundary estimate the same solution.
"""
import sys

import numpy as np
import pytest

from helpers import small_icosphere, star_scene
from wobtoolbox_amd import geometry as G

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _dev(a, gpu):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a)).to(gpu)


def _probe_points(el, rng, n):
    D = el.dim
    V = el.vertices.reshape(-1, D)
    lo, hi = V.min(0), V.max(0)
    span = hi - lo
    x = rng.uniform(lo - 0.3 * span, hi + 0.3 * span, (n, D))
    # on vertices, on edges / faces, and within 1e-6 of them
    k = n // 8
    e = rng.integers(0, el.n, k)
    x[:k] = el.vertices[e, 0]
    w = rng.dirichlet(np.ones(D), k)
    x[k:2 * k] = np.einsum("kd,kdc->kc", w, el.vertices[e])
    x[2 * k:3 * k] = x[k:2 * k] + rng.normal(scale=1e-6, size=(k, D))
    return x.astype(np.float32)


@pytest.mark.parametrize("dim", [2, 3])
def test_closest_points_match_brute_force(dim, gpu, oracle):
    import wobtoolbox_amd as w
    rng = np.random.default_rng(20 + dim)
    el = star_scene(seed=3, n_segments=700, kind=2).elements if dim == 2 else small_icosphere(4, "mixed")[0]
    scene = w.SceneHost(el)
    x = _probe_points(el, rng, 20000)
    e, d, v = [t.cpu().numpy() for t in scene.closest_points(_dev(x, gpu))]
    eo, do, vo = oracle.OracleScene(el).closest_points(x)
    _say(f"dim {dim}: elements equal {np.mean(e == eo):.6f}, max |d - do| {np.abs(d - do).max():.3e}")
    assert np.array_equal(e, eo)
    assert np.array_equal(d, do)
    assert np.array_equal(v, vo)


@pytest.mark.parametrize("dim", [2, 3])
def test_wos_matches_oracle_paths(dim, gpu, oracle):
    import wobtoolbox_amd as w
    rng = np.random.default_rng(30 + dim)
    if dim == 2:
        el = G.circle_dirichlet_2d(n_segments=256, grid_res=8).elements
        r = 0.8 * np.sqrt(rng.uniform(size=200))
        th = rng.uniform(0, 2 * np.pi, 200)
        x = np.stack([r * np.cos(th), r * np.sin(th)], 1)
    else:
        el = small_icosphere(3, "mixed")[0]
        d = rng.normal(size=(200, 3))
        x = d / np.linalg.norm(d, axis=1, keepdims=True) * rng.uniform(0, 0.7, (200, 1))
    x = x.astype(np.float32)
    scene = w.SceneHost(el)
    paths = 32
    s, q = scene.walk_on_spheres(_dev(x, gpu), paths, eps=1e-4, seed=7, path_base=3, stream_id=2)
    so, qo = oracle.OracleScene(el).wos(x, paths, 1e-4, seed=7, path_base=3, stream=2)
    s = s.cpu().numpy()
    frac = np.mean(np.abs(s - so) > 1e-4 * (np.abs(so) + 1))
    _say(f"dim {dim}: points whose sums differ {frac:.4f}")
    assert frac < 0.05
    m, mo = s / paths, so / paths
    se = np.sqrt(np.maximum(qo / paths - mo * mo, 0) / paths)
    assert np.all(np.abs(m - mo) < 6 * se + 1e-5)


def test_wos_disk_harmonic_and_wob_agree(gpu):
    import wobtoolbox_amd as w
    wl = G.circle_dirichlet_2d(n_segments=1024, grid_res=32)
    scene = w.SceneHost(wl.elements)
    pts = wl.points[::4]
    xd = _dev(pts, gpu)
    n = 4096
    s, q = scene.walk_on_spheres(xd, n, eps=1e-4, seed=1)
    m, se = w.mean_and_se(s, q, n)
    m, se = m.cpu().numpy(), se.cpu().numpy()
    exact = wl.exact(pts)[:, 0]
    z = (m - exact) / se
    _say(f"WoS disk: mean |z| {np.abs(z).mean():.3f}, max {np.abs(z).max():.3f}, mean SE {se.mean():.2e}")
    assert np.abs(z).mean() < 1.2 and np.abs(z).max() < 5
    # the double-layer walk on boundary (the app's other mode) estimates the same solution
    est = w.Estimator.from_config(scene, dict(wl.solver, path_length=8))
    s2, q2 = est.compute_domain(xd, 1024, seed=2)
    m2, se2 = [t.cpu().numpy() for t in w.mean_and_se(s2, q2, 1024)]
    z2 = (m2[:, 0] - m) / np.sqrt(se2[:, 0] ** 2 + se ** 2)
    assert np.abs(z2).mean() < 1.2


def test_wos_errors(gpu):
    import wobtoolbox_amd as w
    scene = w.SceneHost(G.circle_dirichlet_2d(n_segments=64, grid_res=8).elements)
    with pytest.raises(w.WobError):
        scene.walk_on_spheres(_dev(np.zeros((4, 2), np.float32), gpu), 4, eps=0.0)
