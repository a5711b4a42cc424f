"""CPU: the oracle's virtual-point-light aggregation (apps/potential_flow_with_vpl.cu:44-71)
pinned against closed forms, and its path-phase helpers against the scene.

Quadrature of layer potentials over a circle / sphere with equal weights reproduces the
classical values, which fix the sign and scale of every term the aggregation adds:
  * single layer of unit density, circle of radius R:  -R log R inside, -R log|x| outside
    (gradient 0 inside, -R x / |x|^2 outside);  sphere: R inside, R^2 / |x| outside;
  * double layer of unit density (Gauss integral): 1 inside, 0 outside, gradient 0;
  * the direct model's Dirichlet / Neumann / Robin records are the single layer times
    domain_type_coefficient, the double layer times it, and -(double layer + alpha single
    layer) times it (bie_models.cuh:428-462).
"""
import numpy as np
import pytest

from wobtoolbox_amd import geometry as G

SL, DL, DIRECT = 0, 1, 2
INTERIOR, EXTERIOR = 0, 1


def circle_records(R, n, center=(0.0, 0.0), types=1, alpha=0.0):
    th = 2 * np.pi * (np.arange(n) + 0.5) / n
    c = np.array(center)
    rec = np.zeros((n, 8), np.float32)
    rec[:, 0] = c[0] + R * np.cos(th)
    rec[:, 1] = c[1] + R * np.sin(th)
    rec[:, 2] = np.cos(th)
    rec[:, 3] = np.sin(th)
    rec[:, 4] = 0.5
    rec[:, 6] = types
    rec[:, 7] = alpha
    return rec


def sphere_records(R, n, types=1, alpha=0.0):
    # Fibonacci lattice: equal-area points
    i = np.arange(n) + 0.5
    z = 1 - 2 * i / n
    r = np.sqrt(1 - z * z)
    ph = np.pi * (1 + 5 ** 0.5) * i
    d = np.stack([r * np.cos(ph), r * np.sin(ph), z], 1)
    rec = np.zeros((n, 10), np.float32)
    rec[:, :3] = R * d
    rec[:, 3:6] = d
    rec[:, 6] = 0.5
    rec[:, 8] = types
    rec[:, 9] = alpha
    return rec


def test_vpl_aggregate_circle_layer_potentials(oracle):
    R, n = 0.8, 4096
    rec = circle_records(R, n)
    w = np.full(n, 1.0 / n, np.float32)
    inv = 2 * np.pi * R
    rng = np.random.default_rng(0)
    ang = rng.uniform(0, 2 * np.pi, 64)
    rin = rng.uniform(0, 0.6 * R, 64)
    rout = rng.uniform(1.5 * R, 3 * R, 64)
    xin = np.stack([rin * np.cos(ang), rin * np.sin(ang)], 1).astype(np.float32)
    xout = np.stack([rout * np.cos(ang), rout * np.sin(ang)], 1).astype(np.float32)
    s, _ = oracle.vpl_aggregate(2, SL, INTERIOR, 1e-4, inv, rec, w, xin)
    np.testing.assert_allclose(s[:, 0], -R * np.log(R), atol=2e-5)
    np.testing.assert_allclose(s[:, 1:], 0.0, atol=2e-5)
    s, _ = oracle.vpl_aggregate(2, SL, INTERIOR, 1e-4, inv, rec, w, xout)
    xo = xout.astype(np.float64)
    r2 = (xo ** 2).sum(1)
    np.testing.assert_allclose(s[:, 0], -R * 0.5 * np.log(r2), atol=2e-5)
    np.testing.assert_allclose(s[:, 1:], -R * xo / r2[:, None], atol=2e-5)
    s, _ = oracle.vpl_aggregate(2, DL, INTERIOR, 1e-4, inv, rec, w, xin)
    np.testing.assert_allclose(s[:, 0], 1.0, atol=2e-5)
    np.testing.assert_allclose(s[:, 1:], 0.0, atol=2e-4)
    s, _ = oracle.vpl_aggregate(2, DL, INTERIOR, 1e-4, inv, rec, w, xout)
    np.testing.assert_allclose(s, 0.0, atol=2e-5)


@pytest.mark.parametrize("domain", [INTERIOR, EXTERIOR])
def test_vpl_aggregate_direct_types(domain, oracle):
    R, n, alpha = 0.8, 4096, 1.7
    w = np.full(n, 1.0 / n, np.float32)
    inv = 2 * np.pi * R
    x = np.array([[0.1, -0.2], [0.3, 0.25], [-0.4, 0.0]], np.float32)
    sgn = 1.0 if domain == INTERIOR else -1.0
    sl, _ = oracle.vpl_aggregate(2, SL, domain, 1e-4, inv, circle_records(R, n), w, x)
    dl, _ = oracle.vpl_aggregate(2, DL, domain, 1e-4, inv, circle_records(R, n), w, x)
    d0, _ = oracle.vpl_aggregate(2, DIRECT, domain, 1e-4, inv, circle_records(R, n, types=0), w, x)
    d1, _ = oracle.vpl_aggregate(2, DIRECT, domain, 1e-4, inv, circle_records(R, n, types=1), w, x)
    d2, _ = oracle.vpl_aggregate(2, DIRECT, domain, 1e-4, inv, circle_records(R, n, types=2, alpha=alpha), w, x)
    np.testing.assert_allclose(d0, sgn * sl, atol=1e-6)
    np.testing.assert_allclose(d1, sgn * dl, atol=1e-6)
    np.testing.assert_allclose(d2, sgn * (dl - alpha * sl), atol=1e-5)
    np.testing.assert_allclose(d2[:, 0], sgn * (1.0 + alpha * R * np.log(R)), atol=1e-4)


def test_vpl_aggregate_sphere_layer_potentials(oracle):
    R, n = 0.9, 40000
    rec = sphere_records(R, n)
    w = np.full(n, 1.0 / n, np.float32)
    inv = 4 * np.pi * R * R
    rng = np.random.default_rng(1)
    d = rng.normal(size=(32, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    xin = (d * rng.uniform(0, 0.5 * R, (32, 1))).astype(np.float32)
    xout = (d * rng.uniform(1.6 * R, 3 * R, (32, 1))).astype(np.float32)
    s, _ = oracle.vpl_aggregate(3, SL, INTERIOR, 1e-4, inv, rec, w, xin)
    np.testing.assert_allclose(s[:, 0], R, rtol=2e-4)
    np.testing.assert_allclose(s[:, 1:], 0.0, atol=2e-4)
    s, _ = oracle.vpl_aggregate(3, SL, INTERIOR, 1e-4, inv, rec, w, xout)
    xo = xout.astype(np.float64)
    r = np.linalg.norm(xo, axis=1)
    np.testing.assert_allclose(s[:, 0], R * R / r, rtol=2e-4)
    np.testing.assert_allclose(s[:, 1:], -R * R * xo / r[:, None] ** 3, atol=2e-4)
    s, _ = oracle.vpl_aggregate(3, DL, INTERIOR, 1e-4, inv, rec, w, xin)
    np.testing.assert_allclose(s[:, 0], 1.0, atol=2e-4)
    s, _ = oracle.vpl_aggregate(3, DL, INTERIOR, 1e-4, inv, rec, w, xout)
    np.testing.assert_allclose(s[:, 0], 0.0, atol=2e-4)


def test_vpl_aggregate_abs_sums_and_coincident_points(oracle):
    """A point on top of a VPL takes the reference's clamped kernel values (r/|r| := r at r = 0)."""
    rec = circle_records(1.0, 16)
    w = np.linspace(-1, 1, 16).astype(np.float32)
    x = rec[:3, :2].copy()
    s, a = oracle.vpl_aggregate(2, SL, INTERIOR, 1e-3, 1.0, rec, w, x)
    assert np.all(np.isfinite(s)) and np.all(a >= np.abs(s) - 1e-12)
    ref = oracle.eval_weights(2, SL, INTERIOR, 1e-3, 1.0, np.repeat(rec[:1], 16, 0), rec, np.ones(16, np.float32),
                              np.repeat(x[:1], 16, 0), np.full(16, 0.5, np.float32))
    np.testing.assert_allclose(s[0, 0], (w.astype(np.float64) * ref[:, 3]).sum(), rtol=1e-6)


def test_vpl_start_points_lie_on_the_boundary(oracle):
    w = G.potential_flow_vpl_2d(grid_res=16)
    os_ = oracle.OracleScene(w.elements)
    seen = set()
    for i in range(200):
        rec, inv = os_.sample_uniform(7, 3, i, 1 << 30)
        assert abs(inv - os_.totals()[0]) < 1e-3
        p = rec[:2].astype(np.float64)
        V = w.elements.vertices.astype(np.float64)
        a, b = V[:, 0], V[:, 1]
        t = np.clip(((p - a) * (b - a)).sum(1) / ((b - a) ** 2).sum(1), 0, 1)
        dist = np.linalg.norm(a + t[:, None] * (b - a) - p, axis=1)
        e = int(np.argmin(dist))
        assert dist[e] < 1e-6
        np.testing.assert_allclose(rec[2:4], w.elements.normals[e], atol=1e-6)
        assert rec[6] == 1 and rec[5] == w.elements.boundary_values[e, 0]
        seen.add(tuple(rec[:2]))
    assert len(seen) == 200
