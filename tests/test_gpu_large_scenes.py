"""GPU parity on the full-size scenes of the benchmark workloads.

The other parity tests use reduced scenes; here the deep trees, long CDFs and guide tables of
the large meshes are checked against the oracle at their real size:
* the 200,000-triangle geodesic sphere (the learned-control-variate and Poisson workloads),
  with its BVH built on the host and on the device: the traversal stack bound both builds
  report, hit sets of 3,000 lines against brute force, and the single-layer RIS-64 walk of
  the control-variate workload path by path against the oracle;
* the 64-scene star batch's largest polygon (14,974 segments): its single-layer RIS walk path
  by path against the oracle;
* the Poisson workload on the same mesh (double layer, Dirichlet, Gaussian-mixture source
  with the inside test) path by path against the oracle, plain and with the harmonic control
  variate, and the control variate on the RIS walk of the control-variate workload.
"""
import sys

import numpy as np
import pytest

from helpers import boundary_records, mean_se
from wobtoolbox_amd import geometry as G

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _dev(a, gpu):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a)).to(gpu)


@pytest.fixture(scope="module")
def geodesic():
    return G.geodesic_mixed_3d(n_points=4096, freq=100)


@pytest.mark.parametrize("device_bvh", [False, True])
def test_geodesic_200k_hit_sets_equal_brute_force(device_bvh, geodesic, gpu, oracle):
    import wobtoolbox_amd as w
    el = geodesic.elements
    scene = w.SceneHost(el, device_bvh=device_bvh)
    info = scene.info()
    _say(f"200k-triangle sphere, {'device' if device_bvh else 'host'} BVH: {info['n_nodes']} nodes, depth "
         f"{info['depth']}, traversal stack bound {info['stack_need']} of 48 entries")
    assert info["stack_need"] <= 48
    sc = oracle.OracleScene(el)
    rng = np.random.default_rng(91)
    n = 3000
    o = geodesic.points[:n].copy()
    o[: n // 3] = boundary_records(el, n // 3, rng)[:, :3]
    rec, inv = sc.sample_lines(geodesic.points[n: n + n // 3], seed=4)
    o[n // 3: 2 * (n // 3)] = np.where((inv > 0)[:, None], rec[:, :3], o[n // 3: 2 * (n // 3)])
    d = rng.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    cnt, ids, ts = [t.cpu().numpy() for t in scene.intersect_lines(_dev(o, gpu), _dev(d, gpu), max_hits=64)]
    rc, rids, rts = sc.intersect_many(o, d, use_bvh=False, cap=64)
    bad = 0
    for i in range(n):
        a = sorted(zip(ids[i, :cnt[i]].tolist(), ts[i, :cnt[i]].view(np.uint32).tolist()))
        b = sorted(zip(rids[i, :rc[i]].tolist(), rts[i, :rc[i]].view(np.uint32).tolist()))
        bad += int(cnt[i] != rc[i] or a != b)
    _say(f"  {bad} of {n} lines differ from brute force; mean hits {cnt.mean():.3f}")
    assert bad == 0


def test_geodesic_200k_ris_walk_matches_oracle(geodesic, gpu, oracle):
    """The control-variate workload's solver (single layer, interior, mixed, RIS 64, c = 2,
    4 bounces) on the full mesh, path for path on the same Philox streams."""
    import wobtoolbox_amd as w
    el = geodesic.elements
    pts = geodesic.points[:64]
    paths = 32
    outs = []
    for device_bvh in (False, True):
        est = w.Estimator.from_config(w.SceneHost(el, device_bvh=device_bvh), geodesic.solver)
        s, q = est.compute_domain(_dev(pts, gpu), paths, seed=17)
        outs.append((s.cpu().numpy(), q.cpu().numpy()))
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
    so, qo = oracle.OracleScene(el).estimate_domain(geodesic.solver, pts, paths, seed=17, pick_mode=1)
    m, se = mean_se(*outs[0], paths)
    mo, seo = mean_se(so, qo, paths)
    diff = np.abs(m - mo)
    frac = (diff > 1e-5 * (np.abs(mo) + se + 1e-3)).mean()
    _say(f"200k sphere RIS walk: points differing > 1e-5 rel {frac:.4f}; median rel diff "
         f"{np.median(diff / (np.abs(mo) + 1e-9)):.2e}")
    assert frac == 0.0


def test_largest_star_ris_walk_matches_oracle(gpu, oracle):
    """The batch workload's largest scene (single layer, exterior, RIS 64) path for path; its 2D
    kernel takes a logarithm (correctly rounded on the device, the host C library's logf in the
    oracle), so a RIS choice decided by the last bit of a weight may differ on a few paths."""
    import wobtoolbox_amd as w
    ws = G.star_exterior_batch_2d()
    big = max(ws, key=lambda x: x.elements.n)
    pts = big.points[:64]
    paths = 64
    est = w.Estimator.from_config(w.SceneHost(big.elements), big.solver)
    s, q = est.compute_domain(_dev(pts, gpu), paths, seed=23)
    so, qo = oracle.OracleScene(big.elements).estimate_domain(big.solver, pts, paths, seed=23, pick_mode=1)
    m, se = mean_se(s.cpu().numpy(), q.cpu().numpy(), paths)
    mo, seo = mean_se(so, qo, paths)
    diff = np.abs(m - mo)
    frac = (diff > 1e-5 * (np.abs(mo) + se + 1e-3)).mean()
    z = diff / np.sqrt(se ** 2 + seo ** 2 + 1e-30)
    _say(f"largest star ({big.elements.n} segments, {big.solver['problem']}): points differing > 1e-5 rel "
         f"{frac:.4f}; max |z| {z.max():.3f}")
    assert frac <= 0.05 and z.max() < 0.5


def _compare(m, se, mo, seo, rel=1e-5):
    diff = np.abs(m - mo)
    frac = (diff > rel * (np.abs(mo) + se + 1e-3)).mean()
    z = diff / np.sqrt(se ** 2 + seo ** 2 + 1e-30)
    return frac, z.max()


def test_poisson_200k_walk_matches_oracle(gpu, oracle):
    """The Poisson workload at its size (200,000 triangles, 16 Gaussians, double layer, path
    length 4, one volume sample per term with the inside test), path for path against the
    oracle, plain and with the harmonic control variate fitted on the device."""
    import wobtoolbox_amd as w
    wl = G.poisson_gaussian_3d(point_sets=[64])
    el = wl.elements
    src = (wl.extra["centers"], wl.extra["widths"], wl.extra["amplitudes"])
    scene = w.SceneHost(el)
    scene.set_volume_gaussians(*src)
    osc = oracle.OracleScene(el)
    osc.set_volume_gaussians(*src)
    rms = scene.fit_harmonic_cv(n_sources=64, n_train=16384, volume_samples=16, eps=wl.solver["eps"], seed=7)
    rec, c0 = scene.harmonic_cv()
    osc.set_harmonic_cv(rec, c0)
    pts = wl.points[:64]
    paths = 32
    for flags in (0, 16):
        solver = dict(wl.solver, flags=flags)
        est = w.Estimator.from_config(scene, solver)
        s, q = est.compute_domain(_dev(pts, gpu), paths, seed=19)
        so, qo = osc.estimate_domain(solver, pts, paths, seed=19, pick_mode=1, threads=8)
        m, se = mean_se(s.cpu().numpy(), q.cpu().numpy(), paths)
        mo, seo = mean_se(so, qo, paths)
        frac, zmax = _compare(m, se, mo, seo)
        _say(f"200k Poisson walk flags {flags}: points differing > 1e-5 rel {frac:.4f}; max |z| {zmax:.4f} "
             f"(fit residual {rms:.4f})")
        assert frac <= 0.02 and zmax < 0.05


def test_geodesic_200k_harmonic_cv_walk_matches_oracle(geodesic, gpu, oracle):
    """The control-variate workload's RIS walk with the harmonic control variate (128 sources)
    on the full mesh, path for path.  u_h's terms take the hardware reciprocal square root on
    the device and an IEEE division in the oracle (a few FP32 ulps apart), and the walk
    multiplies the residual data (about 1e-5 of the data) by weights of order 200, so the two
    sides agree to a small fraction of the standard error rather than to rounding."""
    import wobtoolbox_amd as w
    el = geodesic.elements
    scene = w.SceneHost(el)
    rms = scene.fit_harmonic_cv(n_sources=128, radius_scale=2.0, n_train=32768, seed=4)
    rec, c0 = scene.harmonic_cv()
    osc = oracle.OracleScene(el)
    osc.set_harmonic_cv(rec, c0)
    pts = geodesic.points[:64]
    paths = 16
    solver = dict(geodesic.solver, flags=16)
    est = w.Estimator.from_config(scene, solver)
    s, q = est.compute_domain(_dev(pts, gpu), paths, seed=29)
    so, qo = osc.estimate_domain(solver, pts, paths, seed=29, pick_mode=1, threads=8)
    m, se = mean_se(s.cpu().numpy(), q.cpu().numpy(), paths)
    mo, seo = mean_se(so, qo, paths)
    frac, zmax = _compare(m, se, mo, seo)
    ex = geodesic.exact(pts)[:, 0]
    med = np.median(np.abs(m - mo) / (np.abs(mo) + 1e-9))
    _say(f"200k RIS walk with the harmonic control variate: points differing > 1e-5 rel {frac:.4f}; max |z| "
         f"{zmax:.4f}; median rel diff {med:.2e}; fit residual {rms:.2e}; max |u - exact| "
         f"{np.abs(m[:, 0] - ex).max():.2e} at {paths} paths")
    assert zmax < 0.5 and med < 1e-5
    assert np.abs((m[:, 0] - ex) / se[:, 0]).max() < 4.5
