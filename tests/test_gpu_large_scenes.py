"""GPU parity on the full-size scenes of the benchmark workloads.

The other parity tests use reduced scenes; here the deep trees, long CDFs and guide tables of
the large meshes are checked against the oracle at their real size:
* the 200,000-triangle geodesic sphere (the learned-control-variate and Poisson workloads),
  with its BVH built on the host and on the device: the traversal stack bound both builds
  report, hit sets of 3,000 lines against brute force, and the single-layer RIS-64 walk of
  the control-variate workload path by path against the oracle;
* the 64-scene star batch's largest polygon (14,974 segments): its single-layer RIS walk path
  by path against the oracle.
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
    kernel takes a log (device logf against the host's), so a RIS choice decided by the last
    bit of a weight may differ on a few paths."""
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
