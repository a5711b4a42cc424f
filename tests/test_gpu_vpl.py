"""GPU: the virtual-point-light path (apps/potential_flow_with_vpl.cu) through the C ABI.

* wob_vpl_aggregate against the oracle's FP64 sums of the reference's FP32 terms, for every
  model, both domains, both dimensions, with evaluation points on top of and next to VPLs
  (the kernels' clamps).  Tolerance: 4e-6 (about 32 FP32 ulp) of the sum of the magnitudes
  of each term's pieces before they cancel (the device uses hardware log2 / rcp / rsqrt, a
  different but equally rounded form of r/|r|, and grouped compensated sums).
* the closed-form layer potentials of a circle (no oracle involved).
* wob_vpl_paths: start points bit-identical to the oracle's uniform samples on the same
  Philox streams; contributions equal to the oracle's walks up to FP32 path divergence.
* the whole method on the reference app's scene: its mean over iterations agrees with the
  plain backward estimator of the same problem (same expectation) and, without obstacles,
  with grad u = (1, 0).
"""
import sys

import numpy as np
import pytest

from wobtoolbox_amd import geometry as G

pytestmark = pytest.mark.gpu

SL, DL, DIRECT = 0, 1, 2
MODELS = {SL: "IndirectSingleLayer", DL: "IndirectDoubleLayer", DIRECT: "Direct"}
DOMAINS = {0: "InteriorDomain", 1: "ExteriorDomain"}


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _dev(a, gpu):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a)).to(gpu)


def _estimator(w, scene, bie, domain, eps):
    return w.Estimator(scene, bie_model=MODELS[bie], domain=DOMAINS[domain], estimator="BackwardEstimator",
                       problem="DirichletProblem", eps=eps)


def _random_vpls(dim, n, rng, types=None):
    W = 2 * dim + 4
    rec = np.zeros((n, W), np.float32)
    rec[:, :dim] = rng.uniform(-1, 1, (n, dim))
    nr = rng.normal(size=(n, dim))
    rec[:, dim:2 * dim] = nr / np.linalg.norm(nr, axis=1, keepdims=True)
    rec[:, 2 * dim] = 0.5
    rec[:, 2 * dim + 2] = rng.integers(0, 3, n) if types is None else types
    rec[:, 2 * dim + 3] = rng.uniform(0.5, 2.0, n)
    c = rng.normal(size=n).astype(np.float32)
    return rec, c


def _term_scale(dim, bie, eps, inv, rec, c, pts):
    """FP64 error scale of each aggregated sum: the sum over VPLs of |c| * inv times the
    magnitudes of the pieces of every term before they cancel (G and dG_dny of a Robin record;
    n and D (rhat.n) rhat inside d2G_dxdny; the products inside rhat.n), plus one unit of |G|'s
    absolute rounding (the log of a distance near 1 has an absolute, not a relative, error)."""
    D = dim
    y, ny = rec[:, :D].astype(np.float64), rec[:, D:2 * D].astype(np.float64)
    typ, alpha = rec[:, 2 * D + 2], rec[:, 2 * D + 3].astype(np.float64)
    if bie == SL:
        A, B = np.ones(len(rec)), np.zeros(len(rec))
    elif bie == DL:
        A, B = np.zeros(len(rec)), np.ones(len(rec))
    else:
        A = np.where(typ == 0, 1.0, np.where(typ == 2, alpha, 0.0))
        B = np.where(typ == 0, 0.0, 1.0)
    wc = np.abs(c.astype(np.float64)) * inv
    out = np.zeros((len(pts), D + 1))
    for i0 in range(0, len(pts), 128):
        x = pts[i0:i0 + 128].astype(np.float64)
        r = y[None] - x[:, None]
        z = (r * r).sum(-1)
        nr = np.sqrt(z)
        rh = np.where(z[..., None] > 0, r / np.where(nr > 0, nr, 1.0)[..., None], r)
        rn = (np.abs(rh) * np.abs(ny)[None]).sum(-1)          # dot-product error scale of rhat.n
        if D == 2:
            g = (np.abs(np.log(np.maximum(nr, eps))) + 1.0) / (2 * np.pi)
            dgdx = np.abs(rh) / (2 * np.pi * np.maximum(nr, eps))[..., None]
            dny = rn / (2 * np.pi * np.maximum(nr, eps))
            s1 = 1.0 / (2 * np.pi * np.maximum(z, eps))
        else:
            g = 1.0 / (4 * np.pi * np.maximum(nr, eps))
            dgdx = np.abs(rh) / (4 * np.pi * np.maximum(z, eps))[..., None]
            dny = rn / (4 * np.pi * np.maximum(z, eps))
            s1 = 1.0 / (4 * np.pi * np.maximum(nr * z, eps))
        d2 = s1[..., None] * (np.abs(ny)[None] + D * rn[..., None] * np.abs(rh))
        out[i0:i0 + 128, 0] = ((A * g + B * dny) * wc).sum(-1)
        out[i0:i0 + 128, 1:] = ((A[:, None] * dgdx + B[:, None] * d2) * wc[:, None]).sum(1)
    return out


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("bie", [SL, DL, DIRECT])
@pytest.mark.parametrize("domain", [0, 1])
def test_vpl_aggregate_matches_oracle(dim, bie, domain, gpu, oracle):
    import wobtoolbox_amd as w
    rng = np.random.default_rng(100 * dim + 10 * bie + domain)
    el = G.potential_flow_vpl_2d(grid_res=8).elements if dim == 2 else G.mesh_elements(
        *G.icosphere(1), G.NEUMANN, lambda p, nr, ty: np.zeros(len(p)))
    scene = w.SceneHost(el)
    area = float(scene.info()["total_area"])
    n_vpl, n_pts = 3001, 1500
    rec, c = _random_vpls(dim, n_vpl, rng)
    pts = rng.uniform(-1, 1, (n_pts, dim)).astype(np.float32)
    pts[:20] = rec[:20, :dim]                                        # on top of a VPL (r = 0)
    pts[20:40] = rec[20:40, :dim] + np.float32(3e-5)                 # inside the clamp radius
    pts[40:60] = rec[40:60, :dim] + np.float32(0.02)                 # 2D: sqrt(eps) scale of d2G
    for eps in (1e-4, float(np.finfo(np.float32).eps)):
        est = _estimator(w, scene, bie, domain, eps)
        _say(f"aggregate dim={dim} bie={bie} domain={domain} eps={eps}")
        s = est.aggregate_vpl(_dev(rec, gpu), _dev(c, gpu), _dev(pts, gpu)).cpu().numpy()
        so, a = oracle.vpl_aggregate(dim, bie, domain, eps, area, rec, c, pts)
        scale = _term_scale(dim, bie, eps, area, rec, c, pts)
        assert np.all(scale >= a * (1 - 1e-5))
        err = np.abs(s - so) / scale
        _say(f"  max err / scale = {err.max():.3e}, max err / sum|term| = {(np.abs(s - so) / (a + 1e-30)).max():.3e}")
        assert np.all(np.isfinite(s))
        assert err.max() < 4e-6


def test_vpl_aggregate_accumulates_and_splits(gpu, oracle):
    """Point sets of every size class (one workgroup with one VPL split up to many splits) and
    accumulation into an existing buffer."""
    import wobtoolbox_amd as w
    rng = np.random.default_rng(5)
    el = G.potential_flow_vpl_2d(grid_res=8).elements
    scene = w.SceneHost(el)
    area = float(scene.info()["total_area"])
    est = _estimator(w, scene, SL, 0, 1e-4)
    for n_vpl, n_pts in [(1, 1), (7, 3), (300, 513), (20000, 70)]:
        rec, c = _random_vpls(2, n_vpl, rng)
        pts = rng.uniform(-1, 1, (n_pts, 2)).astype(np.float32)
        s0 = _dev(np.full((n_pts, 3), 2.5), gpu)
        s = est.aggregate_vpl(_dev(rec, gpu), _dev(c, gpu), _dev(pts, gpu), out=s0).cpu().numpy()
        so, a = oracle.vpl_aggregate(2, SL, 0, 1e-4, area, rec, c, pts)
        assert np.abs(s - 2.5 - so).max() <= 4e-6 * a.max() + 1e-12, (n_vpl, n_pts)


def test_vpl_aggregate_circle_closed_form(gpu):
    import wobtoolbox_amd as w
    R, n = 0.8, 8192
    el = G.circle_dirichlet_2d(n_segments=256, grid_res=8).elements
    scene = w.SceneHost(el)
    th = 2 * np.pi * (np.arange(n) + 0.5) / n
    rec = np.zeros((n, 8), np.float32)
    rec[:, 0], rec[:, 1], rec[:, 2], rec[:, 3] = R * np.cos(th), R * np.sin(th), np.cos(th), np.sin(th)
    area = float(scene.info()["total_area"])
    c = np.full(n, (2 * np.pi * R) / n / area, np.float32)        # weights so that c * |boundary| = ds
    x = np.array([[0.0, 0.0], [0.3, -0.2], [2.0, 0.5]], np.float32)
    sl = _estimator(w, scene, SL, 0, 1e-4).aggregate_vpl(_dev(rec, gpu), _dev(c, gpu), _dev(x, gpu)).cpu().numpy()
    dl = _estimator(w, scene, DL, 0, 1e-4).aggregate_vpl(_dev(rec, gpu), _dev(c, gpu), _dev(x, gpu)).cpu().numpy()
    np.testing.assert_allclose(sl[:2, 0], -R * np.log(R), atol=3e-5)
    np.testing.assert_allclose(sl[2, 0], -R * np.log(np.hypot(2.0, 0.5)), atol=3e-5)
    np.testing.assert_allclose(dl[:2, 0], 1.0, atol=3e-5)
    np.testing.assert_allclose(dl[2, 0], 0.0, atol=3e-5)


def test_vpl_paths_match_oracle(gpu, oracle):
    import wobtoolbox_amd as w
    wl = G.potential_flow_vpl_2d(grid_res=8)
    solver = dict(wl.solver, num_resampling_candidates=64)
    scene = w.SceneHost(wl.elements)
    est = w.Estimator.from_config(scene, solver)
    n = 2000
    for path_base in (0, 5):
        rec, c = est.compute_vpl_paths(n, seed=11, path_base=path_base, stream_id=3)
        rec, c = rec.cpu().numpy(), c.cpu().numpy()
        reco, co = oracle.OracleScene(wl.elements).vpl_paths(solver, n, seed=11, path_base=path_base, stream=3)
        assert np.array_equal(rec, reco)
        frac = (np.abs(c - co) > 1e-4 * (np.abs(co) + 1e-3)).mean()
        _say(f"VPL paths path_base={path_base}: contributions differing {frac:.4f}")
        assert frac < 0.05
    with pytest.raises(w.WobError):
        est.compute_vpl_paths(4, stream_id=1 << 30)


def test_vpl_method_matches_plain_estimator(gpu):
    """Mean over iterations of the VPL estimate = the plain backward estimate (same
    expectation) on the reference app's scene; without obstacles grad u = (1, 0)."""
    import torch
    import wobtoolbox_amd as w
    for circles in (True, False):
        wl = G.potential_flow_vpl_2d(grid_res=32, circles=circles)
        solver = dict(wl.solver, num_resampling_candidates=64)
        scene = w.SceneHost(wl.elements)
        est = w.Estimator.from_config(scene, solver)
        pts_np = wl.points[wl.extra["mask"]]
        # keep away from the walls, where the single-layer density's kernel is near-singular
        far = np.all(np.abs(pts_np) < 0.4, axis=1)
        pts = _dev(pts_np[far], gpu)
        iters, n_vpl = 24, 20000
        per = []
        for it in range(iters):
            rec, c = est.compute_vpl_paths(n_vpl, seed=1, path_base=it)
            per.append((est.aggregate_vpl(rec, c, pts) / n_vpl).cpu().numpy())
        per = np.stack(per)
        m_v, se_v = per.mean(0), per.std(0, ddof=1) / np.sqrt(iters)
        s, q = est.compute_domain(pts, 2048, seed=2)
        m_p, se_p = w.mean_and_se(s, q, 2048)
        m_p, se_p = m_p.cpu().numpy(), se_p.cpu().numpy()
        # gradient components (u carries the Neumann problem's free constant in both)
        z = (m_v[:, 1:] - m_p[:, 1:]) / np.sqrt(se_v[:, 1:] ** 2 + se_p[:, 1:] ** 2 + 1e-30)
        _say(f"circles={circles}: |z| mean {np.abs(z).mean():.3f} max {np.abs(z).max():.3f}, "
             f"mean grad {m_v[:, 1:].mean(0)}")
        assert np.abs(z).mean() < 1.2 and np.mean(np.abs(z) > 4) < 0.01
        if not circles:
            np.testing.assert_allclose(m_v[:, 1:].mean(0), [1.0, 0.0], atol=0.05)
