"""GPU parity of the learned harmonic control variate (double-layer Poisson walks and the
single-layer mixed walks of the learned-control-variate workload) through the C ABI.

  * u_h and grad u_h on the device against the oracle and the FP64 closed form;
  * the fit's training targets against the oracle on the same Philox streams (positions bit
    for bit; targets to FP32 rounding of the volume term);
  * the fit recovers a field it represents exactly, and agrees with an FP64 least-squares
    solve of its own targets;
  * the walk with the control variate against the oracle path by path, in single calls,
    batches and the wavefront schedule, with Dirichlet, Neumann and Robin data;
  * unbiasedness against the closed-form solution u = N_f + H and the variance it removes.
"""
import sys

import numpy as np
import pytest

from helpers import hcv_eval_f64, hcv_fit_numpy, hcv_sources, mean_se, small_icosphere
from wobtoolbox_amd import geometry as G

pytestmark = pytest.mark.gpu

DL = dict(bie_model="IndirectDoubleLayer", domain="InteriorDomain", estimator="BackwardEstimator",
          problem="DirichletProblem", eps=1e-4)
SL_MIXED = dict(bie_model="IndirectSingleLayer", domain="InteriorDomain", estimator="BackwardEstimator",
                problem="MixedProblem", path_length=4, eps=1e-5, num_resampling_candidates=64,
                first_kind_equation_scaling_constant=2.0)
HCV = 16


def _dev(a, gpu):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a)).to(gpu)


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _poisson(level=3):
    V, F = G.icosphere(level)
    V = G.displace_radial(V)
    c, wd, a = G.gaussian_sources_3d()
    ex = G.poisson_exact_3d(c, wd, a)
    el = G.mesh_elements(V, F, G.DIRICHLET, lambda p, nr, ty: ex(p)[:, 0])
    return el, (c, wd, a), ex


def _scenes(el, src, oracle):
    import wobtoolbox_amd as wob
    scene = wob.SceneHost(el)
    osc = oracle.OracleScene(el)
    if src is not None:
        scene.set_volume_gaussians(*src)
        osc.set_volume_gaussians(*src)
    return scene, osc


def test_hcv_eval_matches_oracle_and_closed_form(gpu, oracle):
    rng = np.random.default_rng(0)
    for dim in (2, 3):
        zero = lambda p, nr, ty: np.zeros(len(p))  # noqa: E731
        th = 2 * np.pi * np.arange(64) / 64
        el = (G.polygon_elements(np.stack([np.cos(th), np.sin(th)], 1), G.DIRICHLET, zero) if dim == 2 else
              G.mesh_elements(*G.icosphere(1), G.DIRICHLET, zero))
        scene, osc = _scenes(el, None, oracle)
        src = rng.normal(size=(48, dim))
        src = 2.5 * src / np.linalg.norm(src, axis=1, keepdims=True)
        w = rng.uniform(-1, 1, 48)
        rec = np.concatenate([src, w[:, None]], 1).astype(np.float32)
        scene.set_harmonic_cv(rec, 0.25)
        osc.set_harmonic_cv(rec, 0.25)
        got_src, got_c0 = scene.harmonic_cv()
        assert np.array_equal(got_src, rec) and got_c0 == np.float32(0.25)
        pts = rng.uniform(-0.8, 0.8, (2000, dim)).astype(np.float32)
        d = scene.eval_harmonic_cv(_dev(pts, gpu)).cpu().numpy().astype(np.float64)
        o = osc.eval_harmonic_cv(pts).astype(np.float64)
        ref = hcv_eval_f64(pts, rec[:, :dim].astype(np.float64), rec[:, dim].astype(np.float64), np.float32(0.25))
        z = ((pts[:, None, :].astype(np.float64) - rec[None, :, :dim]) ** 2).sum(2)
        mag = 0.25 + (np.abs(w)[None] * (1 / np.sqrt(z) if dim == 3 else np.abs(np.log2(z)))).sum(1)
        gmag = (np.abs(w)[None] / (z if dim == 3 else np.sqrt(z))).sum(1) * (1 if dim == 3 else 2 / np.log(2))
        err_d = np.abs(d[:, 0] - ref[:, 0]) / mag
        err_o = np.abs(d[:, 0] - o[:, 0]) / mag
        gerr = np.abs(d[:, 1:] - ref[:, 1:]) / gmag[:, None]
        _say(f"u_h {dim}D: device vs FP64 {err_d.max():.2e}, vs oracle {err_o.max():.2e} (of the term scale); "
             f"gradient {gerr.max():.2e}")
        assert err_d.max() < 2e-6 and err_o.max() < 2e-6 and gerr.max() < 4e-6


def test_hcv_targets_match_oracle(gpu, oracle):
    el, src, _ = _poisson(3)
    scene, osc = _scenes(el, src, oracle)
    rec = scene.harmonic_cv_targets(4096, volume_samples=4, eps=1e-4, seed=7, stream_id=3).cpu().numpy()
    orec = osc.hcv_targets(4096, seed=7, stream=3, volume_samples=4, eps=1e-4)
    assert np.array_equal(np.delete(rec, 7, 1), np.delete(orec, 7, 1))
    diff = np.abs(rec[:, 7] - orec[:, 7])
    _say(f"targets: max |device - oracle| {diff.max():.2e}, over 1e-5 relative "
         f"{(diff > 1e-5 * (1 + np.abs(orec[:, 7]))).mean()}")
    assert (diff > 1e-5 * (1 + np.abs(orec[:, 7]))).mean() < 0.01 and diff.max() < 1e-3
    # without a volume source the targets are the data at the samples (mixed types)
    el2, _, _ = small_icosphere(3, kind="mixed")
    scene2, osc2 = _scenes(el2, None, oracle)
    r2 = scene2.harmonic_cv_targets(2048, volume_samples=0, seed=7, stream_id=3).cpu().numpy()
    assert np.array_equal(r2, osc2.hcv_targets(2048, seed=7, stream=3, volume_samples=0))


def test_hcv_fit_recovers_a_representable_field(gpu, oracle):
    """Data equal to a combination of the fit's own basis functions: the fit reproduces it."""
    V, F = G.icosphere(3)
    V = G.displace_radial(V)
    el0 = G.mesh_elements(V, F, G.DIRICHLET, lambda p, nr, ty: np.zeros(len(p)))
    src = hcv_sources(el0, 32, 1.5)
    w = np.random.default_rng(3).uniform(-0.3, 0.3, 32)

    def h(p):
        return hcv_eval_f64(p, src, w, 0.5)
    # the data of each sample point is the linear interpolant of h at the element's vertices,
    # so the fit sees h up to that interpolation; a finer mesh keeps it small
    V4, F4 = G.icosphere(5)
    el = G.mesh_elements(G.displace_radial(V4), F4, G.DIRICHLET, lambda p, nr, ty: h(p)[:, 0])
    import wobtoolbox_amd as wob
    scene = wob.SceneHost(el)
    rms = scene.fit_harmonic_cv(n_sources=32, radius_scale=1.5, n_train=8192, volume_samples=0, seed=1)
    pts = G.interior_points_star(512, np.random.default_rng(4), margin=0.05).astype(np.float32)
    got = scene.eval_harmonic_cv(_dev(pts, gpu)).cpu().numpy()
    err = np.abs(got[:, 0] - h(pts)[:, 0]).max()
    _say(f"representable field: fit residual {rms:.2e}, interior error {err:.2e} (|h| ~ {np.abs(h(pts)[:, 0]).max():.2f})")
    assert rms < 1e-4 and err < 2e-4


def test_hcv_fit_matches_f64_least_squares(gpu, oracle):
    el, srcv, _ = _poisson(3)
    import wobtoolbox_amd as wob
    el2, _, _ = small_icosphere(3, kind="mixed")
    el3, _, _ = small_icosphere(3, kind="robin")
    for e, vol in ((el, True), (el2, False), (el3, False)):
        scene = wob.SceneHost(e)
        if vol:
            scene.set_volume_gaussians(*srcv)
        rec = scene.harmonic_cv_targets(8192, volume_samples=8, eps=1e-4, seed=2).cpu().numpy()
        rms = scene.fit_harmonic_cv(n_sources=64, radius_scale=1.5, n_train=8192, volume_samples=8, eps=1e-4, seed=2)
        src = hcv_sources(e, 64, 1.5)
        c0, w = hcv_fit_numpy(rec, src)
        got_src, got_c0 = scene.harmonic_cv()
        assert np.allclose(got_src[:, :3], src, rtol=0, atol=1e-6)
        pts = G.interior_points_star(256, np.random.default_rng(4), margin=0.02).astype(np.float32)
        dev = scene.eval_harmonic_cv(_dev(pts, gpu)).cpu().numpy()[:, 0]
        ref = hcv_eval_f64(pts, src, w, c0)[:, 0]
        _say(f"fit vs FP64 least squares ({'Dirichlet + volume' if vol else 'mixed'}, {e.n} elements, "
             f"Robin {(e.boundary_types == G.ROBIN).any()}): "
             f"max |u_h - u_ref| {np.abs(dev - ref).max():.2e}; residual {rms:.4f}")
        assert np.abs(dev - ref).max() < 1e-4 * (1 + np.abs(ref).max())


def _fit_both(scene, osc, n_sources=64):
    rms = scene.fit_harmonic_cv(n_sources=n_sources, radius_scale=1.5, n_train=16384, volume_samples=16, eps=1e-4,
                                seed=5)
    rec, c0 = scene.harmonic_cv()
    osc.set_harmonic_cv(rec, c0)
    return rms


def test_hcv_walk_matches_oracle(gpu, oracle):
    """Path by path on the same streams, origin element excluded from line queries (so the two
    sides take the same paths to FP32 rounding), and in the reference's own behaviour."""
    import wobtoolbox_amd as wob
    el, srcv, _ = _poisson(3)
    scene, osc = _scenes(el, srcv, oracle)
    _fit_both(scene, osc)
    pts = G.interior_points_star(64, np.random.default_rng(7), margin=0.05).astype(np.float32)
    paths = 256
    for flags in (1 | HCV, HCV):
        solver = dict(DL, path_length=4, num_volume_samples=1, flags=flags)
        est = wob.Estimator.from_config(scene, solver)
        s, q = est.compute_domain(_dev(pts, gpu), paths, seed=11)
        m, se = mean_se(s.cpu().numpy(), q.cpu().numpy(), paths)
        so, qo = osc.estimate_domain(solver, pts, paths, seed=11, pick_mode=1)
        mo, seo = mean_se(so, qo, paths)
        z = np.abs(m - mo) / np.sqrt(se ** 2 + seo ** 2 + 1e-30)
        frac = (np.abs(m - mo) > 1e-4 * (np.abs(mo) + se + 1e-3)).mean()
        _say(f"control-variate walk flags {flags}: max |z| {z.max():.3f}, points over 1e-4 rel {frac:.3f}")
        assert z.max() < 0.05 and frac < 0.02


def test_hcv_batch_and_wavefront_bit_identical(gpu, oracle):
    """Batches and the wavefront schedule give the single call's sums bit for bit."""
    import wobtoolbox_amd as wob
    el, srcv, _ = _poisson(3)
    scene, osc = _scenes(el, srcv, oracle)
    _fit_both(scene, osc, 32)
    sets = [G.interior_points_star(n, np.random.default_rng(20 + n), margin=0.05).astype(np.float32)
            for n in (100, 333, 1000)]
    solver = dict(DL, path_length=4, num_volume_samples=1, flags=HCV)
    est = wob.Estimator.from_config(scene, solver)
    single = [est.compute_domain(_dev(p, gpu), 32, seed=5, stream_id=j) for j, p in enumerate(sets)]
    b = wob.Batch([(est, _dev(p, gpu), 32, j) for j, p in enumerate(sets)], persistent=True)
    b.run(1, seed=5)
    for (s, q), (bs, bq) in zip(single, b.outputs):
        assert np.array_equal(s.cpu().numpy(), bs.cpu().numpy()) and np.array_equal(q.cpu().numpy(), bq.cpu().numpy())
    # wavefront schedule (no volume source: the wavefront walk covers walks without one)
    V, F = G.icosphere(3)
    el2 = G.mesh_elements(G.displace_radial(V), F, G.DIRICHLET, lambda p, nr, ty: G.exact_expcos_xz_3d(p)[:, 0])
    sc2 = wob.SceneHost(el2)
    sc2.fit_harmonic_cv(n_sources=32, n_train=4096, seed=1)
    e2 = wob.Estimator.from_config(sc2, dict(DL, path_length=4, flags=HCV))
    p = _dev(sets[1], gpu)
    ref = e2.compute_domain(p, 64, seed=9)
    for thr in (1, 32):
        e2.schedule, e2.regen_threshold = 2, thr
        got = e2.compute_domain(p, 64, seed=9)
        assert np.array_equal(ref[0].cpu().numpy(), got[0].cpu().numpy())
        assert np.array_equal(ref[1].cpu().numpy(), got[1].cpu().numpy())


def test_hcv_unbiased_and_variance(gpu):
    """Against u = N_f + H on a level-4 icosphere: both estimators within their standard errors,
    the control variate's variance far below plain WoB's."""
    import wobtoolbox_amd as wob
    el, srcv, ex = _poisson(4)
    scene = wob.SceneHost(el)
    scene.set_volume_gaussians(*srcv)
    rms = scene.fit_harmonic_cv(n_sources=64, radius_scale=1.5, n_train=16384, volume_samples=16, eps=1e-4, seed=5)
    pts = G.interior_points_star(256, np.random.default_rng(9), margin=0.05).astype(np.float32)
    n = 8192
    var = {}
    for flags in (0, HCV):
        est = wob.Estimator.from_config(scene, dict(DL, path_length=4, num_volume_samples=1, flags=flags))
        s, q = est.compute_domain(_dev(pts, gpu), n, seed=13)
        m, se = mean_se(s.cpu().numpy(), q.cpu().numpy(), n)
        z = (m - ex(pts)) / se
        var[flags] = float(np.mean(se[:, 0] ** 2))
        _say(f"flags {flags}: max |z| u {np.abs(z[:, 0]).max():.2f}, grad {np.abs(z[:, 1:]).max():.2f}, "
             f"mean z {z[:, 0].mean():+.3f}, per-path variance {var[flags] * n:.4f} (fit residual {rms:.4f})")
        assert np.abs(z[:, 0]).max() < 4.5 and np.abs(z[:, 1:]).max() < 5.0
        assert abs(z[:, 0].mean()) < 4 / np.sqrt(len(pts)) * 1.5
    assert var[0] / var[HCV] > 40, var


def test_hcv_errors(gpu):
    import wobtoolbox_amd as wob
    el, _, _ = _poisson(2)
    scene = wob.SceneHost(el)
    est = wob.Estimator.from_config(scene, dict(DL, path_length=4, flags=HCV))
    pts = _dev(np.zeros((4, 3), np.float32), gpu)
    with pytest.raises(wob.WobError, match="harmonic control variate"):
        est.compute_domain(pts, 4)
    with pytest.raises(wob.WobError, match="volume source"):
        from wobtoolbox_amd import _capi
        import ctypes as C
        d = _capi.HcvDesc(8, 1.5, 64, 4, 1e-4, 0.0, 0, 0)
        _capi.check(_capi.lib().wob_scene_fit_harmonic_cv(scene.handle, C.byref(d), None, None))
    with pytest.raises(wob.WobError, match="radius_scale"):
        scene.fit_harmonic_cv(n_sources=8, radius_scale=0.9, n_train=64)
    scene.fit_harmonic_cv(n_sources=8, n_train=64)
    s, _ = est.compute_domain(pts, 4)
    assert np.isfinite(s.cpu().numpy()).all()
    scene.set_harmonic_cv(None)
    assert scene.harmonic_cv()[0].shape[0] == 0
    with pytest.raises(wob.WobError):
        est.compute_domain(pts, 4)


def test_hcv_single_layer_walk_matches_oracle(gpu, oracle):
    """The single-layer mixed RIS walk with the harmonic control variate, path by path."""
    import wobtoolbox_amd as wob
    el, _, _ = small_icosphere(3, kind="mixed")
    scene, osc = _scenes(el, None, oracle)
    scene.fit_harmonic_cv(n_sources=64, n_train=16384, seed=5)
    rec, c0 = scene.harmonic_cv()
    osc.set_harmonic_cv(rec, c0)
    pts = G.interior_points_star(64, np.random.default_rng(7), margin=0.05).astype(np.float32)
    paths = 128
    solver = dict(SL_MIXED, flags=HCV)
    est = wob.Estimator.from_config(scene, solver)
    s, q = est.compute_domain(_dev(pts, gpu), paths, seed=11)
    m, se = mean_se(s.cpu().numpy(), q.cpu().numpy(), paths)
    so, qo = osc.estimate_domain(solver, pts, paths, seed=11, pick_mode=1)
    mo, seo = mean_se(so, qo, paths)
    z = np.abs(m - mo) / np.sqrt(se ** 2 + seo ** 2 + 1e-30)
    frac = (np.abs(m - mo) > 1e-4 * (np.abs(mo) + se + 1e-3)).mean()
    _say(f"single-layer control-variate walk: max |z| {z.max():.3f}, points over 1e-4 rel {frac:.3f}")
    assert z.max() < 0.05 and frac < 0.02


def test_hcv_single_layer_robin_walk_matches_oracle(gpu, oracle):
    """Robin data (alpha 1 below z = 0, Dirichlet above) with the harmonic control variate:
    the walk path by path against the oracle (the Robin residual g - (du_h/dn + alpha u_h))
    and against the exact solution."""
    import wobtoolbox_amd as wob
    el, _, _ = small_icosphere(3, kind="robin")
    scene, osc = _scenes(el, None, oracle)
    scene.fit_harmonic_cv(n_sources=64, n_train=16384, seed=5)
    rec, c0 = scene.harmonic_cv()
    osc.set_harmonic_cv(rec, c0)
    pts = G.interior_points_star(64, np.random.default_rng(7), margin=0.05).astype(np.float32)
    paths = 128
    solver = dict(SL_MIXED, flags=HCV)
    est = wob.Estimator.from_config(scene, solver)
    s, q = est.compute_domain(_dev(pts, gpu), paths, seed=11)
    m, se = mean_se(s.cpu().numpy(), q.cpu().numpy(), paths)
    so, qo = osc.estimate_domain(solver, pts, paths, seed=11, pick_mode=1)
    mo, seo = mean_se(so, qo, paths)
    z = np.abs(m - mo) / np.sqrt(se ** 2 + seo ** 2 + 1e-30)
    frac = (np.abs(m - mo) > 1e-4 * (np.abs(mo) + se + 1e-3)).mean()
    ex = G.exact_expcos_xz_3d(pts)
    ze = np.abs(m - ex) / se
    _say(f"Robin control-variate walk: max |z| vs oracle {z.max():.3f}, points over 1e-4 rel {frac:.3f}, "
         f"max |z| vs exact {ze.max():.2f}")
    assert z.max() < 0.05 and frac < 0.02
    assert ze[:, 0].max() < 4.5 and ze[:, 1:].max() < 5.0


def test_hcv_single_layer_unbiased_and_variance(gpu):
    """Mixed single-layer workload on a level-5 icosphere against exp(0.6 x) cos(0.6 y) + x z:
    plain WoB and the control variate within their standard errors, the variance far lower;
    the value-table control variate trained on the residual walk combines with it."""
    import wobtoolbox_amd as wob
    el, _, _ = small_icosphere(5, kind="mixed")
    scene = wob.SceneHost(el)
    rms = scene.fit_harmonic_cv(n_sources=128, n_train=32768, seed=5)
    pts = G.interior_points_star(512, np.random.default_rng(9), margin=0.05).astype(np.float32)
    ex = G.exact_expcos_xz_3d(pts)
    var = {}
    for flags, n in ((0, 512), (HCV, 512)):
        est = wob.Estimator.from_config(scene, dict(SL_MIXED, flags=flags))
        s, q = est.compute_domain(_dev(pts, gpu), n, seed=13)
        m, se = mean_se(s.cpu().numpy(), q.cpu().numpy(), n)
        z = (m - ex) / se
        var[flags] = float(np.mean(se[:, 0] ** 2)) * n
        _say(f"single layer flags {flags}: max |z| u {np.abs(z[:, 0]).max():.2f}, grad {np.abs(z[:, 1:]).max():.2f}, "
             f"per-path variance {var[flags]:.4g}, rmse {np.sqrt(np.mean((m[:, 0] - ex[:, 0]) ** 2)):.3g} "
             f"(fit residual {rms:.2e})")
        assert np.abs(z[:, 0]).max() < 4.5 and np.abs(z[:, 1:]).max() < 5.0
    assert var[0] / var[HCV] > 1e4, var


def _star(kind, n_segments=1200):
    from helpers import star_scene
    return star_scene(seed=4, n_segments=n_segments, kind=kind)


def test_hcv_exterior_fit_matches_f64_least_squares(gpu, oracle):
    """The exterior fit (WOB_HCV_EXTERIOR): sources inside the body at 0.75 x the smallest
    centre-to-element distance (FP64 point-segment distances), no constant, and u_h equal to
    an FP64 least-squares solve of the same targets."""
    import wobtoolbox_amd as wob
    for kind in (0, 1, 2):
        w = _star(kind)
        scene = wob.SceneHost(w.elements)
        rec = scene.harmonic_cv_targets(8192, seed=2, exterior=True).cpu().numpy()
        rms = scene.fit_harmonic_cv(n_sources=64, n_train=8192, seed=2, exterior=True)
        src = hcv_sources(w.elements, 64, 0.75, exterior=True)
        c0, wt = hcv_fit_numpy(rec, src, constant=False)
        got, got_c0 = scene.harmonic_cv()
        assert got_c0 == 0.0
        assert np.allclose(got[:, :2], src, rtol=0, atol=1e-6)
        pts = w.points[:256]
        dev = scene.eval_harmonic_cv(_dev(pts, gpu)).cpu().numpy()[:, 0]
        ref = hcv_eval_f64(pts, src, wt, 0.0)[:, 0]
        _say(f"exterior fit kind {kind}: max |u_h - u_ref| {np.abs(dev - ref).max():.2e}, residual {rms:.2e}, "
             f"max |u_h - u| {np.abs(ref - w.exact(pts)[:, 0]).max():.2e}")
        assert np.abs(dev - ref).max() < 1e-5 * (1 + np.abs(ref).max())


def test_hcv_exterior_walk_matches_oracle(gpu, oracle):
    """Exterior single-layer RIS walks (Neumann, Robin, mixed) with the exterior control
    variate, path by path against the oracle, in single calls and in one batch launch."""
    import wobtoolbox_amd as wob
    ests, sets = [], []
    for kind in (0, 1, 2):
        w = _star(kind, 400)
        scene, osc = _scenes(w.elements, None, oracle)
        scene.fit_harmonic_cv(n_sources=64, n_train=8192, seed=3, exterior=True)
        rec, c0 = scene.harmonic_cv()
        osc.set_harmonic_cv(rec, c0)
        pts = w.points[:64]
        paths = 64
        solver = dict(w.solver, flags=HCV)
        est = wob.Estimator.from_config(scene, solver)
        s, q = est.compute_domain(_dev(pts, gpu), paths, seed=11)
        m, se = mean_se(s.cpu().numpy(), q.cpu().numpy(), paths)
        so, qo = osc.estimate_domain(solver, pts, paths, seed=11, pick_mode=1)
        mo, _ = mean_se(so, qo, paths)
        scale = np.abs(mo).max(0) + 1e-6
        frac = (np.abs(m - mo) > 1e-4 * scale).mean()
        _say(f"exterior control-variate walk kind {kind}: max |m - m_oracle| / scale "
             f"{(np.abs(m - mo) / scale).max():.2e}, points over 1e-4 {frac:.3f}")
        assert frac < 0.03
        ests.append((est, s, q))
        sets.append(pts)
    b = wob.Batch([(est, _dev(p, gpu), 64, 0) for (est, _, _), p in zip(ests, sets)], persistent=True)
    b.run(1, seed=11)
    for (_, s, q), (bs, bq) in zip(ests, b.outputs):
        assert np.array_equal(s.cpu().numpy(), bs.cpu().numpy()) and np.array_equal(q.cpu().numpy(), bq.cpu().numpy())


def test_hcv_exterior_errors(gpu):
    import wobtoolbox_amd as wob
    w = _star(0, 300)
    scene = wob.SceneHost(w.elements)
    pts = _dev(w.points[:4], gpu)
    ext = wob.Estimator.from_config(scene, dict(w.solver, flags=HCV))
    interior = wob.Estimator.from_config(scene, dict(w.solver, domain="InteriorDomain", flags=HCV))
    with pytest.raises(wob.WobError, match="radius_scale"):
        scene.fit_harmonic_cv(n_sources=8, radius_scale=1.5, n_train=64, exterior=True)
    scene.fit_harmonic_cv(n_sources=8, n_train=64)   # interior fit
    with pytest.raises(wob.WobError, match="exterior walks need"):
        ext.compute_domain(pts, 4)
    scene.fit_harmonic_cv(n_sources=8, n_train=64, exterior=True)
    with pytest.raises(wob.WobError, match="interior walks need"):
        interior.compute_domain(pts, 4)
    s, _ = ext.compute_domain(pts, 4)
    assert np.isfinite(s.cpu().numpy()).all()
    src, c0 = scene.harmonic_cv()
    scene.set_harmonic_cv(src, c0)   # explicit sources serve either side
    ext.compute_domain(pts, 4)
    interior.compute_domain(pts, 4)


def test_hcv_exterior_3d_walk_matches_oracle(gpu, oracle):
    """3D exterior: Neumann data of the dipole u = z / |x|^3 on a level-3 displaced icosphere
    (single layer, exterior, RIS 16), the exterior fit's sources inside the body: device and
    oracle path by path, and the error against u far below plain WoB's."""
    import wobtoolbox_amd as wob
    V, F = G.icosphere(3)
    V = G.displace_radial(V)

    def dipole(p):
        p = np.asarray(p, np.float64)
        r2 = (p * p).sum(1)
        r3, r5 = r2 ** 1.5, r2 ** 2.5
        return np.stack([p[:, 2] / r3, -3 * p[:, 2] * p[:, 0] / r5, -3 * p[:, 2] * p[:, 1] / r5,
                         1 / r3 - 3 * p[:, 2] ** 2 / r5], 1)
    el = G.mesh_elements(V, F, G.NEUMANN, lambda p, nr, ty: (dipole(p)[:, 1:] * nr).sum(1))
    scene, osc = _scenes(el, None, oracle)
    rms = scene.fit_harmonic_cv(n_sources=64, n_train=16384, seed=3, exterior=True)
    rec, c0 = scene.harmonic_cv()
    assert c0 == 0.0 and np.linalg.norm(rec[:, :3], axis=1).max() < 0.75
    # the inner radius from FP64 point-triangle distances (host) against numpy's
    assert np.allclose(rec[:, :3], hcv_sources(el, 64, 0.75, exterior=True), rtol=0, atol=1e-6)
    osc.set_harmonic_cv(rec, c0)
    rng = np.random.default_rng(8)
    d = rng.normal(size=(64, 3))
    pts = (d / np.linalg.norm(d, axis=1, keepdims=True) * rng.uniform(1.6, 3.0, (64, 1))).astype(np.float32)
    solver = dict(bie_model="IndirectSingleLayer", domain="ExteriorDomain", estimator="BackwardEstimator",
                  problem="NeumannProblem", path_length=6, eps=1e-5, num_resampling_candidates=16)
    paths = 64
    ex = dipole(pts)
    err = {}
    for flags in (0, HCV):
        est = wob.Estimator.from_config(scene, dict(solver, flags=flags))
        s, q = est.compute_domain(_dev(pts, gpu), paths, seed=5)
        m, _ = mean_se(s.cpu().numpy(), q.cpu().numpy(), paths)
        err[flags] = float(np.sqrt(np.mean((m[:, 0] - ex[:, 0]) ** 2)))
        if flags:
            so, qo = osc.estimate_domain(dict(solver, flags=flags), pts, paths, seed=5, pick_mode=1)
            mo, _ = mean_se(so, qo, paths)
            scale = np.abs(mo).max(0) + 1e-6
            frac = (np.abs(m - mo) > 1e-4 * scale).mean()
            _say(f"3D exterior control-variate walk: fit residual {rms:.2e}, points over 1e-4 vs oracle {frac:.3f}")
            assert frac < 0.03
    _say(f"3D exterior Neumann dipole: RMSE u plain {err[0]:.3e}, with the control variate {err[HCV]:.3e}")
    assert err[HCV] < 0.05 * err[0]


def test_hcv_exterior_double_layer_walk_matches_oracle(gpu, oracle):
    """Exterior double-layer Dirichlet walks (line sampling, path length 6) with a zero-charge
    exterior fit: the device fit equals the FP64 constrained least squares (sum_j w_j = 0), the
    walk matches the oracle path by path in 2D (star polygon) and 3D (the dipole z / |x|^3 on a
    displaced icosphere), and the error against the exact solution is far below plain WoB's."""
    import wobtoolbox_amd as wob
    from test_oracle_hcv import _star_dirichlet_exterior
    solver = dict(bie_model="IndirectDoubleLayer", domain="ExteriorDomain", estimator="BackwardEstimator",
                  problem="DirichletProblem", path_length=6, eps=1e-5)
    el2, ex2, pts2 = _star_dirichlet_exterior(400)
    V, F = G.icosphere(3)

    def dipole(p):
        p = np.asarray(p, np.float64)
        r2 = (p * p).sum(1)
        return np.stack([p[:, 2] / r2 ** 1.5, -3 * p[:, 2] * p[:, 0] / r2 ** 2.5, -3 * p[:, 2] * p[:, 1] / r2 ** 2.5,
                         1 / r2 ** 1.5 - 3 * p[:, 2] ** 2 / r2 ** 2.5], 1)
    el3 = G.mesh_elements(G.displace_radial(V), F, G.DIRICHLET, lambda p, nr, ty: dipole(p)[:, 0])
    rng = np.random.default_rng(8)
    d = rng.normal(size=(48, 3))
    pts3 = (d / np.linalg.norm(d, axis=1, keepdims=True) * rng.uniform(1.6, 3.0, (48, 1))).astype(np.float32)
    for el, ex, pts in ((el2, ex2, pts2), (el3, dipole, pts3)):
        D = el.dim
        scene, osc = _scenes(el, None, oracle)
        rms = scene.fit_harmonic_cv(n_sources=64, n_train=8192, seed=3, exterior=True, zero_charge=True)
        rec, c0 = scene.harmonic_cv()
        t = scene.harmonic_cv_targets(8192, seed=3, exterior=True).cpu().numpy()
        src = hcv_sources(el, 64, 0.75, exterior=True)
        _, wt = hcv_fit_numpy(t, src, constant=False, zero_charge=True)
        u_dev = scene.eval_harmonic_cv(_dev(pts, gpu)).cpu().numpy()[:, 0]
        u_ref = hcv_eval_f64(pts, src, wt, 0.0)[:, 0]
        assert c0 == 0.0 and abs(rec[:, D].astype(np.float64).sum()) < 1e-5 * np.abs(rec[:, D]).sum()
        assert np.abs(u_dev - u_ref).max() < 1e-5 * (1 + np.abs(u_ref).max())
        osc.set_harmonic_cv(rec, c0)
        paths = 64
        err = {}
        for flags in (0, HCV):
            est = wob.Estimator.from_config(scene, dict(solver, flags=flags))
            s, q = est.compute_domain(_dev(pts, gpu), paths, seed=5)
            m, _ = mean_se(s.cpu().numpy(), q.cpu().numpy(), paths)
            err[flags] = float(np.sqrt(np.mean((m[:, 0] - ex(pts)[:, 0]) ** 2)))
            if flags:
                so, qo = osc.estimate_domain(dict(solver, flags=flags), pts, paths, seed=5, pick_mode=1)
                mo, _ = mean_se(so, qo, paths)
                scale = np.abs(mo).max(0) + 1e-6
                frac = (np.abs(m - mo) > 1e-4 * scale).mean()
                assert frac < 0.03, frac
        _say(f"{D}D exterior double layer: fit residual {rms:.2e}, RMSE u plain {err[0]:.3e}, "
             f"with the control variate {err[HCV]:.3e}")
        assert err[HCV] < 0.05 * err[0]
    # a fit without the zero-charge constraint is refused for exterior double-layer walks
    scene = wob.SceneHost(el2)
    scene.fit_harmonic_cv(n_sources=16, n_train=256, seed=3, exterior=True)
    est = wob.Estimator.from_config(scene, dict(solver, flags=HCV))
    with pytest.raises(wob.WobError, match="ZERO_CHARGE"):
        est.compute_domain(_dev(pts2, gpu), 4)
    with pytest.raises(wob.WobError, match="needs WOB_HCV_EXTERIOR"):
        scene.fit_harmonic_cv(n_sources=16, n_train=256, zero_charge=True)


def test_hcv_forward_walk_matches_oracle(gpu, oracle):
    """Forward estimators with the control variate (uniform start point carrying the residual
    datum, u_h added): single-layer interior Neumann with line sampling (the reference's
    interior Neumann configs) on the level-3 icosphere, path by path against the oracle, and
    its gradients against the exact ones with the variance far below plain WoB's."""
    import wobtoolbox_amd as wob
    el, _, _ = small_icosphere(3)
    scene, osc = _scenes(el, None, oracle)
    scene.fit_harmonic_cv(n_sources=64, n_train=16384, seed=5)
    rec, c0 = scene.harmonic_cv()
    osc.set_harmonic_cv(rec, c0)
    pts = G.interior_points_star(64, np.random.default_rng(7), margin=0.05).astype(np.float32)
    solver = dict(bie_model="IndirectSingleLayer", domain="InteriorDomain", estimator="ForwardEstimator",
                  problem="NeumannProblem", path_length=7, eps=1e-5)
    ex = G.exact_quadratic_3d(pts)
    paths = 256
    var = {}
    for flags in (0, HCV):
        est = wob.Estimator.from_config(scene, dict(solver, flags=flags))
        s, q = est.compute_domain(_dev(pts, gpu), paths, seed=11)
        m, se = mean_se(s.cpu().numpy(), q.cpu().numpy(), paths)
        var[flags] = float(np.mean(se[:, 1:] ** 2))
        if flags:
            so, qo = osc.estimate_domain(dict(solver, flags=flags), pts, paths, seed=11, pick_mode=1)
            mo, _ = mean_se(so, qo, paths)
            scale = np.abs(mo).max(0) + 1e-6
            frac = (np.abs(m - mo) > 1e-4 * scale).mean()
            z = np.abs(m[:, 1:] - ex[:, 1:]) / se[:, 1:]
            _say(f"forward control-variate walk: points over 1e-4 vs oracle {frac:.3f}, gradient max |z| vs exact "
                 f"{z.max():.2f}")
            assert frac < 0.03 and z.max() < 5.0
    _say(f"forward Neumann: gradient variance plain {var[0]:.3e}, with the control variate {var[HCV]:.3e}")
    assert var[0] / var[HCV] > 100


def test_hcv_direct_model_walk_matches_oracle(gpu, oracle):
    """The direct model with the control variate (the line walk of the direct Neumann
    workload, and the uniform flavour of mixed data): path by path against the oracle in the
    per-lane and the wavefront schedules, and far less variance than plain WoB."""
    import wobtoolbox_amd as wob
    for kind, problem, exact in (("neumann", "NeumannProblem", G.exact_quadratic_3d),
                                 ("mixed", "MixedProblem", G.exact_expcos_xz_3d)):
        el, _, _ = small_icosphere(3, kind=kind)
        scene, osc = _scenes(el, None, oracle)
        scene.fit_harmonic_cv(n_sources=64, n_train=16384, seed=5)
        rec, c0 = scene.harmonic_cv()
        osc.set_harmonic_cv(rec, c0)
        pts = G.interior_points_star(64, np.random.default_rng(7), margin=0.05).astype(np.float32)
        solver = dict(bie_model="Direct", domain="InteriorDomain", estimator="BackwardEstimator", problem=problem,
                      path_length=6, eps=1e-5)
        paths = 128
        var = {}
        for flags in (0, HCV):
            est = wob.Estimator.from_config(scene, dict(solver, flags=flags))
            s, q = est.compute_domain(_dev(pts, gpu), paths, seed=11)
            m, se = mean_se(s.cpu().numpy(), q.cpu().numpy(), paths)
            var[flags] = float(np.mean(se ** 2))
            if flags:
                so, qo = osc.estimate_domain(dict(solver, flags=flags), pts, paths, seed=11, pick_mode=1)
                mo, _ = mean_se(so, qo, paths)
                scale = np.abs(mo).max(0) + 1e-6
                frac = (np.abs(m - mo) > 1e-4 * scale).mean()
                _say(f"direct {problem} control-variate walk: points over 1e-4 vs oracle {frac:.3f}")
                assert frac < 0.05, frac
                if problem == "NeumannProblem":   # the line walk also runs in the wavefront schedule
                    est.schedule = 2
                    sw, qw = est.compute_domain(_dev(pts, gpu), paths, seed=11)
                    assert np.array_equal(s.cpu().numpy(), sw.cpu().numpy())
        _say(f"direct {problem}: variance plain {var[0]:.3e}, with the control variate {var[HCV]:.3e}")
        assert var[0] / var[HCV] > 100
