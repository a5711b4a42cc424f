"""GPU parity of the Poisson volume terms (bie_models.cuh:79-154) through the C ABI.

The device walk with a volume source is compared per point with the oracle run on the
same Philox streams (grid cache sources as in scene.cuh:100-120, Gaussian-mixture sources
with the inside test), for every model / estimator branch that adds a volume term, and
against the closed-form solution u = N_f + H of the Gaussian-source problem.
"""
import sys

import numpy as np
import pytest

from helpers import boundary_records, mean_se
from wobtoolbox_amd import geometry as G

pytestmark = pytest.mark.gpu


def _dev(a, gpu):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a)).to(gpu)


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _mesh(level=3):
    V, F = G.icosphere(level)
    return G.displace_radial(V), F


def _sources(w_scene, o_scene, kind, w):
    if kind == "grid":
        w_scene.set_volume_cache(w.extra["grid_res"], w.extra["domain_size"], w.extra["cache"])
        o_scene.set_volume_grid(w.extra["grid_res"], w.extra["domain_size"], w.extra["cache"])
    else:
        w_scene.set_volume_gaussians(w.extra["centers"], w.extra["widths"], w.extra["amplitudes"])
        o_scene.set_volume_gaussians(w.extra["centers"], w.extra["widths"], w.extra["amplitudes"])


def _mixed_gaussian_workload(problem, level=3):
    """Gaussian source on the displaced icosphere with Dirichlet (z > 0) / Neumann data of u."""
    V, F = _mesh(level)
    base = G.poisson_gaussian_3d(point_sets=[8], V=V, F=F)
    ex = base.exact
    cz = V[F].mean(1)[:, 2]
    if problem == "DirichletProblem":
        types = np.full(len(F), G.DIRICHLET, np.int32)
    elif problem == "NeumannProblem":
        types = np.full(len(F), G.NEUMANN, np.int32)
    else:
        types = np.where(cz > 0, G.DIRICHLET, G.NEUMANN).astype(np.int32)

    def data(p, nr, ty):
        e = ex(p)
        return np.where(ty == G.DIRICHLET, e[:, 0], (e[:, 1:] * nr).sum(1))
    el = G.mesh_elements(V, F, types, data)
    return G.Workload("gauss_" + problem, el, base.points, ex, dict(base.solver, problem=problem), extra=base.extra)


CASES = {
    # name: (workload builder, source kind, solver overrides)
    "grid_dl_reference_reuse": (lambda: G.square_poisson_2d(n_points=48), "grid", dict(flags=1)),
    "grid_dl_independent": (lambda: G.square_poisson_2d(n_points=48), "grid", dict(flags=1 | 8, num_volume_samples=4)),
    "gauss_dl": (lambda: _mixed_gaussian_workload("DirichletProblem"), "gauss", dict(flags=1, num_volume_samples=1)),
    "gauss_sl_dirichlet": (lambda: _mixed_gaussian_workload("DirichletProblem"), "gauss",
                           dict(bie_model="IndirectSingleLayer", num_volume_samples=2,
                                first_kind_equation_scaling_constant=1.0)),
    "gauss_direct_mixed": (lambda: _mixed_gaussian_workload("MixedProblem"), "gauss",
                           dict(bie_model="Direct", num_volume_samples=1, first_kind_equation_scaling_constant=1.0)),
    "gauss_sl_forward_neumann": (lambda: _mixed_gaussian_workload("NeumannProblem"), "gauss",
                                 dict(bie_model="IndirectSingleLayer", estimator="ForwardEstimator", flags=1,
                                      num_volume_samples=1, path_length=3)),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_volume_walk_matches_oracle(case, gpu, oracle):
    import wobtoolbox_amd as wob
    build, kind, over = CASES[case]
    w = build()
    solver = dict(w.solver, **over)
    _say(f"volume case {case}: {solver['bie_model']} {solver['problem']} {solver['estimator']} "
         f"nvs {solver['num_volume_samples']} flags {solver.get('flags', 0)}")
    scene = wob.SceneHost(w.elements)
    osc = oracle.OracleScene(w.elements)
    _sources(scene, osc, kind, w)
    pts = w.points[:48] if kind == "grid" else \
        G.interior_points_star(48, np.random.default_rng(7), margin=0.1).astype(np.float32)
    paths = 256
    est = wob.Estimator.from_config(scene, solver)
    s, q = est.compute_domain(_dev(pts, gpu), paths, seed=11)
    s, q = s.cpu().numpy(), q.cpu().numpy()
    so, qo = osc.estimate_domain(solver, pts, paths, seed=11, pick_mode=1)
    m, se = mean_se(s, q, paths)
    mo, seo = mean_se(so, qo, paths)
    diff = np.abs(m - mo)
    frac = (diff > 1e-4 * (np.abs(mo) + se + 1e-3)).mean()
    z = diff / np.sqrt(se ** 2 + seo ** 2 + 1e-30)
    _say(f"  points differing > 1e-4 rel: {frac:.4f}; max |z| {z.max():.3f}")
    assert frac < 0.1 and z.max() < 0.5


def test_volume_boundary_estimator_matches_oracle(gpu, oracle):
    """On-boundary estimates with a volume term (solution_boundary_known, bie_models.cuh:259-266)."""
    import wobtoolbox_amd as wob
    w = _mixed_gaussian_workload("MixedProblem")
    solver = dict(w.solver, bie_model="IndirectSingleLayer", num_volume_samples=1,
                  first_kind_equation_scaling_constant=1.0)
    scene = wob.SceneHost(w.elements)
    osc = oracle.OracleScene(w.elements)
    _sources(scene, osc, "gauss", w)
    rec = boundary_records(w.elements, 48, np.random.default_rng(8))
    est = wob.Estimator.from_config(scene, solver)
    s, q = est.compute_boundary(_dev(rec, gpu), 256, seed=12)
    s, q = s.cpu().numpy(), q.cpu().numpy()
    so, qo = osc.estimate_boundary(solver, rec, 256, seed=12, pick_mode=1)
    m, se = mean_se(s, q, 256)
    mo, seo = mean_se(so, qo, 256)
    z = np.abs(m - mo) / np.sqrt(se ** 2 + seo ** 2 + 1e-30)
    _say(f"boundary estimator with volume: max |z| {z.max():.3f}")
    assert z.max() < 0.5


def test_volume_gaussian_analytic(gpu):
    """Double layer + Gaussian source on a level-4 icosphere against u = N_f + H."""
    import wobtoolbox_amd as wob
    V, F = _mesh(4)
    w = G.poisson_gaussian_3d(point_sets=[8], V=V, F=F)
    scene = wob.SceneHost(w.elements)
    scene.set_volume_gaussians(w.extra["centers"], w.extra["widths"], w.extra["amplitudes"])
    pts = G.interior_points_star(256, np.random.default_rng(9), margin=0.1).astype(np.float32)
    n = 8192
    est = wob.Estimator.from_config(scene, w.solver)
    s, q = est.compute_domain(_dev(pts, gpu), n, seed=13)
    m, se = mean_se(s.cpu().numpy(), q.cpu().numpy(), n)
    z = (m - w.exact(pts)) / se
    _say(f"Gaussian source analytic: max |z| u {np.abs(z[:, 0]).max():.2f}, grad {np.abs(z[:, 1:]).max():.2f}; "
         f"mean z u {z[:, 0].mean():.3f}; mean se u {se[:, 0].mean():.4f}")
    assert np.abs(z[:, 0]).max() < 5.0 and abs(z[:, 0].mean()) < 0.5
    assert np.abs(z[:, 1:]).max() < 6.0


def test_volume_requires_source_and_batches_match(gpu):
    import torch
    import wobtoolbox_amd as wob
    from wobtoolbox_amd._capi import WobError
    V, F = _mesh(3)
    w = G.poisson_gaussian_3d(point_sets=[40, 70, 25], V=V, F=F)
    plain = wob.SceneHost(w.elements)
    with pytest.raises(WobError):
        wob.Estimator.from_config(plain, w.solver).compute_domain(_dev(w.points[:4], gpu), 4)
    scenes = []
    for j in range(3):   # three sources of different strength on the same mesh
        sc = wob.SceneHost(w.elements)
        sc.set_volume_gaussians(w.extra["centers"], w.extra["widths"], w.extra["amplitudes"] * (1 + j))
        scenes.append(sc)
    sets = [_dev(p, gpu) for p in w.extra["point_sets"]]
    ests = [wob.Estimator.from_config(sc, w.solver, paths_per_lane=8) for sc in scenes]
    single = [e.compute_domain(p, 64, seed=5, stream_id=i) for i, (e, p) in enumerate(zip(ests, sets))]
    for persistent in (0, 1):
        b = wob.Batch([(e, p, 64, i) for i, (e, p) in enumerate(zip(ests, sets))], paths_per_lane=8,
                      persistent=bool(persistent), use_graph=bool(persistent))
        b.run(1, seed=5)
        torch.cuda.synchronize()
        for (s1, q1), (s2, q2) in zip(single, b.outputs):
            assert torch.equal(s1, s2) and torch.equal(q1, q2)
        b.close()
    # a new source on a scene invalidates the batch's table: the next pass uses the new source
    b = wob.Batch([(e, p, 64, i) for i, (e, p) in enumerate(zip(ests, sets))], paths_per_lane=8)
    scenes[0].set_volume_gaussians(w.extra["centers"], w.extra["widths"], w.extra["amplitudes"] * 7)
    b.run(1, seed=5)
    ref = ests[0].compute_domain(sets[0], 64, seed=5, stream_id=0)
    torch.cuda.synchronize()
    assert torch.equal(b.outputs[0][0], ref[0])
    _say("volume batches: bit-identical to single calls; source changes picked up")
