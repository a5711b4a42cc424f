"""CPU pins of the oracle's direct-model rows, exterior walks, forward uniform flavour and the
learned-control-variate workload's solver, each against mathematics rather than a second
restatement.

Green's representation formula.  For u harmonic in the domain (or outside the body, decaying)
and a boundary point x inside a face or segment, where the solid angle is one half,
    u(x) / 2 = +-( integral of G(x, y) du/dn(y) - dG/dn_y(x, y) u(y) dy ),
and for a domain point x the same with u(x).  The direct model (bie_models.cuh:404-588) is one
bounce of exactly this equation: with y uniform on the boundary and the exact Cauchy data,
    E_y[ fredholm_known(x, y) + fredholm_unknown(x, y) mu(y') ] = mu(x)
where mu is du/dn at Dirichlet points and u at Neumann and Robin points (Robin data
du/dn + alpha u), and y' = x when the first-kind split (Dirichlet x, constant c) is taken; and
    E_y[ solution_domain_known(x, y) + solution_domain_unknown(x, y) mu(y) ] = u(x)
with the gradients likewise.  On a polygon or polyhedron the identity is exact (no
discretisation error) when the data are evaluated at the sample points themselves.  Every
(x type, y type) row of the 3 x 3 tables is exercised, interior and exterior, in 2D and 3D.
"""
import numpy as np
import pytest

from helpers import mean_se, small_icosphere
from wobtoolbox_amd import geometry as G

DIRECT = 2


def _exact(dim, exterior):
    """Harmonic test functions: [u, grad u] (FP64).  Exterior: the reference's
    EXTERIOR_SOLUTION1 dipoles x / |x|^2 (general_2d.cu:78-82) and z / |x|^3 (general_3d.cu:80-84)."""
    if not exterior:
        return G.exact_exp_cos_2d if dim == 2 else G.exact_expcos_xz_3d
    if dim == 2:
        def f(p):
            p = np.asarray(p, np.float64)
            r2 = (p ** 2).sum(1)
            return np.stack([p[:, 0] / r2, 1 / r2 - 2 * p[:, 0] ** 2 / r2 ** 2, -2 * p[:, 0] * p[:, 1] / r2 ** 2], 1)
        return f

    def f3(p):
        p = np.asarray(p, np.float64)
        r2 = (p ** 2).sum(1)
        r = np.sqrt(r2)
        return np.stack([p[:, 2] / (r2 * r), -3 * p[:, 2] * p[:, 0] / (r2 ** 2 * r),
                         -3 * p[:, 2] * p[:, 1] / (r2 ** 2 * r), 1 / (r2 * r) - 3 * p[:, 2] ** 2 / (r2 ** 2 * r)], 1)
    return f3


def _geometry(dim, rng):
    """A non-convex star polygon (2D) or the displaced level-3 icosphere (3D) with a random
    Dirichlet / Neumann / Robin type and Robin coefficient per element (the stored data are
    unused: every record below carries exact data at its own point)."""
    if dim == 2:
        th = 2 * np.pi * np.arange(256) / 256
        r = 1 + 0.15 * np.cos(3 * th) + 0.05 * np.sin(5 * th)
        v = np.stack([r * np.cos(th), r * np.sin(th)], 1)
        types = rng.integers(0, 3, len(v))
        alpha = rng.uniform(0.5, 2.0, len(v))
        return G.polygon_elements(v, types, lambda p, nr, ty: np.zeros(len(p)), robin_alpha=alpha)
    V, F = G.icosphere(3)
    V = G.displace_radial(V)
    types = rng.integers(0, 3, len(F))
    alpha = rng.uniform(0.5, 2.0, len(F))
    return G.mesh_elements(V, F, types, lambda p, nr, ty: np.zeros(len(p)), robin_alpha=alpha)


def _areas(el):
    V = el.vertices.astype(np.float64)
    if el.dim == 2:
        return np.linalg.norm(V[:, 1] - V[:, 0], axis=1)
    return 0.5 * np.linalg.norm(np.cross(V[:, 1] - V[:, 0], V[:, 2] - V[:, 0]), axis=1)


def _records(el, e, p, exact):
    """Boundary records [p, n, 0.5, value, type, alpha] at points p of elements e with the exact
    data of the element's type, and the unknown mu there (du/dn at Dirichlet points, else u)."""
    D = el.dim
    p32 = p.astype(np.float32)
    ex = exact(p32.astype(np.float64))
    n = el.normals[e].astype(np.float64)
    dn = (ex[:, 1:] * n).sum(1)
    ty = el.boundary_types[e]
    al = el.robin_alphas[e, 0].astype(np.float64)
    val = np.where(ty == G.DIRICHLET, ex[:, 0], np.where(ty == G.NEUMANN, dn, dn + al * ex[:, 0]))
    mu = np.where(ty == G.DIRICHLET, dn, ex[:, 0])
    rec = np.zeros((len(e), 2 * D + 4), np.float32)
    rec[:, :D] = p32
    rec[:, D:2 * D] = el.normals[e]
    rec[:, 2 * D] = 0.5
    rec[:, 2 * D + 1] = val
    rec[:, 2 * D + 2] = ty
    rec[:, 2 * D + 3] = al
    return rec, mu


def _uniform_samples(el, n, rng, exact):
    """n uniform boundary samples (element by area, uniform in the element), inv_pdf = area."""
    a = _areas(el)
    e = rng.choice(el.n, n, p=a / a.sum())
    V = el.vertices.astype(np.float64)
    if el.dim == 2:
        u = rng.uniform(0, 1, n)[:, None]
        p = V[e, 0] + u * (V[e, 1] - V[e, 0])
    else:
        s = np.sqrt(rng.uniform(0, 1, n))[:, None]
        t = rng.uniform(0, 1, n)[:, None]
        p = (1 - s) * V[e, 0] + s * (1 - t) * V[e, 1] + s * t * V[e, 2]
    rec, mu = _records(el, e, p, exact)
    return rec, mu, np.full(n, a.sum(), np.float32)


def _centroids(el, e):
    return el.vertices[e].astype(np.float64).mean(1)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("exterior", [False, True])
def test_direct_model_fredholm_rows_satisfy_green_identity(dim, exterior, oracle):
    rng = np.random.default_rng(100 * dim + exterior)
    el = _geometry(dim, rng)
    exact = _exact(dim, exterior)
    c = 2.0
    ns = 1 << 18
    zs, worst_wrong = [], 0.0
    for ty in (G.DIRICHLET, G.NEUMANN, G.ROBIN):
        for e in np.flatnonzero(el.boundary_types == ty)[:3]:
            xrec, mux = _records(el, np.array([e]), _centroids(el, [e]), exact)
            yrec, muy, inv = _uniform_samples(el, ns, rng, exact)
            us = rng.uniform(0, 1, ns).astype(np.float32)
            w = oracle.eval_weights(dim, DIRECT, int(exterior), 1e-6, c, np.repeat(xrec, ns, 0), yrec, inv,
                                    np.zeros((ns, dim), np.float32), us).astype(np.float64)
            split = w[:, 2] > 0.5
            val = w[:, 0] + w[:, 1] * np.where(split, mux[0], muy)
            m, se = val.mean(), val.std() / np.sqrt(ns)
            zs.append((m - mux[0]) / se)
            if ty == G.DIRICHLET:
                # a plausible mistake, a sign error on the Dirichlet columns' unknown term
                mu_wrong = np.where(yrec[:, 2 * dim + 2] == G.DIRICHLET, -muy, muy)
                wrong = w[:, 0] + w[:, 1] * np.where(split, mux[0], mu_wrong)
                worst_wrong = max(worst_wrong, abs(wrong.mean() - mux[0]) / (wrong.std() / np.sqrt(ns)))
    zs = np.array(zs)
    print(f"dim {dim} exterior {exterior}: |z| of the 9 boundary rows {np.round(np.abs(zs), 2).tolist()}; "
          f"a sign error on the Dirichlet columns gives |z| up to {worst_wrong:.0f}")
    assert np.abs(zs).max() < 5.0
    assert worst_wrong > 20.0


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("exterior", [False, True])
def test_direct_model_domain_rows_satisfy_green_identity(dim, exterior, oracle):
    rng = np.random.default_rng(200 * dim + exterior)
    el = _geometry(dim, rng)
    exact = _exact(dim, exterior)
    if exterior:
        rr = rng.uniform(1.6, 2.5, 6)
        d = rng.normal(size=(6, dim))
        xd = d / np.linalg.norm(d, axis=1, keepdims=True) * rr[:, None]
    else:
        xd = rng.uniform(-0.45, 0.45, (6, dim))
    xd = xd.astype(np.float32)
    ns = 1 << 18
    zmax = 0.0
    for x in xd:
        yrec, muy, inv = _uniform_samples(el, ns, rng, exact)
        t = oracle.eval_domain_terms(dim, DIRECT, int(exterior), 1e-6, np.repeat(x[None], ns, 0), yrec,
                                     inv).astype(np.float64)
        val = np.concatenate([(t[:, 0] + t[:, 1] * muy)[:, None], t[:, 2:2 + dim] + t[:, 2 + dim:] * muy[:, None]], 1)
        m, se = val.mean(0), val.std(0) / np.sqrt(ns)
        ex = exact(x[None].astype(np.float64))[0]
        zmax = max(zmax, float(np.abs((m - ex) / se).max()))
    print(f"dim {dim} exterior {exterior}: domain rows max |z| {zmax:.2f}")
    assert zmax < 5.0


# ---------------------------------------------------------------------------------------
# Walks against closed-form solutions
# ---------------------------------------------------------------------------------------
def _star_2d(n=512):
    th = 2 * np.pi * np.arange(n) / n
    r = 1 + 0.15 * np.cos(3 * th) + 0.05 * np.sin(5 * th)
    return np.stack([r * np.cos(th), r * np.sin(th)], 1)


def _exterior_points(dim, n, rng, lo, hi):
    d = rng.normal(size=(n, dim))
    return (d / np.linalg.norm(d, axis=1, keepdims=True) * rng.uniform(lo, hi, n)[:, None]).astype(np.float32)


def _walk_z(oracle, el, solver, pts, exact, paths, seed):
    sc = oracle.OracleScene(el)
    s, q = sc.estimate_domain(solver, pts, paths, seed=seed)
    m, se = mean_se(s, q, paths)
    ex = exact(pts.astype(np.float64))
    return (m - ex) / se, np.abs(m - ex)[:, 0].mean() / np.abs(ex[:, 0]).mean()


@pytest.mark.parametrize("dim,bie,problem", [(2, "IndirectDoubleLayer", "DirichletProblem"),
                                             (3, "IndirectDoubleLayer", "DirichletProblem"),
                                             (2, "Direct", "NeumannProblem"),
                                             (3, "Direct", "NeumannProblem")])
def test_exterior_walks_match_reference_exterior_solution(dim, bie, problem, oracle):
    """Exterior double-layer Dirichlet (the reference's exterior Dirichlet configs, path length 6)
    and exterior direct Neumann walks around a non-convex star polygon and the displaced
    icosphere, against EXTERIOR_SOLUTION1 (x / |x|^2, z / |x|^3; general_2d.cu:78-82,
    general_3d.cu:80-84).  An error in the exterior sign or free term makes the estimate miss
    the solution by its own size; the noise here is 5-30 % of it."""
    rng = np.random.default_rng(7 * dim)
    exact = _exact(dim, True)
    pts = _exterior_points(dim, 24, rng, 1.6, 3.0)
    ty = G.DIRICHLET if problem == "DirichletProblem" else G.NEUMANN

    def data(p, nr, t):
        e = exact(p)
        return e[:, 0] if ty == G.DIRICHLET else (e[:, 1:] * nr).sum(1)
    if dim == 2:
        el = G.polygon_elements(_star_2d(), ty, data)
    else:
        V, F = G.icosphere(3)
        el = G.mesh_elements(G.displace_radial(V), F, ty, data)
    solver = dict(bie_model=bie, domain="ExteriorDomain", estimator="BackwardEstimator", problem=problem,
                  path_length=6)
    z, rel = _walk_z(oracle, el, solver, pts, exact, 16384 if dim == 2 else 32768, seed=5)
    print(f"{bie} {problem} exterior {dim}D: max |z| {np.abs(z).max():.2f}, rms z {np.sqrt((z ** 2).mean()):.2f}, "
          f"mean |error| / mean |u| {rel:.3f}")
    assert np.abs(z).max() < 5.0
    assert 0.6 < np.sqrt((z ** 2).mean()) < 1.6


@pytest.mark.parametrize("bie", ["IndirectDoubleLayer", "Direct"])
def test_forward_uniform_walk_matches_harmonic_solution(bie, oracle):
    """The forward estimator's uniform flavour (estimator.cuh:405-454; every forward
    combination but single layer + Neumann) on the unit-circle 256-gon with Dirichlet data
    exp(x) cos(y), path length 4, against the solution."""
    th = 2 * np.pi * np.arange(256) / 256
    v = np.stack([np.cos(th), np.sin(th)], 1)
    el = G.polygon_elements(v, G.DIRICHLET, lambda p, nr, t: G.exact_exp_cos_2d(p)[:, 0])
    pts = np.random.default_rng(2).uniform(-0.5, 0.5, (16, 2)).astype(np.float32)
    solver = dict(bie_model=bie, domain="InteriorDomain", estimator="ForwardEstimator", problem="DirichletProblem",
                  path_length=4, eps=1e-5, first_kind_equation_scaling_constant=2.0)
    z, rel = _walk_z(oracle, el, solver, pts, G.exact_exp_cos_2d, 16384, seed=11)
    print(f"forward uniform {bie}: max |z| {np.abs(z).max():.2f}, mean |error| / mean |u| {rel:.3f}")
    assert np.abs(z).max() < 5.0


@pytest.mark.parametrize("path_length", [4, 8])
def test_control_variate_workload_solver_matches_solution(path_length, oracle):
    """The learned-control-variate workload's solver (single layer, interior, mixed
    Dirichlet z > 0 / Neumann, RIS 64, c = 2; data/config_general_3d_interior_Mixed.json) on the
    level-3 displaced icosphere against exp(0.6 x) cos(0.6 y) + x z, at its path length 4 and
    at 8.  Its per-path standard deviation is about 200 times |u|, so at these path counts the
    check resolves errors of order the solution itself; the truncation bias of path length 4
    is bounded with the GPU at far more paths."""
    el, V, F = small_icosphere(3, "mixed")
    pts = G.interior_points_star(8, np.random.default_rng(5), margin=0.08).astype(np.float32)
    solver = dict(bie_model="IndirectSingleLayer", domain="InteriorDomain", estimator="BackwardEstimator",
                  problem="MixedProblem", path_length=path_length, num_resampling_candidates=64, eps=1e-5,
                  first_kind_equation_scaling_constant=2.0)
    z, rel = _walk_z(oracle, el, solver, pts, G.exact_expcos_xz_3d, 65536, seed=3)
    print(f"SL mixed RIS 64, path length {path_length}: max |z| {np.abs(z).max():.2f}, "
          f"mean |error| / mean |u| {rel:.3f}")
    assert np.abs(z).max() < 5.0
