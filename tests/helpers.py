"""Shared test helpers: small scenes, random kernel inputs, ULP distances."""
from __future__ import annotations

import numpy as np

from wobtoolbox_amd import geometry as G

FLT_EPS = float(np.finfo(np.float32).eps)


def ulp_distance(a, b):
    """|a - b| in units of the FP32 spacing at max(|a|, |b|)."""
    a = np.asarray(a, np.float32).astype(np.float64)
    b = np.asarray(b, np.float32).astype(np.float64)
    scale = np.maximum(np.abs(a), np.abs(b))
    sp = np.spacing(scale.astype(np.float32)).astype(np.float64)
    sp = np.where(sp == 0, np.finfo(np.float32).tiny, sp)
    return np.abs(a - b) / sp


def kernel_scales(dim, y, x):
    """Natural magnitude of each kernel output column (the size of its terms), per sample:
    G, |grad G| for dG_dx and dG_dny, |grad^2 G| for d2G_dxdny and d2G_dnxdny."""
    r = np.linalg.norm(np.asarray(y, np.float64) - np.asarray(x, np.float64), axis=1)
    if dim == 2:
        g0 = np.maximum(1.0, np.abs(np.log(r))) / (2 * np.pi)
        g1 = 1.0 / (2 * np.pi * r)
        g2 = 2.0 / (2 * np.pi * r * r)
    else:
        g0 = 1.0 / (4 * np.pi * r)
        g1 = 1.0 / (4 * np.pi * r * r)
        g2 = 3.0 / (4 * np.pi * r ** 3)
    return np.stack([g0] + [g1] * (dim + 1) + [g2] * (dim + 1), 1)


def random_kernel_inputs(dim, n, rng, min_dist=1e-2):
    y = rng.uniform(-1, 1, (n, dim))
    x = rng.uniform(-1, 1, (n, dim))
    d = np.linalg.norm(y - x, axis=1)
    x[d < min_dist] += min_dist
    ny = rng.normal(size=(n, dim)); ny /= np.linalg.norm(ny, axis=1, keepdims=True)
    nx = rng.normal(size=(n, dim)); nx /= np.linalg.norm(nx, axis=1, keepdims=True)
    return [a.astype(np.float32) for a in (y, x, ny, nx)]


def small_circle(n=64, value=None):
    th = 2 * np.pi * np.arange(n) / n
    v = np.stack([np.cos(th), np.sin(th)], 1)
    fn = (lambda p, nr, ty: np.full(len(p), value)) if value is not None else \
        (lambda p, nr, ty: G.exact_exp_cos_2d(p)[:, 0])
    return G.polygon_elements(v, G.DIRICHLET, fn)


def small_icosphere(level=3, kind="neumann"):
    V, F = G.icosphere(level)
    V = G.displace_radial(V)
    if kind == "neumann":
        return G.mesh_elements(V, F, G.NEUMANN, lambda p, nr, ty: (G.exact_quadratic_3d(p)[:, 1:] * nr).sum(1)), V, F
    cz = V[F].mean(1)[:, 2]
    # "mixed": Dirichlet for z > 0, Neumann below; "robin": Dirichlet for z > 0, Robin (alpha 1,
    # data du/dn + u) below
    types = np.where(cz > 0, G.DIRICHLET, G.ROBIN if kind == "robin" else G.NEUMANN).astype(np.int32)

    def data(p, nr, ty):
        e = G.exact_expcos_xz_3d(p)
        dn = (e[:, 1:] * nr).sum(1)
        return np.where(ty == G.DIRICHLET, e[:, 0], np.where(ty == G.ROBIN, dn + e[:, 0], dn))
    return G.mesh_elements(V, F, types, data, robin_alpha=(types == G.ROBIN).astype(np.float64)), V, F


def star_scene(seed=0, n_segments=400, kind=1):
    """One exterior star polygon scene (kind 0 Neumann, 1 Robin, 2 mixed) with few points."""
    ws = G.star_exterior_batch_2d(n_scenes=1, seed=seed, min_segments=n_segments, max_segments=n_segments,
                                  min_points=64, max_points=64, kinds=[kind])
    return ws[0]


def boundary_records(el, n, rng):
    """Uniform boundary points as BoundaryPoint records [p, n, 0.5, value, type, alpha]."""
    D = el.dim
    areas = []
    for e in range(el.n):
        v = el.vertices[e].astype(np.float64)
        areas.append(np.linalg.norm(v[1] - v[0]) if D == 2 else 0.5 * np.linalg.norm(np.cross(v[1] - v[0], v[2] - v[0])))
    areas = np.array(areas)
    idx = rng.choice(el.n, size=n, p=areas / areas.sum())
    rec = np.zeros((n, 2 * D + 4), np.float32)
    for i, e in enumerate(idx):
        v = el.vertices[e].astype(np.float64)
        if D == 2:
            u = rng.uniform(0.05, 0.95)
            w = np.array([1 - u, u])
        else:
            a, b = rng.uniform(0.05, 0.9, 2)
            if a + b > 0.95:
                a, b = 0.95 - b * 0.5, b * 0.5
            w = np.array([1 - a - b, a, b])
        rec[i, :D] = w @ v
        rec[i, D:2 * D] = el.normals[e]
        rec[i, 2 * D] = 0.5
        rec[i, 2 * D + 1] = w @ el.boundary_values[e]
        rec[i, 2 * D + 2] = el.boundary_types[e]
        rec[i, 2 * D + 3] = w @ el.robin_alphas[e]
    return rec


def mean_se(s, q, n):
    m = s / n
    return m, np.sqrt(np.maximum(q / n - m * m, 0) / n)


def element_distances(el, c):
    """Distance from c to every element (segment or triangle), brute force in FP64 (the
    triangle's closest point as the minimum over a dense barycentric grid refined by
    projection onto the plane where it falls inside)."""
    V = el.vertices.astype(np.float64)
    c = np.asarray(c, np.float64)
    if el.dim == 2:
        a, b = V[:, 0], V[:, 1]
        ab = b - a
        t = np.clip(((c - a) * ab).sum(1) / (ab * ab).sum(1), 0, 1)
        return np.linalg.norm(a + t[:, None] * ab - c, axis=1)
    a, b, t3 = V[:, 0], V[:, 1], V[:, 2]
    nrm = np.cross(b - a, t3 - a)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    q = c - ((c - a) * nrm).sum(1)[:, None] * nrm            # projection onto the plane
    # barycentrics of q; inside -> plane distance, else the closest point on the three edges
    v0, v1, v2 = b - a, t3 - a, q - a
    d00, d01, d11 = (v0 * v0).sum(1), (v0 * v1).sum(1), (v1 * v1).sum(1)
    d20, d21 = (v2 * v0).sum(1), (v2 * v1).sum(1)
    den = d00 * d11 - d01 * d01
    l1, l2 = (d11 * d20 - d01 * d21) / den, (d00 * d21 - d01 * d20) / den
    inside = (l1 >= 0) & (l2 >= 0) & (l1 + l2 <= 1)
    def seg(p0, p1):
        e = p1 - p0
        t = np.clip(((c - p0) * e).sum(1) / (e * e).sum(1), 0, 1)
        return np.linalg.norm(p0 + t[:, None] * e - c, axis=1)
    edge = np.minimum(np.minimum(seg(a, b), seg(b, t3)), seg(t3, a))
    return np.where(inside, np.linalg.norm(q - c, axis=1), edge)


def hcv_sources(el, m, radius_scale, exterior=False):
    """Source positions of the harmonic control variate: a Fibonacci sphere (3D) or equal
    angles on a circle (2D) about the centre c of the vertex box (FP64), of radius
    radius_scale x (largest vertex distance from c) for interior fits, or radius_scale x
    (smallest distance from c to an element) for exterior fits."""
    D = el.dim
    V = el.vertices.reshape(-1, D).astype(np.float64)
    c = 0.5 * (V.min(0) + V.max(0))
    R = radius_scale * (element_distances(el, c).min() if exterior else np.sqrt(((V - c) ** 2).sum(1).max()))
    j = np.arange(m) + 0.5
    if D == 2:
        a = 2 * np.pi * j / m
        return c + R * np.stack([np.cos(a), np.sin(a)], 1)
    z = 1 - 2 * j / m
    r = np.sqrt(np.maximum(0, 1 - z * z))
    ph = np.pi * (1 + np.sqrt(5.0)) * j
    return c + R * np.stack([r * np.cos(ph), r * np.sin(ph), z], 1)


def hcv_rows(rec, src):
    """The fit's rows of BoundaryPoint records [p, n, ia, target, type, alpha] (FP64): the
    datum of each basis function [1, phi_j] of the record's type -- the value (Dirichlet),
    the normal derivative (Neumann) or normal derivative + alpha value (Robin);
    phi = z^(-1/2) (3D) or log2 z (2D) of z = |x - s_j|^2."""
    D = src.shape[1]
    rec = np.asarray(rec, np.float64)
    x, n, ty, al = rec[:, :D], rec[:, D:2 * D], rec[:, 2 * D + 2].astype(int), rec[:, 2 * D + 3]
    d = x[:, None, :] - src[None]
    z = (d ** 2).sum(2)
    dn = (d * n[:, None, :]).sum(2)
    if D == 3:
        val, der = 1 / np.sqrt(z), -dn / z ** 1.5
    else:
        val, der = np.log2(z), 2 / np.log(2) * dn / z
    one = np.ones((len(rec), 1))
    rows_d = np.concatenate([one, val], 1)
    rows_n = np.concatenate([0 * one, der], 1)
    rows_r = np.concatenate([al[:, None], der + al[:, None] * val], 1)
    return np.where((ty == 0)[:, None], rows_d, np.where((ty == 1)[:, None], rows_n, rows_r))


def hcv_eval_f64(pos, src, w, c0):
    """u_h and grad u_h in FP64: (n, dim + 1)."""
    x = np.asarray(pos, np.float64)
    d = x[:, None, :] - src[None]
    z = (d ** 2).sum(2)
    if src.shape[1] == 3:
        u = c0 + (w[None] / np.sqrt(z)).sum(1)
        g = -(w[None, :, None] * d / z[:, :, None] ** 1.5).sum(1)
    else:
        u = c0 + (w[None] * np.log2(z)).sum(1)
        g = (2 / np.log(2)) * (w[None, :, None] * d / z[:, :, None]).sum(1)
    return np.concatenate([u[:, None], g], 1)


def hcv_fit_numpy(rec, src, ridge=1e-10, constant=True, zero_charge=False):
    """Ridge least squares of the records' targets on their rows, unit-RMS columns (FP64):
    returns (c0, w); constant=False (exterior fits) leaves out the constant column (c0 = 0);
    zero_charge=True (with constant=False) keeps sum_j w_j = 0, solving on the columns
    b_j - b_last for all but the last source."""
    B = hcv_rows(rec, src)
    if not constant:
        B = B[:, 1:]
    if zero_charge:
        assert not constant
        B = B[:, :-1] - B[:, -1:]
    t = np.asarray(rec, np.float64)[:, 2 * src.shape[1] + 1]
    ms = (B ** 2).mean(0)
    sc = np.where(ms > 0, 1 / np.sqrt(np.where(ms > 0, ms, 1)), 1.0)   # an all-zero column keeps scale 1
    Bs = B * sc
    A = Bs.T @ Bs
    A += ridge * np.trace(A) / len(A) * np.eye(len(A))
    coef = np.linalg.solve(A, Bs.T @ t) * sc
    if zero_charge:
        coef = np.concatenate([coef, [-coef.sum()]])
    return (coef[0], coef[1:]) if constant else (0.0, coef)
