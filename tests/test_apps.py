"""The reference apps' runner (wobtoolbox_amd/apps/wob_general.cpp): configuration keys, mesh
input, grid masks, analytic fields and output formats of apps/general_2d.cu and
apps/general_3d.cu.  The dry runs need no GPU; the short progressive runs do."""
import json
import os
import subprocess

import numpy as np
import pytest

from wobtoolbox_amd import build as B
from wobtoolbox_amd import geometry as G


def _write_circle_obj(path, r=0.45, n=128):
    th = -2 * np.pi * np.arange(n) / n          # clockwise, as the apps' 2D meshes
    with open(path, "w") as f:
        for t in th:
            f.write(f"v {r * np.cos(t):.9f} {r * np.sin(t):.9f} 0.0\n")
        for i in range(n):
            f.write(f"f {i + 1} {(i + 1) % n + 1}\n")


def _write_sphere_obj(path, level=2):
    V, F = G.icosphere(level)
    V = G.displace_radial(V)
    with open(path, "w") as f:
        for v in V:
            f.write(f"v {v[0]:.9f} {v[1]:.9f} {v[2]:.9f}\n")
        for t in F:
            f.write(f"f {t[0] + 1} {t[1] + 1} {t[2] + 1}\n")


def _config(tmp_path, obj, **kw):
    cfg = {"input_obj_file": str(obj), "output_dir": str(tmp_path / "out"), "path_length": 8,
           "num_sample_paths": 256, "grid_res": 32, "eps": 1e-5, "max_time": 1, "equation_kernel": "PoissonKernel",
           "bie_model": "IndirectDoubleLayer", "domain": "InteriorDomain", "estimator": "BackwardEstimator",
           "problem": "DirichletProblem"}
    cfg.update(kw)
    p = tmp_path / "config.json"
    p.write_text(json.dumps(cfg, indent=4))
    return p


def _read_field(path, comps=1):
    raw = open(path, "rb").read()
    res = int(np.frombuffer(raw[:4], np.uint32)[0])
    return res, np.frombuffer(raw[4:], np.float32).reshape(res, res, comps) if comps > 1 else \
        np.frombuffer(raw[4:], np.float32).reshape(res, res)


def _run(app, cfg, *extra, timeout=300):
    if not os.path.exists(app):
        pytest.fail(f"{app} is not built (run the package build)")
    r = subprocess.run([app, str(cfg), *extra], capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def test_app_2d_dry_run_fields(tmp_path):
    obj = tmp_path / "circle.obj"
    _write_circle_obj(obj)
    cfg = _config(tmp_path, obj)
    out = _run(B.APPS[2], cfg, "--dry-run")
    assert "IndirectDoubleLayer" in out and "DirichletProblem" in out
    od = tmp_path / "out"
    assert json.loads((od / "config.json").read_text())["grid_res"] == 32
    res, u = _read_field(od / "potential_analytical.scalar")
    _, g = _read_field(od / "gradient_analytical.vector2", 2)
    assert res == 32
    x = (np.arange(32) + 0.5) / 32 - 0.5          # utils.cuh:92-99, domain size 1
    X, Y = np.meshgrid(x, x)                      # row = y, column = x
    inside = X ** 2 + Y ** 2 < 0.45 ** 2 * np.cos(np.pi / 128) ** 2 - 1e-6
    outside = X ** 2 + Y ** 2 > 0.45 ** 2 + 1e-6
    assert np.all(np.isfinite(u[inside])) and np.all(np.isnan(u[outside]))
    exact = np.cos(np.pi * X) * np.sinh(np.pi * Y)            # general_2d.cu:86-91
    np.testing.assert_allclose(u[inside], exact[inside], rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(g[inside][:, 1], (np.pi * np.cos(np.pi * X) * np.cosh(np.pi * Y))[inside], rtol=1e-5)


def test_app_3d_dry_run_slice_mask(tmp_path):
    obj = tmp_path / "sphere.obj"
    _write_sphere_obj(obj)
    cfg = _config(tmp_path, obj, bie_model="Direct", problem="NeumannProblem", path_length=6)
    _run(B.APPS[3], cfg, "--dry-run")
    res, u = _read_field(tmp_path / "out" / "potential_analytical.scalar")
    # the mesh is refit to [-1/2, 1/2]^3; the slice x = 0 cuts a disc of radius about 0.4
    frac = np.isfinite(u).mean()
    assert 0.3 < frac < 0.8
    x = (np.arange(res) + 0.5) / res - 0.5
    Z, Yg = np.meshgrid(x, x)                     # column = z, row = y (general_3d.cu:296-297)
    m = np.isfinite(u)
    np.testing.assert_allclose(u[m], (-np.cos(np.pi * Yg) * np.sinh(np.pi * Z))[m], rtol=1e-5, atol=1e-6)


@pytest.mark.gpu
def test_app_2d_progressive_run(tmp_path, gpu):
    obj = tmp_path / "circle.obj"
    _write_circle_obj(obj)
    cfg = _config(tmp_path, obj)
    _run(B.APPS[2], cfg, "--max-iterations", "3")
    od = tmp_path / "out"
    rows = np.loadtxt(od / "rmse_list.csv", delimiter=",", ndmin=2)
    assert rows.shape == (3, 5)
    assert list(rows[:, 0]) == [256, 512, 768]
    print(f"2D app RMSE per pass: {rows[:, 2]}", flush=True)
    # unbiased, noise-limited: RMSE falls like 1 / sqrt(paths) (sqrt(1/3) = 0.58 after three passes)
    assert 0.4 < rows[-1, 2] / rows[0, 2] < 0.8 and rows[-1, 2] < 0.4
    _, u3 = _read_field(od / "potential3.scalar")
    _, ua = _read_field(od / "potential_analytical.scalar")
    assert np.array_equal(np.isnan(u3), np.isnan(ua))


@pytest.mark.gpu
def test_app_3d_progressive_run(tmp_path, gpu):
    """general_3d.cu's progressive loop on the displaced icosphere: sample counts per pass, the
    slice mask of every field, the RMSE row equal to the written field's error, and the error
    falling like 1 / sqrt(paths) (unbiased, noise-limited: sqrt(1/4) = 0.5 after four passes).
    Interior Dirichlet (double layer) is checked on the RMSE itself.  Interior Neumann (Direct)
    shifts every pass by one offset taken, as general_3d.cu:118-133 does, against the analytic
    value at a point off the slice, so there the spread of the error about its mean is checked."""
    obj = tmp_path / "sphere.obj"
    _write_sphere_obj(obj)
    for problem in ("DirichletProblem", "NeumannProblem"):
        bie = "IndirectDoubleLayer" if problem == "DirichletProblem" else "Direct"
        cfg = _config(tmp_path, obj, bie_model=bie, problem=problem, path_length=6, num_sample_paths=128)
        _run(B.APPS[3], cfg, "--max-iterations", "4")
        od = tmp_path / "out"
        rows = np.loadtxt(od / "rmse_list.csv", delimiter=",", ndmin=2)
        assert rows.shape == (4, 6)
        assert list(rows[:, 0]) == [128, 256, 384, 512]
        assert np.all(np.isfinite(rows))
        _, ua = _read_field(od / "potential_analytical.scalar")
        m = np.isfinite(ua)
        err = []
        for it in (1, 4):
            _, u = _read_field(od / f"potential{it}.scalar")
            _, g = _read_field(od / f"gradient{it}.vector3", 3)
            assert g.shape == (32, 32, 3)
            assert np.array_equal(np.isnan(u), np.isnan(ua))
            assert np.all(np.isfinite(g[m])) and np.all(g[~m] == 0)   # gradients outside stay 0
            e = (u[m] - ua[m]).astype(np.float64)
            assert np.sqrt(np.mean(e ** 2)) == pytest.approx(rows[it - 1, 2], rel=1e-3)
            err.append(e)
        spread = [np.sqrt(np.mean(e ** 2)) if problem == "DirichletProblem" else np.std(e) for e in err]
        print(f"3D app {problem}: RMSE u per pass {rows[:, 2]}, error spread pass 1 / 4 {spread}", flush=True)
        assert 0.3 < spread[1] / spread[0] < 0.75


def test_app_harmonic_cv_option_checks_the_solver(tmp_path):
    """--harmonic-cv is refused, before any GPU work, for the solver the control variate does not
    cover (the direct model in an exterior domain)."""
    obj = tmp_path / "sphere.obj"
    _write_sphere_obj(obj)
    for kw in (dict(bie_model="Direct", domain="ExteriorDomain", problem="NeumannProblem"),):
        cfg = _config(tmp_path, obj, **kw)
        r = subprocess.run([B.APPS[3], str(cfg), "--harmonic-cv", "16", "--dry-run"], capture_output=True, text=True,
                           timeout=60)
        assert r.returncode == 2 and "--harmonic-cv needs" in r.stderr, (kw, r.stdout, r.stderr)
    r = subprocess.run([B.APPS[3], str(_config(tmp_path, obj)), "--harmonic-cv", "16", "2.5", "--dry-run"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.gpu
def test_app_harmonic_cv_lowers_rmse(tmp_path, gpu):
    """The 3D app with --harmonic-cv: interior Dirichlet (double layer) and the apps' mixed
    problem (single layer, RIS 16, c = 2) on the displaced icosphere.  The walk estimates the
    residual of a fitted harmonic function, so the RMSE against the apps' analytic solution
    falls far below plain WoB's at the same paths; the seconds column includes the fit."""
    obj = tmp_path / "sphere.obj"
    _write_sphere_obj(obj, level=3)
    for kw in (dict(), dict(bie_model="IndirectSingleLayer", problem="MixedProblem", path_length=4,
                            num_resampling_candidates=16, first_kind_equation_scaling_constant=2.0)):
        cfg = _config(tmp_path, obj, **dict(dict(path_length=6, num_sample_paths=64), **kw))
        rows = {}
        for cv in (False, True):
            extra = ["--harmonic-cv", "64", "2.0"] if cv else []
            out = _run(B.APPS[3], cfg, "--max-iterations", "4", *extra)
            rows[cv] = np.loadtxt(tmp_path / "out" / "rmse_list.csv", delimiter=",", ndmin=2)
            if cv:
                assert "harmonic control variate" in out
        print(f"3D app {kw.get('problem', 'DirichletProblem')}: RMSE u plain {rows[False][:, 2]}, "
              f"with the control variate {rows[True][:, 2]}", flush=True)
        assert list(rows[True][:, 0]) == [64, 128, 192, 256]
        assert rows[True][-1, 2] < 0.1 * rows[False][-1, 2]
        assert rows[True][-1, 2] <= rows[True][0, 2]
        assert rows[True][0, 1] > 0


@pytest.mark.gpu
def test_app_2d_exterior_harmonic_cv(tmp_path, gpu):
    """The 2D app's exterior Neumann problem (single layer, RIS 16) on the circle with
    --harmonic-cv: the sources go inside the circle, and the RMSE against the apps' exterior
    solution x / |x|^2 falls far below plain WoB's."""
    obj = tmp_path / "circle.obj"
    _write_circle_obj(obj)
    cfg = _config(tmp_path, obj, bie_model="IndirectSingleLayer", domain="ExteriorDomain", problem="NeumannProblem",
                  path_length=6, num_sample_paths=64, num_resampling_candidates=16)
    rows = {}
    for cv in (False, True):
        out = _run(B.APPS[2], cfg, "--max-iterations", "3", *(["--harmonic-cv", "32"] if cv else []))
        rows[cv] = np.loadtxt(tmp_path / "out" / "rmse_list.csv", delimiter=",", ndmin=2)
        if cv:
            assert "(exterior)" in out
    print(f"2D app exterior Neumann: RMSE u plain {rows[False][:, 2]}, with the control variate {rows[True][:, 2]}",
          flush=True)
    assert rows[True][-1, 2] < 0.1 * rows[False][-1, 2]
    # exterior Dirichlet with the double layer (the reference's exterior Dirichlet configs):
    # the fit keeps sum_j w_j = 0
    cfg = _config(tmp_path, obj, domain="ExteriorDomain", path_length=6, num_sample_paths=64)
    for cv in (False, True):
        _run(B.APPS[2], cfg, "--max-iterations", "3", *(["--harmonic-cv", "32"] if cv else []))
        rows[cv] = np.loadtxt(tmp_path / "out" / "rmse_list.csv", delimiter=",", ndmin=2)
    print(f"2D app exterior Dirichlet (double layer): RMSE u plain {rows[False][:, 2]}, with the control variate "
          f"{rows[True][:, 2]}", flush=True)
    assert rows[True][-1, 2] < 0.1 * rows[False][-1, 2]
    # interior Neumann with the forward estimator (the reference's interior Neumann configs);
    # the RMSE column compares u after the apps' shift at the centre, so the gradient columns
    # carry the check
    cfg = _config(tmp_path, obj, bie_model="IndirectSingleLayer", estimator="ForwardEstimator",
                  problem="NeumannProblem", path_length=7, num_sample_paths=64)
    for cv in (False, True):
        _run(B.APPS[2], cfg, "--max-iterations", "3", *(["--harmonic-cv", "32"] if cv else []))
        rows[cv] = np.loadtxt(tmp_path / "out" / "rmse_list.csv", delimiter=",", ndmin=2)
    print(f"2D app interior Neumann (forward): RMSE du/dx plain {rows[False][:, 3]}, with the control variate "
          f"{rows[True][:, 3]}", flush=True)
    assert rows[True][-1, 3] < 0.1 * rows[False][-1, 3]


def _vpl_config(tmp_path, **kw):
    cfg = {"output_dir": str(tmp_path / "vpl"), "path_length": 4, "num_sample_paths": 20000,
           "num_resampling_candidates": 64, "grid_res": 64, "eps": 1e-4, "max_time": 1}
    cfg.update(kw)
    p = tmp_path / "vpl.json"
    p.write_text(json.dumps(cfg, indent=4))
    return p


def test_vpl_app_dry_run(tmp_path):
    """apps/potential_flow_with_vpl.cu's scene and mask: 4 + 360 + 180 + 180 + 90 segments and
    the grid points outside the obstacles (the same count as the Python workload's mask)."""
    out = _run(B.VPL_APP, _vpl_config(tmp_path, grid_res=256), "--dry-run")
    assert "elements\t: 814" in out
    n_eval = int(G.potential_flow_vpl_2d(grid_res=256).extra["mask"].sum())
    assert f"evaluation points\t: {n_eval} of 65536" in out
    assert json.load(open(tmp_path / "vpl" / "config.json"))["num_sample_paths"] == 20000


@pytest.mark.gpu
def test_vpl_app_iterations(tmp_path, gpu):
    _run(B.VPL_APP, _vpl_config(tmp_path), "--max-iterations", "3")
    d = tmp_path / "vpl"
    mask = G.potential_flow_vpl_2d(grid_res=64).extra["mask"].reshape(64, 64)
    for it in (1, 2, 3):
        res, u = _read_field(d / f"potential{it}.scalar")
        _, g = _read_field(d / f"gradient{it}.vector2", 2)
        assert res == 64
        assert np.all(np.isnan(u[~mask])) and np.all(np.isfinite(u[mask]))
        assert np.all(g[~mask] == 0) and np.all(np.isfinite(g[mask]))
    # flow from left to right: the mean x-velocity in the free stream is positive, near 1
    c = G.grid_points(64).reshape(64, 64, 2)
    free = mask & (np.abs(c[..., 1]) < 0.45) & (c[..., 0] < -0.4)
    assert 0.5 < g[free][:, 0].mean() < 2.0


def _wos_config(tmp_path, obj, **kw):
    cfg = {"input_obj_file": str(obj), "output_dir": str(tmp_path / "wos"), "path_length": 4,
           "num_sample_paths": 64, "is_wos": True, "max_time": 1, "grid_res": 32, "eps": 1e-4}
    cfg.update(kw)
    p = tmp_path / "wos.json"
    p.write_text(json.dumps(cfg, indent=4))
    return p


def test_wos_app_dry_run(tmp_path):
    """apps/simple_wob_and_wos.cu's mask: the grid points inside the (clockwise) polyline."""
    obj = tmp_path / "circle.obj"
    _write_circle_obj(obj)
    out = _run(B.WOS_APP, _wos_config(tmp_path, obj), "--dry-run")
    assert "#line segments: 128" in out and "is WoS\t: 1" in out
    x = (np.arange(32) + 0.5) / 32 - 0.5
    X, Y = np.meshgrid(x, x)
    inside = (X ** 2 + Y ** 2 < 0.45 ** 2 * np.cos(np.pi / 128) ** 2 - 1e-6).sum()
    edge = ((X ** 2 + Y ** 2 >= 0.45 ** 2 * np.cos(np.pi / 128) ** 2 - 1e-6) & (X ** 2 + Y ** 2 <= 0.45 ** 2 + 1e-6)).sum()
    n = int(out.split("evaluation points\t: ")[1].split(" ")[0])
    assert inside <= n <= inside + edge
    out = _run(B.WOS_APP, _wos_config(tmp_path, obj, evaluate_at_all_points=True), "--dry-run")
    assert "evaluation points\t: 1024 of 1024" in out


@pytest.mark.gpu
def test_wos_app_both_modes_agree(tmp_path, gpu):
    obj = tmp_path / "circle.obj"
    _write_circle_obj(obj)
    _run(B.WOS_APP, _wos_config(tmp_path, obj, num_sample_paths=512), "--max-iterations", "2")
    _, u_wos = _read_field(tmp_path / "wos" / "potential2.scalar")
    (tmp_path / "wob").mkdir()
    _run(B.WOS_APP, _wos_config(tmp_path, obj, is_wos=False, output_dir=str(tmp_path / "wob"),
                                num_sample_paths=512), "--max-iterations", "2")
    _, u_wob = _read_field(tmp_path / "wob" / "potential2.scalar")
    m = np.isfinite(u_wos)
    assert np.array_equal(m, np.isfinite(u_wob)) and m.sum() > 500
    # +-1 data in 40 arc-length pieces: walk-on-spheres estimates lie in [-1, 1]; the two
    # modes estimate the same field (their agreement within standard errors is tested in
    # test_gpu_wos.py; here 1024 paths per point leave noise of about 0.1 in the WoB field)
    assert np.all(np.abs(u_wos[m]) <= 1.0 + 1e-6)
    assert np.corrcoef(u_wos[m], u_wob[m])[0, 1] > 0.8
    assert np.mean(np.abs(u_wos[m] - u_wob[m])) < 0.15


# ---------------------------------------------------------------------------------------
# apps/poisson_equation.cu and apps/boundary_estimate.cu
# ---------------------------------------------------------------------------------------
def _json(tmp_path, name, **cfg):
    p = tmp_path / name
    p.write_text(json.dumps(cfg, indent=4))
    return p


def square_poisson_series(x, y, source=5.0, data=1.0, terms=4001):
    """u with Laplacian(u) = source in (-1, 1)^2 and u = data on the boundary (FP64):
    u = data + source w, w = (x^2 - 1) / 2 + sum_{k odd} b_k cos(k pi x / 2) cosh(k pi y / 2) /
    cosh(k pi / 2), b_k = 2 (-1)^((k-1)/2) / (k pi / 2)^3 (the cosine series of (1 - x^2) / 2,
    which w must cancel on the sides y = +-1); the cosh ratio is evaluated as
    exp(a (|y| - 1)) (1 + exp(-2 a |y|)) / (1 + exp(-2 a))."""
    x, y = np.asarray(x, np.float64), np.abs(np.asarray(y, np.float64))
    w = (x * x - 1) / 2
    for k in range(1, terms + 1, 2):
        a = k * np.pi / 2
        ratio = np.exp(a * (y - 1)) * (1 + np.exp(-2 * a * y)) / (1 + np.exp(-2 * a))
        w = w + 2 * (-1) ** ((k - 1) // 2) / a ** 3 * np.cos(a * x) * ratio
    return data + source * w


def test_square_poisson_series_is_the_solution():
    """The series against its defining properties: boundary values, and the five-point
    Laplacian at interior points."""
    t = np.linspace(-1, 1, 41)
    for bx, by in ((t, np.ones_like(t)), (t, -np.ones_like(t)), (np.ones_like(t), t), (-np.ones_like(t), t)):
        assert np.abs(square_poisson_series(bx, by) - 1).max() < 1e-6
    rng = np.random.default_rng(0)
    x, y = rng.uniform(-0.8, 0.8, (2, 50))
    h = 1e-3
    lap = (square_poisson_series(x + h, y) + square_poisson_series(x - h, y) + square_poisson_series(x, y + h)
           + square_poisson_series(x, y - h) - 4 * square_poisson_series(x, y)) / (h * h)
    assert np.abs(lap - 5).max() < 1e-3


def test_poisson_app_dry_run(tmp_path):
    cfg = _json(tmp_path, "poisson.json", output_dir=str(tmp_path / "p"), path_length=4, num_sample_paths=16,
                num_volume_samples=16, grid_res=24, eps=1e-4, max_time=1)
    out = _run(B.POISSON_APP, cfg, "--dry-run")
    assert "evaluation points\t: 576" in out and "#num volume samples\t: 16" in out
    assert json.loads((tmp_path / "p" / "config.json").read_text())["grid_res"] == 24


def _boundary_cfg(tmp_path, **kw):
    cfg = dict(output_dir=str(tmp_path / "b"), path_length=4, num_sample_paths_domain=64,
               num_sample_paths_boundary=256, grid_res=32, max_time=1)
    cfg.update(kw)
    return _json(tmp_path, "boundary.json", **cfg)


def test_boundary_estimate_app_dry_run(tmp_path):
    out = _run(B.BOUNDARY_APP, _boundary_cfg(tmp_path), "--dry-run")
    c = (np.arange(32) + 0.5) / 32 - 0.5
    x, y = np.meshgrid(c.astype(np.float32), c.astype(np.float32))
    inside = int((np.sqrt(x * x + y * y) <= 0.5).sum())
    assert f"evaluation points\t: {inside} of 1024" in out


@pytest.mark.gpu
def test_poisson_app_matches_the_series_solution(tmp_path, gpu):
    """poisson_equation.cu's setting (square, data 1, source 5 on the grid cache, double layer,
    path length 4, 16 volume samples) against the series solution: the error is noise (its
    mean within a few standard errors of 0, RMSE falling like 1 / sqrt(paths)), and the
    gradient field points along the series' gradient."""
    res = 16
    cfg = _json(tmp_path, "poisson.json", output_dir=str(tmp_path / "p"), path_length=4, num_sample_paths=1024,
                num_volume_samples=16, grid_res=res, eps=1e-4, max_time=1)
    _run(B.POISSON_APP, cfg, "--max-iterations", "4")
    c = ((np.arange(res, dtype=np.float32) + np.float32(0.5)) * np.float32(2.0 / res) - np.float32(1.0))
    x, y = np.meshgrid(c.astype(np.float64), c.astype(np.float64))
    ex = square_poisson_series(x, y)
    errs = []
    for it in (1, 4):
        r, u = _read_field(tmp_path / "p" / f"potential{it}.scalar")
        _, g = _read_field(tmp_path / "p" / f"gradient{it}.vector2", 2)
        assert r == res and np.all(np.isfinite(u)) and np.all(np.isfinite(g))
        errs.append((u - ex).ravel())
    e1, e4 = errs
    rmse1, rmse4 = np.sqrt(np.mean(e1 ** 2)), np.sqrt(np.mean(e4 ** 2))
    print(f"Poisson app: RMSE vs series after 1024 / 4096 paths {rmse1:.4f} / {rmse4:.4f}, "
          f"mean error {e4.mean():+.5f}", flush=True)
    # per-path standard deviation about 7 (the oracle's on this scene): SE 0.11 at 4096 paths
    assert abs(e4.mean()) < 4 * rmse4 / np.sqrt(e4.size)
    assert 0.35 < rmse4 / rmse1 < 0.7 and rmse4 < 0.16
    h = 1e-5
    gx = (square_poisson_series(x + h, y) - square_poisson_series(x - h, y)) / (2 * h)
    gy = (square_poisson_series(x, y + h) - square_poisson_series(x, y - h)) / (2 * h)
    _, g = _read_field(tmp_path / "p" / "gradient4.vector2", 2)
    m = np.hypot(gx, gy) > 2.0
    cosang = (g[..., 0] * gx + g[..., 1] * gy)[m] / (np.hypot(g[..., 0], g[..., 1])[m] * np.hypot(gx, gy)[m])
    print(f"Poisson app: median cosine between estimated and series gradients {np.median(cosang):.3f} "
          f"over {m.sum()} points", flush=True)
    assert np.median(cosang) > 0.8


@pytest.mark.gpu
def test_boundary_estimate_app_matches_the_library_and_its_symmetry(tmp_path, gpu):
    """boundary_estimate.cu's scene and estimators: the app's fields equal the library's
    estimates of the same scene built from the reference's recipe in numpy (same Philox
    streams), and the boundary solution shares the data's symmetries within its noise: the
    data repeat every 240 segments and change sign under a shift of 120, so the boundary
    solution does too, up to the additive constant a Neumann solution leaves free."""
    import torch
    import wobtoolbox_amd as wob
    res, pd, pb = 32, 64, 256
    cfg = _boundary_cfg(tmp_path, grid_res=res, num_sample_paths_domain=pd, num_sample_paths_boundary=pb)
    _run(B.BOUNDARY_APP, cfg, "--max-iterations", "2")
    od = tmp_path / "b"
    # the reference's scene (boundary_estimate.cu:153-172): vertices in double rounded to FP32
    n_seg = 720
    i = np.arange(n_seg + 1)
    v = np.stack([np.float32(0.5) * np.sin(2 * np.pi * i / n_seg), np.float32(0.5) * np.cos(2 * np.pi * i / n_seg)],
                 1).astype(np.float32)
    v1, v2 = v[:-1], v[1:]
    t = v1 - v2
    t = t / np.sqrt(t[:, 0] * t[:, 0] + t[:, 1] * t[:, 1])[:, None]
    nrm = np.stack([t[:, 1], -t[:, 0]], 1).astype(np.float32)
    k = np.arange(n_seg) % 240
    a_val = np.where((k == 0) | (k == 239), 1.0, np.where((k == 120) | (k == 119), -1.0, 0.0))
    b_val = 200.0 * a_val
    el = G.Elements(2, np.stack([v1, v2], 1).astype(np.float64), -nrm.astype(np.float64),
                    np.stack([a_val, b_val], 1), np.full(n_seg, G.NEUMANN, np.int32),
                    np.zeros((n_seg, 2))).astype_contiguous()
    scene = wob.SceneHost(el)
    solver = dict(bie_model="IndirectSingleLayer", domain="InteriorDomain", estimator="ForwardEstimator",
                  problem="NeumannProblem", path_length=4, eps=float(np.finfo(np.float32).eps))
    est = wob.Estimator.from_config(scene, solver)
    c = (np.arange(res, dtype=np.float32) + np.float32(0.5)) * np.float32(1.0 / res) - np.float32(0.5)
    gx, gy = np.meshgrid(c, c)
    inside = np.sqrt(gx * gx + gy * gy) <= np.float32(0.5)
    pts = np.stack([gx[inside], gy[inside]], 1).astype(np.float32)
    s = None
    for it in range(2):
        s, _ = est.compute_domain(torch.as_tensor(pts).to(gpu), pd, seed=0, path_base=it * pd,
                                  out=None if s is None else (s, _))
    u_lib = (s[:, 0].cpu().numpy() / (2 * pd)).astype(np.float32)
    _, u_app = _read_field(od / "potential2.scalar")
    assert np.array_equal(np.isnan(u_app), ~inside)
    assert np.array_equal(u_app[inside], u_lib)
    rec = np.zeros((n_seg, 8), np.float32)
    rec[:, :2] = (v1 + v2) / np.float32(2.0)
    rec[:, 2:4] = nrm
    rec[:, 4] = 0.5
    rec[:, 6] = G.NEUMANN
    sb = qb = None
    for it in range(2):
        sb, qb = est.compute_boundary(torch.as_tensor(rec).to(gpu), pb, seed=0, path_base=it * pb, stream_id=1,
                                      out=None if sb is None else (sb, qb))
    b_lib = (sb.cpu().numpy().ravel() / (2 * pb)).astype(np.float32)
    raw = open(od / "boundary_solution2.scalar", "rb").read()
    assert int(np.frombuffer(raw[:4], np.uint32)[0]) == n_seg
    b_app = np.frombuffer(raw[4:], np.float32)
    assert np.array_equal(b_app, b_lib)
    m = sb.cpu().numpy().ravel() / (2 * pb)
    se = np.sqrt(np.maximum(qb.cpu().numpy().ravel() / (2 * pb) - m * m, 0) / (2 * pb))
    d3 = m - np.roll(m, 240)
    z3 = (d3 - d3.mean()) / np.sqrt(se ** 2 + np.roll(se, 240) ** 2)
    d6 = m + np.roll(m, 120)
    z6 = (d6 - d6.mean()) / np.sqrt(se ** 2 + np.roll(se, 120) ** 2)
    print(f"boundary app: max |z| of the 240-shift difference {np.abs(z3).max():.2f}, of the 120-shift sum "
          f"{np.abs(z6).max():.2f}; boundary solution range {m.min():.3f} .. {m.max():.3f}", flush=True)
    assert np.abs(z3).max() < 5.0 and np.abs(z6).max() < 5.0
