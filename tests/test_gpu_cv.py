"""GPU: BF16 MFMA MLP kernels against the FP64 oracle, and the control-variate walk."""
import os
import sys

import numpy as np
import pytest

from helpers import mean_se, small_icosphere
from wobtoolbox_amd import geometry as G

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as M  # noqa: E402

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _dev(a, gpu):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a)).to(gpu)


def _net(seed=3, dim=3, nout=4, sigma=1.0):
    from wobtoolbox_amd.cv import ControlVariate
    net = ControlVariate(dim, nout, sigma=sigma, seed=seed)
    p, f = net.params()
    rng = np.random.default_rng(seed)
    p = p + rng.normal(0, 0.02, p.shape).astype(np.float32)    # non-zero biases too
    net.set_params(p, f)
    return net


def test_mlp_forward_matches_oracle(gpu):
    net = _net()
    p, f = net.params()
    pos = np.random.default_rng(1).uniform(-1.2, 1.2, (1000, 3)).astype(np.float32)
    out = net.infer(_dev(pos, gpu)).cpu().numpy()
    ref, _ = M.forward(p, f, pos, bf=True)
    ref = ref[:, :4]
    err = np.abs(out - ref).max() / np.sqrt((ref ** 2).mean())
    _say(f"MLP forward: max |gpu - oracle| / rms = {err:.2e}")
    assert err < 2e-2


def test_mlp_gradient_matches_oracle(gpu):
    net = _net()
    p, f = net.params()
    rng = np.random.default_rng(2)
    n = 640
    pos = rng.uniform(-1.2, 1.2, (n, 3)).astype(np.float32)
    idx = rng.integers(0, 4, n).astype(np.int32)
    tgt = rng.normal(size=n).astype(np.float32)
    g, loss = net.gradient(_dev(pos, gpu), _dev(idx, gpu), _dev(tgt, gpu))
    lref, gref = M.loss_and_grad(p, f, pos, idx, tgt, bf=True)
    assert abs(loss - lref) <= 1e-2 * lref
    off = 0
    for l, (o, i) in enumerate(M.layer_shapes()):
        for name, sl in (("W", slice(off, off + o * i)), ("b", slice(off + o * i, off + o * i + o))):
            a, b = g[sl].astype(np.float64), gref[sl]
            if name == "W" and l == 4:
                a, b = a.reshape(16, 128)[:4], b.reshape(16, 128)[:4]   # outputs 4..15 carry no loss
            if name == "b" and l == 4:
                a, b = a[:4], b[:4]
            rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)
            cos = float(a.ravel() @ b.ravel() / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-30))
            _say(f"layer {l} {name}: relative gradient error {rel:.2e}, cosine {cos:.6f}")
            assert rel < 3e-2 and cos > 0.9995
        off += o * i + o


def test_mlp_training_fits_smooth_function(gpu):
    import torch
    net = _net(seed=5)
    rng = np.random.default_rng(5)
    n = 65536
    pos = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    idx = rng.integers(0, 4, n).astype(np.int32)
    tgt = (np.sin(2 * pos[:, 0]) * np.cos(pos[:, 1]) + 0.3 * idx * pos[:, 2]).astype(np.float32)
    losses = net.train(_dev(pos, gpu), _dev(idx, gpu), _dev(tgt, gpu), batch=8192, steps=600, lr=3e-3,
                       record_loss=True)
    torch.cuda.synchronize()
    _say(f"MLP training: loss {losses[0]:.4f} -> {losses[-1]:.6f}")
    assert losses[-1] < 0.02 * losses[0]


def _cv_setup(gpu):
    import wobtoolbox_amd as w
    el, V, F = small_icosphere(3, kind="mixed")
    scene = w.SceneHost(el)
    solver = dict(bie_model="IndirectSingleLayer", domain="InteriorDomain", estimator="BackwardEstimator",
                  problem="MixedProblem", path_length=4, num_resampling_candidates=32, eps=1e-5,
                  first_kind_equation_scaling_constant=2.0)
    pts = G.interior_points_star(96, np.random.default_rng(3), margin=0.1).astype(np.float32)
    return w, el, scene, solver, _dev(pts, gpu), pts


def test_control_variate_zero_table_is_plain_walk(gpu):
    import torch
    from wobtoolbox_amd.cv import bind_table
    w, el, scene, solver, p, _ = _cv_setup(gpu)
    bind_table(scene, torch.zeros((4, el.n, 3), dtype=torch.float32, device=gpu), 4)
    plain = w.Estimator.from_config(scene, solver)
    cv = w.Estimator.from_config(scene, solver, flags=2)
    s0, q0 = plain.compute_domain(p, 64, seed=4)
    s1, q1 = cv.compute_domain(p, 64, seed=4)
    assert torch.equal(s0, s1) and torch.equal(q0, q1)


def test_control_variate_any_table_keeps_expectation(gpu):
    import torch
    from wobtoolbox_amd.cv import bind_table
    w, el, scene, solver, p, _ = _cv_setup(gpu)
    g = torch.Generator(device="cpu").manual_seed(0)
    bind_table(scene, (torch.rand((4, el.n, 3), generator=g) * 4 - 2).to(gpu), 4)
    n = 4096
    s0, q0 = w.Estimator.from_config(scene, solver).compute_domain(p, n, seed=6)
    s1, q1 = w.Estimator.from_config(scene, solver, flags=2).compute_domain(p, n, seed=7)
    m0, se0 = mean_se(s0.cpu().numpy(), q0.cpu().numpy(), n)
    m1, se1 = mean_se(s1.cpu().numpy(), q1.cpu().numpy(), n)
    z = (m1 - m0) / np.sqrt(se0 ** 2 + se1 ** 2)
    _say(f"random table: max |z| vs plain {np.abs(z).max():.2f}, mean z {z.mean(0)}")
    assert np.abs(z).max() < 5.0
    assert np.all(np.abs(z.mean(0)) < 5 / np.sqrt(len(z)))


def test_control_variate_trained_reduces_variance(gpu):
    import torch
    from wobtoolbox_amd.cv import ControlVariate, training_data
    w, el, scene, solver, p, pts = _cv_setup(gpu)
    plain = w.Estimator.from_config(scene, solver)
    pos, idx, tgt = training_data(plain, 65536, seed=1)
    net = ControlVariate(3, 4, sigma=1.0, seed=1)
    losses = net.train(pos, idx, tgt, batch=8192, steps=800, lr=2e-3, record_loss=True)
    net.bind(scene, 4)
    n = 2048
    s0, q0 = plain.compute_domain(p, n, seed=8)
    s1, q1 = w.Estimator.from_config(scene, solver, flags=2).compute_domain(p, n, seed=9)
    torch.cuda.synchronize()
    m0, se0 = mean_se(s0.cpu().numpy(), q0.cpu().numpy(), n)
    m1, se1 = mean_se(s1.cpu().numpy(), q1.cpu().numpy(), n)
    z = (m1 - m0) / np.sqrt(se0 ** 2 + se1 ** 2)
    ratio = float(np.mean(se0[:, 0] ** 2) / np.mean(se1[:, 0] ** 2))
    ex = G.exact_expcos_xz_3d(pts)
    _say(f"trained CV: loss {losses[0]:.3f} -> {losses[-1]:.3f}; variance ratio plain / CV for u {ratio:.2f}; "
         f"max |z| {np.abs(z).max():.2f}; rmse plain {np.sqrt(((m0[:, 0] - ex[:, 0]) ** 2).mean()):.4f} "
         f"CV {np.sqrt(((m1[:, 0] - ex[:, 0]) ** 2).mean()):.4f}")
    assert np.abs(z).max() < 5.0
    assert ratio > 2.0
