"""GPU: the allocation-free domain estimate (wob_estimate_domain_noalloc) through the C ABI.

Its results must be those of wob_estimate_domain, bit for bit (same kernels, the partial sums
in the library's static device array instead of a per-scene buffer), on two calls in a row,
the second with a smaller input; it must leave device memory untouched on a fresh scene and
reject a launch whose partial sums do not fit the array."""
import ctypes as C
import sys

import numpy as np
import pytest

from wobtoolbox_amd import geometry as G

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _call(fn, est, scene, pts, paths, seed, path_base, lc=None):
    import torch
    from wobtoolbox_amd import _capi
    n, V = pts.shape[0], scene.dim + 1
    s = torch.zeros((n, V), dtype=torch.float64, device=pts.device)
    q = torch.zeros_like(s)
    d = est.desc()
    lc = lc if lc is not None else _capi.LaunchConfig(0, 0)
    st = fn(scene.handle, C.byref(d), C.c_void_p(pts.data_ptr()), n, 0, paths, seed, path_base, 0,
            C.c_void_p(s.data_ptr()), C.c_void_p(q.data_ptr()), C.byref(lc),
            C.c_void_p(torch.cuda.current_stream().cuda_stream))
    return st, s, q


@pytest.mark.parametrize("case", ["icosphere_direct_neumann", "circle_dl", "sphere_sl_ris"])
def test_noalloc_equals_estimate_domain(case, gpu):
    import torch
    import wobtoolbox_amd as w
    from wobtoolbox_amd import _capi
    L = _capi.lib()
    if case == "icosphere_direct_neumann":
        x = G.icosphere_neumann_3d(n_points=20000, level=4)
        solver, paths = x.solver, 64
    elif case == "circle_dl":
        x = G.circle_dirichlet_2d(grid_stride=3)
        solver, paths = x.solver, 128
    else:
        x = G.geodesic_mixed_3d(n_points=8000, freq=16)
        solver, paths = x.solver, 16
    scene = w.SceneHost(x.elements)
    est = w.Estimator.from_config(scene, solver)
    pts = torch.as_tensor(x.points).to(gpu)
    # two calls in a row, the second on a smaller point set and with other streams
    for k, (sub, seed, pbase) in enumerate([(pts, 3, 0), (pts[: pts.shape[0] // 3 + 7].contiguous(), 8, paths)]):
        st_a, sa, qa = _call(L.wob_estimate_domain_noalloc, est, scene, sub, paths, seed, pbase)
        assert st_a == 0, L.wob_last_error_message()
        st_b, sb, qb = _call(L.wob_estimate_domain, est, scene, sub, paths, seed, pbase)
        assert st_b == 0, L.wob_last_error_message()
        torch.cuda.synchronize()
        assert torch.equal(sa, sb) and torch.equal(qa, qb), f"{case} call {k}"
        _say(f"{case} call {k}: {sub.shape[0]} points x {paths} paths bit-identical to wob_estimate_domain")


def test_noalloc_allocates_nothing_and_rejects_oversize(gpu):
    import torch
    import wobtoolbox_amd as w
    from wobtoolbox_amd import _capi
    L = _capi.lib()
    x = G.icosphere_neumann_3d(n_points=100000, level=3)
    pts = torch.as_tensor(x.points).to(gpu)
    warm = w.SceneHost(x.elements)
    st, _, _ = _call(L.wob_estimate_domain_noalloc, w.Estimator.from_config(warm, x.solver), warm, pts[:1000], 8, 1, 0)
    assert st == 0
    torch.cuda.synchronize()
    # a fresh scene: the allocating entry point would allocate its partial sums and counter
    scene = w.SceneHost(x.elements)
    est = w.Estimator.from_config(scene, x.solver)
    s = torch.zeros((pts.shape[0], 4), dtype=torch.float64, device=gpu)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    st, _, _ = _call(L.wob_estimate_domain_noalloc, est, scene, pts, 64, 2, 0)
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    assert st == 0
    _say(f"free device memory before / after: {free0} / {free1}")
    assert free1 >= free0
    # one lane per path: 100k points x 64 paths = 6.4M work items x 32 bytes > 64 MiB
    lc = _capi.LaunchConfig(1, 0)
    st, _, _ = _call(L.wob_estimate_domain_noalloc, est, scene, pts, 64, 2, 0, lc)
    msg = L.wob_last_error_message().decode()
    _say(f"oversize request: status {st}: {msg}")
    assert st == 1 and "allocation-free scratch" in msg
    lc = _capi.LaunchConfig(0, 0, _capi.LAUNCH_TIME_WALK)
    st, _, _ = _call(L.wob_estimate_domain_noalloc, est, scene, pts[:100], 8, 2, 0, lc)
    assert st == 1 and b"TIME_WALK" in L.wob_last_error_message()
