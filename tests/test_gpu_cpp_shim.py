"""The C++ host interface (include/wob/wob.hpp) against the Python path, through the C ABI.

tests/cpp/shim_check.cpp builds a SceneHost from ElementXs records, picks the Estimator
instantiation at run time with wob::run (the reference's utils::run, runtime_utils.cuh:106-153)
and accumulates path sums; with the same seed and path numbering the sums must be
bit-identical to those of the Python mirror.
"""
import os
import struct
import subprocess

import numpy as np
import pytest

import helpers
from wobtoolbox_amd import build as B
from wobtoolbox_amd import geometry as G

ENUM = {"IndirectSingleLayer": 0, "IndirectDoubleLayer": 1, "Direct": 2, "InteriorDomain": 0, "ExteriorDomain": 1,
        "BackwardEstimator": 0, "ForwardEstimator": 1, "DirichletProblem": 0, "NeumannProblem": 1,
        "RobinProblem": 2, "MixedProblem": 3}


def test_header_compiles_without_hip_headers():
    """wob.hpp needs only wob_c.h and the C++ standard library."""
    import shutil
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = '#include "wob/wob.hpp"\nint main(){ return wob::string_to_bie_model_type("Direct") == wob::Direct ? 0 : 1; }\n'
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-I", os.path.join(root, "include"),
                        "-x", "c++", "-"], input=src, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def _write_input(path, w, points, paths, seed):
    el = w.elements.astype_contiguous()
    s = w.solver
    hdr = struct.pack("<10I2fQ", el.dim, el.n, len(points), paths, s["path_length"], ENUM[s["bie_model"]],
                      ENUM[s["domain"]], ENUM[s["estimator"]], ENUM[s["problem"]],
                      int(s.get("num_resampling_candidates", 1)), float(s.get("eps", np.finfo(np.float32).eps)),
                      float(s.get("first_kind_equation_scaling_constant", 1.0)), seed)
    with open(path, "wb") as f:
        f.write(hdr)
        for a in (el.vertices, el.normals, el.boundary_values):
            f.write(np.ascontiguousarray(a, np.float32).tobytes())
        f.write(np.ascontiguousarray(el.boundary_types, np.int32).tobytes())
        f.write(np.ascontiguousarray(el.robin_alphas, np.float32).tobytes())
        f.write(np.ascontiguousarray(points, np.float32).tobytes())


@pytest.mark.gpu
@pytest.mark.parametrize("case,device_bvh", [("circle_dl", False), ("star_sl_ris", False),
                                             ("icosphere_direct", False), ("icosphere_direct", True),
                                             ("star_sl_ris", True)])
def test_cpp_shim_matches_python_path(case, device_bvh, gpu, tmp_path):
    import torch
    import wobtoolbox_amd as wob
    if not os.path.exists(B.SHIM_CHECK):
        pytest.fail("wob_shim_check is not built (run the package build)")
    if case == "circle_dl":
        w = G.circle_dirichlet_2d(grid_stride=97)
    elif case == "star_sl_ris":
        w = helpers.star_scene(seed=3, n_segments=300, kind=1)
    else:
        w = G.icosphere_neumann_3d(n_points=500, level=3)
    pts = np.ascontiguousarray(w.points[:500], np.float32)
    paths, seed = 64, 1234
    _write_input(tmp_path / "in.bin", w, pts, paths, seed)
    print(f"shim case {case} (device BVH: {device_bvh}): {len(pts)} points", flush=True)
    args = [B.SHIM_CHECK, str(tmp_path / "in.bin"), str(tmp_path / "out.bin")] + (["device"] if device_bvh else [])
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    V = w.elements.dim + 1
    m = len(pts) * V
    raw = (tmp_path / "out.bin").read_bytes()
    acc = np.frombuffer(raw[:32 * m], np.float64).reshape(4, len(pts), V)
    # the Python side always uses the host-built tree: results do not depend on the tree
    scene = wob.SceneHost(w.elements)
    est = wob.Estimator.from_config(scene, w.solver)
    s, q = est.compute_domain(torch.as_tensor(pts).to(gpu), paths, seed=seed)
    s, q = s.cpu().numpy(), q.cpu().numpy()
    assert np.array_equal(acc[0], s) and np.array_equal(acc[1], q), "templated Estimator differs from the Python path"
    assert np.array_equal(acc[2], s) and np.array_equal(acc[3], q), "run_domain differs from the Python path"
    nodes = struct.unpack("<I", raw[32 * m:32 * m + 4])[0]
    assert nodes == wob.SceneHost(w.elements, device_bvh=device_bvh).info()["n_nodes"]
