"""CPU tests of the C ABI: the library loads, exports every symbol include/wob_c.h declares,
maps the reference's enum strings, and rejects bad requests before touching the GPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "wob_c.h")).read()
    return sorted(set(re.findall(r"\b(wob_[a-z0-9_]+)\s*\(", text)))


def test_exports_every_declared_symbol():
    from wobtoolbox_amd import _capi
    L = _capi.lib()
    syms = declared_symbols()
    assert len(syms) >= 14
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    assert sorted(_capi.EXPORTS) == syms


def test_enum_strings_match_reference():
    from wobtoolbox_amd import _capi
    import wobtoolbox_amd as w
    L = _capi.lib()
    for s, v in [("IndirectSingleLayer", 0), ("IndirectDoubleLayer", 1), ("Direct", 2), ("InteriorDomain", 0),
                 ("ExteriorDomain", 1), ("BackwardEstimator", 0), ("ForwardEstimator", 1), ("DirichletProblem", 0),
                 ("NeumannProblem", 1), ("RobinProblem", 2), ("MixedProblem", 3), ("PoissonKernel", 0)]:
        assert L.wob_string_to_enum(s.encode()) == v
    assert L.wob_string_to_enum(b"NoSuchThing") == -1
    assert L.wob_enum_to_string(b"bie_model", 2) == b"Direct"
    assert L.wob_enum_to_string(b"problem", 7) == b"InvalidProblemType"
    # Python mirror: unknown strings map to the *TypeCount value like the reference
    assert w.string_to_bie_model_type("Direct") == 2 and w.string_to_bie_model_type("x") == 3
    assert w.problem_type_to_string(3) == "MixedProblem"


def test_scene_create_rejects_bad_descriptors():
    from wobtoolbox_amd import _capi
    L = _capi.lib()
    h = C.c_void_p()
    assert L.wob_scene_create(None, C.byref(h)) == 1
    d = _capi.SceneDesc(4, 1, None, None, None, None, None)
    assert L.wob_scene_create(C.byref(d), C.byref(h)) == 1
    assert b"dim" in L.wob_last_error_message()
    v = np.zeros(4, np.float32); n = np.zeros(2, np.float32); t = np.array([5], np.int32)
    d = _capi.SceneDesc(2, 1, v.ctypes.data, n.ctypes.data, n.ctypes.data, t.ctypes.data, None)
    assert L.wob_scene_create(C.byref(d), C.byref(h)) == 1
    assert b"boundary type" in L.wob_last_error_message()
    d = _capi.SceneDesc(2, 0, v.ctypes.data, n.ctypes.data, n.ctypes.data, t.ctypes.data, None)
    assert L.wob_scene_create(C.byref(d), C.byref(h)) == 1
    assert not h.value


def test_estimate_rejects_null_scene_and_bad_args():
    from wobtoolbox_amd import _capi
    L = _capi.lib()
    e = _capi.EstimatorDesc(0, 1, 0, 0, 0, 4, 1, 0, 1e-7, 1.0)
    lc = _capi.LaunchConfig(0, 0)
    assert L.wob_estimate_domain(None, C.byref(e), None, 1, 0, 1, 0, 0, 0, None, None, C.byref(lc), None) == 1
    assert L.wob_eval_kernels(4, 1, None, None, None, None, 1e-7, None, None) == 1
    assert L.wob_eval_weights(2, 7, 0, 1e-7, 1.0, 1, None, None, None, None, None, None, None) == 2
    assert L.wob_philox4x32_10(0, None, None, None, None) == 0   # empty request is a no-op


def test_path_counter_limit_rejected_before_any_work():
    """Every path of a point owns a Philox stream whose path word is 32 bits: a call whose
    path range would pass 2^32 is rejected before the scene is even looked at, and the
    largest allowed range goes on to the ordinary checks."""
    from wobtoolbox_amd import _capi
    L = _capi.lib()
    e = _capi.EstimatorDesc(0, 1, 0, 0, 0, 4, 1, 0, 1e-7, 1.0)
    lc = _capi.LaunchConfig(0, 0)
    top = 1 << 32
    for fn in (L.wob_estimate_domain, L.wob_estimate_boundary):
        assert fn(None, C.byref(e), None, 1, 0, 16, 0, top - 8, 0, None, None, C.byref(lc), None) == 1
        assert b"2^32" in L.wob_last_error_message()
        assert fn(None, C.byref(e), None, 1, 0, 16, 0, top, 0, None, None, C.byref(lc), None) == 1
        assert b"2^32" in L.wob_last_error_message()
        # path_base + paths == 2^32 is the last allowed range: the call fails on the null scene
        assert fn(None, C.byref(e), None, 1, 0, 16, 0, top - 16, 0, None, None, C.byref(lc), None) == 1
        assert b"null scene" in L.wob_last_error_message()
    assert L.wob_wos_estimate(None, None, 1, 0, 4, 0, top - 3, 0, 1e-4, 16, None, None, None) == 1
    assert b"2^32" in L.wob_last_error_message()
    assert L.wob_wos_estimate(None, None, 1, 0, 4, 0, top - 4, 0, 1e-4, 16, None, None, None) == 1
    assert b"null scene" in L.wob_last_error_message()
