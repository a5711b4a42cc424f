"""Build libwob_hip.so (gfx950 HIP kernels + host scene builder + C ABI) in-tree.

Every HIP translation unit is compiled with packed-FP32 instruction selection disabled
(-Xclang -target-feature -Xclang -packed-fp32-ops); the host scene builder is compiled
with FP contraction off so its FP32 areas and Kahan sums keep their written order.
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "libwob_hip.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = "gfx950"

HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-ffp-contract=off"]

SOURCES = [
    ("wob_walk.hip", "hip", ["-DWOB_WALK_DIM=2", "-DWOB_WALK_VOL=0"], "wob_walk_20.o"),
    ("wob_walk.hip", "hip", ["-DWOB_WALK_DIM=2", "-DWOB_WALK_VOL=1"], "wob_walk_21.o"),
    ("wob_walk.hip", "hip", ["-DWOB_WALK_DIM=3", "-DWOB_WALK_VOL=0"], "wob_walk_30.o"),
    ("wob_walk.hip", "hip", ["-DWOB_WALK_DIM=3", "-DWOB_WALK_VOL=1"], "wob_walk_31.o"),
    ("wob_kernels.hip", "hip"),
    ("wob_vpl.hip", "hip"),
    ("wob_wos.hip", "hip"),
    ("wob_bvh.hip", "hip"),
    ("mlp_kernels.hip", "hip"),
    ("wob_capi.cpp", "hip"),
    ("wob_cv.cpp", "hip"),
    ("wob_hcv.cpp", "hip"),
    ("scene_build.cpp", "host"),
]


def _cmd(src, kind, out, defines=()):
    path = os.path.join(CSRC, src)
    if kind == "hip":
        return [HIPCC, *HIP_FLAGS, *defines, "-x", "hip", "-c", path, "-o", out]
    if kind == "hip-host":
        return [HIPCC, "-O2", "-std=c++17", "-fPIC", "-Wall", "-D__HIP_PLATFORM_AMD__", "-c", "-x", "c++", path, "-o", out]
    return ["g++", *HOST_FLAGS, "-c", path, "-o", out]


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    procs = []
    objs = []
    for entry in SOURCES:
        src, kind = entry[0], entry[1]
        defines = entry[2] if len(entry) > 2 else ()
        obj = entry[3] if len(entry) > 3 else os.path.splitext(src)[0] + ".o"
        out = os.path.join(OBJ, obj)
        objs.append(out)
        cmd = _cmd(src, kind, out, defines)
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    errors = []
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            errors.append(f"--- {src} ---\n{out.decode(errors='replace')}")
        elif verbose and out:
            print(out.decode(errors="replace"))
    if errors:
        raise RuntimeError("libwob_hip build failed:\n" + "\n".join(errors))
    tmp = LIB + ".tmp"
    link = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(link, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("libwob_hip link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    _build_shim_check()
    _build_apps()
    return LIB


SHIM_CHECK = os.path.join(LIBDIR, "wob_shim_check")


def _build_shim_check():
    """The C++ host-interface check program (tests/cpp/shim_check.cpp), linked to the library."""
    src = os.path.join(os.path.dirname(PKG), "tests", "cpp", "shim_check.cpp")
    if not os.path.exists(src):
        return
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROCM, "include"), src,
           "-o", SHIM_CHECK, "-L" + LIBDIR, "-lwob_hip", "-L" + os.path.join(ROCM, "lib"), "-lamdhip64",
           "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + os.path.join(ROCM, "lib")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("shim_check build failed:\n" + r.stdout + r.stderr)


APPS = {2: os.path.join(LIBDIR, "wob_general_2d"), 3: os.path.join(LIBDIR, "wob_general_3d")}
VPL_APP = os.path.join(LIBDIR, "wob_vpl")
WOS_APP = os.path.join(LIBDIR, "wob_wos")
POISSON_APP = os.path.join(LIBDIR, "wob_poisson")
BOUNDARY_APP = os.path.join(LIBDIR, "wob_boundary_estimate")


def _build_apps():
    """The reference apps' runners: apps/wob_general.cpp for 2D and 3D, apps/wob_vpl.cpp,
    apps/wob_wos.cpp, apps/wob_poisson.cpp and apps/wob_boundary_estimate.cpp."""
    jobs = [(os.path.join(PKG, "apps", "wob_general.cpp"), out, [f"-DWOB_APP_DIM={dim}"]) for dim, out in APPS.items()]
    jobs.append((os.path.join(PKG, "apps", "wob_vpl.cpp"), VPL_APP, []))
    jobs.append((os.path.join(PKG, "apps", "wob_wos.cpp"), WOS_APP, []))
    jobs.append((os.path.join(PKG, "apps", "wob_poisson.cpp"), POISSON_APP, []))
    jobs.append((os.path.join(PKG, "apps", "wob_boundary_estimate.cpp"), BOUNDARY_APP, []))
    procs = []
    for src, out, defs in jobs:
        cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-D__HIP_PLATFORM_AMD__", *defs,
               "-I" + os.path.join(ROCM, "include"), src, "-o", out, "-L" + LIBDIR, "-lwob_hip",
               "-L" + os.path.join(ROCM, "lib"), "-lamdhip64", "-Wl,-rpath,$ORIGIN",
               "-Wl,-rpath," + os.path.join(ROCM, "lib")]
        procs.append(subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT))
    for p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            raise RuntimeError("app build failed:\n" + out.decode(errors="replace"))


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
