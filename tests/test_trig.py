"""CPU: the walk's double sine / cosine (wobtoolbox_amd/csrc/wob_trig.hpp) against the C
library's, rounded to FP32, for every angle a (0, 1] Philox uniform can give (2^24 values
per argument family: 2D and 3D line directions, Box-Muller angles).  The device compiles the
same IEEE double operations (explicit fused multiply-adds, no contraction), so this checks
the device's line directions against the oracle's over their whole domain."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_small_sincos_rounds_like_the_c_library(tmp_path):
    exe = str(tmp_path / "trig_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fopenmp",
                    os.path.join(ROOT, "tests", "cpp", "trig_check.cpp"), "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    assert [int(v) for v in out] == [0, 0, 0], out
