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


def test_log_cr_is_correctly_rounded(tmp_path):
    """The 2D fundamental solution's logarithm (wobtrig::log_cr) against the C library's
    extended-precision log rounded to float, for every float in [2^-26, 2^12) (about 3.2e8
    values): equal except for two arguments (1.30017221 and 1.54160988) whose logarithm lies
    within 2^-47 of a rounding boundary, where log_cr's double evaluation (relative error about
    2^-47) rounds to the neighbouring float.  Host and device evaluate the same IEEE
    operations."""
    exe = str(tmp_path / "log_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fopenmp",
                    os.path.join(ROOT, "tests", "cpp", "log_check.cpp"), "-o", exe], check=True)
    bad, n = (int(v) for v in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split())
    assert n == 38 << 23 and bad <= 2, (bad, n)
