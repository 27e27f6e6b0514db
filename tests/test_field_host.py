"""The kernels' modular arithmetic (fast_mul_ntt_amd/csrc/fmh_field.h),
compiled for the CPU with g++ and checked against __int128 definitions:
word-by-word Montgomery reduction vs the one-step REDC of the reference
(src/macro.h:109-118), limb reduction, and the lazy DIF/DIT butterflies'
ranges and congruences."""

import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_field_arithmetic_on_cpu(tmp_path):
    exe = tmp_path / "field_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "fast_mul_ntt_amd", "csrc"),
                    os.path.join(ROOT, "tests", "field_check.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe), "300000"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
