"""CPU checks of the boundary: the C-ABI library loads and exports every
symbol include/fmh.h declares (no compute calls: there is no GPU here), the
header keeps the reference's signatures, and bench.py refuses several GPUs."""

import ctypes
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol(fm):
    names = fm.declared_symbols()
    assert {"abs_mul64", "abs_sqr64", "fmh_mul_device", "fmh_sqr_device"} <= set(names)
    L = ctypes.CDLL(fm.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_dropin_signatures_match_reference_header():
    """src/fast_mul.h:39,45: void abs_mul64(const u64*, u64, const u64*, u64, u64*);
    void abs_sqr64(const u64*, u64, u64*); u64 = unsigned long long (src/macro.h:39)."""
    h = open(os.path.join(ROOT, "include", "fmh.h")).read()
    assert re.search(r"typedef unsigned long long u64;", h)
    assert "void abs_mul64(const u64* in1, u64 len1, const u64* in2, u64 len2, u64* out);" in h
    assert "void abs_sqr64(const u64* in1, u64 len1, u64* out);" in h
    assert 'extern "C"' in h


def test_library_is_gfx950_code_object(fm):
    data = open(fm.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode != 0 and "one GPU" in r.stderr


def test_reference_style_c_caller_links(tmp_path):
    """A C program written against the reference's interface (tests/c_caller.c,
    fmh.h in place of fast_mul.h) compiles and links against libfmh.so."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("gcc needed")
    exe = tmp_path / "c_caller"
    subprocess.run(["gcc", "-O2", "-std=c11", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "c_caller.c"), "-L", os.path.join(ROOT, "fast_mul_ntt_amd"),
                    "-lfmh", "-o", str(exe)], check=True)
    assert exe.exists()
