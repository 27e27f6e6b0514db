"""LDS budget of the pass kernels (CPU only).  gfx950 gives a workgroup
160 KiB of LDS, static LDS included; a launch beyond what the runtime grants
fails (the drop-in then aborts with a message, as src/fast_mul.c:320-323
does).  Checked here without a GPU:

* every geometry the tuning table or the autotuner may set is clamped so
  that its tiles stay within the bounds of fmh_internal.h (the library's own
  clamp, through fmh_set_tuning / fmh_get_tuning, which need no device);
* each pass kernel's static LDS, read from the compiled gfx950 code object's
  metadata, plus the largest dynamic tile those bounds allow fits 160 KiB.
"""

import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "fast_mul_ntt_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def _bounds(tmp_path):
    exe = tmp_path / "lds_bounds"
    subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "tests", "emu"), "-I", CSRC,
                    os.path.join(ROOT, "tests", "lds_bounds.cpp"), "-o", str(exe)], check=True)
    per_group, col_max, row_max, col_log, row_log = map(int, subprocess.run([str(exe)], capture_output=True,
                                                                           text=True, check=True).stdout.split())
    return per_group, col_max, row_max, col_log, row_log


def test_tuning_clamp_keeps_tiles_within_bounds(fm, tmp_path):
    from fast_mul_ntt_amd.autotune import CANDIDATES
    per_group, col_max, row_max, col_log, row_log = _bounds(tmp_path)
    specs = [f"{k}={v}" for k, vs in CANDIDATES.items() for v in vs]
    specs += ["col_maxl=13,col_logc=5", "col_maxl=20,col_logc=9", "row_tile=20,row_maxl=20", "row_tile=13,row_maxl=13",
              "col_threads=4096,row_threads=1", "row_radix=7"]
    saved = {(lg, b): fm.get_tuning(lg, batched=b) for lg in (12, 21, 25) for b in (False, True)}
    try:
        for spec in specs:
            for (lg, b) in saved:
                fm.set_tuning(lg, "", batched=b)
                fm.set_tuning(lg, spec, batched=b)
                t = dict(kv.split("=") for kv in fm.get_tuning(lg, batched=b).split(","))
                t = {k: int(v) for k, v in t.items()}
                assert t["col_maxl"] + t["col_logc"] <= col_log, (spec, t)
                assert t["row_maxl"] <= t["row_tile"] <= row_log, (spec, t)
                assert 64 <= t["col_threads"] <= 1024 and 64 <= t["row_threads"] <= 1024, (spec, t)
    finally:
        for (lg, b), spec in saved.items():
            fm.set_tuning(lg, spec, batched=b)


@pytest.mark.skipif(not os.path.exists(HIPCC) or shutil.which("g++") is None, reason="hipcc and g++ needed")
def test_static_plus_dynamic_lds_fits_a_workgroup(tmp_path):
    per_group, col_max, row_max, _, _ = _bounds(tmp_path)
    asm = tmp_path / "fmh_ntt.s"
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "--cuda-device-only", "-S",
                    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                    os.path.join(CSRC, "fmh_ntt.hip"), "-o", str(asm)], check=True, capture_output=True, timeout=600)
    text = asm.read_text()
    # the code object metadata lists, per kernel, its static LDS and its name
    static = {}
    for block in text.split("  - .")[1:]:
        name = re.search(r"\.name:\s+(\S+)", block)
        size = re.search(r"\.group_segment_fixed_size:\s+(\d+)", block)
        if name and size:
            static[name.group(1)] = int(size.group(1))
    col = {k: v for k, v in static.items() if "k_col" in k}
    row = {k: v for k, v in static.items() if "k_row" in k}
    # k_col<FWD, DMA, BIG>: <true|false, false, *>, <true, true, *>; k_row<2|3|4>
    # and k_row_small (the row-only classes of a batch in one launch)
    assert len(col) == 6 and len(row) == 4, static
    for k, v in col.items():
        assert v + col_max <= per_group, (k, v, col_max)
    for k, v in row.items():
        assert v + row_max <= per_group, (k, v, row_max)
    # the other kernels take no dynamic LDS and fit on their own
    for k, v in static.items():
        assert v <= per_group, (k, v)
