"""The library's own kernel sources, compiled for the CPU with g++ and
AddressSanitizer against a small HIP-runtime emulation (tests/emu), run
through the C-ABI: every product, square, spectrum, batch and graph replay
must equal the oracle's, and no kernel may touch memory outside its buffers
(device buffers and the dynamic LDS tile are host allocations of exactly the
requested size)."""

import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _asan_runtime():
    try:
        p = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    except OSError:
        return None
    return p if p and os.path.isabs(p) and os.path.exists(p) else None


needs_asan = pytest.mark.skipif(shutil.which("g++") is None or _asan_runtime() is None,
                                reason="g++ with ASan needed")


@pytest.fixture(scope="module")
def emu_lib(tmp_path_factory):
    out = tmp_path_factory.mktemp("emu")
    subprocess.run([os.path.join(ROOT, "tests", "emu", "build_emu.sh"), str(out)], check=True,
                   capture_output=True, timeout=600)
    return out / "libfmh_emu.so"


@needs_asan
def test_kernels_emulated_on_cpu(emu_lib):
    env = dict(os.environ, LD_PRELOAD=_asan_runtime(), ASAN_OPTIONS="detect_leaks=0",
               # small tiles: transforms above 256 points take two or more
               # column passes, so short operands reach every pass schedule
               FMH_TUNE="col_threads=64,row_threads=64,row_tile=8,row_maxl=7,col_maxl=5",
               # full twiddle tables up to 2^13 points: the first column pass of
               # longer transforms forms its twiddles from two entries
               FMH_TW_FULL_LOG="13")
    r = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "emu", "emu_check.py"), str(emu_lib)],
                       capture_output=True, text=True, env=env, timeout=1200)
    assert r.returncode == 0 and "emu ok" in r.stdout, r.stdout[-3000:] + r.stderr[-3000:]


@needs_asan
def test_batch_of_split_products_emulated_on_cpu(emu_lib):
    """Under the built-in geometry (no $FMH_TUNE): the spectra and pieces of
    split products in merged row launches."""
    env = {k: v for k, v in os.environ.items() if k not in ("FMH_TUNE", "FMH_TUNE_TABLE")}
    env.update(LD_PRELOAD=_asan_runtime(), ASAN_OPTIONS="detect_leaks=0")
    r = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "emu", "emu_batch_split.py"), str(emu_lib)],
                       capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0 and "split batch ok" in r.stdout, r.stdout[-3000:] + r.stderr[-3000:]
