"""Two-level twiddles on the GPU against the CPU oracle.

Column stages of half-span 2^F or more form each twiddle as the product of
an entry of the full table and one of a short extension table (F = 25 by
default, so only transforms above 2^25 points use them; the 2^29- and
2^30-point products of test_gpu_parity.test_largest_transforms_mod_q do).
Here one child process sets $FMH_TW_FULL_LOG=13, the smallest F, so that
every transform above 2^13 points takes that path at sizes the oracle
checks word for word: products of one and two column passes, a square, an
unbalanced product and forward spectra.
"""

import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import sys
import numpy as np
sys.path.insert(0, sys.argv[1])
import fast_mul_ntt_amd as fm
from oracle.oracle import Oracle
orc = Oracle()
for n1, n2 in [(8193, 8000), (65536, 65536), (1 << 20, 1 << 20), (1 << 21, 1 << 21), (1 << 20, 70000)]:
    print("mul", n1, n2, flush=True)
    a = orc.splitmix(n1 + 11, n1)
    b = orc.splitmix(n2 + 13, n2)
    assert np.array_equal(fm.abs_mul64(a, b), orc.mul(a, b)), ("mul", n1, n2)
for n in [9000, 1 << 20]:
    print("sqr", n, flush=True)
    a = orc.splitmix(n + 17, n)
    assert np.array_equal(fm.abs_sqr64(a), orc.sqr(a)), ("sqr", n)
for log_n in [14, 17, 21, 22]:
    print("spectrum", log_n, flush=True)
    n = 1 << log_n
    for prime in range(3):
        limbs = orc.splitmix(log_n * 5 + prime, n // 2)
        assert np.array_equal(fm.forward_spectrum(prime, limbs, n), orc.forward_spectrum(prime, limbs, n)), \
            ("spectrum", log_n, prime)
print("two-level ok", flush=True)
"""


@pytest.mark.timeout(300)
def test_two_level_twiddles_vs_oracle():
    env = dict(os.environ, FMH_TW_FULL_LOG="13")
    r = subprocess.run([sys.executable, "-c", CHILD, ROOT], capture_output=True, text=True, env=env, timeout=280)
    assert r.returncode == 0 and "two-level ok" in r.stdout, r.stdout[-2000:] + r.stderr[-2000:]
