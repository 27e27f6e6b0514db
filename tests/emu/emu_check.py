"""Runs the library's kernels compiled for the CPU (libfmh_emu.so, built by
tests/emu/build_emu.sh) through the C-ABI and compares every result with the
CPU oracle.  Invoked by tests/test_kernels_on_cpu.py in a subprocess (it
needs AddressSanitizer preloaded).  usage: emu_check.py LIBPATH"""

import ctypes
import os
import random
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle.oracle import Oracle  # noqa: E402

L = ctypes.CDLL(sys.argv[1])
P = ctypes.POINTER(ctypes.c_uint64)
u = ctypes.c_uint64
vp = ctypes.c_void_p
L.abs_mul64.argtypes = [P, u, P, u, P]
L.abs_sqr64.argtypes = [P, u, P]
L.fmh_forward_spectrum_device.argtypes = [ctypes.c_int, vp, u, u, vp, vp]
L.fmh_mul_batch_device.argtypes = [vp, u, vp]
L.fmh_batch_create.argtypes = [vp, u, ctypes.POINTER(vp)]
L.fmh_batch_run.argtypes = [vp, vp]
L.fmh_batch_destroy.argtypes = [vp]
orc = Oracle()


class Prod(ctypes.Structure):
    _fields_ = [("in1", vp), ("len1", u), ("in2", vp), ("len2", u), ("out", vp)]


def ptr(a):
    return a.ctypes.data_as(P)


def mul(a, b):
    out = np.zeros(a.size + b.size, np.uint64)
    L.abs_mul64(ptr(a), a.size, ptr(b), b.size, ptr(out))
    return out


def sqr(a):
    out = np.zeros(2 * a.size, np.uint64)
    L.abs_sqr64(ptr(a), a.size, ptr(out))
    return out


def check(cond, what):
    if not cond:
        print("FAIL", what, flush=True)
        sys.exit(1)


# single products: single-pass and multi-pass plans, swapped operands, squares
for n1, n2 in [(1, 1), (1, 5), (5, 1), (3, 17), (64, 64), (255, 257), (1, 2500), (2500, 1), (4097, 1)]:
    a, b = orc.splitmix(n1 * 7 + 1, n1), orc.splitmix(n2 * 11 + 2, n2)
    want = orc.mul(a, b)
    check(np.array_equal(mul(a, b), want), ("mul", n1, n2))
    print("mul", n1, n2, "ok", flush=True)
for n in [1, 2, 33, 2049]:
    a = orc.splitmix(n + 3, n)
    check(np.array_equal(sqr(a), orc.sqr(a)), ("sqr", n))
ones = np.full(300, 0xFFFFFFFFFFFFFFFF, dtype=np.uint64)
check(np.array_equal(mul(ones, ones[:200]), orc.mul(ones, ones[:200])), "all ones")

# forward spectra (device pointers are host pointers here)
for log_n in [0, 3, 13]:
    n = 1 << log_n
    limbs = orc.splitmix(log_n + 100, max(1, n // 2))
    for prime in range(3):
        out = np.zeros(n, np.uint64)
        rc = L.fmh_forward_spectrum_device(prime, limbs.ctypes.data, limbs.size, n, out.ctypes.data, None)
        check(rc == 0 and np.array_equal(out, orc.forward_spectrum(prime, limbs, n)), ("spectrum", log_n, prime))

# a batch of mixed sizes, squares and an empty operand; then a planned batch replayed
rng = random.Random(5)
items, keep = [], []
for i in range(14):
    n1 = rng.choice([0, 1, 7, 64, 300, 2000]) if i != 0 else 0
    sq = rng.random() < 0.3
    n2 = n1 if sq else rng.choice([1, 9, 100, 3000])
    a = orc.splitmix(900 + i, n1)
    b = None if sq else orc.splitmix(950 + i, n2)
    out = np.full(max(1, n1 + n2), 7, np.uint64)
    keep.append((a, b, out, n1 + n2))
    items.append(Prod(a.ctypes.data if n1 else None, n1, None if sq else b.ctypes.data, 0 if sq else n2,
                      out.ctypes.data))
arr = (Prod * len(items))(*items)
check(L.fmh_mul_batch_device(ctypes.cast(arr, vp), len(items), None) == 0, "batch rc")
for i, (a, b, out, n) in enumerate(keep):
    want = orc.sqr(a) if b is None else orc.mul(a, b)
    check(np.array_equal(out[:n], want), ("batch item", i))
h = vp()
check(L.fmh_batch_create(ctypes.cast(arr, vp), len(items), ctypes.byref(h)) == 0, "batch create")
for rep in range(2):
    for (a, b, out, n) in keep:
        a[:] = orc.splitmix(rep * 1000 + a.size, a.size)
        if b is not None:
            b[:] = orc.splitmix(rep * 1000 + 7 + b.size, b.size)
        out[:] = 3
    check(L.fmh_batch_run(h, None) == 0, "batch run")
    for i, (a, b, out, n) in enumerate(keep):
        want = orc.sqr(a) if b is None else orc.mul(a, b)
        check(np.array_equal(out[:n], want), ("replay", rep, i))
check(L.fmh_batch_destroy(h) == 0, "batch destroy")
print("emu ok", flush=True)
