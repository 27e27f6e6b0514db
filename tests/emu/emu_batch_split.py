"""A batch of split products (pieces sharing one spectrum of the short
operand) and small products under the built-in pass geometry, on the CPU
emulation of the kernels (libfmh_emu.so, built by tests/emu/build_emu.sh):
the spectra whose transform is one row tile run in one merged row launch,
the pieces, multiplies and squares of one row tile in another; every
product equals the oracle's.  Invoked by tests/test_kernels_on_cpu.py with
AddressSanitizer preloaded and no $FMH_TUNE.  usage: emu_batch_split.py LIBPATH"""

import ctypes
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle.oracle import Oracle  # noqa: E402

L = ctypes.CDLL(sys.argv[1])
u = ctypes.c_uint64
vp = ctypes.c_void_p
L.fmh_mul_batch_device.argtypes = [vp, u, vp]
L.fmh_timing_enable.argtypes = [ctypes.c_int]
L.fmh_timing_collect.argtypes = [ctypes.c_char_p, u]
L.fmh_timing_collect.restype = u
orc = Oracle()


class Prod(ctypes.Structure):
    _fields_ = [("in1", vp), ("len1", u), ("in2", vp), ("len2", u), ("out", vp)]


def check(cond, what):
    if not cond:
        print("FAIL", what, flush=True)
        sys.exit(1)


# (n1, n2): 20000 x 300 and 300 x 9000 are split into pieces of 2^11 points
# (one row tile under the built-in geometry); the others are whole
# products of one row tile; n2 == 0 is a square
shapes = [(20000, 300), (300, 9000), (9000, 400), (100, 90), (700, 0), (33, 2000), (1500, 1500), (64, 0)]
items, keep = [], []
for i, (n1, n2) in enumerate(shapes):
    a = orc.splitmix(500 + i, n1)
    b = orc.splitmix(600 + i, n2) if n2 else None
    out = np.full(n1 + (n2 or n1), 7, np.uint64)
    keep.append((a, b, out))
    items.append(Prod(a.ctypes.data, n1, b.ctypes.data if n2 else None, n2, out.ctypes.data))
arr = (Prod * len(items))(*items)
L.fmh_timing_collect(None, 0)
L.fmh_timing_enable(1)
check(L.fmh_mul_batch_device(ctypes.cast(arr, vp), len(items), None) == 0, "batch rc")
L.fmh_timing_enable(0)
buf = ctypes.create_string_buffer(4096)
L.fmh_timing_collect(buf, 4096)
launches = {ln.split()[0]: int(ln.split()[1]) for ln in buf.value.decode().splitlines() if ln.strip()}
check(launches.get("row_small") == 2, ("two merged row launches", launches))
for i, (a, b, out) in enumerate(keep):
    want = orc.sqr(a) if b is None else orc.mul(a, b)
    check(np.array_equal(out, want), ("product", i, shapes[i]))
print("split batch ok", launches, flush=True)
