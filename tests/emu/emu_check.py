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
L.fmh_spectrum_create.argtypes = [vp, u, u, vp, ctypes.POINTER(vp)]
L.fmh_spectrum_ntt_len.argtypes = [vp]
L.fmh_spectrum_ntt_len.restype = u
L.fmh_mul_spectrum_device.argtypes = [vp, vp, u, vp, vp]
L.fmh_sqr_spectrum_device.argtypes = [vp, vp, vp]
L.fmh_spectrum_destroy.argtypes = [vp]
L.fmh_server_create.argtypes = [vp, ctypes.POINTER(vp)]
L.fmh_server_submit.argtypes = [vp, vp, u, vp, u, ctypes.POINTER(u)]
L.fmh_server_submit_spectrum.argtypes = [vp, vp, vp, u, ctypes.POINTER(u)]
L.fmh_server_poll.argtypes = [vp, vp, u, ctypes.POINTER(u)]
L.fmh_server_wait.argtypes = [vp, u]
L.fmh_server_result.argtypes = [vp, u, ctypes.POINTER(vp), ctypes.POINTER(u)]
L.fmh_server_release.argtypes = [vp, u]
L.fmh_server_drain.argtypes = [vp]
L.fmh_server_destroy.argtypes = [vp]
L.fmh_server_last_error.restype = ctypes.c_char_p
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


# single products: single-pass and multi-pass plans, swapped operands, squares;
# (3000, 2500) and (2500, 2049) go through the host path that uploads the
# second operand while the first one's spectrum is computed
for n1, n2 in [(1, 1), (1, 5), (5, 1), (3, 17), (64, 64), (255, 257), (1, 2500), (2500, 1), (4097, 1),
               (3000, 2500), (2500, 2049)]:
    a, b = orc.splitmix(n1 * 7 + 1, n1), orc.splitmix(n2 * 11 + 2, n2)
    want = orc.mul(a, b)
    check(np.array_equal(mul(a, b), want), ("mul", n1, n2))
    print("mul", n1, n2, "ok", flush=True)
for n in [1, 2, 33, 2049]:
    a = orc.splitmix(n + 3, n)
    check(np.array_equal(sqr(a), orc.sqr(a)), ("sqr", n))
# unbalanced products: the long operand is cut into pieces (several piece
# counts, a last piece shorter than the others, either operand the long one)
for n1, n2 in [(3000, 100), (100, 3000), (5000, 7), (2600, 300)]:
    a, b = orc.splitmix(n1 + 17, n1), orc.splitmix(n2 + 19, n2)
    check(np.array_equal(mul(a, b), orc.mul(a, b)), ("unbalanced", n1, n2))
print("unbalanced ok", flush=True)
# rounds of at most 2 stages in the row pass (k_row<2>): single-pass and
# multi-pass transforms, multiplies and squares
L.fmh_set_tuning.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
L.fmh_get_tuning.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_char_p, u]
for lg, n1, n2 in [(7, 64, 61), (8, 100, 101), (12, 2048, 2000)]:
    tb = ctypes.create_string_buffer(256)
    L.fmh_get_tuning(lg, 0, tb, 256)
    check(L.fmh_set_tuning(lg, tb.value + b",row_radix=2", 0) == 0, ("radix-2 tuning", lg))
    a, b = orc.splitmix(lg * 5 + 1, n1), orc.splitmix(lg * 5 + 2, n2)
    check(np.array_equal(mul(a, b), orc.mul(a, b)), ("radix-2 mul", lg))
    check(np.array_equal(sqr(a), orc.sqr(a)), ("radix-2 sqr", lg))
    check(L.fmh_set_tuning(lg, tb.value, 0) == 0, ("tuning restored", lg))
print("radix-2 rows ok", flush=True)
ones = np.full(300, 0xFFFFFFFFFFFFFFFF, dtype=np.uint64)
check(np.array_equal(mul(ones, ones[:200]), orc.mul(ones, ones[:200])), "all ones")
# carry chains across 1024-limb chunks: runs of all-ones limbs leave the
# carry-in maps of whole chunks non-constant (chunk scan and fused paths)
big_ones = np.full(3000, 0xFFFFFFFFFFFFFFFF, dtype=np.uint64)
check(np.array_equal(mul(big_ones, big_ones[:2500]), orc.mul(big_ones, big_ones[:2500])), "all ones, 6 chunks")
b = np.zeros(1030, np.uint64)
b[0] = b[-1] = 1
check(np.array_equal(mul(big_ones[:2100], b), orc.mul(big_ones[:2100], b)), "ones x (1 + 2^(64 1029)), 4 chunks")

# fmh_trim from another thread while host-pointer multiplies run: a call that
# uses two streams works in two locked steps, before and after its second
# operand's upload, and keeps its workspace and staging in between (ASan
# reports a freed one as a use after free)
import threading
L.fmh_trim.argtypes = []
stop_trim = threading.Event()


def trim_loop():
    while not stop_trim.is_set():
        L.fmh_trim()


tt = threading.Thread(target=trim_loop)
tt.start()
try:
    a, b = orc.splitmix(61, 3000), orc.splitmix(62, 2500)
    want = orc.mul(a, b)
    for it in range(12):
        check(np.array_equal(mul(a, b), want), ("mul under trim", it))
finally:
    stop_trim.set()
    tt.join()
print("trim ok", flush=True)

# null pointers with nonzero lengths are refused before any launch
L.fmh_mul_device.argtypes = [vp, u, vp, u, vp, vp]
L.fmh_sqr_device.argtypes = [vp, u, vp, vp]
L.fmh_last_error.restype = ctypes.c_char_p
buf = np.zeros(64, np.uint64)
check(L.fmh_mul_device(None, 8, buf.ctypes.data, 8, buf.ctypes.data + 256, None) != 0, "null operand")
check(b"NULL" in L.fmh_last_error(), "null operand message")
check(L.fmh_mul_device(buf.ctypes.data, 8, buf.ctypes.data, 8, None, None) != 0, "null output")
check(L.fmh_forward_spectrum_device(0, None, 8, 16, buf.ctypes.data, None) != 0, "null spectrum input")
h0 = vp()
check(L.fmh_spectrum_create(None, 8, 0, None, ctypes.byref(h0)) != 0, "null spectrum operand")
buf[:] = 7
check(L.fmh_mul_device(None, 0, buf.ctypes.data, 8, buf.ctypes.data + 256, None) == 0, "empty null operand")
check(not buf[32:40].any(), "empty product is zero")
# the mirrored case, with an output of exactly len1 limbs: zeros, not a square
# of the first operand (which would write 2 len1 limbs)
a8 = orc.splitmix(77, 8)
out8 = np.full(8, 7, np.uint64)
check(L.fmh_mul_device(a8.ctypes.data, 8, None, 0, out8.ctypes.data, None) == 0, "empty null second operand")
check(not out8.any(), "a x (empty) is zero")
check(L.fmh_mul_device(a8.ctypes.data, 8, None, 0, None, None) != 0, "null output of an empty product")
check(L.fmh_sqr_device(buf.ctypes.data, 8, None, None) != 0, "null square output")
check(b"NULL" in L.fmh_last_error(), "null square output message")
# operands at 16-byte aligned and at odd word offsets, odd and even lengths,
# products with a column pass: the first column pass fills its tiles by LDS
# DMA from aligned limbs (an odd length's last word read on its own) and
# word by word from the others.  Offset None: an array of its own, whose
# last word ends its allocation (a read past it would stop AddressSanitizer).
# The longer products give the column passes rounds of 3 stages (radix 8).
pool_words = orc.splitmix(4321, 72000)
for off1, n1, off2, n2 in [(0, 4097, 0, 4000), (1, 4097, 0, 3), (0, 4096, 5, 4001), (3, 2999, 7, 2049),
                           (0, 9001, 0, 9000), (1, 9001, 2, 1), (None, 4097, None, 2049), (None, 6001, None, 1),
                           (0, 20001, 1, 20000), (0, 40001, 1, 30000)]:
    a = orc.splitmix(n1, n1) if off1 is None else pool_words[off1:off1 + n1]
    b = orc.splitmix(n2 + 1, n2) if off2 is None else pool_words[off2:off2 + n2]
    out = np.full(n1 + n2, 3, np.uint64)
    check(L.fmh_mul_device(a.ctypes.data, n1, b.ctypes.data, n2, out.ctypes.data, None) == 0, ("offset rc", off1))
    check(np.array_equal(out, orc.mul(a.copy(), b.copy())), ("offset mul", off1, n1, off2, n2))
    out = np.full(2 * n1, 3, np.uint64)
    check(L.fmh_sqr_device(a.ctypes.data, n1, out.ctypes.data, None) == 0, ("offset sqr rc", off1))
    check(np.array_equal(out, orc.sqr(a.copy())), ("offset sqr", off1, n1))
print("operand offsets ok", flush=True)
# lengths beyond the largest transform are refused before any memory is read
# (the buffer is far smaller than the lengths)
L.fmh_max_ntt_len.restype = u
big = (L.fmh_max_ntt_len() >> 1) + 1
check(L.fmh_mul_device(buf.ctypes.data, big, buf.ctypes.data + 8, big, buf.ctypes.data, None) != 0, "too long")
check(b"fmh_max_ntt_len" in L.fmh_last_error(), "too long message")

# the allocation-free variant: the same products as fmh_mul_device (two calls
# in a row, the second smaller; a square; a shape the other entry point
# splits), and refusal of a product longer than its scratch array
L.fmh_init.argtypes = [ctypes.c_int]
L.fmh_mul_device_noalloc.argtypes = [vp, u, vp, u, vp, vp]
check(L.fmh_init(0) == 0, "init")
for n1, n2, sq in [(3000, 2001, False), (700, 100, False), (5000, 7, False), (1500, 0, True)]:
    a = orc.splitmix(n1 + 31, n1)
    b = a if sq else orc.splitmix(n2 + 37, n2)
    nb = b.size
    o1 = np.full(n1 + nb, 9, np.uint64)
    o2 = np.full(n1 + nb, 8, np.uint64)
    check(L.fmh_mul_device_noalloc(a.ctypes.data, n1, b.ctypes.data, nb, o1.ctypes.data, None) == 0,
          ("noalloc rc", n1, n2, L.fmh_last_error()))
    check(L.fmh_mul_device(a.ctypes.data, n1, b.ctypes.data, nb, o2.ctypes.data, None) == 0, ("device rc", n1))
    check(np.array_equal(o1, o2) and np.array_equal(o1, orc.mul(a, b)), ("noalloc", n1, n2, sq))
check(L.fmh_mul_device_noalloc(buf.ctypes.data, 1 << 19, buf.ctypes.data, (1 << 19) + 2, buf.ctypes.data,
                               None) != 0, "noalloc too long")
check(b"FMH_NOALLOC_MAX_NTT_LEN" in L.fmh_last_error(), "noalloc too long message")
check(L.fmh_mul_device_noalloc(None, 8, buf.ctypes.data, 8, buf.ctypes.data, None) != 0, "noalloc null operand")
print("noalloc ok", flush=True)

# forward spectra (device pointers are host pointers here)
for log_n in [0, 3, 13]:
    n = 1 << log_n
    limbs = orc.splitmix(log_n + 100, max(1, n // 2))
    for prime in range(3):
        out = np.zeros(n, np.uint64)
        rc = L.fmh_forward_spectrum_device(prime, limbs.ctypes.data, limbs.size, n, out.ctypes.data, None)
        check(rc == 0 and np.array_equal(out, orc.forward_spectrum(prime, limbs, n)), ("spectrum", log_n, prime))

# resident spectra: default and longer transform lengths, one- and two-pass
# plans; products with several second operands, the square, a product that
# does not fit
for n1, ntt_len, others in [(7, 0, [1, 7]), (7, 2048, [9, 1500]), (600, 0, [500, 600, 1]), (700, 0, [])]:
    a = orc.splitmix(n1 + 500, n1)
    h = vp()
    check(L.fmh_spectrum_create(a.ctypes.data, n1, ntt_len, None, ctypes.byref(h)) == 0, ("spec create", n1))
    N = L.fmh_spectrum_ntt_len(h)
    check(N == (ntt_len or 1 << (2 * n1 - 1).bit_length()), ("spec len", n1, N))
    for n2 in others:
        b = orc.splitmix(n2 + 600, n2)
        out = np.full(n1 + n2, 5, np.uint64)
        check(L.fmh_mul_spectrum_device(h, b.ctypes.data, n2, out.ctypes.data, None) == 0, ("spec mul rc", n1, n2))
        check(np.array_equal(out, orc.mul(a, b)), ("spec mul", n1, n2))
    out = np.full(2 * n1, 5, np.uint64)
    check(L.fmh_sqr_spectrum_device(h, out.ctypes.data, None) == 0, ("spec sqr rc", n1))
    check(np.array_equal(out, orc.sqr(a)), ("spec sqr", n1))
    big = np.ones(N, np.uint64)
    out = np.zeros(N + n1, np.uint64)
    check(L.fmh_mul_spectrum_device(h, big.ctypes.data, N, out.ctypes.data, None) != 0, ("spec too long", n1))
    # a null output is refused, also for an empty second operand (zero product)
    check(L.fmh_mul_spectrum_device(h, big.ctypes.data, 1, None, None) != 0, ("spec null out", n1))
    check(L.fmh_mul_spectrum_device(h, None, 0, None, None) != 0, ("spec null out, empty", n1))
    check(L.fmh_sqr_spectrum_device(h, None, None) != 0, ("spec sqr null out", n1))
    out = np.full(n1, 5, np.uint64)
    check(L.fmh_mul_spectrum_device(h, None, 0, out.ctypes.data, None) == 0 and not out.any(), ("spec x empty", n1))
    check(L.fmh_spectrum_destroy(h) == 0, ("spec destroy", n1))
print("spectra ok", flush=True)

# a batch of mixed sizes, squares and an empty operand; then a planned batch replayed
rng = random.Random(5)
items, keep = [], []
for i in range(14):
    n1 = rng.choice([0, 1, 7, 64, 300, 2000]) if i != 0 else 0
    sq = rng.random() < 0.3
    n2 = n1 if sq else rng.choice([1, 9, 100, 3000, 5])
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

# many small products and squares: the row-only classes (transforms of one
# row tile) go into one launch, up to its class limit; the rest on their own
rng = random.Random(6)
items, keep = [], []
for i in range(70):
    n1 = rng.randint(1, 140)
    sq = i % 3 == 0
    n2 = n1 if sq else rng.randint(1, 140)
    a = orc.splitmix(1200 + i, n1)
    b = None if sq else orc.splitmix(1300 + i, n2)
    out = np.full(n1 + n2, 7, np.uint64)
    keep.append((a, b, out))
    items.append(Prod(a.ctypes.data, n1, None if sq else b.ctypes.data, 0 if sq else n2, out.ctypes.data))
arr = (Prod * len(items))(*items)
L.fmh_timing_enable.argtypes = [ctypes.c_int]
L.fmh_timing_collect.argtypes = [ctypes.c_char_p, u]
L.fmh_timing_collect.restype = u
L.fmh_timing_collect(None, 0)
L.fmh_timing_enable(1)
check(L.fmh_mul_batch_device(ctypes.cast(arr, vp), len(items), None) == 0, "small batch rc")
L.fmh_timing_enable(0)
tbuf = ctypes.create_string_buffer(4096)
L.fmh_timing_collect(tbuf, 4096)
launches = {ln.split()[0]: int(ln.split()[1]) for ln in tbuf.value.decode().splitlines() if ln.strip()}
check(launches.get("row_small") == 1, ("one merged row launch", launches))
for i, (a, b, out) in enumerate(keep):
    want = orc.sqr(a) if b is None else orc.mul(a, b)
    check(np.array_equal(out, want), ("small batch item", i))
print("small batch ok", launches, flush=True)

# request loop: a pool too small to hold every request at once, mixed sizes,
# squares and spectrum requests; every output equals the oracle's
class SrvCfg(ctypes.Structure):
    _fields_ = [("pool_bytes", u), ("streams", ctypes.c_uint), ("max_group", ctypes.c_uint),
                ("max_group_bytes", u)]


cfg = SrvCfg(1 << 20, 3, 8, 0)
srv = vp()
check(L.fmh_server_create(ctypes.byref(cfg), ctypes.byref(srv)) == 0, "server create")
a0 = orc.splitmix(4242, 1000)
sp = vp()
check(L.fmh_spectrum_create(a0.ctypes.data, a0.size, 0, None, ctypes.byref(sp)) == 0, "server spec")
rng = random.Random(11)
reqs = {}
for i in range(24):
    kind = rng.choice(["mul", "mul", "sqr", "spec", "specsqr"])
    rid = u()
    if kind in ("mul", "sqr"):
        n1 = rng.choice([1, 40, 700, 1500])
        a = orc.splitmix(7000 + i, n1)
        b = None if kind == "sqr" else orc.splitmix(8000 + i, rng.choice([1, 50, 1200]))
        rc = L.fmh_server_submit(srv, a.ctypes.data, a.size, None if b is None else b.ctypes.data,
                                 0 if b is None else b.size, ctypes.byref(rid))
        want = orc.sqr(a) if b is None else orc.mul(a, b)
        keep_ops = (a, b)
    else:
        b = None if kind == "specsqr" else orc.splitmix(9000 + i, rng.choice([1, 30, 1000]))
        rc = L.fmh_server_submit_spectrum(srv, sp, None if b is None else b.ctypes.data,
                                          0 if b is None else b.size, ctypes.byref(rid))
        want = orc.sqr(a0) if b is None else orc.mul(a0, b)
        keep_ops = (b,)
    check(rc == 0, ("submit", i, L.fmh_server_last_error()))
    reqs[rid.value] = (want, keep_ops)
check(L.fmh_server_drain(srv) == 0, ("drain", L.fmh_server_last_error()))
ids = (u * 64)()
n = u()
check(L.fmh_server_poll(srv, ids, 64, ctypes.byref(n)) == 0 and n.value == len(reqs), ("poll", n.value))
for rid, (want, _) in reqs.items():
    p, ln = vp(), u()
    check(L.fmh_server_result(srv, rid, ctypes.byref(p), ctypes.byref(ln)) == 0, ("result", rid))
    got = np.ctypeslib.as_array(ctypes.cast(p, P), shape=(ln.value,)).copy()
    check(ln.value == want.size and np.array_equal(got, want), ("served product", rid))
    check(L.fmh_server_release(srv, rid) == 0, ("release", rid))
too_big = np.ones(1 << 17, np.uint64)
check(L.fmh_server_submit(srv, too_big.ctypes.data, too_big.size, too_big.ctypes.data, too_big.size,
                          ctypes.byref(u())) != 0, "footprint above the pool is refused")
check(L.fmh_server_destroy(srv) == 0, "server destroy")
check(L.fmh_spectrum_destroy(sp) == 0, "server spec destroy")
print("server ok", flush=True)
print("emu ok", flush=True)
