"""fast_mul_ntt_amd -- MI355X (gfx950) 3-prime NTT + CRT big-integer multiply.

Python binding (ctypes) of ``libfmh.so``, whose C-ABI is declared in
``include/fmh.h``.  The functions mirror the reference's interface
(HJimmyK/fast_mul_ntt ``src/fast_mul.h:39,45``):

* ``abs_mul64(in1, in2) -> out``  with ``len(out) == len(in1) + len(in2)``
* ``abs_sqr64(in1) -> out``       with ``len(out) == 2 * len(in1)``

over little-endian ``uint64`` limb arrays (numpy), plus device-pointer entry
points.  The compute path is the HIP library only: if ``libfmh.so`` is
missing this module raises instead of falling back to anything else.
"""

from __future__ import annotations

import ctypes
import os
import re

import numpy as np

__all__ = [
    "LIB_PATH",
    "HEADER_PATH",
    "lib",
    "declared_symbols",
    "abs_mul64",
    "abs_sqr64",
    "DeviceBuffer",
    "DeviceView",
    "Stream",
    "mul_device",
    "sqr_device",
    "mul_device_noalloc",
    "forward_spectrum",
    "copy_device",
    "FmhProduct",
    "mul_batch_device",
    "Batch",
    "Spectrum",
    "ServerConfig",
    "ServerInfo",
    "Server",
    "set_tuning",
    "get_tuning",
    "timing_enable",
    "timing_collect",
    "FmhError",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfmh.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "fmh.h")

_u64 = ctypes.c_uint64
_pu64 = ctypes.POINTER(ctypes.c_uint64)
_vp = ctypes.c_void_p
_LIB = None


class FmhError(RuntimeError):
    pass


class FmhProduct(ctypes.Structure):
    """struct fmh_product of include/fmh.h (device pointers)."""

    _fields_ = [("in1", _vp), ("len1", _u64), ("in2", _vp), ("len2", _u64), ("out", _vp)]


class ServerConfig(ctypes.Structure):
    """fmh_server_config of include/fmh.h (zeros pick the defaults)."""

    _fields_ = [("pool_bytes", _u64), ("streams", ctypes.c_uint), ("max_group", ctypes.c_uint),
                ("max_group_bytes", _u64)]


class ServerInfo(ctypes.Structure):
    """fmh_server_info of include/fmh.h."""

    _fields_ = [("submitted", _u64), ("completed", _u64), ("failed", _u64), ("groups", _u64),
                ("mean_group", ctypes.c_double), ("p50_us", ctypes.c_double), ("p99_us", ctypes.c_double),
                ("max_us", ctypes.c_double), ("mean_us", ctypes.c_double),
                ("products_per_s", ctypes.c_double), ("span_s", ctypes.c_double),
                ("pool_bytes", _u64), ("pool_peak_bytes", _u64), ("device_busy_s", ctypes.c_double)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


def declared_symbols(header: str = HEADER_PATH) -> list[str]:
    """Function names declared in include/fmh.h."""
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*[A-Za-z_][\w\s\*]*?\b(\w+)\s*\(", text, flags=re.M)
    return [n for n in names if n not in ("if", "while", "for")]


def lib() -> ctypes.CDLL:
    """Load libfmh.so (raises if it was not built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise FmhError(f"{LIB_PATH} is missing; build it with __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    L.abs_mul64.argtypes = [_pu64, _u64, _pu64, _u64, _pu64]
    L.abs_mul64.restype = None
    L.abs_sqr64.argtypes = [_pu64, _u64, _pu64]
    L.abs_sqr64.restype = None
    L.fmh_mul_device.argtypes = [_vp, _u64, _vp, _u64, _vp, _vp]
    L.fmh_sqr_device.argtypes = [_vp, _u64, _vp, _vp]
    L.fmh_mul_device_noalloc.argtypes = [_vp, _u64, _vp, _u64, _vp, _vp]
    L.fmh_forward_spectrum_device.argtypes = [ctypes.c_int, _vp, _u64, _u64, _vp, _vp]
    L.fmh_max_ntt_len.restype = _u64
    L.fmh_init.argtypes = [ctypes.c_int]
    L.fmh_dev_alloc.argtypes = [ctypes.POINTER(_vp), _u64]
    L.fmh_dev_free.argtypes = [_vp]
    L.fmh_memcpy_h2d.argtypes = [_vp, _vp, _u64, _vp]
    L.fmh_memcpy_d2h.argtypes = [_vp, _vp, _u64, _vp]
    L.fmh_stream_create.argtypes = [ctypes.POINTER(_vp)]
    L.fmh_stream_destroy.argtypes = [_vp]
    L.fmh_stream_synchronize.argtypes = [_vp]
    L.fmh_trim.argtypes = []
    L.fmh_copy_device.argtypes = [_vp, _vp, _u64, _vp]
    L.fmh_mul_batch_device.argtypes = [ctypes.POINTER(FmhProduct), _u64, _vp]
    L.fmh_batch_create.argtypes = [ctypes.POINTER(FmhProduct), _u64, ctypes.POINTER(_vp)]
    L.fmh_batch_run.argtypes = [_vp, _vp]
    L.fmh_batch_destroy.argtypes = [_vp]
    L.fmh_spectrum_create.argtypes = [_vp, _u64, _u64, _vp, ctypes.POINTER(_vp)]
    L.fmh_spectrum_ntt_len.argtypes = [_vp]
    L.fmh_spectrum_ntt_len.restype = _u64
    L.fmh_mul_spectrum_device.argtypes = [_vp, _vp, _u64, _vp, _vp]
    L.fmh_sqr_spectrum_device.argtypes = [_vp, _vp, _vp]
    L.fmh_spectrum_destroy.argtypes = [_vp]
    L.fmh_spectrum_describe.argtypes = [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_u64),
                                        ctypes.POINTER(ctypes.c_int)]
    L.fmh_server_create.argtypes = [ctypes.POINTER(ServerConfig), ctypes.POINTER(_vp)]
    L.fmh_server_submit.argtypes = [_vp, _vp, _u64, _vp, _u64, ctypes.POINTER(_u64)]
    L.fmh_server_submit_spectrum.argtypes = [_vp, _vp, _vp, _u64, ctypes.POINTER(_u64)]
    L.fmh_server_poll.argtypes = [_vp, ctypes.POINTER(_u64), _u64, ctypes.POINTER(_u64)]
    L.fmh_server_wait.argtypes = [_vp, _u64]
    L.fmh_server_result.argtypes = [_vp, _u64, ctypes.POINTER(_vp), ctypes.POINTER(_u64)]
    L.fmh_server_release.argtypes = [_vp, _u64]
    L.fmh_server_drain.argtypes = [_vp]
    L.fmh_server_stats.argtypes = [_vp, ctypes.POINTER(ServerInfo)]
    L.fmh_server_reset_stats.argtypes = [_vp]
    L.fmh_server_class_report.argtypes = [_vp, ctypes.c_char_p, _u64]
    L.fmh_server_class_report.restype = _u64
    L.fmh_server_destroy.argtypes = [_vp]
    L.fmh_server_last_error.restype = ctypes.c_char_p
    L.fmh_last_error.restype = ctypes.c_char_p
    L.fmh_set_tuning.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
    L.fmh_get_tuning.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_char_p, _u64]
    L.fmh_timing_enable.argtypes = [ctypes.c_int]
    L.fmh_timing_only.argtypes = [ctypes.c_char_p]
    L.fmh_timing_collect.argtypes = [ctypes.c_char_p, _u64]
    L.fmh_timing_collect.restype = _u64
    _LIB = L
    return L


def _check(rc: int, what: str) -> None:
    if rc != 0:
        msg = lib().fmh_last_error()
        raise FmhError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


def _as_limbs(x) -> np.ndarray:
    a = np.ascontiguousarray(x, dtype=np.uint64)
    if a.ndim != 1:
        raise ValueError("limb arrays must be one-dimensional")
    return a


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_pu64)


def abs_mul64(in1, in2, out: np.ndarray | None = None) -> np.ndarray:
    """out = in1 * in2 (len(in1) + len(in2) limbs), host arrays, synchronous."""
    a, b = _as_limbs(in1), _as_limbs(in2)
    n = a.size + b.size
    if out is None:
        out = np.empty(n, dtype=np.uint64)
    elif out.dtype != np.uint64 or out.size != n or not out.flags.c_contiguous:
        raise ValueError("out must be a contiguous uint64 array of len(in1)+len(in2) limbs")
    lib().abs_mul64(_ptr(a), a.size, _ptr(b), b.size, _ptr(out))
    return out


def abs_sqr64(in1, out: np.ndarray | None = None) -> np.ndarray:
    """out = in1 ** 2 (2 * len(in1) limbs), host arrays, synchronous."""
    a = _as_limbs(in1)
    if out is None:
        out = np.empty(2 * a.size, dtype=np.uint64)
    elif out.dtype != np.uint64 or out.size != 2 * a.size or not out.flags.c_contiguous:
        raise ValueError("out must be a contiguous uint64 array of 2*len(in1) limbs")
    lib().abs_sqr64(_ptr(a), a.size, _ptr(out))
    return out


class Stream:
    """A HIP stream owned by the library (handle usable as fmh_stream_t)."""

    def __init__(self):
        h = _vp()
        _check(lib().fmh_stream_create(ctypes.byref(h)), "fmh_stream_create")
        self.handle = h.value

    def synchronize(self) -> None:
        _check(lib().fmh_stream_synchronize(self.handle), "fmh_stream_synchronize")

    def close(self) -> None:
        if self.handle is not None:
            _check(lib().fmh_stream_destroy(self.handle), "fmh_stream_destroy")
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class DeviceBuffer:
    """Device memory holding `n` uint64 limbs."""

    def __init__(self, n: int):
        self.n = int(n)
        p = _vp()
        _check(lib().fmh_dev_alloc(ctypes.byref(p), max(1, self.n) * 8), "fmh_dev_alloc")
        self.ptr = p.value

    @classmethod
    def from_host(cls, a, stream: Stream | None = None) -> "DeviceBuffer":
        a = _as_limbs(a)
        d = cls(a.size)
        d.upload(a, stream)
        return d

    def upload(self, a, stream: Stream | None = None) -> None:
        a = _as_limbs(a)
        if a.size > self.n:
            raise ValueError("host array larger than the device buffer")
        if a.size:
            _check(lib().fmh_memcpy_h2d(self.ptr, a.ctypes.data, a.size * 8,
                                        stream.handle if stream else None), "fmh_memcpy_h2d")

    def download(self, n: int | None = None, stream: Stream | None = None) -> np.ndarray:
        n = self.n if n is None else int(n)
        out = np.empty(n, dtype=np.uint64)
        if n:
            _check(lib().fmh_memcpy_d2h(out.ctypes.data, self.ptr, n * 8,
                                        stream.handle if stream else None), "fmh_memcpy_d2h")
        return out

    def view(self, offset: int, n: int) -> "DeviceView":
        """Non-owning view of limbs [offset, offset + n)."""
        if offset < 0 or n < 0 or offset + n > self.n:
            raise ValueError("view outside the buffer")
        return DeviceView(self.ptr + 8 * offset, n)

    def free(self) -> None:
        if self.ptr is not None:
            _check(lib().fmh_dev_free(self.ptr), "fmh_dev_free")
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.free()


class DeviceView:
    """Limbs inside a DeviceBuffer (pointer and length; owns nothing)."""

    def __init__(self, ptr: int, n: int):
        self.ptr = ptr
        self.n = int(n)

    def download(self, n: int | None = None) -> np.ndarray:
        n = self.n if n is None else int(n)
        out = np.empty(n, dtype=np.uint64)
        if n:
            _check(lib().fmh_memcpy_d2h(out.ctypes.data, self.ptr, n * 8, None), "fmh_memcpy_d2h")
        return out


def mul_device(a: DeviceBuffer, na: int, b: DeviceBuffer, nb: int, out: DeviceBuffer,
               stream: Stream | None = None) -> None:
    """Enqueue out = a[:na] * b[:nb] on `stream` (device pointers, asynchronous)."""
    if out.n < na + nb or a.n < na or b.n < nb:
        raise ValueError("device buffers too small")
    _check(lib().fmh_mul_device(a.ptr, na, b.ptr, nb, out.ptr, stream.handle if stream else None),
           "fmh_mul_device")


def mul_device_noalloc(a: DeviceBuffer, na: int, b: DeviceBuffer, nb: int, out: DeviceBuffer,
                       stream: Stream | None = None) -> None:
    """fmh_mul_device_noalloc: out = a[:na] * b[:nb] with the library's fixed
    device scratch (no device allocation; fmh_init first)."""
    if out.n < na + nb or a.n < na or b.n < nb:
        raise ValueError("device buffers too small")
    _check(lib().fmh_mul_device_noalloc(a.ptr, na, b.ptr, nb, out.ptr, stream.handle if stream else None),
           "fmh_mul_device_noalloc")


def sqr_device(a: DeviceBuffer, na: int, out: DeviceBuffer, stream: Stream | None = None) -> None:
    """Enqueue out = a[:na] ** 2 on `stream` (device pointers, asynchronous)."""
    if out.n < 2 * na or a.n < na:
        raise ValueError("device buffers too small")
    _check(lib().fmh_sqr_device(a.ptr, na, out.ptr, stream.handle if stream else None),
           "fmh_sqr_device")


def forward_spectrum(prime: int, limbs, ntt_len: int) -> np.ndarray:
    """Canonical forward spectrum (Montgomery form, reference order) of one operand."""
    a = _as_limbs(limbs)
    with DeviceBuffer.from_host(a) as d, DeviceBuffer(ntt_len) as o:
        _check(lib().fmh_forward_spectrum_device(prime, d.ptr, a.size, ntt_len, o.ptr, None),
               "fmh_forward_spectrum_device")
        return o.download()


def copy_device(dst: DeviceBuffer, src: DeviceBuffer, n: int, stream: Stream | None = None) -> None:
    """Enqueue a plain n-word device copy (kernel "copy")."""
    if dst.n < n or src.n < n:
        raise ValueError("device buffers too small")
    _check(lib().fmh_copy_device(dst.ptr, src.ptr, n, stream.handle if stream else None),
           "fmh_copy_device")


def _products(items) -> ctypes.Array:
    """items: (a, na, b, nb, out) with DeviceBuffers (b None for a square)."""
    arr = (FmhProduct * max(1, len(items)))()
    for i, (a, na, b, nb, out) in enumerate(items):
        nb_eff = na if b is None else nb
        if a.n < na or (b is not None and b.n < nb) or out.n < na + nb_eff:
            raise ValueError(f"batch item {i}: device buffers too small")
        arr[i] = FmhProduct(a.ptr, na, None if b is None else b.ptr, nb_eff if b is not None else 0,
                            out.ptr)
    return arr


def mul_batch_device(items, stream: Stream | None = None) -> None:
    """Enqueue many products (device buffers) as one batch on `stream`."""
    arr = _products(items)
    _check(lib().fmh_mul_batch_device(arr, len(items), stream.handle if stream else None),
           "fmh_mul_batch_device")


class Batch:
    """A planned batch: workspace and descriptors allocated once, launches
    captured in a HIP graph; run() replays it."""

    def __init__(self, items):
        arr = _products(items)
        h = _vp()
        _check(lib().fmh_batch_create(arr, len(items), ctypes.byref(h)), "fmh_batch_create")
        self.handle = h.value

    def run(self, stream: Stream | None = None) -> None:
        _check(lib().fmh_batch_run(self.handle, stream.handle if stream else None), "fmh_batch_run")

    def close(self) -> None:
        if self.handle is not None:
            _check(lib().fmh_batch_destroy(self.handle), "fmh_batch_destroy")
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Spectrum:
    """Resident forward transform of one operand (src/fast_mul.c:135-212
    conv_single_i semantics): products with it transform only the other
    operand, squares of it transform nothing."""

    def __init__(self, a: DeviceBuffer, na: int, ntt_len: int = 0, stream: Stream | None = None):
        if a.n < na:
            raise ValueError("device buffer too small")
        h = _vp()
        _check(lib().fmh_spectrum_create(a.ptr, na, ntt_len, stream.handle if stream else None,
                                         ctypes.byref(h)), "fmh_spectrum_create")
        self.handle = h.value
        self.len = int(na)
        self.ntt_len = int(lib().fmh_spectrum_ntt_len(self.handle))

    def mul(self, b: DeviceBuffer, nb: int, out: DeviceBuffer, stream: Stream | None = None) -> None:
        """Enqueue out = (operand of the spectrum) * b[:nb]."""
        if b.n < nb or out.n < self.len + nb:
            raise ValueError("device buffers too small")
        _check(lib().fmh_mul_spectrum_device(self.handle, b.ptr, nb, out.ptr,
                                             stream.handle if stream else None), "fmh_mul_spectrum_device")

    def sqr(self, out: DeviceBuffer, stream: Stream | None = None) -> None:
        """Enqueue out = (operand of the spectrum) ** 2."""
        if out.n < 2 * self.len:
            raise ValueError("device buffer too small")
        _check(lib().fmh_sqr_spectrum_device(self.handle, out.ptr, stream.handle if stream else None),
               "fmh_sqr_spectrum_device")

    def close(self) -> None:
        if self.handle is not None:
            _check(lib().fmh_spectrum_destroy(self.handle), "fmh_spectrum_destroy")
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _check_srv(rc: int, what: str) -> None:
    if rc != 0:
        msg = lib().fmh_server_last_error()
        raise FmhError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


class Server:
    """Request loop with continuous batching over a device pool (fmh_server_*)."""

    def __init__(self, pool_bytes: int = 0, streams: int = 0, max_group: int = 0, max_group_bytes: int = 0):
        cfg = ServerConfig(pool_bytes, streams, max_group, max_group_bytes)
        h = _vp()
        _check_srv(lib().fmh_server_create(ctypes.byref(cfg), ctypes.byref(h)), "fmh_server_create")
        self.handle = h.value
        self._ids = (_u64 * 4096)()

    def submit(self, a: DeviceBuffer, na: int, b: DeviceBuffer | None = None, nb: int = 0) -> int:
        """Request a[:na] * b[:nb] (b None: the square); returns its id."""
        if a.n < na or (b is not None and b.n < nb):
            raise ValueError("device buffers too small")
        rid = _u64()
        _check_srv(lib().fmh_server_submit(self.handle, a.ptr, na, None if b is None else b.ptr,
                                           nb if b is not None else 0, ctypes.byref(rid)), "fmh_server_submit")
        return rid.value

    def submit_spectrum(self, spec: Spectrum, b: DeviceBuffer | None = None, nb: int = 0) -> int:
        """Request (operand of spec) * b[:nb] (b None: its square); returns its id."""
        rid = _u64()
        _check_srv(lib().fmh_server_submit_spectrum(self.handle, spec.handle, None if b is None else b.ptr,
                                                    nb if b is not None else 0, ctypes.byref(rid)),
                   "fmh_server_submit_spectrum")
        return rid.value

    def poll(self) -> list[int]:
        """Ids completed (or failed) since the last poll."""
        out = []
        while True:
            n = _u64()
            _check_srv(lib().fmh_server_poll(self.handle, self._ids, len(self._ids), ctypes.byref(n)),
                       "fmh_server_poll")
            out.extend(self._ids[i] for i in range(n.value))
            if n.value < len(self._ids):
                return out

    def wait(self, rid: int) -> None:
        _check_srv(lib().fmh_server_wait(self.handle, rid), "fmh_server_wait")

    def result_ptr(self, rid: int) -> tuple[int, int]:
        """(device pointer, limbs) of a completed request's output."""
        p, n = _vp(), _u64()
        _check_srv(lib().fmh_server_result(self.handle, rid, ctypes.byref(p), ctypes.byref(n)),
                   "fmh_server_result")
        return p.value, n.value

    def result(self, rid: int) -> np.ndarray:
        """Copy of a completed request's output (host array)."""
        p, n = self.result_ptr(rid)
        out = np.empty(n, dtype=np.uint64)
        _check(lib().fmh_memcpy_d2h(out.ctypes.data, p, n * 8, None), "fmh_memcpy_d2h")
        return out

    def release(self, rid: int) -> None:
        _check_srv(lib().fmh_server_release(self.handle, rid), "fmh_server_release")

    def drain(self) -> None:
        _check_srv(lib().fmh_server_drain(self.handle), "fmh_server_drain")

    def stats(self) -> dict:
        st = ServerInfo()
        _check_srv(lib().fmh_server_stats(self.handle, ctypes.byref(st)), "fmh_server_stats")
        return st.as_dict()

    def class_report(self) -> list[dict]:
        """Device time per size class: [{kind, log_n, groups, requests, device_ms}]."""
        n = lib().fmh_server_class_report(self.handle, None, 0)
        buf = ctypes.create_string_buffer(int(n))
        lib().fmh_server_class_report(self.handle, buf, len(buf))
        out = []
        for line in buf.value.decode().splitlines():
            k, lg, g, r, ms = line.split()
            out.append({"kind": int(k), "log_n": int(lg), "groups": int(g), "requests": int(r),
                        "device_ms": float(ms)})
        return out

    def reset_stats(self) -> None:
        _check_srv(lib().fmh_server_reset_stats(self.handle), "fmh_server_reset_stats")

    def close(self) -> None:
        if self.handle is not None:
            _check_srv(lib().fmh_server_destroy(self.handle), "fmh_server_destroy")
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def set_tuning(log_n: int, spec: str, batched: bool = False) -> None:
    """Pass geometry of 2^log_n-point transforms (-1: all sizes; "" = defaults),
    for single products or for batched launches of the class."""
    _check(lib().fmh_set_tuning(log_n, spec.encode(), 1 if batched else 0), "fmh_set_tuning")


def get_tuning(log_n: int, batched: bool = False) -> str:
    buf = ctypes.create_string_buffer(256)
    _check(lib().fmh_get_tuning(log_n, 1 if batched else 0, buf, len(buf)), "fmh_get_tuning")
    return buf.value.decode()


def timing_enable(on: bool, only: str | None = None) -> None:
    """Per-kernel HIP-event timing; `only` restricts it to one kernel name."""
    lib().fmh_timing_only(only.encode() if only else None)
    lib().fmh_timing_enable(1 if on else 0)


def timing_collect() -> dict[str, tuple[int, float]]:
    """{kernel name: (launches, total ms)} since the last collect."""
    buf = ctypes.create_string_buffer(1 << 16)
    lib().fmh_timing_collect(buf, len(buf))
    out = {}
    for line in buf.value.decode().splitlines():
        name, cnt, ms = line.split()
        out[name] = (int(cnt), float(ms))
    return out
