"""ctypes binding of the CPU oracle (oracle/liboracle.so) and, where built, of
the reference's own code compiled from its sources (oracle/_ref/*.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() (as
the checker) and bench.py's cpu_baseline leg (as the timed CPU baseline).
The product library never imports it.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_CORE_SO = os.path.join(HERE, "_ref", "libref_core.so")
REF_FFT_SMALL_SO = os.path.join(HERE, "_ref", "libref_fft_small.so")

_u64 = ctypes.c_uint64
_pu64 = ctypes.POINTER(ctypes.c_uint64)

PRIMES = (2485986994308513793, 1945555039024054273, 4179340454199820289)
GENERATORS = (5, 5, 3)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_pu64)


def _limbs(x) -> np.ndarray:
    return np.ascontiguousarray(x, dtype=np.uint64)


class Oracle:
    """The repository's C restatement of src/ (see ntt3_oracle.h)."""

    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        L = ctypes.CDLL(path)
        L.orc_prime_constants.argtypes = [ctypes.c_int, _pu64]
        for fn in ("orc_mont_mul", "orc_mont_mulinto", "orc_mont_pow"):
            getattr(L, fn).argtypes = [ctypes.c_int, _u64, _u64]
            getattr(L, fn).restype = _u64
        for fn in ("orc_tomont", "orc_toint"):
            getattr(L, fn).argtypes = [ctypes.c_int, _u64]
            getattr(L, fn).restype = _u64
        L.orc_twiddle_table.argtypes = [ctypes.c_int, ctypes.c_int, _pu64, _pu64]
        L.orc_dif_raw.argtypes = [ctypes.c_int, _pu64, _u64]
        L.orc_idit_raw.argtypes = [ctypes.c_int, _pu64, _u64]
        L.orc_forward_spectrum.argtypes = [ctypes.c_int, _pu64, _u64, _u64, _pu64]
        L.orc_crt3.argtypes = [_u64, _u64, _u64, _pu64]
        L.orc_abs_mul64.argtypes = [_pu64, _u64, _pu64, _u64, _pu64]
        L.orc_abs_sqr64.argtypes = [_pu64, _u64, _pu64]
        L.orc_schoolbook.argtypes = [_pu64, _u64, _pu64, _u64, _pu64]
        L.orc_mod_q.argtypes = [_pu64, _u64, _u64]
        L.orc_mod_q.restype = _u64
        L.orc_mulmod_q.argtypes = [_u64, _u64, _u64]
        L.orc_mulmod_q.restype = _u64
        L.orc_prefix_mod_q.argtypes = [_pu64, _u64, _u64, _pu64]
        L.orc_splitmix64_fill.argtypes = [_u64, _pu64, _u64, ctypes.c_int]
        L.orc_plant.argtypes = [ctypes.c_int]
        self.L = L

    def plant(self, kind: int) -> None:
        """Plant a mistake in the oracle (tests of the checks; 0 = none)."""
        self.L.orc_plant(kind)

    def constants(self, prime: int) -> list[int]:
        out = np.zeros(17, dtype=np.uint64)
        self.L.orc_prime_constants(prime, _ptr(out))
        return [int(v) for v in out]

    def mont_mul(self, prime, a, b):
        return int(self.L.orc_mont_mul(prime, a, b))

    def mont_mulinto(self, prime, a, b):
        return int(self.L.orc_mont_mulinto(prime, a, b))

    def tomont(self, prime, x):
        return int(self.L.orc_tomont(prime, x))

    def toint(self, prime, x):
        return int(self.L.orc_toint(prime, x))

    def table(self, prime: int, log_len: int):
        n = 1 << log_len
        om, iom = np.zeros(n, np.uint64), np.zeros(n, np.uint64)
        self.L.orc_twiddle_table(prime, log_len, _ptr(om), _ptr(iom))
        return om, iom

    def dif_raw(self, prime: int, data) -> np.ndarray:
        a = _limbs(data).copy()
        self.L.orc_dif_raw(prime, _ptr(a), a.size)
        return a

    def idit_raw(self, prime: int, data) -> np.ndarray:
        a = _limbs(data).copy()
        self.L.orc_idit_raw(prime, _ptr(a), a.size)
        return a

    def forward_spectrum(self, prime: int, limbs, ntt_len: int) -> np.ndarray:
        a = _limbs(limbs)
        out = np.zeros(ntt_len, np.uint64)
        self.L.orc_forward_spectrum(prime, _ptr(a), a.size, ntt_len, _ptr(out))
        return out

    def crt3(self, a, b, c) -> int:
        out = np.zeros(3, np.uint64)
        self.L.orc_crt3(a, b, c, _ptr(out))
        return int(out[0]) | (int(out[1]) << 64) | (int(out[2]) << 128)

    def mul(self, a, b) -> np.ndarray:
        a, b = _limbs(a), _limbs(b)
        out = np.zeros(a.size + b.size, np.uint64)
        self.L.orc_abs_mul64(_ptr(a), a.size, _ptr(b), b.size, _ptr(out))
        return out

    def sqr(self, a) -> np.ndarray:
        a = _limbs(a)
        out = np.zeros(2 * a.size, np.uint64)
        self.L.orc_abs_sqr64(_ptr(a), a.size, _ptr(out))
        return out

    def schoolbook(self, a, b) -> np.ndarray:
        a, b = _limbs(a), _limbs(b)
        out = np.zeros(a.size + b.size, np.uint64)
        self.L.orc_schoolbook(_ptr(a), a.size, _ptr(b), b.size, _ptr(out))
        return out

    def mod_q(self, a, q: int) -> int:
        a = _limbs(a)
        return int(self.L.orc_mod_q(_ptr(a), a.size, q))

    def mulmod_q(self, a: int, b: int, q: int) -> int:
        return int(self.L.orc_mulmod_q(a, b, q))

    def prefix_mod_q(self, a, q: int) -> np.ndarray:
        """Residues mod q of every prefix a[:k], k = 0..len(a)."""
        a = _limbs(a)
        out = np.zeros(a.size + 1, np.uint64)
        self.L.orc_prefix_mod_q(_ptr(a), a.size, q, _ptr(out))
        return out

    @staticmethod
    def slice_mod_q(prefix: np.ndarray, offset: int, n: int, q: int) -> int:
        """Residue mod q of the number a[offset:offset + n], from prefix_mod_q(a, q)."""
        d = (int(prefix[offset + n]) - int(prefix[offset])) % q
        return d * pow(pow(2, 64, q), -offset, q) % q

    def splitmix(self, seed: int, n: int, top_nonzero: bool = True) -> np.ndarray:
        out = np.zeros(n, np.uint64)
        if n:
            self.L.orc_splitmix64_fill(seed, _ptr(out), n, 1 if top_nonzero else 0)
        return out


class RefCore:
    """The reference's src/core.h compiled from its sources (oracle/_ref)."""

    def __init__(self, path: str = REF_CORE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = ctypes.CDLL(path)
        L.ref_constants.argtypes = [ctypes.c_int, _pu64]
        for fn in ("ref_mont_mul", "ref_mont_mulinto"):
            getattr(L, fn).argtypes = [ctypes.c_int, _u64, _u64]
            getattr(L, fn).restype = _u64
        for fn in ("ref_tomont", "ref_toint"):
            getattr(L, fn).argtypes = [ctypes.c_int, _u64]
            getattr(L, fn).restype = _u64
        L.ref_table.argtypes = [ctypes.c_int, ctypes.c_int, _pu64, _pu64]
        L.ref_dif.argtypes = [ctypes.c_int, _pu64, _u64]
        L.ref_idit.argtypes = [ctypes.c_int, _pu64, _u64]
        L.ref_crt3.argtypes = [_u64, _u64, _u64, _pu64]
        self.L = L

    def constants(self, prime: int) -> list[int]:
        out = np.zeros(17, dtype=np.uint64)
        self.L.ref_constants(prime, _ptr(out))
        return [int(v) for v in out]

    def mont_mul(self, prime, a, b):
        return int(self.L.ref_mont_mul(prime, a, b))

    def mont_mulinto(self, prime, a, b):
        return int(self.L.ref_mont_mulinto(prime, a, b))

    def tomont(self, prime, x):
        return int(self.L.ref_tomont(prime, x))

    def toint(self, prime, x):
        return int(self.L.ref_toint(prime, x))

    def table(self, prime: int, log_len: int):
        n = 1 << log_len
        om, iom = np.zeros(n, np.uint64), np.zeros(n, np.uint64)
        self.L.ref_table(prime, log_len, _ptr(om), _ptr(iom))
        return om, iom

    def dif_raw(self, prime: int, data) -> np.ndarray:
        a = _limbs(data).copy()
        if self.L.ref_dif(prime, _ptr(a), a.size) != 0:
            raise ValueError("length above 2^17")
        return a

    def idit_raw(self, prime: int, data) -> np.ndarray:
        a = _limbs(data).copy()
        if self.L.ref_idit(prime, _ptr(a), a.size) != 0:
            raise ValueError("length above 2^17")
        return a

    def crt3(self, a, b, c) -> int:
        out = np.zeros(3, np.uint64)
        self.L.ref_crt3(a, b, c, _ptr(out))
        return int(out[0]) | (int(out[1]) << 64) | (int(out[2]) << 128)


class RefFftSmall:
    """The reference's second library fft_small/ (same abs_mul64 ABI),
    compiled from its sources (oracle/_ref)."""

    def __init__(self, path: str = REF_FFT_SMALL_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        L = ctypes.CDLL(path)
        L.abs_mul64.argtypes = [_pu64, _u64, _pu64, _u64, _pu64]
        L.abs_sqr64.argtypes = [_pu64, _u64, _pu64]
        self.L = L

    def mul(self, a, b) -> np.ndarray:
        a, b = _limbs(a), _limbs(b)
        out = np.zeros(a.size + b.size, np.uint64)
        self.L.abs_mul64(_ptr(a), a.size, _ptr(b), b.size, _ptr(out))
        return out

    def sqr(self, a) -> np.ndarray:
        a = _limbs(a)
        out = np.zeros(2 * a.size, np.uint64)
        self.L.abs_sqr64(_ptr(a), a.size, _ptr(out))
        return out


def limbs_to_int(a) -> int:
    """Little-endian u64 limbs -> Python int."""
    return int.from_bytes(_limbs(a).astype("<u8").tobytes(), "little")


def int_to_limbs(x: int, n: int) -> np.ndarray:
    return np.frombuffer(x.to_bytes(8 * n, "little"), dtype="<u8").astype(np.uint64)
