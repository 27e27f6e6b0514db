"""Checking helpers shared by the GPU tests (not a test module): random
61-bit primes and residue checks of products, run on a host thread pool
(the oracle's C calls release the GIL)."""

from __future__ import annotations

import os
import random
from concurrent.futures import ThreadPoolExecutor

SMALL_PRIMES = (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37)


def primes61(count: int, seed: int) -> list[int]:
    """`count` random primes in [2^60, 2^61) (Miller-Rabin with the first 12
    prime bases, deterministic below 3.3e24)."""
    rng = random.Random(seed)
    out = []
    while len(out) < count:
        q = rng.getrandbits(61) | (1 << 60) | 1
        d, s = q - 1, 0
        while d % 2 == 0:
            d //= 2
            s += 1
        ok = True
        for a in SMALL_PRIMES:
            x = pow(a, d, q)
            if x in (1, q - 1):
                continue
            for _ in range(s - 1):
                x = x * x % q
                if x == q - 1:
                    break
            else:
                ok = False
                break
        if ok:
            out.append(q)
    return out


def pool(workers: int | None = None) -> ThreadPoolExecutor:
    """Host threads for residue checks (at most 16, whatever the host reports:
    a machine with one GPU may give a process only 16 cores)."""
    return ThreadPoolExecutor(max_workers=workers or min(16, os.cpu_count() or 4))
