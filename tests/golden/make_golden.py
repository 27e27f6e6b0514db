#!/usr/bin/env python3
"""Write the golden vectors of tests/golden/ from the reference's own code.

Runs in the development container only (it needs oracle/_ref, which
oracle/Makefile compiles from /root/reference):
  * ref_core_vectors.json   words produced by the reference's src/core.h
                            (data.h constants, Montgomery macros,
                            cover_nttshort_i tables, dif_i / idit_i, crt3)
  * fft_small_products.json products of the reference's fft_small
                            abs_mul64 / abs_sqr64 (its second library,
                            same ABI as src/fast_mul.h)
Operands are splitmix64 streams named by their seed; large outputs are kept
as SHA-256 digests of their little-endian bytes.

    python tests/golden/make_golden.py
"""

from __future__ import annotations

import hashlib
import json
import os
import random
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle.oracle import PRIMES, RefCore, RefFftSmall  # noqa: E402


def splitmix(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.uint64)
    s = seed
    for k in range(n):
        s = (s + 0x9E3779B97F4A7C15) & (2**64 - 1)
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & (2**64 - 1)
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & (2**64 - 1)
        out[k] = z ^ (z >> 31)
    if n and out[-1] == 0:
        out[-1] = 1
    return out


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<u8").tobytes()).hexdigest()


def words(a) -> list[str]:
    return [str(int(v)) for v in a]


def main() -> None:
    ref = RefCore()
    fsd = RefFftSmall()
    rng = random.Random(20260912)
    core = {"primes": [str(p) for p in PRIMES], "constants": {}, "mont": [], "tables": {},
            "dif": [], "crt": []}
    for i in range(3):
        core["constants"][str(i)] = words(ref.constants(i))
        p = PRIMES[i]
        for _ in range(40):
            a, b = rng.randrange(2 * p), rng.randrange(2 * p)
            x = rng.getrandbits(64)
            core["mont"].append([i, str(a), str(b), str(ref.mont_mul(i, a, b)),
                                 str(ref.mont_mulinto(i, a, b)), str(x), str(ref.tomont(i, x)),
                                 str(ref.toint(i, a))])
        om, iom = ref.table(i, 6)
        om17, iom17 = ref.table(i, 17)
        core["tables"][str(i)] = {"log6_omega": words(om), "log6_iomega": words(iom),
                                  "log17_omega_sha256": sha(om17), "log17_iomega_sha256": sha(iom17)}
        for lg in list(range(0, 7)) + [8, 10, 12, 17]:
            n = 1 << lg
            seed = 1000 * i + lg
            x = np.array([ref.tomont(i, int(v)) for v in splitmix(seed, n)], dtype=np.uint64)
            f = ref.dif_raw(i, x)
            g = ref.idit_raw(i, f)
            entry = {"prime": i, "log_n": lg, "seed": seed, "input_sha256": sha(x),
                     "dif_sha256": sha(f), "idit_of_dif_sha256": sha(g)}
            if lg <= 6:
                entry["dif"] = words(f)
                entry["idit_of_dif"] = words(g)
            core["dif"].append(entry)
    for _ in range(200):
        a = rng.randrange(2 * PRIMES[0])
        b = rng.randrange(2 * PRIMES[1])
        c = rng.randrange(2 * PRIMES[2])
        core["crt"].append([str(a), str(b), str(c), str(ref.crt3(a, b, c))])
    with open(os.path.join(HERE, "ref_core_vectors.json"), "w") as f:
        json.dump(core, f, indent=0)

    sizes = [1, 2, 3, 5, 17, 33, 64, 100, 255, 256, 257, 300, 511, 512, 513, 700, 1024, 1100,
             1500, 2048, 2500, 2900]
    prods = []
    seed = 5000
    for si, an in enumerate(sizes):
        for bn in sizes[si::4]:
            a, b = splitmix(seed, an), splitmix(seed + 1, bn)
            c = fsd.mul(a, b)
            e = {"op": "mul", "n1": an, "n2": bn, "seed1": seed, "seed2": seed + 1,
                 "sha256": sha(c)}
            if an + bn <= 16:
                e["out"] = words(c)
            prods.append(e)
            seed += 2
        a = splitmix(seed, an)
        prods.append({"op": "sqr", "n1": an, "seed1": seed, "sha256": sha(fsd.sqr(a))})
        seed += 1
    for an, bn, s in [(70000, 70000, 9001), (100000, 31, 9003), (262144, 262144, 9005),
                      (1 << 20, 1 << 20, 2001)]:
        a, b = splitmix(s, an), splitmix(s + 1, bn)
        prods.append({"op": "mul", "n1": an, "n2": bn, "seed1": s, "seed2": s + 1,
                      "sha256": sha(fsd.mul(a, b))})
    ones = np.full(3000, 0xFFFFFFFFFFFFFFFF, dtype=np.uint64)
    prods.append({"op": "mul_all_ones", "n1": 3000, "n2": 2000, "sha256": sha(fsd.mul(ones, ones[:2000]))})
    with open(os.path.join(HERE, "fft_small_products.json"), "w") as f:
        json.dump(prods, f, indent=0)


if __name__ == "__main__":
    main()
