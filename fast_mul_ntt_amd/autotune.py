"""Autotuner of the NTT pass geometry, one entry per transform length.

    python -m fast_mul_ntt_amd.autotune [--min-log 10] [--max-log 25] [--out PATH]

For each size class 2^L it times a balanced product of two 2^(L-1)-limb
operands (device-resident, seeded operands) under candidate geometries and
keeps the fastest.  The search is coordinate descent over the parameters of a
geometry spec, starting from the current entry:

    row_maxl     points per row-pass line          2^8 .. 2^12
    row_tile     points per row-pass tile          2^10 .. 2^13
    col_maxl     points per column-pass line       2^7 .. 2^11
    col_logc     adjacent lines per column tile    4 .. 16
    col_threads  threads per column workgroup      256 .. 1024
    row_threads  threads per row workgroup         256 .. 1024
    row_radix    stages per row-pass LDS round     3 or 4

Every candidate's product is compared word for word with the product under
the default geometry, so a geometry that computed a different result would
stop the run.  Geometries the kernels cannot launch (too much LDS) are
skipped.  The table is written as "L spec" lines; libfmh.so reads it from
tuning_gfx950.txt beside itself (or $FMH_TUNE_TABLE) at start-up.
"""

from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if os.path.dirname(HERE) not in sys.path:
    sys.path.insert(0, os.path.dirname(HERE))

import fast_mul_ntt_amd as fm  # noqa: E402

CANDIDATES = {
    "row_maxl": [8, 9, 10, 11, 12],
    "row_tile": [10, 11, 12, 13],
    "col_maxl": [7, 8, 9, 10, 11],
    "col_logc": [2, 3, 4],
    "col_threads": [256, 512, 1024],
    "row_threads": [256, 512, 1024],
    "row_radix": [3, 4],
}
DEFAULT_TABLE = os.path.join(HERE, "tuning_gfx950.txt")


def splitmix64(seed: int, n: int) -> np.ndarray:
    gamma = np.uint64(0x9E3779B97F4A7C15)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + gamma * np.arange(1, n + 1, dtype=np.uint64)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def parse(spec: str) -> dict[str, int]:
    return {k: int(v) for k, v in (kv.split("=") for kv in spec.split(",") if kv)}


def fmt(d: dict[str, int]) -> str:
    return ",".join(f"{k}={d[k]}" for k in CANDIDATES)


class Case:
    """A balanced 2^(L-1) x 2^(L-1) multiply, resident on the device."""

    def __init__(self, L: int, stream: fm.Stream):
        self.L = L
        self.n = 1 << (L - 1)
        self.s = stream
        self.a = fm.DeviceBuffer.from_host(splitmix64(1000 + L, self.n), stream)
        self.b = fm.DeviceBuffer.from_host(splitmix64(2000 + L, self.n), stream)
        self.out = fm.DeviceBuffer(2 * self.n)
        self.want = None

    def run_once(self) -> None:
        fm.mul_device(self.a, self.n, self.b, self.n, self.out, self.s)

    def time(self, spec: str, reps: int) -> float | None:
        """Median µs of one product under spec (None: not launchable)."""
        fm.set_tuning(self.L, spec)
        try:
            self.run_once()
            self.s.synchronize()
        except fm.FmhError:
            return None
        got = self.out.download()
        if self.want is None:
            self.want = got
        elif not np.array_equal(got, self.want):
            raise AssertionError(f"2^{self.L}: geometry {spec} changed the product")
        per = max(1, int(2e-3 / max(self._last or 1e-4, 1e-6)))
        samples = []
        for _ in range(reps):
            t0 = time.perf_counter()
            for _ in range(per):
                self.run_once()
            self.s.synchronize()
            samples.append((time.perf_counter() - t0) / per)
        self._last = float(np.median(samples))
        return self._last * 1e6

    _last = None

    def free(self) -> None:
        for d in (self.a, self.b, self.out):
            d.free()


def tune_size(L: int, stream: fm.Stream, reps: int, rounds: int) -> tuple[str, float, float]:
    case = Case(L, stream)
    try:
        fm.set_tuning(L, "")
        base = parse(fm.get_tuning(L))
        default_us = case.time(fmt(base), reps)  # also the reference product
        best, best_us = dict(base), default_us
        for _ in range(rounds):
            improved = False
            for key, values in CANDIDATES.items():
                for v in values:
                    if v == best[key]:
                        continue
                    cand = dict(best)
                    cand[key] = v
                    us = case.time(fmt(cand), reps)
                    if us is not None and us < best_us * 0.99:
                        best, best_us, improved = cand, us, True
            if not improved:
                break
        fm.set_tuning(L, fmt(best))
        return fmt(best), best_us, default_us
    finally:
        case.free()


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--min-log", type=int, default=10)
    ap.add_argument("--max-log", type=int, default=25)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--rounds", type=int, default=2)
    ap.add_argument("--out", default=DEFAULT_TABLE)
    args = ap.parse_args()
    if fm.lib().fmh_init(0) != 0:
        raise fm.FmhError(fm.lib().fmh_last_error().decode())
    stream = fm.Stream()
    lines = ["# NTT pass geometry per transform length 2^L (written by fast_mul_ntt_amd/autotune.py)",
             "# L spec  # us tuned / us default, balanced 2^(L-1) x 2^(L-1) multiply"]
    for L in range(args.min_log, args.max_log + 1):
        spec, us, d_us = tune_size(L, stream, args.reps, args.rounds)
        lines.append(f"{L} {spec}  # {us:.1f} / {d_us:.1f}")
        print(lines[-1], flush=True)
    stream.close()
    with open(args.out, "w") as f:
        f.write("\n".join(lines) + "\n")
    print(f"wrote {args.out}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
