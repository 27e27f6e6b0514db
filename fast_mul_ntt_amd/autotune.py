"""Autotuner of the NTT pass geometry, one entry per transform length.

    python -m fast_mul_ntt_amd.autotune [--min-log 5] [--max-log 25] [--out PATH]
    python -m fast_mul_ntt_amd.autotune --batched-only [--out PATH]

For each size class 2^L it times a balanced product of two 2^(L-1)-limb
operands (device-resident, seeded operands) under candidate geometries and
keeps the fastest; for classes up to 2^21 it also tunes the geometry of
batched launches, timing one batch of many such products (the table's
"b L spec" lines, used when a launch carries several products of a class).
Below about 2^18 points one product takes about as long on the device as
this loop of Python calls takes to issue it, so differences of a few
microseconds there can be hidden; the table's round sizes of 2^12..2^16
came from a comparison driven from C.
The search is coordinate descent over the parameters of a geometry spec,
starting from the class's current entry of the table (with --from-defaults:
from the built-in defaults; batched classes then start from the class's
single-product entry):

    row_maxl     points per row-pass line          2^8 .. 2^12
    row_tile     points per row-pass tile          2^9 .. 2^13
    col_maxl     points per column-pass line       2^7 .. 2^11
    col_logc     adjacent lines per column tile    4 .. 32
    col_threads  threads per column workgroup      64 .. 1024
    row_threads  threads per row workgroup         256 .. 1024
    row_radix    stages per row-pass LDS round     2, 3 or 4
    col_dma      forward column tiles by LDS DMA    0 or 1

A candidate replaces the incumbent only if it is at least 3% faster in
interleaved measurements (candidate and incumbent timed alternately, the
minimum of several bursts each), which keeps clock drift on the device from
choosing geometries.  Every candidate's product is compared word for word
with the product under the default geometry, so a geometry that computed a
different result would stop the run.  Geometries the kernels cannot launch
(too much LDS) are skipped.  libfmh.so reads the table from
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
    "row_tile": [9, 10, 11, 12, 13],
    "col_maxl": [7, 8, 9, 10, 11],
    "col_logc": [2, 3, 4, 5],
    "col_threads": [64, 128, 256, 512, 1024],
    "row_threads": [256, 512, 1024],
    "row_radix": [2, 3, 4],
    "col_dma": [0, 1],
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
    """Balanced 2^(L-1) x 2^(L-1) multiplies resident on the device: one
    product, or (batched) a batch of `count` of them in one launch."""

    def __init__(self, L: int, stream: fm.Stream, count: int = 1):
        self.L = L
        self.n = 1 << (L - 1)
        self.s = stream
        self.batched = count > 1
        self.bufs = []
        self.items = []
        for i in range(count):
            a = fm.DeviceBuffer.from_host(splitmix64(1000 * L + 2 * i, self.n), stream)
            b = fm.DeviceBuffer.from_host(splitmix64(1000 * L + 2 * i + 1, self.n), stream)
            out = fm.DeviceBuffer(2 * self.n)
            self.bufs += [a, b, out]
            self.items.append((a, self.n, b, self.n, out))
        self.want = None
        self.per = 1

    def run_once(self) -> None:
        if self.batched:
            fm.mul_batch_device(self.items, self.s)
        else:
            a, n, b, _, out = self.items[0]
            fm.mul_device(a, n, b, n, out, self.s)

    def check(self, spec: str) -> bool:
        """Runs once under spec; False if the geometry cannot launch."""
        fm.set_tuning(self.L, spec, self.batched)
        try:
            self.run_once()
            self.s.synchronize()
        except fm.FmhError:
            return False
        got = np.concatenate([it[4].download() for it in self.items[:4]])
        if self.want is None:
            self.want = got
        elif not np.array_equal(got, self.want):
            raise AssertionError(f"2^{self.L}: geometry {spec} changed the product")
        return True

    def burst(self, spec: str) -> float:
        """Seconds per launch over one burst of `per` launches under spec."""
        fm.set_tuning(self.L, spec, self.batched)
        t0 = time.perf_counter()
        for _ in range(self.per):
            self.run_once()
        self.s.synchronize()
        return (time.perf_counter() - t0) / self.per

    def compare(self, inc: str, cand: str, reps: int) -> tuple[float, float]:
        """Interleaved minima (µs) of the incumbent and a candidate."""
        ti, tc = [], []
        for _ in range(reps):
            ti.append(self.burst(inc))
            tc.append(self.burst(cand))
        return min(ti) * 1e6, min(tc) * 1e6

    def free(self) -> None:
        for d in self.bufs:
            d.free()


def tune(L: int, stream: fm.Stream, reps: int, rounds: int, count: int = 1,
         from_table: bool = True) -> tuple[str, float, float]:
    case = Case(L, stream, count)
    try:
        if from_table:
            # the incumbent is the table's entry: a change must beat it
            base = parse(fm.get_tuning(L, case.batched))
        else:
            # batched classes start from the class's single-product geometry
            fm.set_tuning(L, "", case.batched)
            base = parse(fm.get_tuning(L, False) if case.batched else fm.get_tuning(L, case.batched))
        case.check(fmt(base))  # the reference product
        t = case.burst(fmt(base))
        case.per = max(1, int(3e-3 / t))
        default_us, _ = case.compare(fmt(base), fmt(base), reps)
        best = dict(base)
        for _ in range(rounds):
            improved = False
            for key, values in CANDIDATES.items():
                for v in values:
                    if v == best[key]:
                        continue
                    cand = dict(best)
                    cand[key] = v
                    if not case.check(fmt(cand)):
                        continue
                    inc_us, cand_us = case.compare(fmt(best), fmt(cand), reps)
                    if cand_us < 0.97 * inc_us:
                        best, improved = cand, True
            if not improved:
                break
        best_us, def_us = case.compare(fmt(best), fmt(base), reps)
        fm.set_tuning(L, fmt(best), case.batched)
        return fmt(best), best_us, def_us
    finally:
        case.free()


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--min-log", type=int, default=5)
    ap.add_argument("--max-log", type=int, default=25)
    ap.add_argument("--max-batch-log", type=int, default=21)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--rounds", type=int, default=2)
    ap.add_argument("--out", default=DEFAULT_TABLE)
    ap.add_argument("--batched-only", action="store_true",
                    help="tune only the batched classes; keep the single-product lines of --out")
    ap.add_argument("--from-defaults", action="store_true",
                    help="start each class from the built-in defaults instead of its table entry")
    args = ap.parse_args()
    if fm.lib().fmh_init(0) != 0:
        raise fm.FmhError(fm.lib().fmh_last_error().decode())
    stream = fm.Stream()
    lines = ["# NTT pass geometry per transform length 2^L (written by fast_mul_ntt_amd/autotune.py)",
             "# L spec      # us tuned / us start, one balanced 2^(L-1) x 2^(L-1) multiply",
             "# b L spec    # the same for one batched launch of max(8, 2^22 / 2^L) such products",
             "# (start: the previous entry of this table, or the defaults with --from-defaults)"]
    if args.batched_only:
        with open(args.out) as f:
            lines += [ln.rstrip("\n") for ln in f if ln[:1].isdigit()]
    else:
        for L in range(args.min_log, args.max_log + 1):
            spec, us, d_us = tune(L, stream, args.reps, args.rounds, from_table=not args.from_defaults)
            lines.append(f"{L} {spec}  # {us:.1f} / {d_us:.1f}")
            print(lines[-1], flush=True)
    for L in range(args.min_log, args.max_batch_log + 1):
        count = max(8, (1 << 22) >> L)
        spec, us, d_us = tune(L, stream, args.reps, args.rounds, count, from_table=not args.from_defaults)
        lines.append(f"b {L} {spec}  # {us:.1f} / {d_us:.1f}")
        print(lines[-1], flush=True)
    stream.close()
    with open(args.out, "w") as f:
        f.write("\n".join(lines) + "\n")
    print(f"wrote {args.out}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
