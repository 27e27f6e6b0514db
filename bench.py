#!/usr/bin/env python3
"""Benchmark: one device-resident 1,048,576 x 1,048,576-limb abs_mul64 per step.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]

Workload: two seeded-uniform operands of 2^20 u64 limbs each (splitmix64,
seeds 2001 / 2002, top limb forced nonzero; NTT length 2^21, three primes),
already resident in HBM.  A step is one fmh_mul_device() call (the whole
pipeline: transforms, pointwise product, inverse transforms, CRT + carry).
Timed: exactly K steps between stream synchronisations.  Prints one JSON line.

A breakdown pass before the timed region records HIP events around every
kernel (on each kernel's launch stream) and picks the dominant kernel; during
the timed steps the library records events around that kernel only, and its
average duration there gives `roofline.achieved`.  The product of the last step
is checked bit for bit against the CPU oracle, which is also timed as the CPU
baseline (single thread, same operands).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from fast_mul_ntt_amd.workloads import (batch_operand, batch_products, serve_requests,  # noqa: E402
                                        serve_source, splitmix64)

METRIC = "µs per 2^20×2^20-limb abs_mul64 (bit-exact); mixed-size products/s + p99"
METRIC_16M = "µs per 2^24×2^24-limb abs_mul64 (bit-exact)"
N1 = N2 = 1 << 20
SEED_A, SEED_B = 2001, 2002
HBM_PEAK_GBS = 8000.0
# Integer VALU: 256 CUs x 64 lanes x 2.4 GHz full-rate lane operations per second.
VALU_LANE_OPS = 256 * 64 * 2.4e9
# HBM bytes per launch (FETCH_SIZE x 2 + WRITE_SIZE, the factors calibrated
# against the copy kernel) from rocprofv3 --pmc runs of this workload, and the
# same runs' VALUBusy per kernel (% of one VALU instruction per SIMD every 4
# cycles); profiles/abs_mul64_1Mx1M_kernels.md holds them.
TRAFFIC: dict[str, float] = {
    "col_fwd": 179470551.0,
    "row_mul": 157022672.0,
    "col_inv": 161079727.0,
    "crt_local": 70193951.0,
    "carry_apply": 2386741.0,
}
VALU_BUSY: dict[str, float] = {"col_fwd": 78.5, "row_mul": 76.4, "col_inv": 52.8, "crt_local": 57.8, "carry_apply": 12.4}
PROFILE = "profiles/abs_mul64_1Mx1M_kernels.md"
# The same for the 2^24 x 2^24 product (NTT 2^25; per launch, averaged over
# the two column passes): profiles/abs_mul64_16Mx16M_kernels.md.
TRAFFIC_16M: dict[str, float] = {
    "col_fwd": 3352976058.0,
    "row_mul": 2423194807.0,
    "col_inv": 2412967116.0,
    "crt_local": 1120974513.0,
    "chunk_scan": 269163.0,
    "carry_apply": 38288730.0,
}
VALU_BUSY_16M: dict[str, float] = {"col_fwd": 77.2, "row_mul": 86.9, "col_inv": 67.5, "crt_local": 109.1,
                                   "chunk_scan": 0.0, "carry_apply": 20.3}
PROFILE_16M = "profiles/abs_mul64_16Mx16M_kernels.md"


def kernel_alg_bytes(n1: int, n2: int, log_n: int, col_passes: int = 1) -> dict[str, float]:
    """Algorithmic HBM bytes per launch of each kernel of one multiply
    (averaged over the launches of a kernel when there are several column
    passes).

    col_fwd     the first pass reads both operands' limbs and writes 3 primes
                x 2 operands x N words; each further pass reads and writes them
    row_mul     reads those 6 residue arrays, writes 3 (the pointwise product,
                inverse row stages done)
    col_inv     each pass reads and writes 3 residue arrays
    crt_local   reads 3 residues per convolution coefficient, writes the
                limbs and one carry-map byte per limb
    carry_apply reads the carry-map bytes, and read-modify-writes the limbs
                whose carry-in map is not constant: for operands like these
                (random words) only each 1024-limb chunk's first limb
    """
    N = 1 << log_n
    conv = n1 + n2 - 1
    out = n1 + n2
    P = max(1, col_passes)
    return {
        "col_fwd": (8.0 * (n1 + n2) + 6 * 8.0 * N + (P - 1) * 12 * 8.0 * N) / P,
        "row_mul": 9 * 8.0 * N,
        "col_inv": 6 * 8.0 * N,
        "crt_local": 3 * 8.0 * conv + 9.0 * out,
        "carry_apply": 1.0 * out + 16.0 * ((out + 1023) // 1024),
    }


def modmul_count(n1: int, n2: int, log_n: int) -> float:
    """64-bit modular products per multiply (radix-2 count, 3 primes)."""
    N = 1 << log_n
    return 3 * (1.5 * N * log_n + N) + 9 * (n1 + n2)


def cpu_baseline(a: np.ndarray, b: np.ndarray, budget_s: float = 12.0):
    """Time the oracle (C restatement of the reference src/ path, 1 thread)."""
    from oracle.oracle import Oracle

    orc = Oracle()
    times = []
    prod = None
    t_start = time.perf_counter()
    while True:
        t0 = time.perf_counter()
        prod = orc.mul(a, b)
        times.append(time.perf_counter() - t0)
        if len(times) >= 3 and time.perf_counter() - t_start > budget_s:
            break
        if len(times) >= 20:
            break
    return prod, times


def cpu_fft_small(a: np.ndarray, b: np.ndarray, want: np.ndarray, budget_s: float = 8.0):
    """Time the reference's second library, fft_small (FLINT small-prime FFT,
    built from the reference's sources into oracle/_ref), on the same operands,
    single thread; None where it was not built."""
    try:
        from oracle.oracle import RefFftSmall
        ref = RefFftSmall()
    except (ImportError, OSError):
        return None
    got = ref.mul(a, b)  # first call: fft_small builds its context
    if not np.array_equal(got, want):
        raise AssertionError("fft_small product differs from the oracle's")
    times = []
    t_start = time.perf_counter()
    while len(times) < 3 or (time.perf_counter() - t_start < budget_s and len(times) < 20):
        t0 = time.perf_counter()
        ref.mul(a, b)
        times.append(time.perf_counter() - t0)
    return times


def run_batch(args, fm) -> dict:
    """4,096 products, operand lengths log-uniform in [64, 65,536] limbs drawn
    independently, 25 % squarings (sampler seed 4001), seeded-uniform limbs:
    one planned batch (graph replay) against 4,096 plain device calls."""
    l1, l2, sq = batch_products()
    count = l1.size
    stream = fm.Stream()
    items, bufs = [], []
    for i in range(count):
        da = fm.DeviceBuffer.from_host(batch_operand(i, 0, int(l1[i])), stream)
        db = None if sq[i] else fm.DeviceBuffer.from_host(batch_operand(i, 1, int(l2[i])), stream)
        n_out = 2 * int(l1[i]) if sq[i] else int(l1[i] + l2[i])
        do = fm.DeviceBuffer(n_out)
        bufs += [da, do] + ([db] if db is not None else [])
        items.append((da, int(l1[i]), db, 0 if sq[i] else int(l2[i]), do))
    limbs = int(sum(2 * l1[i] if sq[i] else l1[i] + l2[i] for i in range(count)))

    def plain():
        for (da, na, db, nb, do) in items:
            if db is None:
                fm.sqr_device(da, na, do, stream)
            else:
                fm.mul_device(da, na, db, nb, do, stream)

    for _ in range(args.warmup):
        plain()
    stream.synchronize()
    t0 = time.perf_counter()
    reps = max(1, args.steps // 4)
    for _ in range(reps):
        plain()
    stream.synchronize()
    t_plain = (time.perf_counter() - t0) / reps
    ref_out = [do.download(na + (nb if db is not None else na)) for (da, na, db, nb, do) in items[:64]]

    batch = fm.Batch(items)
    for _ in range(args.warmup):
        batch.run(stream)
    stream.synchronize()
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        batch.run(stream)
        stream.synchronize()
        lat.append(time.perf_counter() - s0)
    t_batch = (time.perf_counter() - t0) / args.steps
    for (da, na, db, nb, do), want in zip(items[:64], ref_out):
        if not np.array_equal(do.download(want.size), want):
            raise AssertionError("batched product differs from the plain call")
    batch.close()
    for d in bufs:
        d.free()
    stream.close()
    return {
        "metric": "mixed-size products/s (4,096 products, 64..65,536 limbs, 25% squares)",
        "value": round(count / t_batch, 1),
        "unit": "products/s",
        "batch_ms": round(t_batch * 1e3, 3),
        "batch_p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
        "plain_calls_ms": round(t_plain * 1e3, 3),
        "plain_products_per_s": round(count / t_plain, 1),
        "speedup_vs_plain": round(t_plain / t_batch, 2),
        "output_limbs": limbs,
        "steps": args.steps,
        "warmup": args.warmup,
        "higher_is_better": True,
        "dtype": "u64",
        "data": "synthetic (splitmix64, lengths log-uniform, sampler seed 4001)",
        "config": {"workload": "batch of 4096 mixed-size products, hipGraph replay"},
    }


def run_serve(args, fm) -> dict:
    """20,000 multiply/square requests through the request loop (64 GiB pool,
    continuous batching), against plain per-request device calls and against
    the reference's interface (host-pointer abs_mul64 per request, sampled)."""
    count = args.requests
    l1, l2, sq, o1, o2, gaps, src_len = serve_requests(count)
    stream = fm.Stream()
    src_host = serve_source(src_len)
    src = fm.DeviceBuffer.from_host(src_host, stream)
    ops = []
    for i in range(count):
        a = src.view(int(o1[i]), int(l1[i]))
        b = None if sq[i] else src.view(int(o2[i]), int(l2[i]))
        ops.append((a, b))
    out_max = int(max(2 * l1[i] if sq[i] else l1[i] + l2[i] for i in range(count)))
    out = fm.DeviceBuffer(out_max)

    def plain(i):
        a, b = ops[i]
        if b is None:
            fm.sqr_device(a, a.n, out, stream)
        else:
            fm.mul_device(a, a.n, b, b.n, out, stream)

    for i in range(min(8, count)):
        plain(i)
    stream.synchronize()
    t0 = time.perf_counter()
    for i in range(count):
        plain(i)
    stream.synchronize()
    t_plain = time.perf_counter() - t0
    plain_rate = count / t_plain

    # The reference's own interface: host-pointer abs_mul64 / abs_sqr64 per
    # request (PCIe both ways included), on every 40th request.
    host_idx = list(range(0, count, 40))
    t0 = time.perf_counter()
    for i in host_idx:
        a = src_host[o1[i]:o1[i] + l1[i]]
        if sq[i]:
            fm.abs_sqr64(a)
        else:
            fm.abs_mul64(a, src_host[o2[i]:o2[i] + l2[i]])
    host_rate = len(host_idx) / (time.perf_counter() - t0)

    srv = fm.Server(pool_bytes=args.pool_gib << 30, streams=args.serve_streams)
    check_every = max(1, count // 40)
    checked = {}

    def drive(rate: float | None):
        """Submit every request (at Poisson arrival times if rate is given),
        releasing outputs as they complete; returns submission lateness (s)."""
        srv.reset_stats()
        late = []
        ids = {}
        t_start = time.perf_counter()
        due = t_start + (np.cumsum(gaps) / rate if rate else np.zeros(count))
        i = 0
        outstanding = 0
        while i < count or outstanding:
            now = time.perf_counter()
            while i < count and due[i] <= now:
                a, b = ops[i]
                rid = srv.submit(a, a.n, b, 0 if b is None else b.n)
                late.append(time.perf_counter() - due[i])
                ids[rid] = i
                i += 1
                outstanding += 1
            for rid in srv.poll():
                k = ids.pop(rid)
                if rate and k % check_every == 0:
                    checked[k] = srv.result(rid)
                srv.release(rid)
                outstanding -= 1
            if i < count and due[i] > time.perf_counter() + 2e-4:
                time.sleep(1e-4)
        return late

    drive(None)  # closed loop: capacity
    cap = srv.stats()
    cap_classes = srv.class_report()
    rate = args.load * cap["products_per_s"]
    late = drive(rate)
    st = srv.stats()
    classes = srv.class_report()
    srv.close()

    def top_classes(rep, k=8):
        kinds = {0: "mul", 1: "sqr", 2: "mul_spec", 3: "sqr_spec", 4: "split"}
        tot = sum(c["device_ms"] for c in rep) or 1.0
        rep = sorted(rep, key=lambda c: -c["device_ms"])[:k]
        return [{"class": f"{kinds.get(c['kind'], c['kind'])} 2^{c['log_n']}", "requests": c["requests"],
                 "groups": c["groups"], "device_ms": round(c["device_ms"], 1),
                 "share": round(c["device_ms"] / tot, 3)} for c in rep]
    # served outputs equal the products computed alone
    for k, got in checked.items():
        plain(k)
        stream.synchronize()
        if not np.array_equal(out.download(got.size), got):
            raise AssertionError(f"served request {k} differs from the plain call")
    out.free()
    src.free()
    stream.close()
    limbs = int(sum(2 * l1[i] if sq[i] else l1[i] + l2[i] for i in range(count)))
    return {
        "metric": "served products/s (Poisson arrivals) + p99 latency",
        "value": round(st["products_per_s"], 1),
        "unit": "products/s",
        "p50_ms": round(st["p50_us"] / 1e3, 3),
        "p99_ms": round(st["p99_us"] / 1e3, 3),
        "max_ms": round(st["max_us"] / 1e3, 3),
        "offered_rate": round(rate, 1),
        "load": args.load,
        "capacity_products_per_s": round(cap["products_per_s"], 1),
        "capacity_p99_ms": round(cap["p99_us"] / 1e3, 3),
        "plain_products_per_s": round(plain_rate, 1),
        "speedup_vs_plain": round(st["products_per_s"] / plain_rate, 2),
        "capacity_vs_plain": round(cap["products_per_s"] / plain_rate, 2),
        "plain_host_abi_products_per_s": round(host_rate, 1),
        "plain_host_abi_sample": f"every 40th request ({len(host_idx)}), host-pointer abs_mul64/abs_sqr64",
        "speedup_vs_plain_host_abi": round(st["products_per_s"] / host_rate, 2),
        "capacity_vs_plain_host_abi": round(cap["products_per_s"] / host_rate, 2),
        "mean_group": round(st["mean_group"], 2),
        "groups": st["groups"],
        "capacity_mean_group": round(cap["mean_group"], 2),
        "device_busy_frac": round(st["device_busy_s"] / st["span_s"], 3) if st["span_s"] else None,
        "capacity_device_busy_frac": round(cap["device_busy_s"] / cap["span_s"], 3) if cap["span_s"] else None,
        "capacity_classes_by_device_time": top_classes(cap_classes),
        "classes_by_device_time": top_classes(classes),
        "pool_gib": args.pool_gib,
        "serve_streams": args.serve_streams,
        "pool_peak_gib": round(st["pool_peak_bytes"] / 2**30, 3),
        "submit_lateness_p99_ms": round(float(np.percentile(late, 99)) * 1e3, 3),
        "checked_vs_plain": len(checked),
        "requests": count,
        "output_limbs": limbs,
        "higher_is_better": True,
        "dtype": "u64",
        "data": "synthetic (splitmix64 seed 5001 source array, lengths/offsets/arrivals seed 5001)",
        "config": {"workload": f"serve {count} requests, 256..4194304 limbs, 20% squares, "
                               f"Poisson arrivals at {args.load:.0%} of capacity"},
    }


def run_sizes(args, fm) -> dict:
    """Device-resident time of each single-product configuration."""
    cases = [
        ("mul 65536x65536", 65536, 65536, 1001),
        ("sqr 65536", 65536, 0, 1003),
        ("mul 1048576x1048576", 1 << 20, 1 << 20, 2001),
        ("mul 4194304x262144", 4194304, 262144, 2003),
        ("sqr 1048576", 1 << 20, 0, 2005),
        ("mul all-ones 1048576x1048576", 1 << 20, 1 << 20, None),
        ("mul 16777216x16777216", 1 << 24, 1 << 24, 3001),
        ("sqr 16777216", 1 << 24, 0, 3003),
    ]
    stream = fm.Stream()
    res = {}
    # the device leaves its idle clocks first (as in the single-product workloads)
    w = fm.DeviceBuffer.from_host(splitmix64(1, 1 << 20), stream)
    wo = fm.DeviceBuffer(1 << 21)
    w0 = time.perf_counter()
    while time.perf_counter() - w0 < args.min_warmup_s:
        for _ in range(8):
            fm.mul_device(w, 1 << 20, w, 1 << 20, wo, stream)
        stream.synchronize()
    w.free()
    wo.free()
    for name, n1, n2, seed in cases:
        if seed is None:
            a = np.full(n1, 0xFFFFFFFFFFFFFFFF, dtype=np.uint64)
            b = np.full(n2, 0xFFFFFFFFFFFFFFFF, dtype=np.uint64)
        else:
            a = splitmix64(seed, n1)
            b = splitmix64(seed + 1, n2) if n2 else None
        da = fm.DeviceBuffer.from_host(a, stream)
        db = fm.DeviceBuffer.from_host(b, stream) if b is not None else None
        do = fm.DeviceBuffer(n1 + (n2 if n2 else n1))

        def one():
            if db is None:
                fm.sqr_device(da, n1, do, stream)
            else:
                fm.mul_device(da, n1, db, n2, do, stream)

        # each case warms up for at least a quarter of --min-warmup-s: after
        # the short products the device needs that long to come back to its
        # clocks under load (2^20 x 2^20 read 240-245 us after 10 steps)
        w0 = time.perf_counter()
        while True:
            for _ in range(args.warmup):
                one()
            stream.synchronize()
            if time.perf_counter() - w0 >= args.min_warmup_s / 4:
                break
        reps = max(3, args.steps)
        t0 = time.perf_counter()
        for _ in range(reps):
            one()
        stream.synchronize()
        res[name] = round((time.perf_counter() - t0) / reps * 1e6, 1)
        for d in (da, db, do):
            if d is not None:
                d.free()
    stream.close()
    return {"metric": "µs per product, device-resident", "unit": "µs", "values": res,
            "steps": args.steps, "warmup": args.warmup, "higher_is_better": False, "dtype": "u64",
            "data": "synthetic (splitmix64, seeds 1000*cfg+k)", "config": {"workload": "single products"}}


def run_single(args, fm, n1: int, n2: int, seed_a: int, seed_b: int, metric: str,
               traffic: dict[str, float], valu_busy: dict[str, float], profile: str) -> dict:
    """One device-resident n1 x n2 multiply per step (the headline shape by
    default); per-kernel breakdown, roofline of the dominant kernel, host-ABI
    time and the CPU baselines on the same operands."""
    a = splitmix64(seed_a, n1)
    b = splitmix64(seed_b, n2)
    log_n = (n1 + n2 - 2).bit_length()
    stream = fm.Stream()
    da = fm.DeviceBuffer.from_host(a, stream)
    db = fm.DeviceBuffer.from_host(b, stream)
    dout = fm.DeviceBuffer(n1 + n2)

    # Warm-up: at least W steps and at least --min-warmup-s seconds of them,
    # so that the device has left its idle clocks before anything is timed
    # (a few steps take about a millisecond).
    warm_steps = 0
    w0 = time.perf_counter()
    while warm_steps < args.warmup or time.perf_counter() - w0 < args.min_warmup_s:
        for _ in range(8):
            fm.mul_device(da, n1, db, n2, dout, stream)
        stream.synchronize()
        warm_steps += 8
    warm_s = time.perf_counter() - w0
    fm.timing_collect()

    # Breakdown pass (not the timed region): events around every kernel.
    n_break = max(3, args.steps // 4)
    fm.timing_enable(True)
    for _ in range(n_break):
        fm.mul_device(da, n1, db, n2, dout, stream)
    stream.synchronize()
    fm.timing_enable(False)
    breakdown = fm.timing_collect()
    passes = round(breakdown.get("col_fwd", (n_break, 0.0))[0] / n_break)
    alg = kernel_alg_bytes(n1, n2, log_n, passes)
    dom = max((k for k in breakdown if k in alg),
              key=lambda k: breakdown[k][1] / breakdown[k][0])

    # Timed region: K steps; events only around the dominant kernel, so that
    # they add no gaps between the other launches.
    fm.timing_enable(True, only=dom)
    stream.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        fm.mul_device(da, n1, db, n2, dout, stream)
    stream.synchronize()
    t1 = time.perf_counter()
    fm.timing_enable(False)
    kt = fm.timing_collect()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    got = dout.download(n1 + n2)
    if args.copy_calib:
        words = 1 << 25
        with fm.DeviceBuffer(words) as cs, fm.DeviceBuffer(words) as cd:
            fm.copy_device(cd, cs, words, stream)
            stream.synchronize()

    # Host-pointer drop-in (PCIe included), reported beside the device number:
    # a fresh output array per call, as the reference's bench allocates one
    # (bench/bench.c:70-90); two untimed calls first (the library's staging
    # buffers and host threads are made on first use).
    for _ in range(2):
        fm.abs_mul64(a, b)
    # then timed: fresh output arrays (their pages are first touched inside
    # the call), and one output array reused
    t_host = []
    for _ in range(7):
        h0 = time.perf_counter()
        host_out = fm.abs_mul64(a, b)
        t_host.append(time.perf_counter() - h0)
    if not np.array_equal(host_out, got):
        raise AssertionError("host-pointer abs_mul64 differs from the device-pointer result")
    t_host_reuse = []
    for _ in range(7):
        h0 = time.perf_counter()
        fm.abs_mul64(a, b, host_out)
        t_host_reuse.append(time.perf_counter() - h0)
    if not np.array_equal(host_out, got):
        raise AssertionError("host-pointer abs_mul64 differs from the device-pointer result")

    per_kernel = {}
    for name, (cnt, tot_ms) in breakdown.items():
        avg_us = tot_ms * 1e3 / cnt
        entry = {"launches_per_step": cnt / n_break, "avg_us": round(avg_us, 3)}
        if name in alg:
            entry["alg_bytes"] = alg[name]
            entry["GBps"] = round(alg[name] / (avg_us * 1e-6) / 1e9, 1)
        if name in traffic:
            entry["traffic_bytes"] = traffic[name]
        if name in valu_busy:
            entry["valu_busy_pct"] = valu_busy[name]
        per_kernel[name] = entry
    kernel_sum_us = sum(v["avg_us"] * v["launches_per_step"] for v in per_kernel.values())
    dom_cnt, dom_ms = kt[dom]
    dom_us = dom_ms * 1e3 / dom_cnt
    dom_gbs = alg[dom] / (dom_us * 1e-6) / 1e9
    mm = modmul_count(n1, n2, log_n)

    result = {
        "metric": metric,
        "value": round(ms_per_step * 1e3, 2),
        "unit": "µs",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "warmup_steps_run": warm_steps,
        "warmup_s": round(warm_s, 3),
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": f"synthetic (splitmix64 seeds {seed_a}/{seed_b}, top limb nonzero)",
        "config": {
            "workload": f"abs_mul64 {n1}x{n2} limbs, device-resident",
            "ntt_len": 1 << log_n,
            "column_passes": passes,
            "primes": 3,
            "products_per_step": 1,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "kernel_avg_us": round(dom_us, 3),
            "launches_timed": dom_cnt,
            "achieved": round(dom_gbs, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(dom_gbs / HBM_PEAK_GBS, 4),
            "traffic": traffic.get(dom),
        },
        # The binding resource is the integer VALU, not HBM (the roofline block
        # above keeps the HBM form): the dominant kernel's VALUBusy as measured
        # by rocprofv3 at this geometry, with its source.
        "roofline_binding": "valu",
        "valu": {
            "kernel": dom,
            "busy_frac": round(valu_busy[dom] / 100.0, 3) if dom in valu_busy else None,
            "busy_source": f"{profile} (rocprofv3 --pmc VALUBusy, same command)",
            "modmuls_per_product": mm,
            "Gmodmul_per_s": round(mm / (ms_per_step * 1e-3) / 1e9, 2),
            "lane_ops_per_modmul_at_peak": round(VALU_LANE_OPS * ms_per_step * 1e-3 / mm, 1),
        },
        "kernels": per_kernel,
        "kernel_sum_us": round(kernel_sum_us, 2),
        "host_abi_us": round(min(t_host) * 1e6, 1),
        "host_abi_median_us": round(float(np.median(t_host)) * 1e6, 1),
        "host_abi_reused_out_us": round(min(t_host_reuse) * 1e6, 1),
        "host_abi_reused_out_median_us": round(float(np.median(t_host_reuse)) * 1e6, 1),
    }

    if not args.no_cpu_baseline:
        want, times = cpu_baseline(a, b)
        if not np.array_equal(got, want):
            raise AssertionError("GPU product differs from the CPU oracle")
        result["bit_exact_vs_oracle"] = True
        result["cpu_baseline"] = {
            "value": round(float(np.mean(times)) * 1e6, 1),
            "median": round(float(np.median(times)) * 1e6, 1),
            "unit": "µs",
            "cores": 1,
            "kind": "port",
            "sample": f"{len(times)} x the same {n1}x{n2}-limb abs_mul64 (oracle, 1 thread)",
        }
        try:
            with open("/proc/cpuinfo") as f:
                model = next(l.split(":", 1)[1].strip() for l in f if l.startswith("model name"))
            result["cpu_baseline"]["cpu"] = model
        except Exception:
            pass
        fs = cpu_fft_small(a, b, want)
        if fs is not None:
            result["cpu_fft_small"] = {
                "value": round(float(np.mean(fs)) * 1e6, 1),
                "median": round(float(np.median(fs)) * 1e6, 1),
                "unit": "µs",
                "cores": 1,
                "kind": "reference",
                "sample": f"{len(fs)} x the same {n1}x{n2}-limb abs_mul64 of the reference's "
                          "fft_small (built from its sources, 1 thread, after one warm-up call)",
            }
    return result


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--min-warmup-s", type=float, default=2.0,
                    help="warm-up runs at least --warmup steps and at least this many seconds")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calib", action="store_true",
                    help="also run one 256 MiB copy kernel (profiler byte-counter calibration)")
    ap.add_argument("--workload", default="mul1m", choices=["mul1m", "mul16m", "batch", "sizes", "serve"],
                    help="mul1m: the 2^20 x 2^20 headline (default); mul16m: the 2^24 x 2^24 product "
                         "measured the same way; batch: 4,096 mixed products; sizes: every "
                         "single-product configuration; serve: 20,000 requests through the request loop")
    ap.add_argument("--requests", type=int, default=20000, help="serve: number of requests")
    ap.add_argument("--load", type=float, default=0.7, help="serve: offered rate / capacity")
    ap.add_argument("--pool-gib", type=int, default=64, help="serve: device pool (GiB)")
    ap.add_argument("--serve-streams", type=int, default=4, help="serve: streams on which groups overlap")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus})",
              file=sys.stderr)
        return 2

    import fast_mul_ntt_amd as fm

    L = fm.lib()
    if L.fmh_init(0) != 0:
        raise RuntimeError(L.fmh_last_error().decode())
    if args.workload == "batch":
        print(json.dumps(run_batch(args, fm)))
        return 0
    if args.workload == "sizes":
        print(json.dumps(run_sizes(args, fm)))
        return 0
    if args.workload == "serve":
        print(json.dumps(run_serve(args, fm)))
        return 0
    if args.workload == "mul16m":
        res = run_single(args, fm, 1 << 24, 1 << 24, 3001, 3002, METRIC_16M, TRAFFIC_16M, VALU_BUSY_16M, PROFILE_16M)
    else:
        res = run_single(args, fm, N1, N2, SEED_A, SEED_B, METRIC, TRAFFIC, VALU_BUSY, PROFILE)
    print(json.dumps(res))
    return 0


if __name__ == "__main__":
    sys.exit(main())
