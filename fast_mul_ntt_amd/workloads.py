"""Seeded synthetic workloads of the benchmark configurations, shared by
bench.py and the GPU tests so that both run exactly the same inputs.

* ``splitmix64(seed, n)``: uniform u64 limbs (Steele, Lea, Flood 2014),
  top limb forced nonzero.
* ``batch_products()``: 4,096 products, operand lengths log-uniform in
  [64, 65,536] limbs drawn independently, 25 % squares (sampler seed 4001);
  operand i uses limb seeds 40,000 + 2 i and 40,001 + 2 i.
* ``serve_requests()``: 20,000 requests, lengths log-uniform in
  [256, 4,194,304] limbs drawn independently, 20 % squares (seed 5001);
  operands are slices at seeded offsets of one source array of
  2 x 4,194,304 limbs (splitmix64 seed 5001), and arrival gaps are unit-rate
  exponential (scaled by 1 / rate for Poisson arrivals).
"""

from __future__ import annotations

import numpy as np

BATCH_COUNT = 4096
BATCH_SEED = 4001
BATCH_LIMB_SEED = 40_000
SERVE_COUNT = 20_000
SERVE_SEED = 5001
SERVE_MAX = 4_194_304


def splitmix64(seed: int, n: int, top_nonzero: bool = True) -> np.ndarray:
    """splitmix64 stream: state += golden gamma, then the mix of Vigna's
    reference generator; the top limb is forced to 1 if it came out 0."""
    gamma = np.uint64(0x9E3779B97F4A7C15)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + gamma * np.arange(1, n + 1, dtype=np.uint64)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    if top_nonzero and n and z[-1] == 0:
        z[-1] = 1
    return z


def log_uniform_lengths(rng: np.random.Generator, count: int, lo: int, hi: int) -> np.ndarray:
    return np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), size=count))).astype(np.int64)


def batch_products(count: int = BATCH_COUNT, seed: int = BATCH_SEED):
    """(len1, len2, square) arrays of the batched configuration; len2 of a
    square is unused (the square of operand 1)."""
    rng = np.random.default_rng(seed)
    l1 = log_uniform_lengths(rng, count, 64, 65536)
    l2 = log_uniform_lengths(rng, count, 64, 65536)
    sq = rng.random(count) < 0.25
    return l1, l2, sq


def batch_operand(i: int, which: int, n: int) -> np.ndarray:
    """Limbs of operand `which` (0 or 1) of batched product i."""
    return splitmix64(BATCH_LIMB_SEED + 2 * i + which, n)


def serve_requests(count: int = SERVE_COUNT, seed: int = SERVE_SEED):
    """(len1, len2, square, offset1, offset2, gaps, source length) of the
    serving configuration."""
    rng = np.random.default_rng(seed)
    l1 = log_uniform_lengths(rng, count, 256, SERVE_MAX)
    l2 = log_uniform_lengths(rng, count, 256, SERVE_MAX)
    sq = rng.random(count) < 0.20
    src_len = 2 * SERVE_MAX
    o1 = (rng.random(count) * (src_len - l1 + 1)).astype(np.int64)
    o2 = (rng.random(count) * (src_len - l2 + 1)).astype(np.int64)
    gaps = rng.exponential(1.0, size=count)
    return l1, l2, sq, o1, o2, gaps, src_len


def serve_source(src_len: int, seed: int = SERVE_SEED) -> np.ndarray:
    return splitmix64(seed, src_len)
