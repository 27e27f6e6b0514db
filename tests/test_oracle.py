"""Pins the CPU oracle (oracle/ntt3_oracle.c) before anything is checked against it.

Each oracle function is checked against something other than its own formula:
  * golden words written by the reference's own code (tests/golden/*.json:
    src/core.h compiled from its sources, and fft_small's abs_mul64);
  * the live reference libraries in oracle/_ref when they are built;
  * mathematics: primality and root orders, congruences, an O(N^2) DFT,
    closed-form spectra of unit vectors, Python big-int CRT and products,
    (B^a - 1)(B^b - 1), mod-q residues.
A last group plants a one-bit mistake in an oracle output and shows that the
checks used by the GPU tests catch it.
"""

import hashlib
import json
import os
import random

import numpy as np
import pytest

from oracle.oracle import GENERATORS, PRIMES, RefCore, RefFftSmall, int_to_limbs, limbs_to_int

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
R = 1 << 64
CK = [(69, 55), (27, 56), (29, 57)]


def load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


CORE = load("ref_core_vectors.json")
PRODS = load("fft_small_products.json")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<u8").tobytes()).hexdigest()


def have_ref():
    try:
        RefCore()
        return True
    except FileNotFoundError:
        return False


def bitrev(x, bits):
    return int(format(x, f"0{bits}b")[::-1], 2) if bits else 0


def is_prime(n):
    if n < 2:
        return False
    for q in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        if n % q == 0:
            return n == q
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


# ------------------------------------------------------------------- inputs

def test_splitmix64_known_answer(orc):
    # Published first outputs of splitmix64 from state 0 (Vigna, splitmix64.c).
    assert [int(v) for v in orc.splitmix(0, 3, False)] == [
        0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


# ---------------------------------------------------------------- constants

@pytest.mark.parametrize("i", [0, 1, 2])
def test_primes_and_roots(i):
    p, g = PRIMES[i], GENERATORS[i]
    c, k = CK[i]
    assert p == c * 2**k + 1 and is_prime(p) and 4 * p < R
    # g generates the multiplicative group: g^((p-1)/q) != 1 for every prime q | p-1
    factors = {2} | {q for q in (3, 23, 29) if c % q == 0}
    for q in factors:
        assert pow(g, (p - 1) // q, p) != 1
    assert (p - 2) * p % R == R - 1  # -p^-1 mod 2^64 = p - 2


@pytest.mark.parametrize("i", [0, 1, 2])
def test_constants_match_reference_data_h(orc, i):
    assert [str(v) for v in orc.constants(i)] == CORE["constants"][str(i)]


@pytest.mark.parametrize("i", [0, 1, 2])
def test_constants_identities(orc, i):
    p = PRIMES[i]
    k = orc.constants(i)
    rinv = pow(R, -1, p)
    val = lambda m: m * rinv % p  # Montgomery form -> value
    assert k[3] == R * R % p and k[8] == R % p
    assert val(k[6]) == GENERATORS[i] and val(k[7]) * GENERATORS[i] % p == 1
    w4, w4i = val(k[9]), val(k[10])
    assert w4 * w4 % p == p - 1 and w4 * w4i % p == 1
    w8 = val(k[11])
    assert pow(w8, 2, p) == w4 and pow(w8, 4, p) == p - 1
    assert [val(k[11 + j]) for j in range(3)] == [pow(w8, j + 1, p) for j in range(3)]
    assert [val(k[14 + j]) * pow(w8, j + 1, p) % p for j in range(3)] == [1, 1, 1]


# --------------------------------------------------------------- Montgomery

def test_montgomery_ops_match_reference_words(orc):
    for i, a, b, mm, mi, x, tm, ti in CORE["mont"]:
        a, b, x = int(a), int(b), int(x)
        assert orc.mont_mul(i, a, b) == int(mm)
        assert orc.mont_mulinto(i, a, b) == int(mi)
        assert orc.tomont(i, x) == int(tm)
        assert orc.toint(i, a) == int(ti)


def test_montgomery_ops_congruences(orc):
    rng = random.Random(5)
    for i in range(3):
        p = PRIMES[i]
        rinv = pow(R, -1, p)
        for _ in range(500):
            a, b = rng.randrange(2 * p), rng.randrange(2 * p)
            m = orc.mont_mul(i, a, b)
            assert m % p == a * b * rinv % p and m < 2 * p
            assert orc.mont_mulinto(i, a, b) == a * b * rinv % p
            x = rng.getrandbits(64)
            assert orc.tomont(i, x) == x * R % p
            assert orc.toint(i, a) == a * rinv % p


# ----------------------------------------------------------------- twiddles

@pytest.mark.parametrize("i", [0, 1, 2])
def test_twiddle_table_matches_reference(orc, i):
    om, iom = orc.table(i, 6)
    g = CORE["tables"][str(i)]
    assert [str(int(v)) for v in om] == g["log6_omega"]
    assert [str(int(v)) for v in iom] == g["log6_iomega"]
    om, iom = orc.table(i, 17)
    assert sha(om) == g["log17_omega_sha256"] and sha(iom) == g["log17_iomega_sha256"]
    # values: omega[L/2 + t] = w_L^t, w_L = g^((p-1)/L)
    p = PRIMES[i]
    rinv = pow(R, -1, p)
    for L in (2, 4, 64, 1 << 17):
        wL = pow(GENERATORS[i], (p - 1) // L, p)
        for t in {0, min(1, L // 2 - 1), L // 2 - 1}:
            assert int(om[L // 2 + t]) * rinv % p == pow(wL, t, p)
            assert int(iom[L // 2 + t]) * rinv % p == pow(wL, -t, p)


# --------------------------------------------------------------- transforms

def test_dif_idit_raw_match_reference_words(orc):
    for e in CORE["dif"]:
        i, n = e["prime"], 1 << e["log_n"]
        x = np.array([orc.tomont(i, int(v)) for v in orc.splitmix(e["seed"], n)], dtype=np.uint64)
        assert sha(x) == e["input_sha256"]
        f = orc.dif_raw(i, x)
        g = orc.idit_raw(i, f)
        assert sha(f) == e["dif_sha256"], (i, n)
        assert sha(g) == e["idit_of_dif_sha256"], (i, n)
        if "dif" in e:
            assert [str(int(v)) for v in f] == e["dif"]


@pytest.mark.skipif(not have_ref(), reason="oracle/_ref not built (reference sources absent)")
def test_dif_idit_raw_match_live_reference(orc):
    ref = RefCore()
    rng = np.random.default_rng(9)
    for i in range(3):
        p = PRIMES[i]
        for lg in range(0, 18):
            x = (rng.integers(0, 2**63, size=1 << lg, dtype=np.uint64) % np.uint64(p))
            assert np.array_equal(orc.dif_raw(i, x), ref.dif_raw(i, x)), (i, lg)
            assert np.array_equal(orc.idit_raw(i, x), ref.idit_raw(i, x)), (i, lg)


@pytest.mark.parametrize("log_n", [0, 1, 2, 3, 4, 5, 6])
def test_forward_spectrum_is_the_dft_in_bitreversed_order(orc, log_n):
    n = 1 << log_n
    for i in range(3):
        p = PRIMES[i]
        limbs = orc.splitmix(77 + log_n, max(1, n // 2 + (n > 2)))[: n]
        w = pow(GENERATORS[i], (p - 1) // n, p)
        xs = [int(v) % p for v in limbs] + [0] * (n - len(limbs))
        X = [sum(xs[t] * pow(w, t * k, p) for t in range(n)) % p for k in range(n)]
        got = orc.forward_spectrum(i, limbs, n)
        assert [int(v) for v in got] == [X[bitrev(j, log_n)] * R % p for j in range(n)]


@pytest.mark.parametrize("log_n", [17, 18, 20])
def test_forward_spectrum_unit_vectors_closed_form(orc, log_n):
    """x = e_m has X_k = w^(m k): pins the split-radix path above 2^17 points."""
    n = 1 << log_n
    rng = random.Random(log_n)
    for i in range(3):
        p = PRIMES[i]
        w = pow(GENERATORS[i], (p - 1) // n, p)
        m = rng.randrange(n // 2)
        limbs = np.zeros(m + 1, dtype=np.uint64)
        limbs[m] = 1
        got = orc.forward_spectrum(i, limbs, n)
        for j in [0, 1, n - 1] + [rng.randrange(n) for _ in range(20)]:
            k = bitrev(j, log_n)
            assert int(got[j]) == pow(w, m * k, p) * R % p, (i, j)


# ---------------------------------------------------------------------- CRT

def test_crt3_matches_reference_words_and_math(orc):
    P = PRIMES[0] * PRIMES[1] * PRIMES[2]
    rinv = [pow(R, -1, p) for p in PRIMES]
    for a, b, c, want in CORE["crt"]:
        a, b, c = int(a), int(b), int(c)
        got = orc.crt3(a, b, c)
        assert got == int(want)
        # math: got is the unique value in [0, P) with got = res_i * R^-1 (mod p_i)
        assert 0 <= got < P
        for r, p, ri in zip((a, b, c), PRIMES, rinv):
            assert got % p == r * ri % p


@pytest.mark.skipif(not have_ref(), reason="oracle/_ref not built (reference sources absent)")
def test_crt3_matches_live_reference(orc):
    ref = RefCore()
    rng = random.Random(11)
    for _ in range(2000):
        a, b, c = (rng.randrange(2 * p) for p in PRIMES)
        assert orc.crt3(a, b, c) == ref.crt3(a, b, c)


# ----------------------------------------------------------------- products

SIZES = [1, 2, 3, 5, 17, 33, 64, 100, 255, 256, 257, 300, 511, 512, 513, 700, 1024, 1100]


def test_products_vs_schoolbook(orc):
    seed = 40
    for si, an in enumerate(SIZES):
        for bn in SIZES[si::2]:
            a, b = orc.splitmix(seed, an), orc.splitmix(seed + 1, bn)
            seed += 2
            want = orc.schoolbook(a, b)
            assert np.array_equal(orc.mul(a, b), want), (an, bn)
            assert np.array_equal(orc.mul(b, a), want)
        a = orc.splitmix(seed, an)
        assert np.array_equal(orc.sqr(a), orc.schoolbook(a, a))


def test_schoolbook_vs_python_bigints(orc):
    rng = random.Random(3)
    for _ in range(30):
        an, bn = rng.randint(1, 40), rng.randint(1, 40)
        a, b = orc.splitmix(rng.getrandbits(32), an), orc.splitmix(rng.getrandbits(32), bn)
        assert limbs_to_int(orc.schoolbook(a, b)) == limbs_to_int(a) * limbs_to_int(b)


def test_products_match_reference_fft_small_golden(orc):
    for e in PRODS:
        if e["op"] == "mul":
            if e["n1"] * e["n2"] > 1 << 36:
                continue  # the 2^20 x 2^20 vector is checked in the slow test below
            a, b = orc.splitmix(e["seed1"], e["n1"]), orc.splitmix(e["seed2"], e["n2"])
            c = orc.mul(a, b)
            if "out" in e:
                assert [str(int(v)) for v in c] == e["out"]
        elif e["op"] == "sqr":
            c = orc.sqr(orc.splitmix(e["seed1"], e["n1"]))
        else:
            ones = np.full(e["n1"], 0xFFFFFFFFFFFFFFFF, dtype=np.uint64)
            c = orc.mul(ones, ones[: e["n2"]])
        assert sha(c) == e["sha256"], e


def test_headline_product_matches_reference_fft_small_golden(orc):
    e = next(x for x in PRODS if x["op"] == "mul" and x["n1"] == 1 << 20)
    a, b = orc.splitmix(e["seed1"], e["n1"]), orc.splitmix(e["seed2"], e["n2"])
    assert sha(orc.mul(a, b)) == e["sha256"]


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(GOLDEN), "..", "oracle", "_ref",
                                                    "libref_fft_small.so")),
                    reason="oracle/_ref not built")
def test_products_match_live_fft_small(orc):
    fsd = RefFftSmall()
    rng = random.Random(17)
    for _ in range(20):
        an, bn = rng.randint(1, 5000), rng.randint(1, 5000)
        a, b = orc.splitmix(rng.getrandbits(40), an), orc.splitmix(rng.getrandbits(40), bn)
        assert np.array_equal(orc.mul(a, b), fsd.mul(a, b))
        assert np.array_equal(orc.sqr(a), fsd.sqr(a))


@pytest.mark.parametrize("n1,n2", [(1, 1), (5, 9), (1000, 1), (2048, 2049), (40000, 30000)])
def test_all_ones_closed_form(orc, n1, n2):
    """(B^a - 1)(B^b - 1) = B^(a+b) - B^a - B^b + 1, B = 2^64."""
    a = np.full(n1, 0xFFFFFFFFFFFFFFFF, dtype=np.uint64)
    b = np.full(n2, 0xFFFFFFFFFFFFFFFF, dtype=np.uint64)
    v = (1 << (64 * (n1 + n2))) - (1 << (64 * n1)) - (1 << (64 * n2)) + 1
    assert np.array_equal(orc.mul(a, b), int_to_limbs(v, n1 + n2))


@pytest.mark.parametrize("k", [0, 1, 7, 300])
def test_shift_by_limbs(orc, k):
    a = orc.splitmix(123, 999)
    b = np.zeros(k + 1, dtype=np.uint64)
    b[k] = 1  # B^k
    want = np.zeros(a.size + k + 1, dtype=np.uint64)
    want[k:k + a.size] = a
    assert np.array_equal(orc.mul(a, b), want)


def test_mod_q_spot_checks_large(orc):
    rng = random.Random(2)
    qs = []
    while len(qs) < 3:
        q = rng.getrandbits(61) | (1 << 60) | 1
        if is_prime(q):
            qs.append(q)
    for n1, n2 in [(1 << 18, 1 << 18), (300000, 5000)]:
        a, b = orc.splitmix(n1 + 5, n1), orc.splitmix(n2 + 6, n2)
        c = orc.mul(a, b)
        for q in qs:
            assert orc.mod_q(c, q) == orc.mulmod_q(orc.mod_q(a, q), orc.mod_q(b, q), q)


def test_mod_q_matches_python(orc):
    rng = random.Random(4)
    for _ in range(20):
        a = orc.splitmix(rng.getrandbits(30), rng.randint(1, 50))
        q = rng.getrandbits(61) | 1
        assert orc.mod_q(a, q) == limbs_to_int(a) % q


def test_slice_residues_from_prefixes(orc):
    """Residues of slices of one array from its prefix residues (the serving
    workload's operands are slices of one source array) equal Python big-int
    residues of the slices, and direct mod_q."""
    rng = random.Random(9)
    src = orc.splitmix(123, 5000)
    for _ in range(6):
        q = rng.getrandbits(61) | (1 << 60) | 1
        pre = orc.prefix_mod_q(src, q)
        assert pre[0] == 0 and int(pre[-1]) == limbs_to_int(src) % q
        for _ in range(10):
            o = rng.randint(0, 4999)
            n = rng.randint(0, 5000 - o)
            want = limbs_to_int(src[o:o + n]) % q
            assert orc.slice_mod_q(pre, o, n, q) == want == orc.mod_q(src[o:o + n], q), (o, n)


def test_zero_length_operands(orc):
    a = orc.splitmix(1, 4)
    assert np.array_equal(orc.mul(a, np.zeros(0, np.uint64)), np.zeros(4, np.uint64))
    assert orc.sqr(np.zeros(0, np.uint64)).size == 0


# ------------------------------------------------------------ planted mistakes

def test_planted_mistakes_are_caught(orc):
    """A one-bit error in a product or a spectrum fails the checks the GPU tests use."""
    a, b = orc.splitmix(8, 20000), orc.splitmix(9, 20000)
    c = orc.mul(a, b)
    bad = c.copy()
    bad[12345] ^= np.uint64(1 << 17)
    q = (1 << 61) - 1
    assert orc.mod_q(bad, q) != orc.mulmod_q(orc.mod_q(a, q), orc.mod_q(b, q), q)
    assert not np.array_equal(bad, c)
    e = next(x for x in PRODS if x["op"] == "mul" and x["n1"] == 2900)
    c2 = orc.mul(orc.splitmix(e["seed1"], e["n1"]), orc.splitmix(e["seed2"], e["n2"]))
    c2[0] ^= np.uint64(1)
    assert sha(c2) != e["sha256"]
    s = orc.forward_spectrum(0, a[:100], 256)
    s[3] = (int(s[3]) + 1) % PRIMES[0]
    assert not np.array_equal(s, orc.forward_spectrum(0, a[:100], 256))


# Plausible mistakes planted in the oracle itself: each one must fail at least
# one of the checks that pin the oracle (schoolbook products, the O(N^2) DFT
# of the spectrum, the CRT's defining congruences).
PLANTS = {1: "dropped CRT term", 2: "root of the wrong sign", 3: "twiddle index off by one",
          4: "operand lengths transposed", 5: "carry word dropped"}


def _checks_that_fail(orc):
    failed = []
    a, b = orc.splitmix(501, 300), orc.splitmix(502, 700)
    if not np.array_equal(orc.mul(a, b), orc.schoolbook(a, b)):
        failed.append("schoolbook")
    n, i = 64, 1
    p = PRIMES[i]
    limbs = orc.splitmix(503, 40)
    w = pow(GENERATORS[i], (p - 1) // n, p)
    xs = [int(v) % p for v in limbs] + [0] * (n - 40)
    X = [sum(xs[t] * pow(w, t * k, p) for t in range(n)) % p for k in range(n)]
    if [int(v) for v in orc.forward_spectrum(i, limbs, n)] != [X[bitrev(j, 6)] * R % p for j in range(n)]:
        failed.append("dft")
    rng = random.Random(504)
    for _ in range(50):
        res = [rng.randrange(p) for p in PRIMES]
        got = orc.crt3(*res)
        if any(got % q != r * pow(R, -1, q) % q for r, q in zip(res, PRIMES)):
            failed.append("crt")
            break
    return failed


@pytest.mark.parametrize("kind", sorted(PLANTS))
def test_planted_oracle_mistakes_fail_a_check(orc, kind):
    assert _checks_that_fail(orc) == []
    orc.plant(kind)
    try:
        failed = _checks_that_fail(orc)
    finally:
        orc.plant(0)
    assert failed, f"planted mistake not caught: {PLANTS[kind]}"
    assert _checks_that_fail(orc) == []
