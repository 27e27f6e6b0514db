/*
 * ntt3_oracle.h -- CPU oracle for the 3-prime Montgomery NTT + CRT multiply.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or as the timed CPU baseline.  The product library (libfmh.so) never links,
 * loads or calls it.
 *
 * This is a plain-C restatement of the algorithm of HJimmyK/fast_mul_ntt src/:
 *   - Montgomery arithmetic with R = 2^64      (src/macro.h:74-160)
 *   - the three primes and their constants     (src/data.h:69-157)
 *   - twiddle table, radix-2^2 DIF / DIT, tails (src/core.h:106-509)
 *   - CRT of three residues                    (src/core.h:77-104)
 *   - convolution drivers incl. the split-radix
 *     recursion above 2^17 points              (src/fast_mul.c:51-288)
 *   - abs_mul64 / abs_sqr64 bodies             (src/fast_mul.c:299-460)
 * The lazy (non-canonical) intermediate values follow the reference step by
 * step, so raw transform outputs can be compared word for word with the
 * reference's own core.h (oracle/_ref) where that library is built.
 *
 * How each function is pinned is listed in the header comment of
 * ntt3_oracle.c.
 */
#ifndef FMH_NTT3_ORACLE_H
#define FMH_NTT3_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef unsigned long long u64;

/* Derived constants of prime i (0,1,2), in this order:
 * [0] p  [1] 2p  [2] -p^-1 mod 2^64  [3] R^2 mod p  [4] g  [5] g^-1 mod p
 * [6] mont(g)  [7] mont(g^-1)  [8] mont(1)  [9] w4  [10] w4^-1
 * [11..13] w8^1..3  [14..16] w8^-1..-3    (w* in Montgomery form, canonical)
 * Returns the number of words written (17). */
int orc_prime_constants(int prime, u64 out[17]);

/* Word-level Montgomery operations of prime i, as src/macro.h defines them. */
u64 orc_mont_mul(int prime, u64 a, u64 b);     /* lazy: no final subtract   */
u64 orc_mont_mulinto(int prime, u64 a, u64 b); /* one conditional subtract  */
u64 orc_tomont(int prime, u64 x);
u64 orc_toint(int prime, u64 x);
u64 orc_mont_pow(int prime, u64 base_mont, u64 e);

/* Twiddle table of src/core.h:214-238 for lengths up to 2^log_len:
 * omega[L/2 + t] = w_L^t, iomega[L/2 + t] = w_L^-t (Montgomery, lazy chain),
 * omega[0] = iomega[0] = 0.  Both arrays hold 2^log_len words. */
void orc_twiddle_table(int prime, int log_len, u64* omega, u64* iomega);

/* In-place radix-2^2 DIF / DIT of src/core.h:443-509 on Montgomery inputs,
 * len a power of two <= 2^17.  Raw lazy outputs. */
void orc_dif_raw(int prime, u64* data, u64 len);
void orc_idit_raw(int prime, u64* data, u64 len);

/* Forward transform of the zero-padded limb vector (len limbs, ntt_len a
 * power of two >= len) for one prime, exactly as abs_mul64 would run it
 * (to-Montgomery, then the DIF of conv_rec incl. the split-radix step above
 * 2^17).  Output: ntt_len canonical Montgomery residues in the reference's
 * in-place order, position j = X[bitrev(j)]. */
void orc_forward_spectrum(int prime, const u64* limbs, u64 len, u64 ntt_len, u64* out);

/* CRT of src/core.h:77-104: residues a,b,c (Montgomery, lazy) -> u192. */
void orc_crt3(u64 a, u64 b, u64 c, u64 out[3]);

/* The products.  out has len1+len2 (sqr: 2*len1) limbs.  A zero-length
 * operand gives the all-zero product (the reference aborts or overruns). */
void orc_abs_mul64(const u64* in1, u64 len1, const u64* in2, u64 len2, u64* out);
void orc_abs_sqr64(const u64* in1, u64 len1, u64* out);

/* Planted mistakes, for the tests that show the oracle's checks catch them
 * (0 = none): 1 drops a CRT term, 2 swaps the forward and inverse twiddle
 * tables (wrong sign of the root), 3 shifts a DIF twiddle index by one,
 * 4 pads the second operand to the first one's length (transposed lengths,
 * reads kept in bounds), 5 drops the top word of the running carry. */
void orc_plant(int kind);

/* Independent checks (not part of the reference algorithm). */
void orc_schoolbook(const u64* a, u64 na, const u64* b, u64 nb, u64* out);
u64 orc_mod_q(const u64* a, u64 n, u64 q);          /* value mod q, q < 2^63 */
u64 orc_mulmod_q(u64 a, u64 b, u64 q);
/* out[k] = (a[0] + a[1] 2^64 + ... + a[k-1] 2^(64 (k-1))) mod q for k = 0..n
 * (n + 1 words): residues of every slice a[o, o + l) follow as
 * (out[o + l] - out[o]) 2^(-64 o) mod q. */
void orc_prefix_mod_q(const u64* a, u64 n, u64 q, u64* out);
/* splitmix64 stream starting at state `seed`; if top_nonzero and out[n-1]==0,
 * out[n-1] is set to 1. */
void orc_splitmix64_fill(u64 seed, u64* out, u64 n, int top_nonzero);

#ifdef __cplusplus
}
#endif

#endif
