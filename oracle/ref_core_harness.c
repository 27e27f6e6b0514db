/*
 * ref_core_harness.c -- thin C wrappers around the reference's own header
 * library src/core.h (+ data.h, macro.h), compiled from /root/reference/src
 * by oracle/Makefile into oracle/_ref/libref_core.so.
 *
 * TEST INFRASTRUCTURE ONLY: used to pin the oracle (ntt3_oracle.c) word for
 * word against the reference's twiddle tables, radix-2^2 DIF/DIT, Montgomery
 * macros and crt3, and to write tests/golden/ref_core_vectors.json.
 *
 * The reference's src/fast_mul.c (conv_rec_i, abs_mul64) is not built: it
 * needs the MSVC CRT's _aligned_malloc/_aligned_free, which this image does
 * not provide.  core.h, data.h and macro.h need nothing outside libc.
 */
#include "core.h" /* -I /root/reference/src */

#define PRIME_SWITCH(i, stmt1, stmt2, stmt3) \
    do {                                     \
        switch (i) {                         \
            case 0: stmt1; break;            \
            case 1: stmt2; break;            \
            default: stmt3; break;           \
        }                                    \
    } while (0)

int ref_constants(int i, u64 out[17]) {
#define FILL(k)                                                                         \
    do {                                                                                \
        u64 v[17] = {g_mod(k),       g_mod2(k),      g_modInvNeg(k), g_r2(k),           \
                     g_root(k),      g_rootinv(k),   g_mont_root(k), g_mont_rootinv(k), \
                     g_one(k),       g_w41(k),       g_w41inv(k),    g_w1(k),           \
                     g_w2(k),        g_w3(k),        g_w1inv(k),     g_w2inv(k),        \
                     g_w3inv(k)};                                                       \
        for (int j = 0; j < 17; j++) out[j] = v[j];                                     \
    } while (0)
    PRIME_SWITCH(i, FILL(1), FILL(2), FILL(3));
#undef FILL
    return 17;
}

u64 ref_mont_mul(int i, u64 a, u64 b) {
    u64 r = 0;
    PRIME_SWITCH(i, _mont_mul_func(r, a, b, 1), _mont_mul_func(r, a, b, 2), _mont_mul_func(r, a, b, 3));
    return r;
}

u64 ref_mont_mulinto(int i, u64 a, u64 b) {
    PRIME_SWITCH(i, _mont_mulinto_func(a, b, 1), _mont_mulinto_func(a, b, 2), _mont_mulinto_func(a, b, 3));
    return a;
}

u64 ref_tomont(int i, u64 x) {
    PRIME_SWITCH(i, _mont_tomont_func(x, 1), _mont_tomont_func(x, 2), _mont_tomont_func(x, 3));
    return x;
}

u64 ref_toint(int i, u64 x) {
    PRIME_SWITCH(i, _mont_toint_func(x, 1), _mont_toint_func(x, 2), _mont_toint_func(x, 3));
    return x;
}

/* Twiddle table for lengths up to 2^log_len (both arrays 2^log_len words). */
void ref_table(int i, int log_len, u64* omega, u64* iomega) {
    ntt_short t;
    t.ntt_len = 1ull << log_len;
    t.log_len = (size_t)log_len;
    t.omega = omega;
    t.iomega = iomega;
    PRIME_SWITCH(i, cover_nttshort_func(t.log_len, &t, 1), cover_nttshort_func(t.log_len, &t, 2),
                 cover_nttshort_func(t.log_len, &t, 3));
}

static int table_for(int i, u64 len, ntt_short* t) {
    u64 tl = len < long_threshold ? len : long_threshold;
    if (tl == 0) tl = 1;
    t->ntt_len = tl;
    t->log_len = log2_64(tl);
    t->omega = (u64*)malloc(tl * sizeof(u64));
    t->iomega = (u64*)malloc(tl * sizeof(u64));
    if (!t->omega || !t->iomega) return -1;
    ref_table(i, (int)t->log_len, t->omega, t->iomega);
    return 0;
}

/* dif_i / idit_i (core.h:443-509) on len <= 2^17 Montgomery words, in place. */
int ref_dif(int i, u64* data, u64 len) {
    ntt_short t;
    if (len > long_threshold || table_for(i, len, &t)) return -1;
    PRIME_SWITCH(i, dif_func(data, &t, len, 1), dif_func(data, &t, len, 2), dif_func(data, &t, len, 3));
    free(t.omega);
    free(t.iomega);
    return 0;
}

int ref_idit(int i, u64* data, u64 len) {
    ntt_short t;
    if (len > long_threshold || table_for(i, len, &t)) return -1;
    PRIME_SWITCH(i, idit_func(data, &t, len, 1), idit_func(data, &t, len, 2), idit_func(data, &t, len, 3));
    free(t.omega);
    free(t.iomega);
    return 0;
}

void ref_crt3(u64 a, u64 b, u64 c, u64 out[3]) {
    u192 r = {0, 0, 0};
    crt3(a, b, c, r);
    out[0] = r[0];
    out[1] = r[1];
    out[2] = r[2];
}
