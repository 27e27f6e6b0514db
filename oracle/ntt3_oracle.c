/*
 * ntt3_oracle.c -- CPU oracle: restatement of HJimmyK/fast_mul_ntt src/.
 *
 * TEST INFRASTRUCTURE ONLY (see ntt3_oracle.h).  Nothing in the product
 * library links or calls this file.
 *
 * How each part is pinned (tests/test_oracle.py):
 *   constants        vs the reference's data.h/core.h words (golden file
 *                    tests/golden/ref_core_vectors.json, written by the
 *                    reference's own core.h compiled in oracle/_ref) and vs
 *                    their defining identities (w4^2 = -1, g^((p-1)/2) = -1 ...)
 *   mont_* / tomont  vs the reference's macros (same golden file) and vs
 *                    Python big-int arithmetic (a*b*2^-64 mod p)
 *   twiddle table    word for word vs the reference's cover_nttshort_i
 *   dif/idit raw     word for word vs the reference's dif_i / idit_i
 *                    (N <= 2^12 in the golden file, up to 2^17 live when
 *                    oracle/_ref is built), and vs an O(N^2) DFT
 *   forward spectrum vs the O(N^2) DFT (small N) and the reference dif_i
 *   crt3             vs the reference's crt3 and vs Python CRT
 *   abs_mul/abs_sqr  vs schoolbook, closed forms ((B^a-1)(B^b-1)), mod-q
 *                    checks with random 61-bit primes, and word for word vs
 *                    the reference's second library (fft_small abs_mul64)
 *                    compiled from its sources in oracle/_ref
 */
#define _POSIX_C_SOURCE 200112L /* posix_memalign */
#include "ntt3_oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

#define LONG_THRESHOLD 131072ULL /* src/core.h:428-431: iterative below 2^17 */

/* ---------------------------------------------------------------- primes */

typedef struct {
    u64 p, p2, pinvneg, r2;
    u64 g, ginv;
    u64 mont_g, mont_ginv, one;
    u64 w4, w4inv, w8[3], w8inv[3];
} prime_t;

/* src/data.h:69-81: the three primes and their primitive roots. */
static const u64 PRIME_P[3] = {2485986994308513793ull, 1945555039024054273ull,
                               4179340454199820289ull};
static const u64 PRIME_G[3] = {5, 5, 3};

static prime_t PR[3];
static int g_ready = 0;

/* CRT constants of src/core.h:78-84, derived in orc_init. */
static u64 CRT_P123[3];
static u64 CRT_M[3][2];    /* M_i = P / p_i as u128 words: p2p3, p1p3, p1p2 */
static u64 CRT_MINV[3];    /* mont((P/p_i)^-1 mod p_i) */

static inline u64 mul_hi(u64 a, u64 b) { return (u64)(((u128)a * b) >> 64); }

/* macro.h:109-118  _mont_mul: REDC without the final subtract.  The 128-bit
 * sum a*b + m*p wraps mod 2^128 exactly as _u128add does. */
static inline u64 mm_lazy(const prime_t* P, u64 a, u64 b) {
    u128 ab = (u128)a * b;
    u64 m = (u64)ab * P->pinvneg;
    u128 s = (u128)m * P->p + ab;
    return (u64)(s >> 64);
}

/* macro.h:131-140  _mont_mulinto: REDC with one conditional subtract. */
static inline u64 mm_into(const prime_t* P, u64 a, u64 b) {
    u64 t = mm_lazy(P, a, b);
    return t < P->p ? t : t - P->p;
}

/* macro.h:142-151 */
static inline u64 to_mont(const prime_t* P, u64 x) { return mm_into(P, x, P->r2); }

/* macro.h:153-160  _mont_toint: REDC(x * 1). */
static inline u64 to_int(const prime_t* P, u64 x) {
    u64 m = x * P->pinvneg;
    u128 s = (u128)m * P->p + x;
    u64 t = (u64)(s >> 64);
    return t < P->p ? t : t - P->p;
}

/* core.h:46-59  square-and-multiply in the Montgomery domain. */
static u64 mpow(const prime_t* P, u64 m, u64 n) {
    u64 r = P->one;
    while (n > 0) {
        if (n & 1) r = mm_into(P, r, m);
        m = mm_into(P, m, m);
        n >>= 1;
    }
    return r;
}

/* macro.h:74-98: lazy additive operations over [0,2p) / [0,4p). */
static inline u64 m_add(const prime_t* P, u64 x, u64 y) {
    u64 r = x + y;
    return r < P->p2 ? r : r - P->p2;
}
static inline u64 m_sub(const prime_t* P, u64 x, u64 y) {
    u64 d = x - y;
    return d > x ? d + P->p2 : d;
}
static inline u64 r_sub(const prime_t* P, u64 x, u64 y) { return x - y + P->p2; }
static inline u64 m_norm2(const prime_t* P, u64 x) { return x >= P->p2 ? x - P->p2 : x; }

static u64 plain_powmod(u64 b, u64 e, u64 p) {
    u128 r = 1, x = b % p;
    while (e) {
        if (e & 1) r = (r * x) % p;
        x = (x * x) % p;
        e >>= 1;
    }
    return (u64)r;
}

static void orc_init(void) {
    if (g_ready) return;
    for (int i = 0; i < 3; i++) {
        prime_t* P = &PR[i];
        P->p = PRIME_P[i];
        P->p2 = 2 * P->p;
        /* -p^-1 mod 2^64 by Newton iteration (p odd). */
        u64 inv = P->p;
        for (int k = 0; k < 6; k++) inv *= 2 - P->p * inv;
        P->pinvneg = (u64)0 - inv;
        u64 r1 = (u64)((((u128)1) << 64) % P->p);
        P->r2 = (u64)(((u128)r1 * r1) % P->p);
        P->g = PRIME_G[i];
        P->ginv = plain_powmod(P->g, P->p - 2, P->p);
        P->one = r1;
        P->mont_g = (u64)((((u128)P->g) << 64) % P->p);
        P->mont_ginv = (u64)((((u128)P->ginv) << 64) % P->p);
        /* data.h:110-152: w4 = qpow(mont(g),(p-1)/4), w8^k = qpow(w8,k). */
        P->w4 = mpow(P, P->mont_g, (P->p - 1) / 4);
        P->w4inv = mpow(P, P->mont_ginv, (P->p - 1) / 4);
        u64 w8 = mpow(P, P->mont_g, (P->p - 1) / 8);
        u64 w8i = mpow(P, P->mont_ginv, (P->p - 1) / 8);
        for (int k = 0; k < 3; k++) {
            P->w8[k] = mpow(P, w8, (u64)k + 1);
            P->w8inv[k] = mpow(P, w8i, (u64)k + 1);
        }
    }
    /* CRT: P = p1 p2 p3, M_i = P/p_i, MINV_i = mont(M_i^-1 mod p_i). */
    u128 m23 = (u128)PRIME_P[1] * PRIME_P[2];
    u128 m13 = (u128)PRIME_P[0] * PRIME_P[2];
    u128 m12 = (u128)PRIME_P[0] * PRIME_P[1];
    u128 ms[3] = {m23, m13, m12};
    for (int i = 0; i < 3; i++) {
        CRT_M[i][0] = (u64)ms[i];
        CRT_M[i][1] = (u64)(ms[i] >> 64);
        u64 red = (u64)(ms[i] % PRIME_P[i]);
        u64 inv = plain_powmod(red, PRIME_P[i] - 2, PRIME_P[i]);
        CRT_MINV[i] = (u64)((((u128)inv) << 64) % PRIME_P[i]);
    }
    /* P = m12 * p3 as 192 bits */
    u128 lo = (u128)(u64)m12 * PRIME_P[2];
    u128 hi = (u128)(u64)(m12 >> 64) * PRIME_P[2];
    CRT_P123[0] = (u64)lo;
    u128 mid = (lo >> 64) + (u64)hi;
    CRT_P123[1] = (u64)mid;
    CRT_P123[2] = (u64)(hi >> 64) + (u64)(mid >> 64);
    g_ready = 1;
}

static const prime_t* prime(int i) {
    orc_init();
    return &PR[i];
}

int orc_prime_constants(int i, u64 out[17]) {
    const prime_t* P = prime(i);
    u64 v[17] = {P->p,    P->p2,       P->pinvneg,  P->r2,       P->g,        P->ginv,
                 P->mont_g, P->mont_ginv, P->one,    P->w4,       P->w4inv,    P->w8[0],
                 P->w8[1], P->w8[2],    P->w8inv[0], P->w8inv[1], P->w8inv[2]};
    memcpy(out, v, sizeof(v));
    return 17;
}

u64 orc_mont_mul(int i, u64 a, u64 b) { return mm_lazy(prime(i), a, b); }
u64 orc_mont_mulinto(int i, u64 a, u64 b) { return mm_into(prime(i), a, b); }
u64 orc_tomont(int i, u64 x) { return to_mont(prime(i), x); }
u64 orc_toint(int i, u64 x) { return to_int(prime(i), x); }
u64 orc_mont_pow(int i, u64 b, u64 e) { return mpow(prime(i), b, e); }

/* ------------------------------------------------------------ twiddles */

/* core.h:214-238 (cover_nttshort_i): for L = 2^0..2^log_len, the entries
 * [L/2, L) are w_L^0, w_L^1, ... by a chained lazy Montgomery product. */
static void build_table(const prime_t* P, int log_len, u64* om, u64* iom) {
    om[0] = 0;
    iom[0] = 0;
    for (int lg = 0; lg <= log_len; lg++) {
        u64 L = 1ull << lg, cnt = L / 2;
        u64 root1 = mpow(P, P->mont_g, (P->p - 1) / L);
        u64 root2 = mpow(P, P->mont_ginv, (P->p - 1) / L);
        u64 a = P->one, b = P->one;
        for (u64 t = 0; t < cnt; t++) {
            om[L / 2 + t] = a;
            a = mm_lazy(P, a, root1);
            iom[L / 2 + t] = b;
            b = mm_lazy(P, b, root2);
        }
    }
}

void orc_twiddle_table(int i, int log_len, u64* omega, u64* iomega) {
    build_table(prime(i), log_len, omega, iomega);
}

/* ---------------------------------------------------------- butterflies */

/* core.h:106-111 _transform2 */
#define T2(P, s, d)                   \
    do {                              \
        u64 _t = (s), _u = (d);       \
        (s) = m_add((P), _t, _u);     \
        (d) = m_sub((P), _t, _u);     \
    } while (0)

/* core.h:152-159 _dif_butterfly2 (Gentleman-Sande) */
static inline void bf_dif2(const prime_t* P, u64* r0, u64* r1, u64 w) {
    u64 x = m_add(P, *r0, *r1);
    u64 y = r_sub(P, *r0, *r1);
    *r0 = x;
    *r1 = mm_lazy(P, y, w);
}

/* core.h:143-150 _dit_butterfly2 (Cooley-Tukey) */
static inline void bf_dit2(const prime_t* P, u64* r0, u64* r1, u64 w) {
    u64 x = m_norm2(P, *r0);
    u64 y = mm_lazy(P, *r1, w);
    *r0 = x + y;
    *r1 = r_sub(P, x, y);
}

/* core.h:116-127 _dif_butterfly244 (split-radix L butterfly) */
static inline void bf_dif244(const prime_t* P, u64* r0, u64* r1, u64* r2, u64* r3) {
    u64 t0 = *r0 + *r2;
    u64 t2 = m_sub(P, *r0, *r2);
    u64 t1 = *r1 + *r3;
    u64 t3 = r_sub(P, *r1, *r3);
    t3 = mm_lazy(P, t3, P->w4);
    *r0 = m_norm2(P, t0);
    *r1 = m_norm2(P, t1);
    *r2 = t2 + t3;
    *r3 = r_sub(P, t2, t3);
}

/* core.h:129-141 _idit_butterfly244 */
static inline void bf_idit244(const prime_t* P, u64* r0, u64* r1, u64* r2, u64* r3) {
    u64 t0 = m_norm2(P, *r0);
    u64 t1 = m_norm2(P, *r1);
    u64 t2 = m_add(P, *r2, *r3);
    u64 t3 = r_sub(P, *r2, *r3);
    t3 = mm_lazy(P, t3, P->w4inv);
    *r0 = t0 + t2;
    *r2 = r_sub(P, t0, t2);
    *r1 = t1 + t3;
    *r3 = r_sub(P, t1, t3);
}

/* core.h:316-333 4-point tails */
static inline void dif4(const prime_t* P, u64* a) {
    u64 t0 = a[0], t1 = a[1], t2 = a[2], t3 = a[3];
    T2(P, t0, t2);
    T2(P, t1, t3);
    t3 = mm_lazy(P, t3, P->w4);
    a[0] = m_add(P, t0, t1);
    a[1] = m_sub(P, t0, t1);
    a[2] = m_add(P, t2, t3);
    a[3] = m_sub(P, t2, t3);
}
static inline void idit4(const prime_t* P, u64* a) {
    u64 t0 = a[0], t1 = a[1], t2 = a[2], t3 = a[3];
    T2(P, t0, t1);
    T2(P, t2, t3);
    t3 = mm_lazy(P, t3, P->w4inv);
    a[0] = m_add(P, t0, t2);
    a[1] = m_add(P, t1, t3);
    a[2] = m_sub(P, t0, t2);
    a[3] = m_sub(P, t1, t3);
}

/* core.h:357-411 8-point tails */
static inline void dif8(const prime_t* P, u64* a) {
    u64 t0 = a[0], t1 = a[1], t2 = a[2], t3 = a[3], t4 = a[4], t5 = a[5], t6 = a[6], t7 = a[7];
    T2(P, t0, t4);
    T2(P, t1, t5);
    T2(P, t2, t6);
    T2(P, t3, t7);
    t5 = mm_lazy(P, t5, P->w8[0]);
    t6 = mm_lazy(P, t6, P->w8[1]);
    t7 = mm_lazy(P, t7, P->w8[2]);
    T2(P, t0, t2);
    T2(P, t1, t3);
    T2(P, t4, t6);
    T2(P, t5, t7);
    t3 = mm_lazy(P, t3, P->w4);
    t7 = mm_lazy(P, t7, P->w4);
    a[0] = m_add(P, t0, t1);
    a[1] = m_sub(P, t0, t1);
    a[2] = m_add(P, t2, t3);
    a[3] = m_sub(P, t2, t3);
    a[4] = m_add(P, t4, t5);
    a[5] = m_sub(P, t4, t5);
    a[6] = m_add(P, t6, t7);
    a[7] = m_sub(P, t6, t7);
}
static inline void idit8(const prime_t* P, u64* a) {
    u64 t0 = a[0], t1 = a[1], t2 = a[2], t3 = a[3], t4 = a[4], t5 = a[5], t6 = a[6], t7 = a[7];
    T2(P, t0, t1);
    T2(P, t2, t3);
    T2(P, t4, t5);
    T2(P, t6, t7);
    t3 = mm_lazy(P, t3, P->w4inv);
    t7 = mm_lazy(P, t7, P->w4inv);
    T2(P, t0, t2);
    T2(P, t1, t3);
    T2(P, t4, t6);
    T2(P, t5, t7);
    t5 = mm_lazy(P, t5, P->w8inv[0]);
    t6 = mm_lazy(P, t6, P->w8inv[1]);
    t7 = mm_lazy(P, t7, P->w8inv[2]);
    a[0] = m_add(P, t0, t4);
    a[1] = m_add(P, t1, t5);
    a[2] = m_add(P, t2, t6);
    a[3] = m_add(P, t3, t7);
    a[4] = m_sub(P, t0, t4);
    a[5] = m_sub(P, t1, t5);
    a[6] = m_sub(P, t2, t6);
    a[7] = m_sub(P, t3, t7);
}

static inline unsigned log2u(u64 n) { return n ? 63u - (unsigned)__builtin_clzll(n) : 0u; }

/* Planted mistake in force (tests only; 0 in every normal use). */
static int g_plant = 0;
void orc_plant(int kind) { g_plant = kind; }

/* core.h:443-474 dif_i: radix-2^2 passes from rank=len down to 16, then a
 * 4-point tail (even log2) or 8-point tail (odd log2).  The first block of
 * the tail is the "_len" variant that shrinks to 2 / 1 points for len < 4/8. */
static void dif_core(const prime_t* P, u64* a, const u64* om, u64 len) {
    u64 rank = len;
    for (; rank >= 16; rank /= 4) {
        u64 gap = rank / 4;
        const u64* w = om + rank / 2;
        const u64* wl = om + rank / 4;
        for (u64 jj = 0; jj < len; jj += rank) {
            u64* b = a + jj;
            for (u64 ii = 0; ii < gap; ii++) {
                u64 t0 = b[ii], t1 = b[gap + ii], t2 = b[2 * gap + ii], t3 = b[3 * gap + ii];
                u64 o = wl[ii];
                if (g_plant == 3 && ii + 1 < gap) o = wl[ii + 1];
                bf_dif2(P, &t0, &t2, w[ii]);
                bf_dif2(P, &t1, &t3, w[gap + ii]);
                bf_dif2(P, &t0, &t1, o);
                bf_dif2(P, &t2, &t3, o);
                b[ii] = t0;
                b[gap + ii] = t1;
                b[2 * gap + ii] = t2;
                b[3 * gap + ii] = t3;
            }
        }
    }
    if (log2u(rank) % 2 == 0) {
        if (len >= 4) {
            for (u64 ii = 0; ii < len; ii += 4) dif4(P, a + ii);
        } else if (len >= 2) {
            T2(P, a[0], a[1]);
        }
    } else {
        if (len >= 8) {
            for (u64 ii = 0; ii < len; ii += 8) dif8(P, a + ii);
        } else if (len >= 4) {
            dif4(P, a);
        } else if (len >= 2) {
            T2(P, a[0], a[1]);
        }
    }
}

/* core.h:476-509 idit_i: the tail first, then radix-2^2 DIT passes. */
static void idit_core(const prime_t* P, u64* a, const u64* iom, u64 len) {
    u64 rank;
    if (log2u(len) % 2 == 0) {
        if (len >= 4) {
            for (u64 ii = 0; ii < len; ii += 4) idit4(P, a + ii);
        } else if (len >= 2) {
            T2(P, a[0], a[1]);
        }
        rank = 16;
    } else {
        if (len >= 8) {
            for (u64 ii = 0; ii < len; ii += 8) idit8(P, a + ii);
        } else if (len >= 4) {
            idit4(P, a);
        } else if (len >= 2) {
            T2(P, a[0], a[1]);
        }
        rank = 32;
    }
    for (; rank <= len; rank *= 4) {
        u64 gap = rank / 4;
        const u64* w = iom + rank / 2;
        const u64* wl = iom + rank / 4;
        for (u64 jj = 0; jj < len; jj += rank) {
            u64* b = a + jj;
            for (u64 ii = 0; ii < gap; ii++) {
                u64 t0 = b[ii], t1 = b[gap + ii], t2 = b[2 * gap + ii], t3 = b[3 * gap + ii];
                u64 o = wl[ii];
                bf_dit2(P, &t0, &t1, o);
                bf_dit2(P, &t2, &t3, o);
                bf_dit2(P, &t0, &t2, w[ii]);
                bf_dit2(P, &t1, &t3, w[gap + ii]);
                b[ii] = t0;
                b[gap + ii] = t1;
                b[2 * gap + ii] = t2;
                b[3 * gap + ii] = t3;
            }
        }
    }
}

typedef struct {
    u64 len; /* table covers lengths up to len */
    u64* om;
    u64* iom;
} table_t;

static int table_alloc(const prime_t* P, table_t* T, u64 ntt_len) {
    T->len = ntt_len < LONG_THRESHOLD ? ntt_len : LONG_THRESHOLD;
    if (T->len == 0) T->len = 1;
    T->om = (u64*)malloc(T->len * sizeof(u64));
    T->iom = (u64*)malloc(T->len * sizeof(u64));
    if (!T->om || !T->iom) return -1;
    build_table(P, (int)log2u(T->len), T->om, T->iom);
    if (g_plant == 2) {
        u64* t = T->om;
        T->om = T->iom;
        T->iom = t;
    }
    return 0;
}
static void table_free(table_t* T) {
    free(T->om);
    free(T->iom);
}

void orc_dif_raw(int i, u64* data, u64 len) {
    const prime_t* P = prime(i);
    table_t T;
    if (table_alloc(P, &T, len)) abort();
    dif_core(P, data, T.om, len);
    table_free(&T);
}

void orc_idit_raw(int i, u64* data, u64 len) {
    const prime_t* P = prime(i);
    table_t T;
    if (table_alloc(P, &T, len)) abort();
    idit_core(P, data, T.iom, len);
    table_free(&T);
}

/* ------------------------------------------------------- convolutions */

static u64 inv_len_mont(const prime_t* P, u64 n) {
    /* fast_mul.c:59-61: inv_len = qpow(tomont(n), p-2) */
    return mpow(P, to_mont(P, n), P->p - 2);
}

/* Split-radix forward step of fast_mul.c:74-95 on one array. */
static void fwd_split_step(const prime_t* P, u64* a, u64 n) {
    const u64 q = n / 4;
    u64 u1 = mpow(P, to_mont(P, P->g), (P->p - 1) / n);
    u64 u3 = mpow(P, u1, 3);
    u64 o1 = P->one, o3 = P->one;
    for (u64 ii = 0; ii < q; ii++) {
        u64 t0 = a[ii], t1 = a[q + ii], t2 = a[2 * q + ii], t3 = a[3 * q + ii];
        bf_dif244(P, &t0, &t1, &t2, &t3);
        a[ii] = t0;
        a[q + ii] = t1;
        a[2 * q + ii] = mm_lazy(P, t2, o1);
        a[3 * q + ii] = mm_lazy(P, t3, o3);
        o1 = mm_into(P, o1, u1);
        o3 = mm_into(P, o3, u3);
    }
}

/* Split-radix inverse step of fast_mul.c:99-132 (N^-1 folded in if norm). */
static void inv_split_step(const prime_t* P, u64* out, u64 n, int norm) {
    const u64 q = n / 4;
    u64 u1 = mpow(P, to_mont(P, P->ginv), (P->p - 1) / n);
    u64 u3 = mpow(P, u1, 3);
    if (norm) {
        u64 il = inv_len_mont(P, n);
        u64 o1 = il, o3 = il;
        for (u64 ii = 0; ii < q; ii++) {
            u64 t0 = mm_lazy(P, out[ii], il);
            u64 t1 = mm_lazy(P, out[q + ii], il);
            u64 t2 = mm_lazy(P, out[2 * q + ii], o1);
            u64 t3 = mm_lazy(P, out[3 * q + ii], o3);
            bf_idit244(P, &t0, &t1, &t2, &t3);
            out[ii] = t0;
            out[q + ii] = t1;
            out[2 * q + ii] = t2;
            out[3 * q + ii] = t3;
            o1 = mm_into(P, o1, u1);
            o3 = mm_into(P, o3, u3);
        }
    } else {
        u64 o1 = P->one, o3 = P->one;
        for (u64 ii = 0; ii < q; ii++) {
            u64 t0 = out[ii], t1 = out[q + ii];
            u64 t2 = mm_lazy(P, out[2 * q + ii], o1);
            u64 t3 = mm_lazy(P, out[3 * q + ii], o3);
            bf_idit244(P, &t0, &t1, &t2, &t3);
            out[ii] = t0;
            out[q + ii] = t1;
            out[2 * q + ii] = t2;
            out[3 * q + ii] = t3;
            o1 = mm_into(P, o1, u1);
            o3 = mm_into(P, o3, u3);
        }
    }
}

/* fast_mul.c:51-133 conv_rec_i: cyclic convolution mod p of two arrays. */
static void conv_rec(const prime_t* P, u64* in1, u64* in2, u64* out, const table_t* T, u64 n, int norm) {
    if (n <= LONG_THRESHOLD) {
        dif_core(P, in1, T->om, n);
        dif_core(P, in2, T->om, n);
        if (norm) {
            u64 il = inv_len_mont(P, n);
            for (u64 ii = 0; ii < n; ii++) out[ii] = mm_into(P, mm_lazy(P, in1[ii], in2[ii]), il);
        } else {
            for (u64 ii = 0; ii < n; ii++) out[ii] = mm_lazy(P, in1[ii], in2[ii]);
        }
        idit_core(P, out, T->iom, n);
        return;
    }
    const u64 q = n / 4;
    fwd_split_step(P, in1, n);
    fwd_split_step(P, in2, n);
    conv_rec(P, in1, in2, out, T, n / 2, 0);
    conv_rec(P, in1 + 2 * q, in2 + 2 * q, out + 2 * q, T, n / 4, 0);
    conv_rec(P, in1 + 3 * q, in2 + 3 * q, out + 3 * q, T, n / 4, 0);
    inv_split_step(P, out, n, norm);
}

/* fast_mul.c:214-288 conv_sqr_i */
static void conv_sqr(const prime_t* P, u64* in1, u64* out, const table_t* T, u64 n, int norm) {
    if (n <= LONG_THRESHOLD) {
        dif_core(P, in1, T->om, n);
        if (norm) {
            u64 il = inv_len_mont(P, n);
            for (u64 ii = 0; ii < n; ii++) out[ii] = mm_into(P, mm_lazy(P, in1[ii], in1[ii]), il);
        } else {
            for (u64 ii = 0; ii < n; ii++) out[ii] = mm_lazy(P, in1[ii], in1[ii]);
        }
        idit_core(P, out, T->iom, n);
        return;
    }
    const u64 q = n / 4;
    fwd_split_step(P, in1, n);
    conv_sqr(P, in1, out, T, n / 2, 0);
    conv_sqr(P, in1 + 2 * q, out + 2 * q, T, n / 4, 0);
    conv_sqr(P, in1 + 3 * q, out + 3 * q, T, n / 4, 0);
    inv_split_step(P, out, n, norm);
}

/* Forward half of conv_rec on one array (used for the spectrum check). */
static void fwd_rec(const prime_t* P, u64* a, const table_t* T, u64 n) {
    if (n <= LONG_THRESHOLD) {
        dif_core(P, a, T->om, n);
        return;
    }
    const u64 q = n / 4;
    fwd_split_step(P, a, n);
    fwd_rec(P, a, T, n / 2);
    fwd_rec(P, a + 2 * q, T, n / 4);
    fwd_rec(P, a + 3 * q, T, n / 4);
}

void orc_forward_spectrum(int i, const u64* limbs, u64 len, u64 ntt_len, u64* out) {
    const prime_t* P = prime(i);
    for (u64 k = 0; k < ntt_len; k++) out[k] = k < len ? to_mont(P, limbs[k]) : 0;
    table_t T;
    if (table_alloc(P, &T, ntt_len)) abort();
    fwd_rec(P, out, &T, ntt_len);
    table_free(&T);
    for (u64 k = 0; k < ntt_len; k++) out[k] %= P->p;
}

/* ------------------------------------------------------------------ CRT */

static inline void add192(u64 r[3], const u64 x[3]) {
    u128 s = (u128)r[0] + x[0];
    r[0] = (u64)s;
    s = (s >> 64) + r[1] + x[1];
    r[1] = (u64)s;
    r[2] = r[2] + x[2] + (u64)(s >> 64);
}
static inline int lt192(const u64 a[3], const u64 b[3]) {
    if (a[2] != b[2]) return a[2] < b[2];
    if (a[1] != b[1]) return a[1] < b[1];
    return a[0] < b[0];
}
static inline void sub192(u64 r[3], const u64 x[3]) {
    u128 d = (u128)r[0] - x[0];
    r[0] = (u64)d;
    u64 b = (u64)(d >> 64) & 1;
    d = (u128)r[1] - x[1] - b;
    r[1] = (u64)d;
    b = (u64)(d >> 64) & 1;
    r[2] = r[2] - x[2] - b;
}
/* u128 x u64 -> u192 (macro.h:162-170) */
static inline void mul128x64(u64 r[3], const u64 a[2], u64 b) {
    u128 lo = (u128)a[0] * b;
    u128 hi = (u128)a[1] * b;
    r[0] = (u64)lo;
    u128 mid = (lo >> 64) + (u64)hi;
    r[1] = (u64)mid;
    r[2] = (u64)(hi >> 64) + (u64)(mid >> 64);
}

/* core.h:77-104 crt3 */
void orc_crt3(u64 a, u64 b, u64 c, u64 out[3]) {
    orc_init();
    u64 x[3] = {a, b, c};
    out[0] = out[1] = out[2] = 0;
    for (int i = 0; i < 3; i++) {
        if (g_plant == 1 && i == 2) continue;
        u64 v = to_int(&PR[i], mm_lazy(&PR[i], x[i], CRT_MINV[i]));
        u64 t[3];
        mul128x64(t, CRT_M[i], v);
        add192(out, t);
    }
    if (!lt192(out, CRT_P123)) sub192(out, CRT_P123);
    if (!lt192(out, CRT_P123)) sub192(out, CRT_P123);
}

/* ------------------------------------------------------------ products */

/* fast_mul.c:299-306 */
static u64 ceil_pow2(u64 n) {
    n--;
    for (int s = 1; s < 64; s *= 2) n |= n >> s;
    return n + 1;
}

/* fast_mul.c:378-389: serial 192-bit carry over the CRT coefficients. */
static void crt_carry(u64* const bufs[3], u64 conv_len, u64* out) {
    u64 acc[3] = {0, 0, 0};
    for (u64 k = 0; k < conv_len; k++) {
        u64 t[3];
        orc_crt3(bufs[0][k], bufs[1][k], bufs[2][k], t);
        add192(acc, t);
        out[k] = acc[0];
        acc[0] = acc[1];
        acc[1] = g_plant == 5 ? 0 : acc[2];
        acc[2] = 0;
    }
    out[conv_len] = acc[0];
}

static u64* xalloc(u64 n) {
    u64* p = NULL;
    if (posix_memalign((void**)&p, 64, n * sizeof(u64)) != 0 || p == NULL) {
        fprintf(stderr, "Memory allocation failed.\n");
        abort();
    }
    return p;
}

/* fast_mul.c:308-396 abs_mul64 */
void orc_abs_mul64(const u64* in1, u64 len1, const u64* in2, u64 len2, u64* out) {
    orc_init();
    if (len1 == 0 || len2 == 0) {
        memset(out, 0, (len1 + len2) * sizeof(u64));
        return;
    }
    const u64 out_len = len1 + len2, conv_len = out_len - 1;
    const u64 n = ceil_pow2(conv_len);
    u64* bufs[3] = {xalloc(n), xalloc(n), xalloc(n)};
    u64* tmp = xalloc(n);
    for (int i = 0; i < 3; i++) {
        const prime_t* P = &PR[i];
        for (u64 k = 0; k < n; k++) bufs[i][k] = k < len1 ? to_mont(P, in1[k]) : 0;
    }
    for (int i = 0; i < 3; i++) {
        const prime_t* P = &PR[i];
        table_t T;
        if (table_alloc(P, &T, n)) abort();
        const u64 pad2 = g_plant == 4 ? (len1 < len2 ? len1 : len2) : len2;
        for (u64 k = 0; k < n; k++) tmp[k] = k < pad2 ? to_mont(P, in2[k]) : 0;
        conv_rec(P, bufs[i], tmp, bufs[i], &T, n, 1);
        table_free(&T);
    }
    free(tmp);
    crt_carry(bufs, conv_len, out);
    for (int i = 0; i < 3; i++) free(bufs[i]);
}

/* fast_mul.c:398-460 abs_sqr64 */
void orc_abs_sqr64(const u64* in1, u64 len1, u64* out) {
    orc_init();
    if (len1 == 0) return;
    const u64 out_len = 2 * len1, conv_len = out_len - 1;
    const u64 n = ceil_pow2(conv_len);
    u64* bufs[3] = {xalloc(n), xalloc(n), xalloc(n)};
    for (int i = 0; i < 3; i++) {
        const prime_t* P = &PR[i];
        for (u64 k = 0; k < n; k++) bufs[i][k] = k < len1 ? to_mont(P, in1[k]) : 0;
        table_t T;
        if (table_alloc(P, &T, n)) abort();
        conv_sqr(P, bufs[i], bufs[i], &T, n, 1);
        table_free(&T);
    }
    crt_carry(bufs, conv_len, out);
    for (int i = 0; i < 3; i++) free(bufs[i]);
}

/* ---------------------------------------------------- independent checks */

void orc_schoolbook(const u64* a, u64 na, const u64* b, u64 nb, u64* out) {
    memset(out, 0, (na + nb) * sizeof(u64));
    for (u64 i = 0; i < na; i++) {
        u64 carry = 0;
        for (u64 j = 0; j < nb; j++) {
            u128 t = (u128)a[i] * b[j] + out[i + j] + carry;
            out[i + j] = (u64)t;
            carry = (u64)(t >> 64);
        }
        out[i + nb] = carry;
    }
}

u64 orc_mod_q(const u64* a, u64 n, u64 q) {
    u128 r = 0;
    for (u64 k = n; k-- > 0;) r = ((r << 64) | a[k]) % q;
    return (u64)r;
}

u64 orc_mulmod_q(u64 a, u64 b, u64 q) { return (u64)(((u128)a * b) % q); }

void orc_prefix_mod_q(const u64* a, u64 n, u64 q, u64* out) {
    const u64 b = (u64)((((u128)1) << 64) % q); /* 2^64 mod q */
    u64 acc = 0, pw = 1 % q;                    /* pw = 2^(64 k) mod q */
    out[0] = 0;
    for (u64 k = 0; k < n; k++) {
        acc = (u64)(((u128)acc + (u128)(a[k] % q) * pw) % q);
        pw = (u64)(((u128)pw * b) % q);
        out[k + 1] = acc;
    }
}

void orc_splitmix64_fill(u64 seed, u64* out, u64 n, int top_nonzero) {
    u64 s = seed;
    for (u64 k = 0; k < n; k++) {
        s += 0x9E3779B97F4A7C15ull;
        u64 z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        out[k] = z ^ (z >> 31);
    }
    if (top_nonzero && n > 0 && out[n - 1] == 0) out[n - 1] = 1;
}
