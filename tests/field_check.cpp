// CPU check of the device arithmetic in fast_mul_ntt_amd/csrc/fmh_field.h:
// the same functions, compiled with g++, against __int128 definitions.
// Exit status 0 = all checks passed.
#include <stdio.h>
#include <stdlib.h>

#include <initializer_list>

#include "fmh_field.h"

typedef unsigned __int128 u128;
using namespace fmh;

static u64 s_state = 0x1234567890ABCDEFull;
static u64 next64() {  // splitmix64
    u64 z = (s_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Reference REDC (src/macro.h:109-118): (a b + m p) / 2^64,
// m = a b (-p^-1) mod 2^64.
static u64 redc_ref(u64 a, u64 b, u64 p) {
    u64 inv = p;
    for (int i = 0; i < 6; i++) inv *= 2 - p * inv;
    const u128 x = (u128)a * b;
    const u64 m = (u64)x * (0 - inv);
    const u128 s = x + (u128)m * p;  // < 2^128 for the inputs used below
    return (u64)(s >> 64);
}

// The kernels' subtractive REDC equals the reference's, plus p when
// a b = 0 (mod 2^64) (the quotient digit is then 0 in both forms).
static u64 redc_expected(u64 a, u64 b, u64 p) {
    const u64 r = redc_ref(a, b, p);
    return (u64)((u128)a * b) == 0 ? r + p : r;
}

static int fails = 0;
#define CHECK(c)                                                    \
    do {                                                            \
        if (!(c)) {                                                 \
            if (fails++ < 10) fprintf(stderr, "FAIL %s:%d %s\n", __FILE__, __LINE__, #c); \
        }                                                           \
    } while (0)

template <int P>
static void check_prime(long iters) {
    const u64 p = Prime<P>::p, p2 = 2 * p, p4 = 4 * p;
    const u64 edge[] = {0, 1, 2, p - 1, p, p + 1, p2 - 1, p2, p4 - 1, 0xFFFFFFFFull, 0x100000000ull,
                        ~0ull, ~0ull - 1, (u64)1 << 63};
    // redc on every domain the kernels use: (x < 2^64, w < p) and (x, y < 4p with xy < p 2^64)
    for (u64 a : edge)
        for (u64 b : edge) {
            if (b < p) CHECK(redc<P>(a, b) == redc_expected(a, b, p) && redc<P>(a, b) < p2);
            if ((u128)a * b < (u128)p << 64) CHECK(redc<P>(a, b) == redc_expected(a, b, p) && redc<P>(a, b) < p2);
        }
    for (long i = 0; i < iters; i++) {
        const u64 a = next64(), w = next64() % p;
        const u64 r = redc<P>(a, w);
        CHECK(r == redc_expected(a, w, p));
        CHECK(r < p2);
        const u64 x = next64() % p2, y = next64() % p2;
        CHECK(redc<P>(x, y) == redc_expected(x, y, p));
        CHECK(redc<P>(x, y) < p2);
        const u64 z = next64() % p4;
        CHECK(redc<P>(z, w) == redc_expected(z, w, p) && redc<P>(z, w) < p2);
        // low word of the product zero: the one case where the forms differ
        const u64 za = next64() << 32, zb = (next64() % p) & ~0xFFFFFFFFull;
        CHECK(redc<P>(za, zb) == redc_ref(za, zb, p) + p && redc<P>(za, zb) < p2);
        // reduce_word = x mod p
        const u64 v = next64();
        CHECK(reduce_word<P>(v) == v % p);
        // butterflies: ranges and congruences (values mod p)
        u64 bx = next64() % p2, by = next64() % p2;
        const u64 ox = bx, oy = by;
        bf_dif<P>(bx, by, w);
        CHECK(bx < p2 && by < p2);
        CHECK(bx % p == (u64)(((u128)ox + oy) % p));
        const u64 rinv_w = redc_ref(1, w, p) % p;  // w * 2^-64 mod p
        CHECK(by % p == (u64)((u128)((ox + p2 - oy) % p) * rinv_w % p));
        u64 dx = next64() % p4, dy = next64() % p4;
        const u64 ex = dx, ey = dy;
        bf_dit<P>(dx, dy, w);
        CHECK(dx < p4 && dy < p4);
        const u64 t = (u64)((u128)(ey % p) * rinv_w % p);
        CHECK(dx % p == (ex % p + t) % p);
        CHECK(dy % p == (ex % p + p - t) % p);
        // the w = 1 butterflies: same values mod p as the plain ones with the
        // Montgomery form of 1, same ranges
        const u64 one_m = (u64)(((u128)1 << 64) % p);
        u64 ux = next64() % p2, uy = next64() % p2, vx = ux, vy = uy;
        bf_dif_one<P>(ux, uy);
        bf_dif<P>(vx, vy, one_m);
        CHECK(ux < p2 && uy < p2 && ux % p == vx % p && uy % p == vy % p);
        u64 sx = next64() % p4, sy = next64() % p4, tx = sx, ty = sy;
        bf_dit_one<P>(sx, sy);
        bf_dit<P>(tx, ty, one_m);
        CHECK(sx < p4 && sy < p4 && sx % p == tx % p && sy % p == ty % p);
        const u64 q = next64() % p4;
        CHECK(norm2p<P>(q) == (q >= p2 ? q - p2 : q));
        const u64 c2 = next64() % p2;
        CHECK(canon<P>(c2) == c2 % p);
    }
    for (u64 v = 0; v < 4096; v++) {  // reduce_word around every multiple of 2^k
        const u64 base = (v << Prime<P>::k) >> 2;
        for (u64 d : {0ull, 1ull, 2ull, 63ull}) {
            CHECK(reduce_word<P>(base + d) == (base + d) % p);
            CHECK(reduce_word<P>(base - d) == (base - d) % p);
        }
    }
}

int main(int argc, char** argv) {
    const long iters = argc > 1 ? atol(argv[1]) : 200000;
    check_prime<0>(iters);
    check_prime<1>(iters);
    check_prime<2>(iters);
    printf("%s (%d failures)\n", fails ? "FAILED" : "ok", fails);
    return fails ? 1 : 0;
}
