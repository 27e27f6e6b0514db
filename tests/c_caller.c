/* A reference-style C caller of the drop-in ABI: includes fmh.h in place of
 * src/fast_mul.h, links libfmh.so, and checks abs_mul64 / abs_sqr64 against
 * a schoolbook product computed here.  Prints "c caller ok" on success. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fmh.h"

typedef unsigned __int128 u128;

static u64 state = 12345;
static u64 next64(void) { /* splitmix64 */
    u64 z = (state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static void schoolbook(const u64* a, u64 na, const u64* b, u64 nb, u64* out) {
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

static int check(u64 na, u64 nb) {
    u64* a = malloc(na * sizeof(u64));
    u64* b = malloc(nb * sizeof(u64));
    u64* got = malloc((na + nb) * sizeof(u64));
    u64* want = malloc((na + nb) * sizeof(u64));
    for (u64 i = 0; i < na; i++) a[i] = next64();
    for (u64 i = 0; i < nb; i++) b[i] = next64();
    abs_mul64(a, na, b, nb, got);
    schoolbook(a, na, b, nb, want);
    int ok = memcmp(got, want, (na + nb) * sizeof(u64)) == 0;
    u64* sq = malloc(2 * na * sizeof(u64));
    u64* sqw = malloc(2 * na * sizeof(u64));
    abs_sqr64(a, na, sq);
    schoolbook(a, na, a, na, sqw);
    ok = ok && memcmp(sq, sqw, 2 * na * sizeof(u64)) == 0;
    /* the output may alias an input */
    u64* alias = malloc((na + nb) * sizeof(u64));
    memcpy(alias, a, na * sizeof(u64));
    abs_mul64(alias, na, b, nb, alias);
    ok = ok && memcmp(alias, want, (na + nb) * sizeof(u64)) == 0;
    free(a), free(b), free(got), free(want), free(sq), free(sqw), free(alias);
    if (!ok) fprintf(stderr, "mismatch at %llu x %llu\n", na, nb);
    return ok;
}

int main(void) {
    const u64 sizes[][2] = {{1, 1}, {3, 2}, {100, 7}, {1000, 1000}, {3000, 2500}, {5000, 40}};
    for (unsigned i = 0; i < sizeof(sizes) / sizeof(sizes[0]); i++)
        if (!check(sizes[i][0], sizes[i][1])) return 1;
    printf("c caller ok\n");
    return 0;
}
