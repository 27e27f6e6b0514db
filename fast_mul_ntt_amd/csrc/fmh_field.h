// fmh_field.h -- arithmetic modulo the three NTT primes, shared by the gfx950
// kernels and by host code (the CPU checks of this arithmetic compile the
// same functions with g++).
//
// The primes, roots and R = 2^64 are those of src/data.h:69-81 in the
// reference; every prime has the form p = c * 2^k + 1 with k >= 55:
//   p0 = 69 * 2^55 + 1   g = 5
//   p1 = 27 * 2^56 + 1   g = 5
//   p2 = 29 * 2^57 + 1   g = 3
// so p = PHI * 2^32 + 1 with PHI = p >> 32, and 4p < 2^64.
//
// Montgomery reduction (R = 2^64) is done one 32-bit word at a time in its
// subtractive form: for p = 1 (mod 2^32), p^-1 = 1 (mod 2^32), so the
// per-word quotient digit is the word itself and each word costs one
// 32x32+64 multiply-add with 2^32 - PHI.  redc returns
//     (x - m p) / 2^64 + p,   m = x p^-1 mod 2^64,
// which lies in (0, 2p) and is congruent to x 2^-64.  It equals the
// reference's REDC (src/macro.h:109-118), (x + m' p) / 2^64 with
// m' = -x p^-1 mod 2^64, except when x = 0 (mod 2^64), where it is p more.
//
// Value ranges used by the kernels:
//   * redc(x, y) < 2p whenever x * y < p * 2^64, because the result is below
//     x*y/2^64 + p.  This holds for any 64-bit x with y < p (a canonical
//     twiddle), and for x, y < 2p (4p^2 < 2^64 p).
//   * forward (DIF) data stay in [0, 2p): (x + y) is reduced once, and
//     x - y + 2p lies in (0, 4p), which fits 64 bits.
//   * inverse (DIT) data stay in [0, 4p) (Harvey's lazy butterfly): the upper
//     input is reduced below 2p, the product below 2p, so both outputs lie
//     below 4p < 2^64.
#pragma once

#include <stdint.h>

typedef unsigned long long u64;
typedef unsigned int u32;
typedef unsigned char u8;

#if defined(__HIPCC__)
#define FMH_HD __host__ __device__ __forceinline__
#else
#define FMH_HD static inline
#endif

namespace fmh {

constexpr int kNumPrimes = 3;

template <int P>
struct Prime;
template <>
struct Prime<0> {
    static constexpr u64 p = 2485986994308513793ull;
    static constexpr u64 c = 69;
    static constexpr int k = 55;
    static constexpr u64 g = 5;
};
template <>
struct Prime<1> {
    static constexpr u64 p = 1945555039024054273ull;
    static constexpr u64 c = 27;
    static constexpr int k = 56;
    static constexpr u64 g = 5;
};
template <>
struct Prime<2> {
    static constexpr u64 p = 4179340454199820289ull;
    static constexpr u64 c = 29;
    static constexpr int k = 57;
    static constexpr u64 g = 3;
};

static_assert(Prime<0>::p == (Prime<0>::c << Prime<0>::k) + 1, "p0 form");
static_assert(Prime<1>::p == (Prime<1>::c << Prime<1>::k) + 1, "p1 form");
static_assert(Prime<2>::p == (Prime<2>::c << Prime<2>::k) + 1, "p2 form");
static_assert(Prime<2>::p < (~0ull) / 4, "4p < 2^64 for the largest prime");

// Subtractive Montgomery reduction of x = a b in two word steps (see above).
// With x = hi 2^64 + x1 2^32 + x0:
//   word 0: m0 = x0;  (x - m0 p) / 2^32 = hi 2^32 + D,  D = x1 - x0 PHI
//   word 1: m1 = D mod 2^32;  result = hi + floor(D / 2^32) - m1 PHI (+ p)
// t = x0 (2^32 - PHI) + x1 = D + x0 2^32 is exact in 64 bits, so
// m1 = lo(t) and e = -floor(D / 2^32) = x0 - hi(t), 0 <= e < 2^30 (PHI <
// 2^30 for these primes).  The result is hi + p - e - m1 PHI.
//
// The words are formed with 32 x 32 + 64-bit multiply-adds, arranged so that
// each addend is a whole previous result wherever possible (an addend made
// of one 32-bit word and a zero costs a register move on gfx950):
//   p00 = a0 b0,  p01 = a0 b1 + hi(p00)
//   q   = a1 b0 + p01 = (a1 b0 + lo(p01)) + hi(p01) 2^32, carry c out of 64 bits:
//         lo(q) = x1, and hi(p01) + hi(a1 b0 + lo(p01)) = hi(q) + c 2^32
//   hi  = a1 b1 + hi(p01) + hi(a1 b0 + lo(p01)) = a1 b1 + hi(q) + c 2^32
//   u   = x0 NPHI + q (mod 2^64) = t + hi(q) 2^32: lo(u) = m1, hi(u) = hi(t) + hi(q)
//   f   = 1 - e = hi(u) - (x0 + hi(q) - 1)  (mod 2^32, as a signed word in (-2^30, 1])
//   acc = a1 b1 + (hi(q) + (PHI + c) 2^32) + f = hi + p - e
//   result = acc - m1 PHI = (acc - m1 2^32) + m1 (2^32 - PHI)  (mod 2^64)
// The host build runs the same steps in plain C++.
template <int P>
FMH_HD u64 redc(u64 a, u64 b) {
    constexpr u32 PHI = (u32)(Prime<P>::p >> 32);
    constexpr u32 NPHI = 0u - PHI;  // 2^32 - PHI
    static_assert(((u64)PHI << 32) + 1 == Prime<P>::p, "p = PHI 2^32 + 1");
    static_assert(PHI < (1u << 30), "e < 2^30");
    const u32 a0 = (u32)a, a1 = (u32)(a >> 32), b0 = (u32)b, b1 = (u32)(b >> 32);
    const u64 p00 = (u64)a0 * b0;
    const u64 p01 = (u64)a0 * b1 + (p00 >> 32);
    const u32 x0 = (u32)p00;
#if defined(__HIP_DEVICE_COMPILE__)
    // Written out so that each multiply-add takes a whole 64-bit addend and
    // the carry c goes straight into the next word (the compiler otherwise
    // splits the sums into 64-bit add chains).  The unused carry outputs go
    // to vcc, so no SGPRs are taken for them.
    u64 q, u, acc, r;
    u32 cp;  // PHI + c
    // the carry-in read of c (vcc) waits two states after the write
    asm("v_mad_u64_u32 %0, vcc, %2, %3, %4\n\ts_nop 1\n\tv_addc_co_u32_e32 %1, vcc, 0, %5, vcc"
        : "=&v"(q), "=v"(cp)
        : "v"(a1), "v"(b0), "v"(p01), "v"(PHI)
        : "vcc");
    const u32 hq = (u32)(q >> 32);
    asm("v_mad_u64_u32 %0, vcc, %1, %2, %3" : "=v"(u) : "v"(x0), "s"(NPHI), "v"(q) : "vcc");
    const u32 m1 = (u32)u;
    u32 s, f, fh;
    asm("v_add3_u32 %0, %1, %2, -1" : "=v"(s) : "v"(x0), "v"(hq));
    asm("v_sub_u32 %0, %1, %2" : "=v"(f) : "v"((u32)(u >> 32)), "v"(s));
    asm("v_ashrrev_i32 %0, 31, %1" : "=v"(fh) : "v"(f));
    asm("v_mad_u64_u32 %0, vcc, %1, %2, %3" : "=v"(acc) : "v"(a1), "v"(b1), "v"(((u64)cp << 32) | hq) : "vcc");
    acc += ((u64)fh << 32) | f;
    u32 acc_hi;
    asm("v_sub_u32 %0, %1, %2" : "=v"(acc_hi) : "v"((u32)(acc >> 32)), "v"(m1));
    asm("v_mad_u64_u32 %0, vcc, %1, %2, %3" : "=v"(r) : "v"(m1), "s"(NPHI), "v"(((u64)acc_hi << 32) | (u32)acc) : "vcc");
    return r;
#else
    const u64 lo10 = (u64)a1 * b0;
    const u64 q = lo10 + p01;
    const u32 c = q < p01 ? 1u : 0u;
    const u32 hq = (u32)(q >> 32);
    const u64 u = (u64)x0 * NPHI + q;
    const u32 m1 = (u32)u;
    const u32 f = (u32)(u >> 32) - (x0 + hq - 1u);
    const u64 fs = (u64)(long long)(int)f;
    u64 acc = (u64)a1 * b1 + ((((u64)(PHI + c)) << 32) | hq) + fs;
    acc -= (u64)m1 << 32;
    return acc + (u64)m1 * NPHI;
#endif
}

// 64 x 64 -> 128-bit product from 32-bit multiply-adds.
FMH_HD void mul_64x64(u64 a, u64 b, u64& lo, u64& hi) {
    const u32 a0 = (u32)a, a1 = (u32)(a >> 32), b0 = (u32)b, b1 = (u32)(b >> 32);
    const u64 p00 = (u64)a0 * b0;
    const u64 p01 = (u64)a0 * b1 + (p00 >> 32);
    const u64 p10 = (u64)a1 * b0 + (u32)p01;
    hi = (u64)a1 * b1 + (p01 >> 32) + (p10 >> 32);
    lo = (p10 << 32) | (u32)p00;
}

// x - 2p if x >= 2p (x < 4p): the unsigned minimum of x and x - 2p.
template <int P>
FMH_HD u64 norm2p(u64 x) {
    constexpr u64 p2 = 2 * Prime<P>::p;
    const u64 d = x - p2;
    return d < x ? d : x;
}

template <int P>
FMH_HD u64 add2p(u64 a, u64 b) {
    return norm2p<P>(a + b);
}

template <int P>
FMH_HD u64 canon(u64 a) {  // [0, 2p) -> [0, p)
    constexpr u64 p = Prime<P>::p;
    const u64 d = a - p;
    return d < a ? d : a;
}

// Any 64-bit word x -> x mod p in [0, p), using p = c 2^k + 1:
// with q = floor((x >> k) / c), x - q p = ((x>>k) - q c) 2^k + (x mod 2^k) - q
// lies in [-q, p - 2]; a negative result is fixed by adding p.
template <int P>
FMH_HD u64 reduce_word(u64 x) {
    constexpr u64 p = Prime<P>::p;
    const u32 xh = (u32)(x >> Prime<P>::k);
    const u32 q = xh / (u32)Prime<P>::c;
    u64 r = x - (u64)q * p;
    return ((long long)r < 0) ? r + p : r;
}

// Gentleman-Sande (DIF) butterfly, x, y in [0, 2p): (x + y, (x - y) w).
template <int P>
FMH_HD void bf_dif(u64& x, u64& y, u64 w) {
    constexpr u64 p2 = 2 * Prime<P>::p;
    const u64 s = add2p<P>(x, y);
    y = redc<P>(x + p2 - y, w);
    x = s;
}

// The DIF butterfly with w = 1: (x + y, x - y), outputs in [0, 2p).
template <int P>
FMH_HD void bf_dif_one(u64& x, u64& y) {
    constexpr u64 p2 = 2 * Prime<P>::p;
    const u64 s = add2p<P>(x, y);
    y = norm2p<P>(x + p2 - y);
    x = s;
}

// Cooley-Tukey (DIT) butterfly, x, y in [0, 4p): (x + y w, x - y w), outputs
// in [0, 4p).
template <int P>
FMH_HD void bf_dit(u64& x, u64& y, u64 w) {
    constexpr u64 p2 = 2 * Prime<P>::p;
    const u64 xr = norm2p<P>(x);
    const u64 t = redc<P>(y, w);
    x = xr + t;
    y = xr + p2 - t;
}

// The DIT butterfly with w = 1, same ranges.
template <int P>
FMH_HD void bf_dit_one(u64& x, u64& y) {
    constexpr u64 p2 = 2 * Prime<P>::p;
    const u64 xr = norm2p<P>(x);
    const u64 t = norm2p<P>(y);
    x = xr + t;
    y = xr + p2 - t;
}

}  // namespace fmh
