// fmh_field.h -- arithmetic modulo the three NTT primes, for gfx950 device
// code and for the host planner.
//
// The primes, roots and R = 2^64 are those of src/data.h:69-81 in the
// reference; every prime has the special form p = c * 2^k + 1:
//   p0 = 69 * 2^55 + 1   g = 5
//   p1 = 27 * 2^56 + 1   g = 5
//   p2 = 29 * 2^57 + 1   g = 3
// so -p^-1 mod 2^64 = p - 2 (src/data.h:106-108) and 4p < 2^64.
//
// Value ranges used by the kernels (all proofs are one line):
//   * redc(x, w) = x * w * 2^-64 mod p lies in [0, 2p) for any 64-bit x when
//     w < p, because (x*w + m*p) / 2^64 < (2^64 * p + 2^64 * p) / 2^64.
//   * redc(x, y) lies in [0, 2p) when x, y < 2p, because 4p^2 < 2^64 * p.
//   * add2p(x, y) of x, y in [0, 2p) lies in [0, 2p); x - y + 2p lies in
//     (0, 4p) and fits 64 bits since 4p < 2^64.
// Data words between stages therefore always lie in [0, 2p); twiddles are
// canonical Montgomery forms (w * 2^64 mod p, in [0, p)).
#pragma once

#include <stdint.h>

typedef unsigned long long u64;
typedef unsigned int u32;

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

#if defined(__HIPCC__)

// Montgomery reduction of the 128-bit product a*b: (a*b + m*p) / 2^64 with
// m = (a*b mod 2^64) * (p - 2) mod 2^64.  The low words of a*b and m*p sum to
// 0 or 2^64, so the carry into the high word is (lo != 0).
template <int P>
__device__ __forceinline__ u64 redc(u64 a, u64 b) {
    constexpr u64 p = Prime<P>::p;
    const u64 lo = a * b;
    const u64 hi = __umul64hi(a, b);
    const u64 m = lo * (p - 2);
    const u64 mh = __umul64hi(m, p);
    return hi + mh + (lo != 0 ? 1ull : 0ull);
}

template <int P>
__device__ __forceinline__ u64 add2p(u64 a, u64 b) {
    constexpr u64 p2 = 2 * Prime<P>::p;
    const u64 s = a + b;
    return s >= p2 ? s - p2 : s;
}

template <int P>
__device__ __forceinline__ u64 norm2p(u64 a) {
    constexpr u64 p2 = 2 * Prime<P>::p;
    return a >= p2 ? a - p2 : a;
}

template <int P>
__device__ __forceinline__ u64 canon(u64 a) {  // [0, 2p) -> [0, p)
    constexpr u64 p = Prime<P>::p;
    return a >= p ? a - p : a;
}

// Any 64-bit word x -> x mod p in [0, p), using p = c 2^k + 1:
// with q = floor((x >> k) / c), x - q p = ((x>>k) - q c) 2^k + (x mod 2^k) - q
// lies in [-q, p - 2]; a negative result is fixed by adding p.
template <int P>
__device__ __forceinline__ u64 reduce_word(u64 x) {
    constexpr u64 p = Prime<P>::p;
    const u32 xh = (u32)(x >> Prime<P>::k);
    const u32 q = xh / (u32)Prime<P>::c;
    u64 r = x - (u64)q * p;
    return ((long long)r < 0) ? r + p : r;
}

// Gentleman-Sande (DIF) butterfly: (x, y) -> (x + y, (x - y) w).
template <int P>
__device__ __forceinline__ void bf_dif(u64& x, u64& y, u64 w) {
    constexpr u64 p2 = 2 * Prime<P>::p;
    const u64 s = add2p<P>(x, y);
    y = redc<P>(x - y + p2, w);
    x = s;
}

// Cooley-Tukey (DIT) butterfly: (x, y) -> (x + y w, x - y w).
template <int P>
__device__ __forceinline__ void bf_dit(u64& x, u64& y, u64 w) {
    constexpr u64 p2 = 2 * Prime<P>::p;
    const u64 t = redc<P>(y, w);
    y = norm2p<P>(x - t + p2);
    x = add2p<P>(x, t);
}

#endif  // __HIPCC__

}  // namespace fmh
