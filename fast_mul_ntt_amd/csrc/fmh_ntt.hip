// fmh_ntt.hip -- gfx950 kernels of the 3-prime NTT multiply and their
// launchers.
//
// Transform.  Every transform is the in-place radix-2 decimation-in-
// frequency NTT of length N = 2^n (stage s pairs x[i] and x[i + h] with
// h = 2^(n-1-s) and multiplies the difference by w_{2h}^(i mod 2h)).  Its
// output at position j is X[bitrev_n(j)], the order of the reference's dif_i
// and of its split-radix recursion (src/core.h:443-474,
// src/fast_mul.c:74-98).  The inverse runs the Cooley-Tukey stages in the
// opposite order with inverse twiddles (src/core.h:476-509); the factor N it
// leaves is removed in the CRT constants.
//
// Passes.  The n stages are cut into consecutive groups.  A group [s0, s1)
// only combines indices that differ in bits n-s1 .. n-s0-1, so a workgroup
// loads one "line" per (block, offset) pair: 2^(s1-s0) words at stride
// B = 2^(n-s1), a few adjacent lines at once so every row of the tile is one
// coalesced segment.  The group's stages then run as radix-2^2 rounds (two
// stages per LDS round trip, the reference's radix-2^2 pairing); the first
// round reads its words from global memory and the last writes them there,
// so the tile is never copied in or out on its own.  The
// last group (stride 1) is the row pass; for products it is fused with the
// pointwise product and the first inverse stages, so the spectra never
// travel to HBM.
//
// Twiddles.  One table per prime and direction, tw[h + t] = w_{2h}^t in
// Montgomery form (the layout of the reference's NTTshort, src/core.h:
// 163-170, 214-238, but canonical and covering every h < 2^log_t); a stage
// of half-span h at global index i reads tw[h + (i mod h)].
//
// Data representation.  Residues are plain values x mod p kept in [0, 2p)
// (see fmh_field.h); multiplying by a Montgomery-form twiddle keeps them
// plain.  The pointwise Montgomery product leaves a factor 2^-64 that, like
// N, is folded into the CRT constants.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <dlfcn.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "fmh_internal.h"

namespace fmh {

// ------------------------------------------------------------------ host math

typedef unsigned __int128 u128;

static u64 h_powmod(u64 b, u64 e, u64 p) {
    u128 r = 1, x = b % p;
    while (e) {
        if (e & 1) r = (r * x) % p;
        x = (x * x) % p;
        e >>= 1;
    }
    return (u64)r;
}
static u64 h_mont(u64 x, u64 p) { return (u64)((((u128)(x % p)) << 64) % p); }

static const u64 kP[3] = {Prime<0>::p, Prime<1>::p, Prime<2>::p};
static const u64 kG[3] = {Prime<0>::g, Prime<1>::g, Prime<2>::g};

// Garner constants of the CRT for transform length 2^log_n.  The inverse
// transform leaves v_i = C * 2^log_n * 2^-64 (mod p_i) for coefficient C; with
// redc(v, K) = v K 2^-64:
//   x0 = C mod p0                        = redc(v0, k00)
//   y1 = (C - x0) / p0 mod p1            = redc(v1, k11) - redc(x0, k10)
//   y2 = (C - x0 - p0 y1) / (p0 p1) mod p2
//                                        = redc(v2, k22) - redc(x0, k20) - redc(y1, k21)
//   C  = x0 + p0 y1 + p0 p1 y2           (< p0 p1 p2, no final reduction)
CrtConsts crt_consts(int log_n) {
    auto inv = [](u64 a, u64 p) { return h_powmod(a % p, p - 2, p); };
    auto mulm = [](u64 a, u64 b, u64 p) { return (u64)(((u128)a * b) % p); };
    u64 ninv[3], r[3], r2[3];
    for (int i = 0; i < 3; i++) {
        ninv[i] = inv(h_powmod(2, (u64)log_n, kP[i]), kP[i]);
        r[i] = (u64)((((u128)1) << 64) % kP[i]);
        r2[i] = mulm(r[i], r[i], kP[i]);
    }
    const u64 p0 = kP[0], p1 = kP[1], p2 = kP[2];
    const u64 p01_mod2 = (u64)(((u128)p0 * p1) % p2);
    CrtConsts c;
    c.k00 = mulm(ninv[0], r2[0], p0);
    c.k11 = mulm(mulm(ninv[1], r2[1], p1), inv(p0, p1), p1);
    c.k10 = mulm(inv(p0, p1), r[1], p1);
    c.k22 = mulm(mulm(ninv[2], r2[2], p2), inv(p01_mod2, p2), p2);
    c.k20 = mulm(inv(p01_mod2, p2), r[2], p2);
    c.k21 = mulm(inv(p1, p2), r[2], p2);
    const u128 m01 = (u128)p0 * p1;
    c.m01[0] = (u64)m01;
    c.m01[1] = (u64)(m01 >> 64);
    return c;
}

// Parses "key=value,..." into t; returns false on a malformed spec.
static bool parse_tuning(const std::string& spec, Tuning& t) {
    size_t pos = 0;
    while (pos < spec.size()) {
        size_t end = spec.find(',', pos);
        if (end == std::string::npos) end = spec.size();
        const std::string kv = spec.substr(pos, end - pos);
        pos = end + 1;
        if (kv.empty()) continue;
        const size_t eq = kv.find('=');
        if (eq == std::string::npos) return false;
        const std::string k = kv.substr(0, eq);
        const int v = atoi(kv.c_str() + eq + 1);
        if (k == "col_logc") t.col_logc = v;
        else if (k == "col_maxl") t.col_maxl = v;
        else if (k == "row_maxl") t.row_maxl = v;
        else if (k == "row_tile") t.row_tile = v;
        else if (k == "col_threads") t.col_threads = v;
        else if (k == "row_threads") t.row_threads = v;
        else if (k == "row_radix") t.row_radix = v;
        else if (k == "col_dma") t.col_dma = v;
        else return false;
    }
    return true;
}

// Keeps the geometry inside what the kernels assume.
static Tuning clamp_tuning(Tuning t) {
    t.col_logc = std::max(0, std::min(t.col_logc, 5));
    t.col_maxl = std::max(4, std::min(t.col_maxl, 13));  // <= 7 column passes (Plan::col_s)
    t.row_tile = std::max(2, std::min(t.row_tile, kMaxRowTileLog));
    t.row_maxl = std::max(t.col_logc, std::min(t.row_maxl, t.row_tile));
    if (t.col_maxl + t.col_logc > kMaxColTileLog) t.col_maxl = kMaxColTileLog - t.col_logc;
    t.col_threads = std::max(64, std::min(t.col_threads, kMaxPassThreads)) & ~63;
    t.row_threads = std::max(64, std::min(t.row_threads, kMaxPassThreads)) & ~63;
    t.row_radix = t.row_radix <= 2 ? 2 : t.row_radix == 3 ? 3 : 4;
    t.col_dma = t.col_dma ? 1 : 0;
    return t;
}

static std::mutex g_tune_mu;
static Tuning g_tune[2][kMaxLogN + 1];  // [batched][log_n]
// Whether a batched entry was given (table line or set_tuning); a class
// without one launches batches with its single-product geometry (the table
// tunes batches only up to 2^17, and a longer transform fills the GPU alone).
static bool g_tune_batched_set[kMaxLogN + 1];
static bool g_tune_ready = false;
static std::string g_tune_env;  // $FMH_TUNE, applied last

// Directory of this shared library (for the tuning table beside it).
static std::string library_dir() {
    Dl_info info{};
    if (dladdr((const void*)&library_dir, &info) && info.dli_fname) {
        std::string f(info.dli_fname);
        const size_t slash = f.rfind('/');
        if (slash != std::string::npos) return f.substr(0, slash);
    }
    return ".";
}

static void tuning_init_locked() {
    if (g_tune_ready) return;
    g_tune_ready = true;
    const char* tab = getenv("FMH_TUNE_TABLE");
    const std::string path = tab ? std::string(tab) : library_dir() + "/tuning_gfx950.txt";
    if (FILE* f = fopen(path.c_str(), "r")) {
        char line[512];
        while (fgets(line, sizeof(line), f)) {
            int log_n = -1;
            char spec[480] = {0};
            const int b = line[0] == 'b' ? 1 : 0;
            if (line[0] == '#' || sscanf(line + b, "%d %479s", &log_n, spec) != 2) continue;
            if (log_n < 0 || log_n > kMaxLogN) continue;
            Tuning t;
            if (parse_tuning(spec, t)) {
                g_tune[b][log_n] = t;
                if (b) g_tune_batched_set[log_n] = true;
            }
        }
        fclose(f);
    }
    if (const char* env = getenv("FMH_TUNE")) g_tune_env = env;
}

Tuning tuning(int log_n, bool batched) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    tuning_init_locked();
    const int l = std::max(0, std::min(log_n, kMaxLogN));
    Tuning t = g_tune[batched && g_tune_batched_set[l] ? 1 : 0][l];
    if (!g_tune_env.empty()) (void)parse_tuning(g_tune_env, t);
    return clamp_tuning(t);
}

int set_tuning(int log_n, const char* spec, bool batched) {
    if (log_n < -1 || log_n > kMaxLogN) return -1;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    tuning_init_locked();
    Tuning* tab = g_tune[batched ? 1 : 0];
    for (int l = (log_n < 0 ? 0 : log_n); l <= (log_n < 0 ? kMaxLogN : log_n); l++) {
        // a batched spec without an entry of its own starts from the
        // single-product geometry the class used until now
        Tuning t = (spec && *spec) ? (batched && !g_tune_batched_set[l] ? g_tune[0][l] : tab[l]) : Tuning();
        if (spec && *spec && !parse_tuning(spec, t)) return -1;
        tab[l] = t;
        if (batched) g_tune_batched_set[l] = spec && *spec;
    }
    return 0;
}

Plan make_plan(int log_n) { return make_plan(log_n, tuning(log_n)); }

Plan make_plan(int log_n, const Tuning& tu) {
    Plan pl{};
    pl.log_n = log_n;
    if (log_n <= tu.row_tile) {
        pl.n_col = 0;
        pl.row_s = 0;
        return pl;
    }
    pl.row_s = log_n - tu.row_maxl;
    const int r = pl.row_s;
    const int ncol = (r + tu.col_maxl - 1) / tu.col_maxl;
    pl.n_col = ncol;
    int s = 0;
    for (int i = 0; i < ncol; i++) {
        pl.col_s[i] = s;
        s += (r - s) / (ncol - i);
    }
    pl.col_s[ncol] = r;
    return pl;
}

static int ceil_log2(u64 x) {
    int l = 0;
    while ((1ull << l) < x) l++;
    return l;
}

// ---------------------------------------------------------------- timing

static std::mutex g_tmu;
static bool g_timing = false;
static std::string g_timing_only;  // empty: time every kernel
struct TimedLaunch {
    std::string name;
    hipEvent_t a, b;
};
static std::vector<TimedLaunch> g_launches;

void timing_enable(bool on) {
    std::lock_guard<std::mutex> lk(g_tmu);
    g_timing = on;
}
bool timing_enabled() {
    std::lock_guard<std::mutex> lk(g_tmu);
    return g_timing;
}
void timing_only(const char* name) {
    std::lock_guard<std::mutex> lk(g_tmu);
    g_timing_only = name ? name : "";
}
hipError_t timing_begin(const char* name, hipStream_t s, hipEvent_t* start) {
    *start = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_tmu);
        if (!g_timing || (!g_timing_only.empty() && g_timing_only != name)) return hipSuccess;
    }
    hipError_t e = hipEventCreate(start);
    if (e != hipSuccess) return e;
    return hipEventRecord(*start, s);
}
hipError_t timing_end(const char* name, hipStream_t s, hipEvent_t start) {
    if (start == nullptr) return hipSuccess;
    hipEvent_t stop;
    hipError_t e = hipEventCreate(&stop);
    if (e != hipSuccess) return e;
    e = hipEventRecord(stop, s);
    std::lock_guard<std::mutex> lk(g_tmu);
    g_launches.push_back({name, start, stop});
    return e;
}

// Sums the recorded launches per kernel name; returns them as
// "name count total_ms" lines and forgets them.
std::string timing_collect() {
    std::vector<TimedLaunch> v;
    {
        std::lock_guard<std::mutex> lk(g_tmu);
        v.swap(g_launches);
    }
    std::vector<std::pair<std::string, std::pair<long, double>>> agg;
    for (auto& t : v) {
        float ms = 0.f;
        (void)hipEventSynchronize(t.b);
        (void)hipEventElapsedTime(&ms, t.a, t.b);
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
        bool found = false;
        for (auto& a : agg)
            if (a.first == t.name) {
                a.second.first++;
                a.second.second += ms;
                found = true;
            }
        if (!found) agg.push_back({t.name, {1L, (double)ms}});
    }
    std::string out;
    char line[256];
    for (auto& a : agg) {
        snprintf(line, sizeof(line), "%s %ld %.6f\n", a.first.c_str(), a.second.first, a.second.second);
        out += line;
    }
    return out;
}

#define FMH_LAUNCH(name, kern, grid, block, lds, stream, ...)                    \
    do {                                                                         \
        hipEvent_t ev_;                                                          \
        hipError_t e_ = timing_begin(name, stream, &ev_);                        \
        if (e_ != hipSuccess) return e_;                                         \
        hipLaunchKernelGGL(kern, grid, block, lds, stream, __VA_ARGS__);         \
        e_ = hipGetLastError();                                                  \
        if (e_ != hipSuccess) return e_;                                         \
        e_ = timing_end(name, stream, ev_);                                      \
        if (e_ != hipSuccess) return e_;                                         \
    } while (0)

// ------------------------------------------------------- global memory access
//
// Operand and residue pointers reach the kernels through descriptors, so the
// compiler cannot tell that they point to global memory and would emit flat
// loads and stores; a wave's wait for its LDS accesses then also waits for
// every flat access in flight.  gmem() states the address space.
#if defined(__HIP_DEVICE_COMPILE__)
#define FMH_GLOBAL __attribute__((address_space(1)))
#else
#define FMH_GLOBAL
#endif
template <class T>
__device__ __forceinline__ FMH_GLOBAL T* gmem(T* p) {
    return (FMH_GLOBAL T*)p;
}

// Copies into / out of a tile move kBatch words per thread per step: every
// load of a step is issued before the first one is used, so a wave waits for
// one memory latency per step instead of one per word.  The loads of a step
// are unconditional (an index past the end is clamped to the last word and
// its value dropped), which keeps them in one basic block.
constexpr int kBatch = 8;

// put(e, get(e)) for every tile word e in [0, n) of this thread (e = tid +
// k nthr).  get() must be a plain load that is valid for any e < n: the
// step's loads all run first (past n, get(n - 1) is loaded and dropped).
template <class Get, class Put>
__device__ __forceinline__ void batched_copy(int n, int nthr, Get get, Put put) {
    for (int e0 = threadIdx.x; e0 < n; e0 += kBatch * nthr) {
        u64 v[kBatch];
#pragma unroll
        for (int k = 0; k < kBatch; k++) v[k] = get(min(e0 + k * nthr, n - 1));
#pragma unroll
        for (int k = 0; k < kBatch; k++)
            if (e0 + k * nthr < n) put(e0 + k * nthr, v[k]);
    }
}

// LDS DMA (global_load_lds_dwordx4): each lane of a wave moves the 16 bytes
// at its own global address src into LDS at dst + 16 lane, where dst is the
// wave's (uniform) LDS address; the data never passes through registers, so
// a wave can have a whole tile in flight.  lds_dma_wait() waits for this
// wave's transfers (vmcnt counts them); a barrier then publishes them to the
// workgroup.
__device__ __forceinline__ void lds_dma16(const u64* src, u64* dst) {
#if defined(__HIP_DEVICE_COMPILE__)
    __builtin_amdgcn_global_load_lds((const FMH_GLOBAL void*)src, (__attribute__((address_space(3))) void*)dst, 16,
                                     0, 0);
#else
    memcpy(dst + 2 * (threadIdx.x & 63), src, 16);
#endif
}
__device__ __forceinline__ void lds_dma_wait() {
#if defined(__HIP_DEVICE_COMPILE__)
    __builtin_amdgcn_s_waitcnt(0);
#endif
}
__device__ __forceinline__ int wave_uniform(int v) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_readfirstlane(v);
#else
    return v;
#endif
}

// ---------------------------------------------------------------- LDS tiles
//
// A column tile holds C = 2^logC lines of L = 2^logL words: word (c, j) at
// t[(j << logC) + c], so every tile row is C adjacent words.  Line c has
// stride offset b = b0 + c within the global stride B = 2^logB.  For a stage
// of local half-span hl the global half-span is H = hl B, and the twiddle of
// the lower element at local j is W[H + (j mod hl) B + b] = W[((hl + j mod hl)
// << logB) + b0 + c].  Consecutive threads take consecutive c, then
// consecutive j: a wave reads whole tile rows.

// Per group the twiddle rows are read as uniform row pointers W + (k hl <<
// logB) plus one per-lane 32-bit byte offset ((u << logB) + b0 + c) 8, so
// the loads use a scalar base and no 64-bit address arithmetic per lane
// (offsets within a stage stay below its half-span H <= N/2 words, so byte
// offsets stay below 4N <= 2^32 for transforms of up to 2^30 points).
__device__ __forceinline__ u64 load_at(const u64* base, u32 byte_off) {
    return *(const FMH_GLOBAL u64*)((const FMH_GLOBAL char*)gmem(base) + byte_off);
}

// One round of a column tile: R = 2 (radix-2^2, half-spans 2 hl and hl) or
// R = 1 (radix-2, half-span hl) over every group of 2^R words (c, j0 + q hl),
// q < 2^R, of line c.  ld(e) reads tile word e = j C + c, fix(e, v) gives its
// value and st(e, v) writes it back: to and from the tile in LDS, or the
// transform's words in global memory for the first and the last round of a
// tile.  G groups of a thread issue their loads together (an index past the
// end is clamped and its result dropped).
// Twiddle t of the row of half-span 2^lh >= 2^F (lk = lh - F + 1): the
// product of W[2^(F-1) + (t >> lk)] and X[2^lk + (t mod 2^lk)] (see Twiddles
// in fmh_internal.h), canonical like the table entries.
template <int P>
__device__ __forceinline__ u64 tw_two(const u64* __restrict__ W, const u64* __restrict__ X, int F, int lk, u32 t) {
    const u64 hi = gmem(W)[(1u << (F - 1)) + (t >> lk)];
    const u64 lo = gmem(X)[(1u << lk) + (t & ((1u << lk) - 1))];
    return canon<P>(redc<P>(hi, lo));
}

// BIG: the launch has rows of half-span 2^F or more (transforms above 2^F
// points); the rounds that reach them form those twiddles with tw_two.
template <int P, int R, bool DIF, int G, bool BIG, class Ld, class Fix, class St>
__device__ __forceinline__ void col_round(int logL, int logC, int logB, u32 b0, const u64* __restrict__ W,
                                          const u64* __restrict__ X, int F, int lq, int nthr, Ld ld, Fix fix,
                                          St st) {
    constexpr int K = 1 << R;
    const int C = 1 << logC;
    const int hl = 1 << lq;
    const int nbf = 1 << (logL - R + logC);
    const int sp = hl << logC;  // tile distance of words q and q + 1 of a group
    // global half-spans 2^lhc (the round's last stage) and 2^(lhc + 1)
    const int lhc = lq + logB;
    const bool two_c = BIG && lhc >= F;
    const bool two_a = BIG && R == 2 && lhc + 1 >= F;
    const u64* __restrict__ Wa = two_a ? W : W + ((u64)(2 * hl) << logB);
    const u64* __restrict__ Wb = two_a ? W : W + ((u64)(3 * hl) << logB);
    const u64* __restrict__ Wc = two_c ? W : W + ((u64)hl << logB);
    for (int beta0 = threadIdx.x; beta0 < nbf; beta0 += G * nthr) {
        u64 x[G][K], w[G][R == 2 ? 3 : 1];
        int i0[G];
#pragma unroll
        for (int k = 0; k < G; k++) {
            const int beta = G > 1 ? min(beta0 + k * nthr, nbf - 1) : beta0;
            const int c = beta & (C - 1);
            const int r = beta >> logC;
            const int u = r & (hl - 1);
            const int j0 = ((r - u) << R) + u;  // (r / hl) 2^R hl + u
            const u32 t = ((u32)u << logB) + b0 + (u32)c;
            const u32 off = t << 3;
            if (two_c)
                w[k][0] = tw_two<P>(W, X, F, lhc - F + 1, t);
            else
                w[k][0] = load_at(Wc, off);
            if (R == 2) {
                if (two_a) {
                    w[k][R == 2 ? 1 : 0] = tw_two<P>(W, X, F, lhc - F + 2, t);
                    w[k][R == 2 ? 2 : 0] = tw_two<P>(W, X, F, lhc - F + 2, t + ((u32)hl << logB));
                } else {
                    w[k][R == 2 ? 1 : 0] = load_at(Wa, off);
                    w[k][R == 2 ? 2 : 0] = load_at(Wb, off);
                }
            }
            i0[k] = (j0 << logC) + c;
#pragma unroll
            for (int q = 0; q < K; q++) x[k][q] = ld(i0[k] + q * sp);
        }
#pragma unroll
        for (int k = 0; k < G; k++) {
            if (G > 1 && beta0 + k * nthr >= nbf) break;
#pragma unroll
            for (int q = 0; q < K; q++) x[k][q] = fix(i0[k] + q * sp, x[k][q]);
            u64* y = x[k];
            if (R == 1) {
                if (DIF)
                    bf_dif<P>(y[0], y[1], w[k][0]);
                else
                    bf_dit<P>(y[0], y[1], w[k][0]);
            } else if (DIF) {
                bf_dif<P>(y[0], y[K > 2 ? 2 : 0], w[k][R == 2 ? 1 : 0]);
                bf_dif<P>(y[1], y[K > 2 ? 3 : 1], w[k][R == 2 ? 2 : 0]);
                bf_dif<P>(y[0], y[1], w[k][0]);
                bf_dif<P>(y[K > 2 ? 2 : 0], y[K > 2 ? 3 : 1], w[k][0]);
            } else {
                bf_dit<P>(y[0], y[1], w[k][0]);
                bf_dit<P>(y[K > 2 ? 2 : 0], y[K > 2 ? 3 : 1], w[k][0]);
                bf_dit<P>(y[0], y[K > 2 ? 2 : 0], w[k][R == 2 ? 1 : 0]);
                bf_dit<P>(y[1], y[K > 2 ? 3 : 1], w[k][R == 2 ? 2 : 0]);
            }
#pragma unroll
            for (int q = 0; q < K; q++) st(i0[k] + q * sp, x[k][q]);
        }
    }
}

// Round `first` of a tile reads the transform (ld, fix), round `last` writes
// it (st); the rounds between go through the tile t in LDS.  IN_LDS: the
// tile already holds the transform's words (filled by LDS DMA), and ld reads
// them from there.
template <int P, int R, bool DIF, bool IN_LDS, bool BIG, class Ld, class Fix, class St>
__device__ __forceinline__ void col_round_io(bool first, bool last, u64* __restrict__ t, int logL, int logC, int logB,
                                             u32 b0, const u64* __restrict__ W, const u64* __restrict__ X, int F,
                                             int lq, int nthr, Ld ld, Fix fix, St st) {
    // four data words of the first round in flight per thread: eight (two
    // radix-4 groups) took the kernel from 42 to 94 VGPRs and ran slower
    constexpr int GI = IN_LDS ? 1 : 4 >> R;
    auto tld = [&](int e) { return t[e]; };
    auto same = [](int, u64 v) { return v; };
    auto tst = [&](int e, u64 v) { t[e] = v; };
    if (first && last)
        col_round<P, R, DIF, GI, BIG>(logL, logC, logB, b0, W, X, F, lq, nthr, ld, fix, st);
    else if (first)
        col_round<P, R, DIF, GI, BIG>(logL, logC, logB, b0, W, X, F, lq, nthr, ld, fix, tst);
    else if (last)
        col_round<P, R, DIF, 1, BIG>(logL, logC, logB, b0, W, X, F, lq, nthr, tld, same, st);
    else
        col_round<P, R, DIF, 1, BIG>(logL, logC, logB, b0, W, X, F, lq, nthr, tld, same, tst);
    if (!last) __syncthreads();
}

// The stages of a column tile (logL >= 1): DIF as radix-2^2 rounds over
// hl = L/2 .. 1 with one radix-2 stage last when logL is odd; DIT as the
// radix-2 stage first when logL is odd, then radix-2^2 rounds up to hl = L/2.
template <int P, bool DIF, bool IN_LDS, bool BIG, class Ld, class Fix, class St>
__device__ __forceinline__ void tile_rounds(u64* __restrict__ t, int logL, int logC, int logB, u32 b0,
                                            const u64* __restrict__ W, const u64* __restrict__ X, int F, int nthr,
                                            Ld ld, Fix fix, St st) {
    const int nr = (logL >> 1) + (logL & 1);
    const bool odd = logL & 1;
    for (int i = 0; i < nr; i++) {
        const bool first = i == 0, last = i == nr - 1;
        const bool r1 = odd && (DIF ? last : first);
        if (r1) {
            col_round_io<P, 1, DIF, IN_LDS, BIG>(first, last, t, logL, logC, logB, b0, W, X, F, 0, nthr, ld, fix,
                                                 st);
        } else {
            const int lq = DIF ? logL - 2 * (i + 1) : (logL & 1) + 2 * (i - (odd ? 1 : 0));
            col_round_io<P, 2, DIF, IN_LDS, BIG>(first, last, t, logL, logC, logB, b0, W, X, F, lq, nthr, ld, fix,
                                                 st);
        }
    }
}

// ------------------------------------------------------------ column pass
//
// Kernels read their products through descriptors: `items` (device array of
// a batch) or, for a single product, the copy `one` passed by value.

__device__ __forceinline__ const ProductDesc& product(const ProductDesc* items, const ProductDesc& one,
                                                      u32 i) {
    return items ? items[i] : one;
}

struct ColArgs {
    const ProductDesc* items;
    ProductDesc one;
    const u64* tw[3];
    const u64* twx[3];  // twiddles of half-spans >= 2^tw_full (launches with BIG set)
    int tw_full;
    int log_n, s0, s1;
    int logC;  // 2^logC adjacent lines per tile
    int from_limbs;
    int nops;     // operands transformed per product: 2 (multiply) or 1
    int threads;  // workgroup size
};

// Workgroups are dispatched to the 8 XCDs of the chip round-robin by
// workgroup id; each XCD has its own L2.  xcd_slot maps id x of n so that
// XCD k receives the contiguous slot range [k n / 8, (k + 1) n / 8), in
// dispatch order: neighbouring slots run on one XCD at about the same time.
constexpr u32 kXcds = 8;
__device__ __forceinline__ u32 xcd_slot(u32 x, u32 n) {
    const u32 xcd = x % kXcds, local = x / kXcds;
    const u32 per = n / kXcds, rem = n % kXcds;
    return xcd < rem ? xcd * (per + 1) + local : rem * (per + 1) + (xcd - rem) * per + local;
}

// Fills the column tile (n words, t[e] = word at(e) of src) by LDS DMA.
// Tile words 2i and 2i + 1 lie in one tile row (C >= 2 adjacent words), so
// each lane moves one such pair, and each wave instruction fills 128 words
// (1 KiB) of the tile.  Pairs for which skip(e) holds are not loaded.
template <class At, class Skip>
__device__ __forceinline__ void col_dma_fill(u64* t, int n, int nthr, const u64* src, At at, Skip skip) {
    const int lane2 = (int)(threadIdx.x & 63) << 1;
    for (int k = wave_uniform((int)threadIdx.x >> 6) << 7; k < n; k += nthr << 1) {
        const int e = k + lane2;
        if (e < n && !skip(e)) lds_dma16(src + at(e), t + k);
    }
    lds_dma_wait();
    __syncthreads();
}

template <int P, bool FWD, bool DMA, bool BIG>
__device__ void col_body(const ColArgs& A, u32 tile, int op) {
    extern __shared__ u64 lds[];
    const u32 prod = blockIdx.z;
    const ProductDesc& D = product(A.items, A.one, prod);
    const int logL = A.s1 - A.s0;
    const int logB = A.log_n - A.s1;
    const int logC = A.logC;
    const int C = 1 << logC;
    const int n = C << logL;
    const u64 lam0 = (u64)tile << logC;
    const u64 a = lam0 >> logB;
    const u32 b0 = (u32)(lam0 & ((1ull << logB) - 1));
    const u64 base = (a << (logL + logB)) + b0;
    const int nthr = A.threads;
    FMH_GLOBAL u64* __restrict__ res = gmem(D.res + ((u64)(op * 3 + P) << A.log_n));
    const FMH_GLOBAL u64* __restrict__ limbs = gmem(op ? D.b : D.a);
    const u64 len = op ? D.nb : D.na;
    // tile word e = j C + c is word base + j B + c of the transform
    auto at = [&](int e) { return base + ((u64)(e >> logC) << logB) + (u64)(e & (C - 1)); };
    auto put = [&](int e, u64 v) { res[at(e)] = v; };
    auto same = [](int, u64 v) { return v; };
    if (DMA) {
        // the tile is filled by DMA, then every round runs from LDS
        auto tld = [&](int e) { return lds[e]; };
        if (!A.from_limbs) {
            col_dma_fill(lds, n, nthr, (const u64*)res, at, [](int) { return false; });
            tile_rounds<P, FWD, true, BIG>(lds, logL, logC, logB, b0, A.tw[P], A.twx[P], A.tw_full, nthr, tld, same,
                                           put);
        } else {
            // Pairs reaching past the operand's last word are not loaded: the
            // words past it are zero, and an odd length's last word (the
            // first of its pair) is read directly.  The host sends only
            // operands whose limbs are 16-byte aligned here.
            col_dma_fill(lds, n, nthr, op ? D.b : D.a, at, [&](int e) { return at(e) + 1 >= len; });
            auto val = [&](int e, u64 v) {
                const u64 g = at(e);
                if (g >= len) return 0ull;
                return reduce_word<P>((g | 1) >= len ? limbs[g] : v);
            };
            tile_rounds<P, FWD, true, BIG>(lds, logL, logC, logB, b0, A.tw[P], A.twx[P], A.tw_full, nthr, tld, val,
                                           put);
        }
    } else if (!A.from_limbs) {
        auto get = [&](int e) { return res[at(e)]; };
        tile_rounds<P, FWD, false, BIG>(lds, logL, logC, logB, b0, A.tw[P], A.twx[P], A.tw_full, nthr, get, same,
                                        put);
    } else {
        // an empty operand reads the first twiddle and drops it
        const FMH_GLOBAL u64* __restrict__ src = len ? limbs : gmem(A.tw[P]);
        const u64 last = len ? len - 1 : 0;
        auto get = [&](int e) {
            const u64 g = at(e);
            return src[g < last ? g : last];
        };
        auto val = [&](int e, u64 v) { return at(e) < len ? reduce_word<P>(v) : 0ull; };
        tile_rounds<P, FWD, false, BIG>(lds, logL, logC, logB, b0, A.tw[P], A.twx[P], A.tw_full, nthr, get, val,
                                        put);
    }
}

// slot = (tile * 3 + prime) * nops + operand, placed by xcd_slot: the
// operands of a tile share its twiddle rows, the three primes of a tile read
// the same limbs (first forward pass), and neighbouring tiles share 128-byte
// lines of data and twiddles (a tile row is 2^logC words); all of them run
// on one XCD at about the same time, so the shared reads hit its L2.
// DMA: the tile is filled by LDS DMA (residues, or limbs that are 16-byte
// aligned; tiles of at least 2 adjacent lines).  Forward passes only: at
// 2^25 points it took the forward passes from 1050 and 918 us to 1012 and
// 859 us, and the inverse passes from 451 and 542 us to 445 and 559 us.
// BIG: the pass has stages of half-span 2^tw_full or more (only the first
// column pass of transforms above 2^tw_full points).
template <bool FWD, bool DMA, bool BIG>
__global__ __launch_bounds__(kMaxPassThreads) void k_col(ColArgs A) {
    const u32 slot = xcd_slot(blockIdx.x, gridDim.x);
    const u32 per_tile = 3u * (u32)A.nops;
    const u32 tile = slot / per_tile;
    const u32 rem = slot - tile * per_tile;
    const u32 prime = rem / (u32)A.nops;
    const int op = (int)(rem - prime * (u32)A.nops);
    switch (prime) {
        case 0: col_body<0, FWD, DMA, BIG>(A, tile, op); break;
        case 1: col_body<1, FWD, DMA, BIG>(A, tile, op); break;
        default: col_body<2, FWD, DMA, BIG>(A, tile, op); break;
    }
}

// --------------------------------------------------------------- row pass
//
// A row tile holds C lines of L = 2^logL contiguous words.  Line lambda of a
// launch is line (lambda mod 2^s0) of product (lambda >> s0): for a large
// product the lines of a tile are neighbouring blocks of the same transform,
// for small products (s0 = 0) every line is a whole product, so one
// workgroup carries C small transforms at once.

// Row modes: kRowMul, kRowSqr and kRowMulSpec / kRowSqrSpec (the product
// kinds of the same names) finish the forward transform of the loaded
// operand(s), form the pointwise product and run the first inverse stages;
// kRowFwdOnly exports the spectrum in canonical Montgomery form (parity
// checks); kRowFwdKeep stores it as the plain values [0, p) of a resident
// spectrum.
enum RowMode { kRowMul = 0, kRowSqr = 1, kRowMulSpec = 2, kRowSqrSpec = 3, kRowFwdOnly = 4, kRowFwdKeep = 5 };
constexpr int kMaxRowLines = 64;

struct RowArgs {
    const ProductDesc* items;
    ProductDesc one;
    u32 count;  // products in the launch
    const u64* tw[3];
    const u64* itw[3];
    u64 r2[3];  // 2^128 mod p (spectrum export to Montgomery form)
    int log_n, s0;
    int logC;  // 2^logC lines per tile
    int from_limbs;
    int mode;
    int threads;  // workgroup size
};

// Where a line's operand words come from: src[op][min(g0 + j, last[op])],
// reduced and zero from len[op] on when the pass starts from limbs.  A line
// with nothing to read (inactive, empty operand) points at the first twiddle
// with last = 0, so every load of the pass goes to a valid address.
struct RowLine {
    const u64* src[2];  // limbs, residues of the column passes, or the spectrum
    u64 last[2];
    u64 len[2];
    u64* ra;            // residues of this prime, written back
    const u64* rs;      // resident spectrum of this prime (spec modes)
    u64 g0;             // offset of the line inside its product's arrays
    u32 active;
};

// Row-tile LDS layout.  Word i of a tile lives at row_swz(i) = i XOR bits
// 4-7 of i: the words of each aligned 16-word block are permuted by the
// block's index.  With 64 four-byte banks serving 32 lanes of 8-byte words,
// the row pass's access patterns then cost at most two passes:
//   * the global <-> LDS copies and rounds with half-span hl >= 16 read 16
//     or 32 consecutive aligned words per group of lanes (one pass);
//   * the fused round reads word 8k + q in lane k, and rounds with hl = 8
//     read u + 8q: consecutive lanes land in different 16-word blocks, which
//     the XOR spreads over the banks (two passes per 32 lanes).
// A parity-based variant that is conflict-free for 16-lane groups cost
// four more VALU operations per address and ran the 2^21-point row pass
// 7% slower (116.5 against 108.9 us).
//
// row_swz is linear over XOR, and a group's word offsets q 2^lq occupy bits
// that its base leaves zero, so row_swz(base + (q << lq)) =
// row_swz(base) ^ row_swz(q << lq): one per-lane value per group and a
// wave-uniform one per word.
__device__ __forceinline__ int row_swz(int i) {
    return i ^ ((i >> 4) & 15);
}

// One round of R consecutive stages on one operand tile (word c * L + j),
// half-spans hl 2^(R-1) .. hl for DIF (decreasing) or hl .. hl 2^(R-1) for
// DIT (increasing), hl = 2^lq.  Group g holds the 2^R words
// base + q hl of one line; the twiddle of the pair (q, q + hq) at stage
// half-span H = hq hl is W[H + (j mod H)] = W[(hq + q mod hq) hl + u].
// Tile word of the first element of group g (of one operand).
template <int R>
__device__ __forceinline__ int row_group_base(int logL, int lq, int g) {
    const int u = g & ((1 << lq) - 1);
    const int rest = g >> lq;
    const int lb = logL - lq - R;  // log2 of groups per line
    const int blk = rest & ((1 << lb) - 1);
    const int c = rest >> lb;
    return (c << logL) + (blk << (lq + R)) + u;
}

// The R stages of a group in registers; u = base mod hl.
template <int P, int R, bool DIF>
__device__ __forceinline__ void row_group_stages(u64 (&x)[1 << R], int lq, int u, const u64* __restrict__ W) {
    constexpr int Q = 1 << R;
#pragma unroll
    for (int st = 0; st < R; st++) {
        const int hq = DIF ? (1 << (R - 1 - st)) : (1 << st);
        u64 w[Q / 2];
#pragma unroll
        for (int r = 0; r < hq; r++) w[r] = W[((hq + r) << lq) + u];
#pragma unroll
        for (int q = 0; q < Q; q++) {
            if (q & hq) continue;
            if (DIF)
                bf_dif<P>(x[q], x[q + hq], w[q & (hq - 1)]);
            else
                bf_dit<P>(x[q], x[q + hq], w[q & (hq - 1)]);
        }
    }
}

template <int P, int R, bool DIF>
__device__ __forceinline__ void row_round_group(u64* __restrict__ t, int logL, int lq, int g,
                                                const u64* __restrict__ W) {
    constexpr int Q = 1 << R;
    const int base = row_group_base<R>(logL, lq, g);
    u64 x[Q];
    const int sb = row_swz(base);
#pragma unroll
    for (int q = 0; q < Q; q++) x[q] = t[sb ^ row_swz(q << lq)];
    row_group_stages<P, R, DIF>(x, lq, base & ((1 << lq) - 1), W);
#pragma unroll
    for (int q = 0; q < Q; q++) t[sb ^ row_swz(q << lq)] = x[q];
}

// A round over the tiles of nt operands (t1 = second operand or null).
// RMAX: that of the calling kernel instance.  This function is not inlined;
// a copy per instance lets each be register-allocated for that instance
// alone (shared by k_row<3> and k_row<4>, it made k_row<4> take 120 VGPRs
// and spill; its own copy takes 109 and none: row_mul 104-107 -> 101-102 us
// at 2^21, 1,284-1,301 -> 1,218 us at 2^25).
template <int P, int R, bool DIF, int RMAX>
__device__ void row_round(u64* t0, u64* t1, int nt, int n, int logL, int lq, const u64* W, int nthr) {
    const int per = n >> R;
    for (int g = threadIdx.x; g < nt * per; g += nthr) {
        if (g < per)
            row_round_group<P, R, DIF>(t0, logL, lq, g, W);
        else
            row_round_group<P, R, DIF>(t1, logL, lq, g - per, W);
    }
    __syncthreads();
}

template <int P, bool DIF, int RMAX>
__device__ void row_round_r(int r, u64* t0, u64* t1, int nt, int n, int logL, int lq, const u64* W, int nthr) {
    switch (r) {
        case 1: row_round<P, 1, DIF, RMAX>(t0, t1, nt, n, logL, lq, W, nthr); break;
        case 2: row_round<P, 2, DIF, RMAX>(t0, t1, nt, n, logL, lq, W, nthr); break;
        case 3: row_round<P, 3, DIF, RMAX>(t0, t1, nt, n, logL, lq, W, nthr); break;
        default: row_round<P, (RMAX > 3 ? 4 : 3), DIF, RMAX>(t0, t1, nt, n, logL, lq, W, nthr); break;
    }
}

// The first forward round of a row tile, read straight from the operands'
// sources instead of a filled tile: group g (of operand g / per) loads its
// 2^R words of line c, runs the R DIF stages and writes them to the tile.
// The G groups of one step issue all their loads first (about kBatch words
// in flight per thread); a group index past the end is clamped and its
// result dropped.  Word j of line c of operand op is
// src[op][min(g0 + j, last[op])], zero when the line is inactive or j is
// past len[op], reduced mod p when the sources are limbs.
template <int P, int R>
__device__ __forceinline__ void row_round_in(const RowLine* s_line, bool reduce, u64* t0, u64* t1, int nt, int n,
                                             int logL, int lq, const u64* __restrict__ W, int nthr) {
    constexpr int Q = 1 << R;
    constexpr int G = Q >= kBatch ? 1 : kBatch / Q;
    const int per = n >> R;
    const int total = nt * per;
    const int L = 1 << logL;
    for (int g0 = threadIdx.x; g0 < total; g0 += G * nthr) {
        u64 x[G][Q];
        int base[G], op[G];
#pragma unroll
        for (int k = 0; k < G; k++) {
            const int g = min(g0 + k * nthr, total - 1);
            op[k] = g >= per ? 1 : 0;
            base[k] = row_group_base<R>(logL, lq, g - op[k] * per);
            const RowLine& li = s_line[base[k] >> logL];
            const u64 j0 = li.g0 + (u64)(base[k] & (L - 1));
            const u64 last = li.last[op[k]];
            const FMH_GLOBAL u64* src = gmem(li.src[op[k]]);
#pragma unroll
            for (int q = 0; q < Q; q++) {
                const u64 j = j0 + ((u64)q << lq);
                x[k][q] = src[j < last ? j : last];
            }
        }
#pragma unroll
        for (int k = 0; k < G; k++) {
            if (g0 + k * nthr >= total) break;
            const RowLine& li = s_line[base[k] >> logL];
            const u64 j0 = li.g0 + (u64)(base[k] & (L - 1));
            const u64 len = li.active ? li.len[op[k]] : 0;
#pragma unroll
            for (int q = 0; q < Q; q++) {
                const u64 j = j0 + ((u64)q << lq);
                x[k][q] = j >= len ? 0ull : reduce ? reduce_word<P>(x[k][q]) : x[k][q];
            }
            row_group_stages<P, R, true>(x[k], lq, base[k] & ((1 << lq) - 1), W);
            u64* __restrict__ t = op[k] ? t1 : t0;
            const int sb = row_swz(base[k]);
#pragma unroll
            for (int q = 0; q < Q; q++) t[sb ^ row_swz(q << lq)] = x[k][q];
        }
    }
    __syncthreads();
}

// The last inverse round of a row tile, written straight to the residues of
// each active line instead of back to the tile.
template <int P, int R>
__device__ __forceinline__ void row_round_out(const RowLine* s_line, const u64* __restrict__ t, int n, int logL,
                                              int lq, const u64* __restrict__ W, int nthr) {
    constexpr int Q = 1 << R;
    const int L = 1 << logL;
    for (int g = threadIdx.x; g < (n >> R); g += nthr) {
        const int base = row_group_base<R>(logL, lq, g);
        u64 x[Q];
        const int sb = row_swz(base);
#pragma unroll
        for (int q = 0; q < Q; q++) x[q] = t[sb ^ row_swz(q << lq)];
        row_group_stages<P, R, false>(x, lq, base & ((1 << lq) - 1), W);
        const RowLine& li = s_line[base >> logL];
        if (!li.active) continue;
        FMH_GLOBAL u64* out = gmem(li.ra + li.g0 + (u64)(base & (L - 1)));
#pragma unroll
        for (int q = 0; q < Q; q++) out[q << lq] = x[q];
    }
}

template <int P, int RMAX>
__device__ void row_round_in_r(int r, const RowLine* s_line, bool reduce, u64* t0, u64* t1, int nt, int n,
                               int logL, int lq, const u64* W, int nthr) {
    switch (r) {
        case 1: row_round_in<P, 1>(s_line, reduce, t0, t1, nt, n, logL, lq, W, nthr); break;
        case 2: row_round_in<P, 2>(s_line, reduce, t0, t1, nt, n, logL, lq, W, nthr); break;
        case 3: row_round_in<P, 3>(s_line, reduce, t0, t1, nt, n, logL, lq, W, nthr); break;
        default: row_round_in<P, (RMAX > 3 ? 4 : 3)>(s_line, reduce, t0, t1, nt, n, logL, lq, W, nthr); break;
    }
}

template <int P, int RMAX>
__device__ void row_round_out_r(int r, const RowLine* s_line, const u64* t, int n, int logL, int lq,
                                const u64* W, int nthr) {
    switch (r) {
        case 1: row_round_out<P, 1>(s_line, t, n, logL, lq, W, nthr); break;
        case 2: row_round_out<P, 2>(s_line, t, n, logL, lq, W, nthr); break;
        default: row_round_out<P, 3>(s_line, t, n, logL, lq, W, nthr); break;
    }
}

// The fused middle of the row pass: for each group of 2^R consecutive words
// (the last R forward stages, half-spans 2^(R-1) .. 1), finish the forward
// transform of operand A (and B), form the pointwise product in registers,
// and run the first R inverse stages, then store A.  Twiddles here do not
// depend on the lane.  Modes kRowFwdOnly / kRowFwdKeep stop after the
// forward stages.
template <int P, int R>
__device__ __forceinline__ void row_fused(const RowArgs& A, const RowLine* s_line, u64* tA, u64* tB, int n, int logL) {
    constexpr int Q = 1 << R;
    const u64* __restrict__ W = A.tw[P];
    const u64* __restrict__ IW = A.itw[P];
    const bool two = A.mode == kRowMul;
    const bool fwd = A.mode != kRowSqrSpec;
    const bool inv = A.mode == kRowMul || A.mode == kRowSqr || A.mode == kRowMulSpec || A.mode == kRowSqrSpec;
    for (int g = threadIdx.x; g < (n >> R); g += A.threads) {
        const int base = g << R;
        const int sb = row_swz(base);  // word base + q is at sb ^ q
        u64 x[Q], y[Q];
#pragma unroll
        for (int q = 0; q < Q; q++) {
            x[q] = tA[sb ^ q];
            if (two) y[q] = tB[sb ^ q];
        }
        if (fwd) {
#pragma unroll
            for (int st = 0; st < R; st++) {
                const int hq = 1 << (R - 1 - st);
#pragma unroll
                for (int q = 0; q < Q; q++) {
                    if (q & hq) continue;
                    if ((q & (hq - 1)) == 0) {  // twiddle w_{2 hq}^0 = 1
                        bf_dif_one<P>(x[q], x[q + hq]);
                        if (two) bf_dif_one<P>(y[q], y[q + hq]);
                        continue;
                    }
                    const u64 w = W[hq + (q & (hq - 1))];
                    bf_dif<P>(x[q], x[q + hq], w);
                    if (two) bf_dif<P>(y[q], y[q + hq], w);
                }
            }
        }
        if (!inv) {
#pragma unroll
            for (int q = 0; q < Q; q++) tA[sb ^ q] = x[q];
            continue;
        }
        if (A.mode == kRowMulSpec) {
            // an inactive line (never stored) reads line 0's spectrum
            const RowLine& li = s_line[base >> logL];
            const FMH_GLOBAL u64* rs =
                gmem(li.active ? li.rs + li.g0 + (u64)(base & ((1 << logL) - 1)) : s_line[0].rs);
#pragma unroll
            for (int q = 0; q < Q; q++) x[q] = redc<P>(x[q], rs[q]);
        } else {
#pragma unroll
            for (int q = 0; q < Q; q++) x[q] = redc<P>(x[q], two ? y[q] : x[q]);
        }
#pragma unroll
        for (int st = 0; st < R; st++) {
            const int hq = 1 << st;
#pragma unroll
            for (int q = 0; q < Q; q++) {
                if (q & hq) continue;
                if ((q & (hq - 1)) == 0)
                    bf_dit_one<P>(x[q], x[q + hq]);
                else
                    bf_dit<P>(x[q], x[q + hq], IW[hq + (q & (hq - 1))]);
            }
        }
#pragma unroll
        for (int q = 0; q < Q; q++) tA[sb ^ q] = x[q];
    }
    __syncthreads();
}

// RMAX: the most stages one round takes (4: radix-16 rounds for two
// operands, more registers; 3: radix-8 rounds, fewer registers and more
// waves per SIMD; 2: radix-4 rounds, for transforms up to 2^17 points, whose
// few workgroups run each round's code once: there the time goes to fetching
// that code into the instruction cache (77 misses per dispatch for a
// 128-point product with radix-16 rounds), and the short rounds execute
// less of it: 2^7 points 20.1 -> 16.8 us, 2^13..2^17 2.5 us less).
template <int P, int RMAX>
__device__ __forceinline__ void row_body(const RowArgs& A, RowLine* s_line, u32 tile) {
    extern __shared__ u64 lds[];
    const int logL = A.log_n - A.s0;
    const int logC = A.logC;
    const int C = 1 << logC;
    const int n = C << logL;
    const int nl = n < 16 ? 16 : n;  // words per operand tile in LDS (row_swz stays below 16)
    const bool two = (A.mode == kRowMul);
    const bool spec_in = (A.mode == kRowSqrSpec);  // the tile starts as a spectrum
    u64* __restrict__ tA = lds;
    u64* __restrict__ tB = lds + nl;
    if ((int)threadIdx.x < C) {
        const u64 lambda = ((u64)tile << logC) + threadIdx.x;
        const u64 prod = lambda >> A.s0;
        RowLine li{};
        li.active = prod < A.count ? 1u : 0u;
        for (int op = 0; op < 2; op++) {
            li.src[op] = A.tw[P];
            li.last[op] = 0;
            li.len[op] = ~0ull;
        }
        if (li.active) {
            const ProductDesc& D = product(A.items, A.one, (u32)prod);
            li.ra = D.res + ((u64)P << A.log_n);
            li.rs = D.spec ? D.spec + ((u64)P << A.log_n) : nullptr;
            li.g0 = (lambda & ((1ull << A.s0) - 1)) << logL;
            for (int op = 0; op < (two ? 2 : 1); op++) {
                if (spec_in) {
                    li.src[op] = li.rs;
                    li.last[op] = ~0ull;
                } else if (A.from_limbs) {
                    const u64 len = op ? D.nb : D.na;
                    li.len[op] = len;
                    if (len) {
                        li.src[op] = op ? D.b : D.a;
                        li.last[op] = len - 1;
                    }
                } else {
                    li.src[op] = D.res + ((u64)(3 * op + P) << A.log_n);
                    li.last[op] = ~0ull;
                }
            }
        }
        s_line[threadIdx.x] = li;
    }
    __syncthreads();
    // Lines at least one workgroup long are walked line by line (the line's
    // description is then uniform); shorter lines element by element.
    const int nthr = A.threads;
    const int L = 1 << logL;
    const bool by_line = L >= nthr;
    auto word_ptr = [&](const RowLine& li, int op, u64 g) {
        const u64 l = li.last[op];
        return gmem(li.src[op] + (g < l ? g : l));
    };
    auto word_val = [&](const RowLine& li, int op, u64 g, u64 v) -> u64 {
        if (!li.active || g >= li.len[op]) return 0;
        return A.from_limbs && !spec_in ? reduce_word<P>(v) : v;
    };
    auto fill = [&](u64* t, int op) {
        if (by_line) {
            for (int c = 0; c < C; c++) {
                const RowLine li = s_line[c];
                const int e0 = c << logL;
                batched_copy(
                    L, nthr, [&](int j) { return *word_ptr(li, op, li.g0 + (u64)j); },
                    [&](int j, u64 v) { t[row_swz(e0 + j)] = word_val(li, op, li.g0 + (u64)j, v); });
            }
        } else {
            batched_copy(
                n, nthr,
                [&](int e) {
                    const RowLine& li = s_line[e >> logL];
                    return *word_ptr(li, op, li.g0 + (u64)(e & (L - 1)));
                },
                [&](int e, u64 v) {
                    const RowLine& li = s_line[e >> logL];
                    t[row_swz(e)] = word_val(li, op, li.g0 + (u64)(e & (L - 1)), v);
                });
        }
    };
    // Stage schedule: the last RF forward stages, the pointwise product and
    // the first RF inverse stages run fused in registers; the other stages
    // run as rounds of up to 4 stages (both operands) or 3 (one operand).
    // The first forward round reads the sources itself and the last inverse
    // round writes the residues itself; a tile with no such round is filled
    // and stored by plain copies.
    constexpr int RM3 = RMAX < 3 ? RMAX : 3;  // most stages of a one-operand round and of the fused round
    const int RF = logL < RM3 ? logL : RM3;
    const int nt = two ? 2 : 1;
    const bool fold_in = !spec_in && logL > RF;
    if (!fold_in) {
        fill(tA, 0);
        if (two) fill(tB, 1);
        __syncthreads();
    }
    if (!spec_in) {
        const int rmax = two ? RMAX : RM3;
        int top = logL;
        while (top > RF) {
            const int left = top - RF;
            const int r = left % rmax ? left % rmax : rmax;
            if (top == logL)
                row_round_in_r<P, RMAX>(r, s_line, A.from_limbs != 0, tA, tB, nt, n, logL, top - r, A.tw[P], nthr);
            else
                row_round_r<P, true, RMAX>(r, tA, tB, nt, n, logL, top - r, A.tw[P], nthr);
            top -= r;
        }
    }
    switch (RF) {
        case 0: row_fused<P, 0>(A, s_line, tA, tB, n, logL); break;
        case 1: row_fused<P, 1>(A, s_line, tA, tB, n, logL); break;
        case 2: row_fused<P, 2>(A, s_line, tA, tB, n, logL); break;
        default: row_fused<P, 3>(A, s_line, tA, tB, n, logL); break;
    }
    if (A.mode == kRowFwdOnly || A.mode == kRowFwdKeep) {
        const u64 r2 = A.r2[P];
        const bool mont = A.mode == kRowFwdOnly;
        batched_copy(
            n, nthr, [&](int e) { return tA[row_swz(e)]; },
            [&](int e, u64 v) {
                const RowLine& li = s_line[e >> logL];
                if (li.active)
                    gmem(li.ra)[li.g0 + (e & (L - 1))] = canon<P>(mont ? redc<P>(v, r2) : norm2p<P>(v));
            });
        return;
    }
    for (int lq = RF; lq < logL;) {
        const int r = logL - lq < RM3 ? logL - lq : RM3;
        if (lq + r == logL)
            row_round_out_r<P, RMAX>(r, s_line, tA, n, logL, lq, A.itw[P], nthr);
        else
            row_round_r<P, false, RMAX>(r, tA, nullptr, 1, n, logL, lq, A.itw[P], nthr);
        lq += r;
    }
    if (logL > RF) return;
    if (by_line) {
        for (int c = 0; c < C; c++) {
            const RowLine li = s_line[c];
            if (!li.active) continue;
            FMH_GLOBAL u64* out = gmem(li.ra + li.g0);
            const int e0 = c << logL;
            batched_copy(
                L, nthr, [&](int j) { return tA[row_swz(e0 + j)]; }, [&](int j, u64 v) { out[j] = v; });
        }
    } else {
        batched_copy(
            n, nthr, [&](int e) { return tA[row_swz(e)]; },
            [&](int e, u64 v) {
                const RowLine& li = s_line[e >> logL];
                if (li.active) gmem(li.ra)[li.g0 + (e & (L - 1))] = v;
            });
    }
}

// Every tile the tuning clamp allows fits a workgroup's LDS together with the
// kernels' static LDS (the column kernel has none; the row kernel has the line
// table).
static_assert(col_lds_bytes(kMaxColTileLog, 0) <= kLdsPerGroup, "column tile LDS");
static_assert(kMaxLogN <= 30, "column-pass twiddle byte offsets are 32-bit");
static_assert(row_lds_bytes(2, kMaxRowTileLog, 0) + sizeof(RowLine) * kMaxRowLines <= kLdsPerGroup,
              "row tile LDS");

// slot = tile * 3 + prime, placed by xcd_slot: the three primes of a tile
// run on one XCD at about the same time, so a pass that starts from limbs
// (small products) reads them from HBM once.
template <int RMAX>
__global__ __launch_bounds__(kMaxPassThreads) void k_row(RowArgs A) {
    __shared__ RowLine s_line[kMaxRowLines];
    const u32 slot = xcd_slot(blockIdx.x, gridDim.x);
    const u32 tile = slot / 3u;
    switch (slot - 3u * tile) {
        case 0: row_body<0, RMAX>(A, s_line, tile); break;
        case 1: row_body<1, RMAX>(A, s_line, tile); break;
        default: row_body<2, RMAX>(A, s_line, tile); break;
    }
}

// The row-only classes of a batch (transforms of at most one row tile: every
// line a whole product) in one launch instead of one per class.  Workgroup
// tiles are numbered across the classes; class c owns tiles [tile0[c],
// tile0[c + 1]) and runs as its own k_row launch would, with its own
// arguments.  Batches of small products of many sizes (64..65,536 limbs)
// have a dozen such classes, most of a few dozen workgroups, which one at a
// time leave the GPU nearly idle.  The arguments of every class travel
// whole in the argument block (14 x 264 bytes of the 4 KiB a launch takes):
// a workgroup reads its class in place, where a copy with the class's fields
// patched in would go to scratch memory.
constexpr int kMaxMergedClasses = 14;
struct RowMultiArgs {
    u32 ncls;
    u32 tile0[kMaxMergedClasses];
    RowArgs cls[kMaxMergedClasses];
};
static_assert(sizeof(RowMultiArgs) <= 4096, "kernel argument block");

__global__ __launch_bounds__(kMaxPassThreads) void k_row_small(RowMultiArgs M) {
    __shared__ RowLine s_line[kMaxRowLines];
    const u32 slot = xcd_slot(blockIdx.x, gridDim.x);
    const u32 tile = slot / 3u;
    u32 c = 0;
    while (c + 1 < M.ncls && tile >= M.tile0[c + 1]) c++;
    const RowArgs& A = M.cls[c];
    const u32 t = tile - M.tile0[c];
    switch (slot - 3u * tile) {
        case 0: row_body<0, 4>(A, s_line, t); break;
        case 1: row_body<1, 4>(A, s_line, t); break;
        default: row_body<2, 4>(A, s_line, t); break;
    }
}

// Row-only classes gathered for one k_row_small launch.
struct MergedRows {
    RowMultiArgs M{};
    size_t lds = 0;
    int threads = 0;
    u64 tiles = 0;
    // Takes the class's row pass (as launch_class would set it up, with the
    // tiles numbered on); false if the class does not fit this launch.
    bool add(const Job& j, size_t first, u32 count, const Tuning& tu, const Twiddles& tw);
    // Launches the classes gathered so far, if any, and starts over.
    hipError_t launch(hipStream_t s);
};

// ------------------------------------------------------- CRT + carry scan
//
// Coefficient k of the product is c_k = CRT(x_0, x_1, x_2) in [0, P) with
// x_i = redc(res_i[k], K_i) (src/core.h:77-104 computes the same value).
// The limbs are out_k = (s_k + carry_k) mod 2^64 with
// s_k = c_k[0] + c_{k-1}[1] + c_{k-2}[2] < 3 * 2^64 and carry_k <= 2
// (c_k < 2^184).  The carry out of limb k is a map f_k: {0..3} -> {0..3},
// f_k(c) = hi(s_k) + [lo(s_k) + c >= 2^64]; maps compose associatively, so
// the serial loop of src/fast_mul.c:378-389 becomes a scan over maps.  In
// registers a map is four bytes, byte c = m(c), so that g after f is one byte
// permute (v_perm_b32 selects byte f(c) of g for byte c); in fmap[] a limb's
// map is packed into one byte as four 2-bit values.
//
// Three kernels, no inter-workgroup waiting and no contended atomics:
//   k_crt_local   one chunk of limbs per workgroup: CRT, s_k, for each limb
//                 the map from the chunk's carry-in to the limb's carry-in
//                 into fmap[], and the chunk's composed map into
//                 chunk_map[]; writes lo(s_k) into out[], plus the carry-in
//                 where that map is constant (the limb is then final)
//   k_chunk_scan  one workgroup: the carry into every chunk, chunk_cin[]
//   k_carry_apply adds fmap[k](chunk_cin) into each limb of out[] whose map
//                 is not constant (in practice the chunk's first limb).

struct CarryArgs {
    const ProductDesc* items;
    ProductDesc one;
    const u32* chunk_item;  // crt entry of each chunk (nullptr: entry 0)
    u8* fmap;               // per limb: chunk carry-in -> limb carry-in
    u32* chunk_map;         // composed map of each chunk
    u32* chunk_cin;         // carry into each chunk
};

__device__ __forceinline__ u32 map_compose(u32 g, u32 f) {  // g after f
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_perm(g, g, f);
#else
    u32 r = 0;
    for (int c = 0; c < 4; c++) r |= ((g >> (8 * ((f >> (8 * c)) & 3u))) & 0xFFu) << (8 * c);
    return r;
#endif
}
__device__ __forceinline__ u32 map_apply(u32 m, u32 c) { return (m >> (8 * c)) & 0xFFu; }
constexpr u32 kMapId = 0x03020100u;  // 0 -> 0, 1 -> 1, 2 -> 2, 3 -> 3
// fmap[] byte <-> register form
__device__ __forceinline__ u32 map_pack(u32 m) {
    return (m & 3u) | ((m >> 6) & 0xCu) | ((m >> 12) & 0x30u) | ((m >> 18) & 0xC0u);
}
__device__ __forceinline__ u32 map_unpack(u32 b) {
    return (b & 3u) | ((b & 0xCu) << 6) | ((b & 0x30u) << 12) | ((b & 0xC0u) << 18);
}
// A packed map that sends every carry to the same value.
__device__ __forceinline__ bool map_const(u32 b) { return b == (b & 3u) * 0x55u; }

// Residues of coefficient k (zero beyond the convolution).  For a split
// product, coefficient k sums those of piece i = k / m at offset k - i m and
// of piece i - 1 at offset k - i m + m, where that offset is inside its
// cyclic length (the sum stays below p0 p1 p2, so adding residues is exact).
// Inverse transforms leave residues below 4p; each is reduced below 2p so
// that the sum of two fits 64 bits.
__device__ __forceinline__ void crt_load(const ProductDesc& D, u64 k, u64& v0, u64& v1, u64& v2) {
    v0 = v1 = v2 = 0;
    if (k >= D.na + D.nb - 1) return;
    const u64 N = 1ull << D.log_n;
    if (!D.split_n) {
        const FMH_GLOBAL u64* q = gmem(D.res + k);
        v0 = q[0];
        v1 = q[N];
        v2 = q[2 * N];
        return;
    }
    const u32 m = (u32)D.split_m;
    const u32 i = (u32)k / m;  // k < 2^30
    const u64 r = k - (u64)i * m;
    if (i < D.split_n) {
        const FMH_GLOBAL u64* q = gmem(D.res + (u64)i * 3 * N + r);
        v0 = norm2p<0>(q[0]);
        v1 = norm2p<1>(q[N]);
        v2 = norm2p<2>(q[2 * N]);
    }
    if (i >= 1 && i - 1 < D.split_n && r + m < N) {
        const FMH_GLOBAL u64* q = gmem(D.res + (u64)(i - 1) * 3 * N + r + m);
        v0 += norm2p<0>(q[0]);
        v1 += norm2p<1>(q[N]);
        v2 += norm2p<2>(q[2 * N]);
    }
}

// Coefficient from its residues as three words (a zero coefficient has zero
// residues and gives zero).
__device__ __forceinline__ void crt_coef(const CrtConsts& g, u64 v0, u64 v1, u64 v2, u64& w0, u64& w1,
                                         u64& w2) {
    constexpr u64 p0 = Prime<0>::p, p1 = Prime<1>::p, p2 = Prime<2>::p;
    const u64 x0 = canon<0>(redc<0>(v0, g.k00));
    const u64 y1 = canon<1>(norm2p<1>(redc<1>(v1, g.k11) + 2 * p1 - redc<1>(x0, g.k10)));
    u64 t = norm2p<2>(redc<2>(v2, g.k22) + 2 * p2 - redc<2>(x0, g.k20));
    const u64 y2 = canon<2>(norm2p<2>(t + 2 * p2 - redc<2>(y1, g.k21)));
    // C = x0 + p0 y1 + (p0 p1) y2
    u64 a_lo, a_hi, b_lo, b_hi, c_lo, c_hi;
    mul_64x64(p0, y1, a_lo, a_hi);
    mul_64x64(g.m01[0], y2, b_lo, b_hi);
    mul_64x64(g.m01[1], y2, c_lo, c_hi);
    const u128 s0 = (u128)x0 + a_lo + b_lo;
    const u128 s1 = (u128)(u64)(s0 >> 64) + a_hi + b_hi + c_lo;
    w0 = (u64)s0;
    w1 = (u64)s1;
    w2 = c_hi + (u64)(s1 >> 64);
}

// Exclusive scan of one map per thread over the workgroup (order matters);
// returns the composition of the maps of lower threads, *agg gets all of them.
__device__ __forceinline__ u32 block_excl_scan_maps(u32 F, u32* s_wave, u32* agg) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nw = blockDim.x >> 6;
    u32 incl = F;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const u32 o = __shfl_up(incl, d, 64);
        if (lane >= d) incl = map_compose(incl, o);
    }
    u32 excl = __shfl_up(incl, 1, 64);
    if (lane == 0) excl = kMapId;
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    u32 wpre = kMapId;
    for (int w = 0; w < wave; w++) wpre = map_compose(s_wave[w], wpre);
    u32 all = kMapId;
    for (int w = 0; w < nw; w++) all = map_compose(s_wave[w], all);
    *agg = all;
    return map_compose(excl, wpre);
}

// Chunk layout: limb i = e kCrtThreads + t of a chunk (e < kCrtPerThread) is
// handled by thread t, so every load and store of a row e is coalesced.  Limb
// i needs words 1 and 2 of coefficients i - 1 and i - 2: lanes below in the
// same wave (shuffles), or, for lanes 0 and 1, the last two lanes of the
// previous 64-limb segment (through LDS; the chunk's first segment takes
// coefficients k0 - 2 and k0 - 1, computed by threads 0 and 1).  The carry
// maps are scanned per segment in the wave, then across the segments in
// limb order.
__device__ __forceinline__ u64 shfl_up_u64(u64 v, unsigned d) {
    const u32 lo = __shfl_up((unsigned)v, d, 64), hi = __shfl_up((unsigned)(v >> 32), d, 64);
    return ((u64)hi << 32) | lo;
}

__global__ __launch_bounds__(kCrtThreads) void k_crt_local(CarryArgs A) {
    constexpr int E = kCrtPerThread;
    constexpr int CH = kCarryChunk;
    constexpr int NW = kCrtThreads / 64;
    constexpr int NS = E * NW;  // 64-limb segments per chunk, segment e NW + wave
    // s_edge[s + 1]: words 1 and 2 of the coefficients of lanes 62 and 63 of
    // segment s; s_edge[0]: those of coefficients k0 - 2 and k0 - 1
    __shared__ u64 s_edge[NS + 1][4];
    __shared__ u32 s_agg[NS];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const u64 tile = blockIdx.x;
    const ProductDesc& D = product(A.items, A.one, A.chunk_item ? A.chunk_item[tile] : 0u);
    const u64 k0 = (tile - D.chunk0) * (u64)CH;
    const u64 out_len = D.na + D.nb;

    u64 v[E][3];
    // threads 0 and 1 also take coefficient k0 - 2 + tid
    const bool before = tid < 2 && k0 >= (u64)(2 - tid);
    u64 x[3] = {0, 0, 0};
    if (!D.split_n) {
        // all 3E + 3 loads in flight at once: indices outside the product
        // are clamped to its last coefficient and their values dropped
        const u64 kend = D.na + D.nb - 1;
        const u64 N = 1ull << D.log_n;
        const FMH_GLOBAL u64* res = gmem(D.res);
#pragma unroll
        for (int e = 0; e < E; e++) {
            const u64 k = k0 + (u64)(e * kCrtThreads + tid);
            const u64 kc = k < kend ? k : kend - 1;
            v[e][0] = res[kc];
            v[e][1] = res[kc + N];
            v[e][2] = res[kc + 2 * N];
        }
        const u64 kx = before ? k0 - 2 + tid : kend - 1;
        x[0] = res[kx];
        x[1] = res[kx + N];
        x[2] = res[kx + 2 * N];
#pragma unroll
        for (int e = 0; e < E; e++)
            if (k0 + (u64)(e * kCrtThreads + tid) >= kend) v[e][0] = v[e][1] = v[e][2] = 0;
        if (!before) x[0] = x[1] = x[2] = 0;
    } else {
        if (before) crt_load(D, k0 - 2 + tid, x[0], x[1], x[2]);
#pragma unroll
        for (int e = 0; e < E; e++) crt_load(D, k0 + (u64)(e * kCrtThreads + tid), v[e][0], v[e][1], v[e][2]);
    }
    u64 w0[E], w1[E], w2[E];
#pragma unroll
    for (int e = 0; e < E; e++) {
        crt_coef(D.crt, v[e][0], v[e][1], v[e][2], w0[e], w1[e], w2[e]);
        if (lane >= 62) {
            s_edge[e * NW + wave + 1][2 * (lane - 62)] = w1[e];
            s_edge[e * NW + wave + 1][2 * (lane - 62) + 1] = w2[e];
        }
    }
    if (tid < 2) {
        u64 c0, c1, c2;
        crt_coef(D.crt, x[0], x[1], x[2], c0, c1, c2);
        s_edge[0][2 * tid] = c1;
        s_edge[0][2 * tid + 1] = c2;
    }
    __syncthreads();

    u32 f[E], incl[E];
    u64 lo[E];
#pragma unroll
    for (int e = 0; e < E; e++) {
        const int sg = e * NW + wave;
        u64 u1 = shfl_up_u64(w1[e], 1), u2 = shfl_up_u64(w2[e], 2);
        if (lane == 0) {
            u1 = s_edge[sg][2];
            u2 = s_edge[sg][1];
        } else if (lane == 1) {
            u2 = s_edge[sg][3];
        }
        const u128 sum = (u128)w0[e] + u1 + u2;
        const u64 l = (u64)sum;
        const u32 h = (u32)(sum >> 64);
        // byte c: h + [l + c >= 2^64]
        f[e] = h * 0x01010101u + ((l == ~0ull ? 1u : 0u) << 8) + ((l >= ~0ull - 1 ? 1u : 0u) << 16) +
               ((l >= ~0ull - 2 ? 1u : 0u) << 24);
        lo[e] = l;
        // inclusive scan of the segment's maps in lane order
        u32 m = f[e];
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const u32 o = __shfl_up(m, d, 64);
            if (lane >= d) m = map_compose(m, o);
        }
        incl[e] = m;
        if (lane == 63) s_agg[sg] = m;
    }
    __syncthreads();
    // maps of the segments before each of this thread's segments
    u32 run = kMapId, pre[E];
#pragma unroll
    for (int sg = 0; sg < NS; sg++) {
        if (sg % NW == wave) pre[sg / NW] = run;
        run = map_compose(s_agg[sg], run);
    }
    // carry-in of each limb as a map of the chunk's carry-in; where that map
    // is constant (every limb past the first whose low word cannot overflow)
    // the carry is added here and the limb is final
    FMH_GLOBAL u8* fm = gmem(A.fmap + tile * CH);
    FMH_GLOBAL u64* out = gmem(D.out);
#pragma unroll
    for (int e = 0; e < E; e++) {
        u32 ex = __shfl_up(incl[e], 1, 64);
        if (lane == 0) ex = kMapId;
        const u32 mp = map_compose(ex, pre[e]);
        const u32 pk = map_pack(mp);
        fm[e * kCrtThreads + tid] = (u8)pk;
        const u64 k = k0 + (u64)(e * kCrtThreads + tid);
        if (k < out_len) out[k] = lo[e] + (map_const(pk) ? map_apply(mp, 0) : 0u);
    }
    if (tid == 0) A.chunk_map[tile] = run;
}

// One workgroup per product: the carry into each of its chunks.  The chunk
// maps pass through LDS in tiles of kScanTile (coalesced loads); within a
// tile, thread t owns a contiguous segment, and the carry out of each tile
// enters the next.
#ifndef FMH_SCAN_TILE
#define FMH_SCAN_TILE 8192  // the CPU emulation builds with a small tile to reach the multi-tile path
#endif
constexpr int kScanTile = FMH_SCAN_TILE;

__global__ __launch_bounds__(kScanThreads) void k_chunk_scan(CarryArgs A) {
    __shared__ u32 s_wave[kScanThreads / 64];
    __shared__ u32 s_m[kScanTile];
    const int tid = threadIdx.x;
    const ProductDesc& D = product(A.items, A.one, blockIdx.x);
    const u32 n = D.nchunks;
    const FMH_GLOBAL u32* maps = gmem(A.chunk_map + D.chunk0);
    FMH_GLOBAL u32* cin = gmem(A.chunk_cin + D.chunk0);
    u32 carry = 0;  // carry into the current tile
    for (u32 base = 0; base < n; base += kScanTile) {
        const u32 cnt = min(n - base, (u32)kScanTile);
        batched_copy(
            (int)cnt, kScanThreads, [&](int i) { return (u64)maps[base + i]; }, [&](int i, u64 v) { s_m[i] = (u32)v; });
        __syncthreads();
        const u32 per = (cnt + kScanThreads - 1) / kScanThreads;
        const u32 beg = min(cnt, (u32)tid * per), end = min(cnt, beg + per);
        u32 seg = kMapId;
        for (u32 i = beg; i < end; i++) seg = map_compose(s_m[i], seg);
        u32 all;
        const u32 excl = block_excl_scan_maps(seg, s_wave, &all);
        u32 c = map_apply(excl, carry);
        for (u32 i = beg; i < end; i++) {
            cin[base + i] = c;
            c = map_apply(s_m[i], c);
        }
        carry = map_apply(all, carry);
        __syncthreads();  // s_m and s_wave are reused by the next tile
    }
}

// FUSED: the workgroup composes the maps of the chunks before its own
// instead of reading k_chunk_scan's result (products of at most
// kFusedChunks chunks; each workgroup then reads at most that many maps).
#ifndef FMH_FUSED_CHUNKS
#define FMH_FUSED_CHUNKS 2048  // the CPU emulation builds with 4 to reach both paths with short products
#endif
constexpr u32 kFusedChunks = FMH_FUSED_CHUNKS;

template <bool FUSED>
__global__ __launch_bounds__(kThreads) void k_carry_apply(CarryArgs A) {
    constexpr int E = kCarryPerThread;
    constexpr int CH = kCarryChunk;
    const int tid = threadIdx.x;
    const u64 tile = blockIdx.x;
    const ProductDesc& D = product(A.items, A.one, A.chunk_item ? A.chunk_item[tile] : 0u);
    const u64 k0 = (tile - D.chunk0) * (u64)CH;
    // the chunk starts inside the product: k0 < out_len
    const u64 lim = D.na + D.nb - k0;
    const FMH_GLOBAL u8* fm = gmem(A.fmap + tile * CH);
    u32 cin;
    if (FUSED) {
        // maps of chunks chunk0 .. tile - 1, thread t composing a contiguous
        // run of them in order
        // (before < kFusedChunks: every load of a thread is issued at once)
        constexpr u32 kPer = (kFusedChunks + kThreads - 1) / kThreads;
        __shared__ u32 s_wave[kThreads / 64];
        const u32 before = (u32)(tile - D.chunk0);
        const u32 per = (before + kThreads - 1) / kThreads;
        const u32 beg = min(before, (u32)tid * per), end = min(before, beg + per);
        const FMH_GLOBAL u32* maps = gmem(A.chunk_map + D.chunk0);
        u32 mv[kPer];
#pragma unroll
        for (u32 k = 0; k < kPer; k++) mv[k] = beg + k < end ? maps[beg + k] : kMapId;
        u32 seg = kMapId;
#pragma unroll
        for (u32 k = 0; k < kPer; k++) seg = map_compose(mv[k], seg);
        u32 all;
        (void)block_excl_scan_maps(seg, s_wave, &all);
        cin = map_apply(all, 0);
    } else {
        cin = A.chunk_cin[tile];
    }
    // limbs whose carry-in map is constant were finished by the CRT kernel
    // (in practice all but the chunk's first)
    FMH_GLOBAL u64* out = gmem(D.out + k0);
    u32 b[E];
#pragma unroll
    for (int e = 0; e < E; e++) b[e] = fm[e * kThreads + tid];
#pragma unroll
    for (int e = 0; e < E; e++) {
        const u64 i = (u64)(e * kThreads + tid);
        if (map_const(b[e]) || i >= lim) continue;
        const u32 c = map_apply(map_unpack(b[e]), cin);
        if (c != 0) out[i] += c;
    }
}

// ---------------------------------------------------------- small kernels

__global__ void k_fill_u32(u32* p, u64 n, u32 v) {
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_fill_u64(u64* p, u64 n, u64 v) {
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Twiddle tables.  Top section: tw[T/2 + t] = root^t for t < T/2, computed
// by square-and-multiply (canonical Montgomery products); lower sections are
// subsamples: tw[h + t] = tw[T/2 + t * (T / 2h)].
template <int P>
__device__ u64 mont_pow_canon(u64 base_m, u64 e, u64 one_m) {
    u64 r = one_m;
    while (e) {
        if (e & 1) r = canon<P>(redc<P>(r, base_m));
        base_m = canon<P>(redc<P>(base_m, base_m));
        e >>= 1;
    }
    return r;
}

struct TwArgs {
    u64* tab[6];   // [dir * 3 + prime]
    u64 root[6];   // Montgomery root of order T (dir 0) or its inverse (dir 1)
    u64 one[3];
    int log_t;
};

__global__ void k_tw_top(TwArgs A) {
    const u64 half = 1ull << (A.log_t - 1);
    const u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    const int idx = blockIdx.y;  // dir * 3 + prime
    if (t >= half) return;
    const int P = idx % 3;
    u64 v;
    switch (P) {
        case 0: v = mont_pow_canon<0>(A.root[idx], t, A.one[0]); break;
        case 1: v = mont_pow_canon<1>(A.root[idx], t, A.one[1]); break;
        default: v = mont_pow_canon<2>(A.root[idx], t, A.one[2]); break;
    }
    A.tab[idx][half + t] = v;
}

__global__ void k_tw_low(TwArgs A) {
    const u64 half = 1ull << (A.log_t - 1);
    const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;  // 0 .. half-1
    const int idx = blockIdx.y;
    if (i >= half) return;
    u64* tab = A.tab[idx];
    if (i == 0) {
        tab[0] = 0;
        return;
    }
    const int lh = 63 - __clzll((long long)i);  // h = 2^lh <= i < 2h
    const u64 h = 1ull << lh;
    const u64 t = i - h;
    tab[i] = tab[half + t * (half >> lh)];
}

int tw_full_log() {
    static const int F = [] {
        const char* v = getenv("FMH_TW_FULL_LOG");
        const int f = v && *v ? atoi(v) : kTwFullLog;
        return std::max(kMaxRowTileLog, std::min(f, kMaxLogN));
    }();
    return F;
}

// Builds the tables for transforms of up to 2^log_t points and then swaps
// them in.  The full tables (2^min(log_t, F) words, built on the device) are
// rebuilt only when they grow; the extension tables of the half-spans >= 2^F
// (2^(log_t - F + 1) words, at most 2^(31 - F): 64 words per table with
// F = 25) are computed on the host.  Tables that are replaced are never
// freed: work already enqueued, and graphs captured with them, keep valid
// pointers (tables only grow, so at most half of the memory held is
// retired).
hipError_t build_twiddles(Twiddles& tw, int log_t, hipStream_t s) {
    if (log_t < 1) log_t = 1;
    const int F = tw_full_log();
    const int lf = std::min(log_t, F);
    u64* fwd[3] = {tw.fwd[0], tw.fwd[1], tw.fwd[2]};
    u64* inv[3] = {tw.inv[0], tw.inv[1], tw.inv[2]};
    u64* fwd_x[3] = {nullptr, nullptr, nullptr};
    u64* inv_x[3] = {nullptr, nullptr, nullptr};
    std::vector<void*> made;
    auto fail = [&](hipError_t e) {
        (void)hipStreamSynchronize(s);  // table kernels already launched finish first
        for (void* q : made) (void)hipFree(q);
        return e;
    };
    if (tw.log_full != lf) {
        const u64 T = 1ull << lf;
        TwArgs A{};
        A.log_t = lf;
        for (int i = 0; i < 3; i++) {
            hipError_t e = hipMalloc(&fwd[i], T * sizeof(u64));
            if (e == hipSuccess) made.push_back(fwd[i]);
            if (e == hipSuccess) e = hipMalloc(&inv[i], T * sizeof(u64));
            if (e == hipSuccess) made.push_back(inv[i]);
            if (e != hipSuccess) return fail(e);
            const u64 p = kP[i];
            const u64 root = h_powmod(kG[i], (p - 1) >> lf, p);
            const u64 iroot = h_powmod(root, p - 2, p);
            A.tab[i] = fwd[i];
            A.tab[3 + i] = inv[i];
            A.root[i] = h_mont(root, p);
            A.root[3 + i] = h_mont(iroot, p);
            A.one[i] = h_mont(1, p);
        }
        const u64 half = T / 2;
        dim3 grid((unsigned)((half + 255) / 256), 6);
        hipLaunchKernelGGL(k_tw_top, grid, dim3(256), 0, s, A);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return fail(e);
        hipLaunchKernelGGL(k_tw_low, grid, dim3(256), 0, s, A);
        e = hipGetLastError();
        if (e != hipSuccess) return fail(e);
    }
    if (log_t > F) {
        // entry [K + t] = w_{2h}^t (t < K) for h = K 2^(F-1), K = 2 .. 2^(log_t - F)
        const u64 X = 2ull << (log_t - F);
        std::vector<u64> hx(X), hix(X);
        for (int i = 0; i < 3; i++) {
            const u64 p = kP[i];
            hx[0] = hx[1] = hix[0] = hix[1] = 0;
            for (u64 K = 2; K < X; K <<= 1) {
                const int order_log = F + ceil_log2(K);  // 2h = 2^(F + log K)
                const u64 root = h_powmod(kG[i], (p - 1) >> order_log, p);
                const u64 iroot = h_powmod(root, p - 2, p);
                u64 r = 1, ir = 1;
                for (u64 t = 0; t < K; t++) {
                    hx[K + t] = h_mont(r, p);
                    hix[K + t] = h_mont(ir, p);
                    r = (u64)(((u128)r * root) % p);
                    ir = (u64)(((u128)ir * iroot) % p);
                }
            }
            hipError_t e = hipMalloc(&fwd_x[i], X * sizeof(u64));
            if (e == hipSuccess) made.push_back(fwd_x[i]);
            if (e == hipSuccess) e = hipMalloc(&inv_x[i], X * sizeof(u64));
            if (e == hipSuccess) made.push_back(inv_x[i]);
            if (e == hipSuccess) e = hipMemcpy(fwd_x[i], hx.data(), X * sizeof(u64), hipMemcpyHostToDevice);
            if (e == hipSuccess) e = hipMemcpy(inv_x[i], hix.data(), X * sizeof(u64), hipMemcpyHostToDevice);
            if (e != hipSuccess) return fail(e);
        }
    }
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) return fail(e);
    for (int i = 0; i < 3; i++) {
        tw.fwd[i] = fwd[i];
        tw.inv[i] = inv[i];
        tw.fwd_x[i] = fwd_x[i];
        tw.inv_x[i] = inv_x[i];
    }
    tw.log_full = lf;
    tw.log_t = log_t;
    return hipSuccess;
}

// ------------------------------------------------------------- pipelines

// Garner constants for every transform length, computed once.
static const CrtConsts& crt_table(int log_n) {
    static CrtConsts tab[kMaxLogN + 1];
    static std::once_flag once;
    std::call_once(once, [] {
        for (int l = 0; l <= kMaxLogN; l++) tab[l] = crt_consts(l);
    });
    return tab[log_n];
}

// Tiles above 64 KiB of LDS need the per-kernel limit raised (gfx950 has
// 160 KiB of LDS per workgroup).
// Largest dynamic LDS each pass kernel may be launched with: gfx950 offers
// 160 KiB per workgroup, static LDS included; kernels default to 64 KiB
// unless the limit is raised.  Launches that would exceed it are refused
// with an error instead of being sent.
// k_col<true, false, false>, k_col<false, false, false>, k_row<4>, k_row<3>,
// k_col<true, true, false>, the BIG instances of the three k_col, k_row_small,
// k_row<2>
static size_t g_max_dyn[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};

static size_t raise_lds_limit(const void* f) {
    constexpr size_t kDefault = 64 * 1024;
    hipFuncAttributes attr{};
    size_t stat = 0;
    if (hipFuncGetAttributes(&attr, f) == hipSuccess) stat = attr.sharedSizeBytes;
    size_t allowed = kLdsPerGroup - stat;
    if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)allowed) != hipSuccess)
        allowed = kDefault > stat ? kDefault - stat : 0;
    (void)hipGetLastError();  // a refused attribute must not surface at the next launch
    return allowed;
}

void prepare_kernels() {
    (void)crt_table(0);
    static std::once_flag once;
    std::call_once(once, [] {
        g_max_dyn[0] = raise_lds_limit((const void*)k_col<true, false, false>);
        g_max_dyn[1] = raise_lds_limit((const void*)k_col<false, false, false>);
        g_max_dyn[2] = raise_lds_limit((const void*)k_row<4>);
        g_max_dyn[3] = raise_lds_limit((const void*)k_row<3>);
        g_max_dyn[4] = raise_lds_limit((const void*)k_col<true, true, false>);
        g_max_dyn[5] = raise_lds_limit((const void*)k_col<true, false, true>);
        g_max_dyn[6] = raise_lds_limit((const void*)k_col<false, false, true>);
        g_max_dyn[7] = raise_lds_limit((const void*)k_col<true, true, true>);
        g_max_dyn[8] = raise_lds_limit((const void*)k_row_small);
        g_max_dyn[9] = raise_lds_limit((const void*)k_row<2>);
    });
}

// Split plan of an unbalanced product n1 x n2 (n1 >= n2, balanced length
// 2^L): pieces of m limbs of the long operand, each multiplied at 2^k points
// (m + n2 - 1 <= 2^k) with one spectrum of the short operand.  Cost model:
// transforms of N points cost N (log N + 4); the split runs 2 K + 1 of them at
// 2^k, the balanced product 3 at 2^L.  Returns k, or 0 for no split.
constexpr u32 kMaxPieces = 4096;
constexpr u32 kSpecPart = ~0u;  // part index of a split product's spectrum item
constexpr int kMinPieceLog = 10;
static int split_plan(u64 n1, u64 n2, int L, u64* m_out, u32* pieces) {
    if (L <= tuning(L).row_tile) return 0;  // one row pass already
    const double balanced = 3.0 * (double)(1ull << L) * (L + 4);
    double best = 0.75 * balanced;  // split only for a clear gain
    int best_k = 0;
    for (int k = std::max(kMinPieceLog, ceil_log2(2 * n2)); k < L; k++) {
        const u64 m = (1ull << k) - n2 + 1;
        const u64 K = (n1 + m - 1) / m;
        if (K < 2 || K > kMaxPieces) continue;
        const double cost = (2.0 * (double)K + 1.0) * (double)(1ull << k) * (k + 4);
        if (cost < best) {
            best = cost;
            best_k = k;
            *m_out = m;
            *pieces = (u32)K;
        }
    }
    return best_k;
}

bool products_valid(const ProductIn* in, size_t count) {
    for (size_t i = 0; i < count; i++) {
        const ProductIn& p = in[i];
        // the operand that is transformed (for a spectrum product: the other
        // one; a null b means a square), and the output (the lengths as in
        // job_build)
        if (p.na && !p.a) return false;
        const u64 out_len = p.spec ? (p.a ? p.na + p.nb : 2 * p.nb) : (p.b ? p.na + p.nb : 2 * p.na);
        if (out_len && !p.out) return false;
    }
    return true;
}

int job_build(Job& j, const ProductIn* in, size_t count, bool allow_split) {
    j = Job();
    std::vector<ProductDesc> items;
    std::vector<u32> owner, part;  // crt entry and piece of each transform item
    for (size_t i = 0; i < count; i++) {
        const ProductIn& p = in[i];
        ProductDesc d{};
        if (p.spec) {
            d.kind = p.a ? kKindMulSpec : kKindSqrSpec;
            d.a = p.a ? p.a : nullptr;
            d.b = nullptr;
            d.spec = p.spec;
            d.na = p.a ? p.na : p.nb;
            d.nb = p.nb;
        } else {
            const bool square = p.b == nullptr || (p.b == p.a && p.nb == p.na);
            d.kind = square ? kKindSqr : kKindMul;
            d.a = p.a;
            d.b = square ? p.a : p.b;
            d.na = p.na;
            d.nb = square ? p.na : p.nb;
        }
        d.out = p.out;
        if (d.na == 0 || d.nb == 0) {
            if (d.na + d.nb) j.zeros.push_back({p.out, d.na + d.nb});
            continue;
        }
        const u64 out_len = d.na + d.nb;
        int log_n = ceil_log2(out_len - 1);
        if (log_n > kMaxLogN) return -1;
        if (p.spec) {
            // the product is taken at the spectrum's length (any cyclic
            // length >= na + nb - 1 gives the same coefficients)
            if (log_n > p.spec_log_n) return -2;
            log_n = p.spec_log_n;
        }
        d.nchunks = (u32)((out_len + kCarryChunk - 1) / kCarryChunk);
        const u32 ci = (u32)j.crt.size();
        u64 m = 0;
        u32 K = 0;
        const int k = allow_split && d.kind == kKindMul
                          ? split_plan(std::max(d.na, d.nb), std::min(d.na, d.nb), log_n, &m, &K)
                          : 0;
        if (k) {
            // long operand a, short operand b
            const u64* la = d.na >= d.nb ? d.a : d.b;
            const u64* sb = d.na >= d.nb ? d.b : d.a;
            const u64 ln = std::max(d.na, d.nb), sn = std::min(d.na, d.nb);
            ProductDesc f{};
            f.kind = kKindSpecFwd;
            f.a = f.b = sb;
            f.na = f.nb = sn;
            f.log_n = (u32)k;
            items.push_back(f);
            owner.push_back(ci);
            part.push_back(kSpecPart);
            for (u32 q = 0; q < K; q++) {
                ProductDesc c{};
                c.kind = kKindMulSpec;
                c.a = la + (u64)q * m;
                c.na = std::min<u64>(m, ln - (u64)q * m);
                c.nb = sn;
                c.log_n = (u32)k;
                c.crt = crt_table(k);
                items.push_back(c);
                owner.push_back(ci);
                part.push_back(q);
            }
            d.log_n = (u32)k;
            d.split_m = m;
            d.split_n = K;
            d.crt = crt_table(k);
        } else {
            d.log_n = (u32)log_n;
            d.crt = crt_table(log_n);
            items.push_back(d);
            owner.push_back(ci);
            part.push_back(0);
        }
        j.crt.push_back(d);
    }
    // transforms sorted by class, spectra of split products first; the
    // pieces of one product stay adjacent
    std::vector<u32> order(items.size());
    for (u32 i = 0; i < (u32)order.size(); i++) order[i] = i;
    auto key = [&](u32 x) { return items[x].kind == kKindSpecFwd ? -1 : (int)items[x].kind; };
    std::stable_sort(order.begin(), order.end(), [&](u32 x, u32 y) {
        return key(x) != key(y) ? key(x) < key(y) : items[x].log_n < items[y].log_n;
    });
    j.crt_res_off.assign(j.crt.size(), 0);
    j.items.resize(items.size());
    j.res_off.resize(items.size());
    j.item_owner.resize(items.size());
    for (size_t i = 0; i < order.size(); i++) {
        const ProductDesc& d = items[order[i]];
        j.items[i] = d;
        j.res_off[i] = j.res_words;
        j.item_owner[i] = owner[order[i]];
        if (part[order[i]] == 0) j.crt_res_off[owner[order[i]]] = j.res_words;
        j.res_words += (u64)(d.kind == kKindMul ? 6 : 3) << d.log_n;
        j.max_log_n = std::max(j.max_log_n, (int)d.log_n);
    }
    j.item_spec.assign(items.size(), 0);
    for (size_t i = 0; i < order.size(); i++) j.item_spec[i] = part[order[i]] == kSpecPart;
    for (size_t c = 0; c < j.crt.size(); c++) {
        ProductDesc& d = j.crt[c];
        d.chunk0 = j.chunks;
        j.chunks += d.nchunks;
        for (u32 q = 0; q < d.nchunks; q++) j.chunk_item.push_back((u32)c);
    }
    return 0;
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

size_t job_desc_bytes(const Job& j) {
    if (j.items.size() <= 1 && j.crt.size() <= 1) return 0;
    return align256(j.items.size() * sizeof(ProductDesc)) + align256(j.crt.size() * sizeof(ProductDesc)) +
           align256(j.chunk_item.size() * sizeof(u32));
}

size_t job_ws_bytes(const Job& j) {
    size_t b = align256(j.res_words * sizeof(u64));
    b += align256((size_t)j.chunks * kCarryChunk);
    b += 2 * align256((size_t)j.chunks * sizeof(u32));
    b += job_desc_bytes(j);
    return b;
}

void job_bind(Job& j, void* ws) {
    char* p = (char*)ws;
    u64* res = (u64*)p;
    p += align256(j.res_words * sizeof(u64));
    // residues; the pieces of a split product read the spectrum its
    // kKindSpecFwd item writes
    std::vector<const u64*> spec_of(j.crt.size(), nullptr);
    for (size_t c = 0; c < j.crt.size(); c++) j.crt[c].res = res + j.crt_res_off[c];
    for (size_t i = 0; i < j.items.size(); i++) {
        j.items[i].res = res + j.res_off[i];
        if (j.item_spec[i]) spec_of[j.item_owner[i]] = j.items[i].res;
    }
    for (size_t i = 0; i < j.items.size(); i++)
        if (!j.item_spec[i] && spec_of[j.item_owner[i]]) j.items[i].spec = spec_of[j.item_owner[i]];
    j.fmap = (u8*)p;
    p += align256((size_t)j.chunks * kCarryChunk);
    j.chunk_map = (u32*)p;
    p += align256((size_t)j.chunks * sizeof(u32));
    j.chunk_cin = (u32*)p;
    p += align256((size_t)j.chunks * sizeof(u32));
    if (job_desc_bytes(j)) {
        j.d_desc = p;
        j.d_items = (ProductDesc*)p;
        p += align256(j.items.size() * sizeof(ProductDesc));
        j.d_crt = (ProductDesc*)p;
        p += align256(j.crt.size() * sizeof(ProductDesc));
        j.d_chunk_item = (u32*)p;
    } else {
        j.d_desc = nullptr;
        j.d_items = nullptr;
        j.d_crt = nullptr;
        j.d_chunk_item = nullptr;
    }
}

void job_pack(const Job& j, void* host) {
    char* p = (char*)host;
    memcpy(p, j.items.data(), j.items.size() * sizeof(ProductDesc));
    p += align256(j.items.size() * sizeof(ProductDesc));
    memcpy(p, j.crt.data(), j.crt.size() * sizeof(ProductDesc));
    p += align256(j.crt.size() * sizeof(ProductDesc));
    memcpy(p, j.chunk_item.data(), j.chunk_item.size() * sizeof(u32));
}

hipError_t job_upload(const Job& j, hipStream_t s) {
    const size_t n = job_desc_bytes(j);
    if (!n) return hipSuccess;
    std::vector<char> host(n);
    job_pack(j, host.data());
    hipError_t e = hipMemcpyAsync(j.d_desc, host.data(), n, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(s);  // the host copy goes out of scope
}

// Whether a forward column pass over the products d[0, count) may fill its
// tiles by LDS DMA: tiles of at least two adjacent lines, and for the first
// pass every operand it reads (nops of each product) 16-byte aligned.
static bool col_dma_ok(const Tuning& tu, int logC, bool from_limbs, const ProductDesc* d, size_t count, int nops) {
    if (!tu.col_dma || logC < 1) return false;
    if (!from_limbs) return true;
    for (size_t i = 0; i < count; i++)
        for (int op = 0; op < nops; op++) {
            const u64 len = op ? d[i].nb : d[i].na;
            const uintptr_t p = (uintptr_t)(op ? d[i].b : d[i].a);
            if (len && (p & 15)) return false;
        }
    return true;
}

// One column pass: the k_col instance for its direction, its fill (LDS DMA or
// not) and whether its stages reach half-spans of 2^tw.log_full or more.
static hipError_t launch_col(ColArgs& C, const Twiddles& tw, bool fwd, bool dma, dim3 grid, size_t lds,
                             hipStream_t s) {
    for (int i = 0; i < 3; i++) {
        C.tw[i] = fwd ? tw.fwd[i] : tw.inv[i];
        C.twx[i] = fwd ? tw.fwd_x[i] : tw.inv_x[i];
    }
    C.tw_full = tw.log_full;
    const bool big = C.log_n - 1 - C.s0 >= tw.log_full;  // largest half-span 2^(log_n - 1 - s0)
    if (big && !C.twx[0]) return hipErrorInvalidValue;
    const int k = fwd ? (dma ? 4 : 0) : 1;
    if (lds > g_max_dyn[big ? (k == 4 ? 7 : 5 + k) : k]) return hipErrorInvalidValue;
    const char* name = fwd ? "col_fwd" : "col_inv";
    const dim3 block(C.threads);
    if (fwd && dma) {
        if (big)
            FMH_LAUNCH(name, (k_col<true, true, true>), grid, block, lds, s, C);
        else
            FMH_LAUNCH(name, (k_col<true, true, false>), grid, block, lds, s, C);
    } else if (fwd) {
        if (big)
            FMH_LAUNCH(name, (k_col<true, false, true>), grid, block, lds, s, C);
        else
            FMH_LAUNCH(name, (k_col<true, false, false>), grid, block, lds, s, C);
    } else {
        if (big)
            FMH_LAUNCH(name, (k_col<false, false, true>), grid, block, lds, s, C);
        else
            FMH_LAUNCH(name, (k_col<false, false, false>), grid, block, lds, s, C);
    }
    return hipSuccess;
}

// One row pass: the k_row instance of the geometry's round size (2, 3 or 4
// stages per LDS round), 3 workgroups (one per prime) per tile.
static hipError_t launch_row(int radix, const char* name, u64 tiles, int threads, size_t lds, hipStream_t s,
                             const RowArgs& R) {
    const int slot = radix == 2 ? 9 : radix == 3 ? 3 : 2;
    if (lds > g_max_dyn[slot]) return hipErrorInvalidValue;
    const dim3 grid((unsigned)(tiles * 3), 1, 1);
    if (radix == 2)
        FMH_LAUNCH(name, k_row<2>, grid, dim3(threads), lds, s, R);
    else if (radix == 3)
        FMH_LAUNCH(name, k_row<3>, grid, dim3(threads), lds, s, R);
    else
        FMH_LAUNCH(name, k_row<4>, grid, dim3(threads), lds, s, R);
    return hipSuccess;
}

bool MergedRows::add(const Job& j, size_t first, u32 count, const Tuning& tu, const Twiddles& tw) {
    if (M.ncls >= (u32)kMaxMergedClasses || (M.ncls && tu.row_threads != threads)) return false;
    const ProductDesc& d0 = j.items[first];
    const int logL = (int)d0.log_n;
    int logC = std::max(0, std::min(tu.row_tile - logL, 6));
    while (logC > 0 && (1ull << (logC - 1)) >= count) logC--;
    RowArgs& R = M.cls[M.ncls];
    R = RowArgs{};
    R.items = j.d_items + first;
    R.count = count;
    for (int i = 0; i < 3; i++) {
        R.tw[i] = tw.fwd[i];
        R.itw[i] = tw.inv[i];
    }
    R.log_n = logL;
    R.s0 = 0;
    R.from_limbs = 1;
    R.mode = d0.kind == kKindMul      ? kRowMul
             : d0.kind == kKindSqr    ? kRowSqr
             : d0.kind == kKindMulSpec ? kRowMulSpec
                                       : kRowFwdKeep;
    R.threads = tu.row_threads;
    R.logC = logC;
    M.tile0[M.ncls++] = (u32)tiles;
    tiles += ((u64)count + (1ull << logC) - 1) >> logC;
    lds = std::max(lds, row_lds_bytes(d0.kind == kKindMul ? 2 : 1, logL, logC));
    threads = tu.row_threads;
    return true;
}

hipError_t MergedRows::launch(hipStream_t s) {
    if (!M.ncls) return hipSuccess;
    if (lds > g_max_dyn[8] || tiles * 3 > 0x7fffffffull) return hipErrorInvalidValue;
    const RowMultiArgs args = M;  // a launch argument, not a member: this object starts over below
    FMH_LAUNCH("row_small", k_row_small, dim3((unsigned)(tiles * 3), 1, 1), dim3(threads), lds, s, args);
    *this = MergedRows();
    return hipSuccess;
}

// Transforms of one size class: items [first, first + count) share log_n and
// the square flag.
static hipError_t launch_class(const Job& j, size_t first, u32 count, const Twiddles& tw, hipStream_t s) {
    const ProductDesc& d0 = j.items[first];
    const int log_n = (int)d0.log_n;
    const Tuning tu = tuning(log_n, count >= kBatchedClass);
    const u32 kind = d0.kind;
    const u64 N = 1ull << log_n;
    const Plan pl = make_plan(log_n, tu);
    const ProductDesc* items = j.items.size() > 1 ? j.d_items + first : nullptr;
    const int nops = kind == kKindMul ? 2 : kind == kKindSqrSpec ? 0 : 1;  // operands to transform
    for (int pi = 0; nops > 0 && pi < pl.n_col; pi++) {
        ColArgs C{};
        C.items = items;
        C.one = d0;
        C.log_n = log_n;
        C.s0 = pl.col_s[pi];
        C.s1 = pl.col_s[pi + 1];
        C.from_limbs = (pi == 0);
        C.nops = nops;
        C.threads = tu.col_threads;
        const int logL = C.s1 - C.s0;
        C.logC = tu.col_logc;
        const u64 tiles = N >> (logL + C.logC);
        const size_t lds = col_lds_bytes(logL, C.logC);
        const bool dma = col_dma_ok(tu, C.logC, C.from_limbs != 0, &j.items[first], count, nops);
        const hipError_t e = launch_col(C, tw, true, dma, dim3((unsigned)(tiles * 3 * nops), 1, count), lds, s);
        if (e != hipSuccess) return e;
    }
    {
        RowArgs R{};
        R.items = items;
        R.one = d0;
        R.count = count;
        for (int i = 0; i < 3; i++) {
            R.tw[i] = tw.fwd[i];
            R.itw[i] = tw.inv[i];
        }
        R.log_n = log_n;
        R.s0 = pl.row_s;
        R.from_limbs = (pl.n_col == 0);
        static const char* const kRowName[5] = {"row_mul", "row_sqr", "row_mul_spec", "row_sqr_spec",
                                                "row_fwd_keep"};
        static const int kRowModeOf[5] = {kRowMul, kRowSqr, kRowMulSpec, kRowSqrSpec, kRowFwdKeep};
        R.mode = kRowModeOf[kind];
        R.threads = tu.row_threads;
        const int logL = log_n - pl.row_s;
        const u64 lines = (u64)count << pl.row_s;
        int logC = std::max(0, std::min(tu.row_tile - logL, 6));  // kMaxRowLines = 64
        while (logC > 0 && (1ull << (logC - 1)) >= lines) logC--;
        R.logC = logC;
        const u64 tiles = (lines + (1ull << logC) - 1) >> logC;
        const size_t lds = row_lds_bytes(kind == kKindMul ? 2 : 1, logL, logC);
        const hipError_t e = launch_row(tu.row_radix, kRowName[kind], tiles, tu.row_threads, lds, s, R);
        if (e != hipSuccess) return e;
    }
    for (int pi = pl.n_col - 1; kind != kKindSpecFwd && pi >= 0; pi--) {
        ColArgs C{};
        C.items = items;
        C.one = d0;
        C.log_n = log_n;
        C.s0 = pl.col_s[pi];
        C.s1 = pl.col_s[pi + 1];
        C.from_limbs = 0;
        C.nops = 1;
        C.threads = tu.col_threads;
        const int logL = C.s1 - C.s0;
        C.logC = tu.col_logc;
        const u64 tiles = N >> (logL + C.logC);
        const size_t lds = col_lds_bytes(logL, C.logC);
        const hipError_t e = launch_col(C, tw, false, false, dim3((unsigned)(tiles * 3), 1, count), lds, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

bool job_one_multiply_with_col(const Job& j) {
    return j.items.size() == 1 && j.zeros.empty() && j.items[0].kind == kKindMul &&
           make_plan((int)j.items[0].log_n, tuning((int)j.items[0].log_n)).n_col > 0;
}

hipError_t job_launch(const Job& j, const Twiddles& tw, hipStream_t s) {
    prepare_kernels();
    if (j.max_log_n > tw.log_t) return hipErrorInvalidValue;
    for (const auto& z : j.zeros) {
        hipError_t e = enqueue_zero(z.first, z.second, s);
        if (e != hipSuccess) return e;
    }
    if (j.items.empty()) return hipSuccess;
    if (job_desc_bytes(j) && !j.d_desc) return hipErrorInvalidValue;
    constexpr u32 kMaxClass = 32767;  // grid.z holds the products of a class
    // Row-only classes of a batch go into merged k_row_small launches, with
    // radix-16 rounds whatever round size their geometry names (the
    // products are the same): the spectra of split products (kKindSpecFwd)
    // in one, launched before the first class of another kind (the pieces
    // read those spectra), and the multiplies, squares and pieces
    // (kKindMulSpec) in another after every other class (classes of these
    // kinds do not depend on each other).  A class with another workgroup
    // size than the first of its group, or beyond a group's 14 classes,
    // launches on its own.
    MergedRows spec_rows, rows;
    size_t first = 0;
    while (first < j.items.size()) {
        size_t last = first;
        while (last < j.items.size() && last - first < kMaxClass && j.items[last].log_n == j.items[first].log_n &&
               j.items[last].kind == j.items[first].kind)
            last++;
        const u32 count = (u32)(last - first);
        const ProductDesc& d0 = j.items[first];
        const Tuning tu = tuning((int)d0.log_n, count >= kBatchedClass);
        const bool row_only = j.items.size() > 1 && (int)d0.log_n <= tu.row_tile;
        MergedRows* g = nullptr;
        if (row_only && d0.kind == kKindSpecFwd)
            g = &spec_rows;
        else if (row_only && (d0.kind == kKindMul || d0.kind == kKindSqr || d0.kind == kKindMulSpec))
            g = &rows;
        if (d0.kind != kKindSpecFwd) {
            hipError_t e = spec_rows.launch(s);
            if (e != hipSuccess) return e;
        }
        if (!(g && g->add(j, first, count, tu, tw))) {
            hipError_t e = launch_class(j, first, count, tw, s);
            if (e != hipSuccess) return e;
        }
        first = last;
    }
    hipError_t e = spec_rows.launch(s);
    if (e == hipSuccess) e = rows.launch(s);
    if (e != hipSuccess) return e;
    CarryArgs K{};
    K.items = j.crt.size() > 1 ? j.d_crt : nullptr;
    K.one = j.crt[0];
    K.chunk_item = j.crt.size() > 1 ? j.d_chunk_item : nullptr;
    K.fmap = j.fmap;
    K.chunk_map = j.chunk_map;
    K.chunk_cin = j.chunk_cin;
    FMH_LAUNCH("crt_local", k_crt_local, dim3(j.chunks), dim3(kCrtThreads), 0, s, K);
    u32 most = 0;  // chunks of the job's longest product
    for (const auto& d : j.crt) most = std::max(most, d.nchunks);
    if (most <= kFusedChunks) {
        FMH_LAUNCH("carry_apply", k_carry_apply<true>, dim3(j.chunks), dim3(kThreads), 0, s, K);
    } else {
        FMH_LAUNCH("chunk_scan", k_chunk_scan, dim3((unsigned)j.crt.size()), dim3(kScanThreads), 0, s, K);
        FMH_LAUNCH("carry_apply", k_carry_apply<false>, dim3(j.chunks), dim3(kThreads), 0, s, K);
    }
    return hipSuccess;
}

__global__ void k_copy(u64* __restrict__ dst, const u64* __restrict__ src, u64 n) {
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

hipError_t enqueue_copy(u64* dst, const u64* src, u64 n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    u64 blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    FMH_LAUNCH("copy", k_copy, dim3((unsigned)blocks), dim3(256), 0, s, dst, src, n);
    return hipSuccess;
}

hipError_t enqueue_zero(u64* out, u64 n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    FMH_LAUNCH("fill_out", k_fill_u64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, out, n, 0ull);
    return hipSuccess;
}

size_t spectrum_ws_bytes(int log_n) { return (size_t)3 << log_n << 3; }

// Forward transform of one operand into the 3 residue arrays at res, ending
// with the row pass in `mode` (kRowFwdOnly or kRowFwdKeep).
static hipError_t enqueue_forward(const u64* a, u64 na, int log_n, u64* res, int mode, const Twiddles& tw,
                                  hipStream_t s) {
    if (log_n > kMaxLogN || log_n > tw.log_t) return hipErrorInvalidValue;
    prepare_kernels();
    const u64 N = 1ull << log_n;
    const Tuning tu = tuning(log_n);
    const Plan pl = make_plan(log_n, tu);
    ProductDesc d{};
    d.a = d.b = a;
    d.na = d.nb = na;
    d.res = res;
    d.log_n = (u32)log_n;
    d.kind = kKindSqr;
    for (int pi = 0; pi < pl.n_col; pi++) {
        ColArgs C{};
        C.one = d;
        C.log_n = log_n;
        C.s0 = pl.col_s[pi];
        C.s1 = pl.col_s[pi + 1];
        C.from_limbs = (pi == 0);
        C.nops = 1;
        C.threads = tu.col_threads;
        const int logL = C.s1 - C.s0;
        C.logC = tu.col_logc;
        const u64 tiles = N >> (logL + C.logC);
        const size_t lds = col_lds_bytes(logL, C.logC);
        const bool dma = col_dma_ok(tu, C.logC, C.from_limbs != 0, &d, 1, 1);
        const hipError_t e = launch_col(C, tw, true, dma, dim3((unsigned)(tiles * 3), 1, 1), lds, s);
        if (e != hipSuccess) return e;
    }
    RowArgs R{};
    R.one = d;
    R.count = 1;
    for (int i = 0; i < 3; i++) {
        R.tw[i] = tw.fwd[i];
        R.itw[i] = tw.inv[i];
        const u64 p = kP[i];
        const u64 r = (u64)((((u128)1) << 64) % p);
        R.r2[i] = (u64)(((u128)r * r) % p);
    }
    R.log_n = log_n;
    R.s0 = pl.row_s;
    R.from_limbs = (pl.n_col == 0);
    R.mode = mode;
    R.threads = tu.row_threads;
    const int logL = log_n - pl.row_s;
    const int logC = std::max(0, std::min({pl.row_s, tu.row_tile - logL, 6}));
    R.logC = logC;
    const u64 tiles = N >> (logL + logC);
    const size_t lds = row_lds_bytes(1, logL, logC);
    return launch_row(tu.row_radix, mode == kRowFwdKeep ? "row_fwd_keep" : "row_fwd", tiles, tu.row_threads, lds, s,
                      R);
}

hipError_t enqueue_spectrum(int prime, const u64* a, u64 na, int log_n, u64* out, void* ws,
                            const Twiddles& tw, hipStream_t s) {
    if (prime < 0 || prime > 2) return hipErrorInvalidValue;
    u64* res = (u64*)ws;
    hipError_t e = enqueue_forward(a, na, log_n, res, kRowFwdOnly, tw, s);
    if (e != hipSuccess) return e;
    return hipMemcpyAsync(out, res + ((u64)prime << log_n), (sizeof(u64) << log_n), hipMemcpyDeviceToDevice, s);
}

hipError_t enqueue_resident_spectrum(const u64* a, u64 na, int log_n, u64* spec, const Twiddles& tw,
                                     hipStream_t s) {
    return enqueue_forward(a, na, log_n, spec, kRowFwdKeep, tw, s);
}

}  // namespace fmh
