// fmh_internal.h -- launchers shared by the kernels (fmh_ntt.hip) and the
// host layer (fmh_api.hip).  Not part of the public C-ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "fmh_field.h"

namespace fmh {

// Largest transform length handled (2^28 points per prime: products of up to
// 2^27 + 2^27 limbs).  The reference's own limit is 2^55 (src/data.h:69).
constexpr int kMaxLogN = 28;

// LDS tile geometry.
constexpr int kThreads = 256;        // threads per workgroup for every pass
constexpr int kColLogC = 3;          // column passes: 8 adjacent lines
constexpr int kColMaxLogL = 10;      // column passes: at most 2^10 points/line
constexpr int kRowMaxLogL = 11;      // row pass: at most 2^11 points/line
constexpr int kRowTileLog = 12;      // row pass: 2^12 points per operand tile
constexpr int kCarryPerThread = 8;   // limbs per thread in the carry scan
constexpr int kCarryChunk = kThreads * kCarryPerThread;

// One transform plan: stages [0, n) of the in-place radix-2 DIF split into
// column passes [s_i, s_{i+1}) and a final row pass [s_row, n).
struct Plan {
    int log_n;
    int n_col;           // number of column passes
    int col_s[8];        // column pass i covers stages [col_s[i], col_s[i+1])
    int row_s;           // row pass covers stages [row_s, log_n)
};

Plan make_plan(int log_n);

// Per-call CRT constants: K_i = N^-1 * R^2 * (P/p_i)^-1 mod p_i.
struct CrtConsts {
    u64 k[3];
};
CrtConsts crt_consts(int log_n);

// Twiddle tables: tw[prime] (forward) and itw[prime] (inverse), each
// 2^log_t words, entry [h + t] = w_{2h}^{+-t} * 2^64 mod p for 1 <= h < 2^log_t.
struct Twiddles {
    int log_t = -1;
    u64* fwd[3] = {nullptr, nullptr, nullptr};
    u64* inv[3] = {nullptr, nullptr, nullptr};
};

hipError_t build_twiddles(Twiddles& tw, int log_t, hipStream_t s);

// Workspace layout for one product of transform length N = 2^log_n:
//   res[0..2]  residues of operand A (and the product), N words each
//   res[3..5]  residues of operand B (multiply only), N words each
//   flags      carry look-back: 1 counter + one word per chunk (u32)
size_t workspace_bytes(int log_n, bool square, u64 out_len);

// Enqueue the whole device pipeline on stream s.  All pointers are device
// pointers; `ws` must hold workspace_bytes(...) bytes.
hipError_t enqueue_mul(const u64* a, u64 na, const u64* b, u64 nb, u64* out, void* ws,
                       const Twiddles& tw, hipStream_t s);
hipError_t enqueue_sqr(const u64* a, u64 na, u64* out, void* ws, const Twiddles& tw,
                       hipStream_t s);
hipError_t enqueue_zero(u64* out, u64 n, hipStream_t s);
// Plain device copy of n words (HBM calibration kernel for profilers).
hipError_t enqueue_copy(u64* dst, const u64* src, u64 n, hipStream_t s);
// Forward spectrum of one operand for one prime, canonical Montgomery form,
// reference (bit-reversed) order.  ws holds at least 2^log_n words.
hipError_t enqueue_spectrum(int prime, const u64* a, u64 na, int log_n, u64* out, void* ws,
                            const Twiddles& tw, hipStream_t s);

// Optional per-kernel timing with HIP events on the launch stream.
// timing_collect() waits for the recorded launches and returns one
// "name count total_ms" line per kernel name.
std::string timing_collect();
void timing_enable(bool on);
bool timing_enabled();
void timing_record(const char* name, hipEvent_t start, hipEvent_t stop);
hipError_t timing_begin(hipStream_t s, hipEvent_t* start);
hipError_t timing_end(const char* name, hipStream_t s, hipEvent_t start);

}  // namespace fmh
