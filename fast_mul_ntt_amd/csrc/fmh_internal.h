// fmh_internal.h -- launchers shared by the kernels (fmh_ntt.hip) and the
// host layer (fmh_api.hip).  Not part of the public C-ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

#include "fmh_field.h"

namespace fmh {

// Largest transform length handled (2^30 points per prime: products of up to
// 2^29 + 2^29 limbs, about 58 GiB of device memory with the twiddle tables).
// Column-pass twiddle offsets are 32-bit byte offsets below 4N; the reference's
// own limit is 2^55 (src/data.h:69).
constexpr int kMaxLogN = 30;

// LDS of the pass kernels.  gfx950 gives a workgroup 160 KiB, static LDS
// included.  The tuning clamp keeps every tile within the bounds below, and
// static_asserts beside the kernels check that bound plus each kernel's
// static LDS against kLdsPerGroup (a CPU test checks the compiled static
// sizes as well); a launch over the limit the runtime grants is refused.
constexpr size_t kLdsPerGroup = 160 * 1024;
constexpr int kMaxColTileLog = 14;  // words per column tile: 2^(col_maxl + col_logc) <= 2^14
constexpr int kMaxRowTileLog = 13;  // points per operand row tile: 2^row_tile <= 2^13
// Dynamic LDS of a column tile of 2^logC lines of 2^logL words, and of a row
// tile of `operands` tiles of 2^logL x 2^logC words (at least 16 words each:
// the row-tile swizzle permutes 16-word blocks).
constexpr size_t col_lds_bytes(int logL, int logC) { return sizeof(unsigned long long) << (logL + logC); }
constexpr size_t row_lds_bytes(int operands, int logL, int logC) {
    return (size_t)operands * sizeof(unsigned long long) * (logL + logC < 4 ? 16 : (size_t)1 << (logL + logC));
}

constexpr int kThreads = 256;         // threads per workgroup of the carry kernels
constexpr int kScanThreads = 512;     // threads of the chunk scan (one workgroup per product)
constexpr int kMaxPassThreads = 1024; // launch bound of the NTT pass kernels
constexpr int kCarryPerThread = 4;    // limbs per thread in the carry kernels
constexpr int kCarryChunk = kThreads * kCarryPerThread;  // limbs per chunk
// The CRT kernel covers a chunk with more threads and fewer coefficients
// each (fewer registers, more waves): 2^21-point products 25.2 -> 22.5 us.
constexpr int kCrtThreads = 512;
constexpr int kCrtPerThread = kCarryChunk / kCrtThreads;

// Tile geometry of the NTT passes, per transform length 2^log_n.  Sources,
// later ones overriding earlier ones:
//   * the defaults below;
//   * a tuning table, one "log_n spec" line per size class (and "b log_n
//     spec" lines for batched launches), read once from
//     $FMH_TUNE_TABLE or else tuning_gfx950.txt beside libfmh.so (written by
//     fast_mul_ntt_amd/autotune.py);
//   * fmh_set_tuning(log_n, spec) at run time;
//   * $FMH_TUNE (one spec for every size; for sweeps and tests).
// A spec reads "col_logc=3,col_maxl=10,row_maxl=11,row_tile=12,
// col_threads=512,row_threads=512,row_radix=4,col_dma=1"; missing keys keep
// their value.
struct Tuning {
    int col_logc = 3;       // column passes: 2^col_logc adjacent lines per tile
    int col_maxl = 10;      // column passes: at most 2^col_maxl points per line
    int row_maxl = 11;      // row pass: at most 2^row_maxl points per line
    int row_tile = 12;      // row pass: 2^row_tile points per operand tile; a
                            // transform of at most 2^row_tile points is one pass
    int col_threads = 512;  // threads per workgroup, column passes
    int row_threads = 512;  // threads per workgroup, row pass
    int row_radix = 4;      // row pass: at most row_radix stages per LDS round (2, 3 or 4)
    int col_dma = 1;        // forward column passes: fill tiles by LDS DMA where possible (0 or 1)
};
// batched: the geometry for a launch that carries several products of the
// class (table lines "b <log_n> <spec>"); single products use the plain
// entries, and so do batches of a class that has no batched entry.
Tuning tuning(int log_n, bool batched = false);
// Applies spec to the class log_n (-1: every class); "" restores the
// defaults of that class.  Returns 0, or -1 for a malformed spec.
int set_tuning(int log_n, const char* spec, bool batched = false);
// Launches of at least this many products of one class use the batched table.
constexpr unsigned kBatchedClass = 4;

// One transform plan: stages [0, n) of the in-place radix-2 DIF split into
// column passes [s_i, s_{i+1}) and a final row pass [s_row, n).
struct Plan {
    int log_n;
    int n_col;           // number of column passes
    int col_s[8];        // column pass i covers stages [col_s[i], col_s[i+1])
    int row_s;           // row pass covers stages [row_s, log_n)
};

Plan make_plan(int log_n);
Plan make_plan(int log_n, const Tuning& tu);

// Per-call CRT constants (Garner form, see crt_consts in fmh_ntt.hip).
struct CrtConsts {
    u64 k00, k11, k10, k22, k20, k21;
    u64 m01[2];  // p0 * p1 as two words
};
CrtConsts crt_consts(int log_n);

// Twiddle tables covering transforms of up to 2^log_t points, per prime:
//   fwd / inv   2^log_full words (log_full = min(log_t, F)), entry [h + t] =
//               w_{2h}^{+-t} * 2^64 mod p for 1 <= h < 2^log_full;
//   fwd_x/inv_x for the longer half-spans h = 2^e, F <= e < log_t (transforms
//               above 2^F points): 2^(log_t - F + 1) words, entry [K + t] =
//               w_{2h}^{+-t} * 2^64 mod p for t < K = h / 2^(F-1).
// A twiddle of such a row is the product of two entries: with t = t_hi K +
// t_lo, w_{2h}^t = w_{2^F}^{t_hi} w_{2h}^{t_lo} = fwd[2^(F-1) + t_hi] *
// fwd_x[K + t_lo].  So the memory held stays at 48 * 2^F bytes however long
// the transform (1.5 GiB with F = 25, instead of 48 GiB at 2^30 points),
// and only column stages of half-span >= 2^F pay the extra product.  F is
// kTwFullLog, or $FMH_TW_FULL_LOG (at least kMaxRowTileLog: row-pass
// twiddles always come from the full table).
constexpr int kTwFullLog = 25;
struct Twiddles {
    int log_t = -1;
    int log_full = -1;
    u64* fwd[3] = {nullptr, nullptr, nullptr};
    u64* inv[3] = {nullptr, nullptr, nullptr};
    u64* fwd_x[3] = {nullptr, nullptr, nullptr};
    u64* inv_x[3] = {nullptr, nullptr, nullptr};
};
// F of the tables this process builds.
int tw_full_log();

hipError_t build_twiddles(Twiddles& tw, int log_t, hipStream_t s);

// What a product transforms.  A resident spectrum is the complete forward
// transform of an operand (3 arrays of 2^log_n words, values in [0, p), the
// layout the row pass produces), kept in device memory for reuse.
enum ProductKind : u32 {
    kKindMul = 0,     // a * b: both operands transformed
    kKindSqr = 1,     // a * a: one forward transform
    kKindMulSpec = 2, // a * (operand of spec): a transformed, spec read
    kKindSqrSpec = 3, // (operand of spec)^2: no forward transform
    kKindSpecFwd = 4, // forward transform of a into a spectrum at res (the
                      // short operand of a split product); launched first
};

// One product as launched: device pointers and its share of the workspace.
// The same struct describes a transform (job items) and a CRT + carry unit
// (job crt entries); for an unsplit product both are the same values.
struct ProductDesc {
    const u64* a;
    const u64* b;     // equals a for a square
    const u64* spec;  // resident spectrum (kinds kKindMulSpec, kKindSqrSpec)
    u64 na, nb;       // nb: length of the spectrum's operand for the spec kinds
    u64* out;         // na + nb limbs
    u64* res;         // residues: 3 arrays of 2^log_n words (6 for kKindMul)
    u32 log_n;
    u32 kind;
    u32 chunk0;       // first carry chunk of this product in the job's chunk arrays
    u32 nchunks;
    // Split product (CRT entries only): the long operand was cut into
    // split_n pieces of split_m limbs, each multiplied at 2^log_n points;
    // piece j's residues are at res + 3 j 2^log_n, and coefficient k of the
    // product sums those of pieces k / split_m and k / split_m - 1.
    u64 split_m;
    u32 split_n;      // 0: not split
    u32 pad_;
    CrtConsts crt;
};

// A product as requested.  Without a spectrum: b == nullptr (or b == a with
// nb == na) is a square.  With spec (a resident spectrum of 2^spec_log_n
// points of an operand of nb limbs): a == nullptr is the square of that
// operand, otherwise the product a * that operand.
struct ProductIn {
    const u64* a;
    u64 na;
    const u64* b;
    u64 nb;
    u64* out;
    const u64* spec = nullptr;
    int spec_log_n = 0;
};

// A set of products launched together: one product, or a batch of products
// of any sizes.  Transforms are grouped into size classes (kind, log_n) and
// each class runs as one launch per pass; the CRT and carry kernels cover all
// products at once.  An unbalanced product may be split: its long operand is
// cut into pieces that each multiply a spectrum of the short one (a
// kKindSpecFwd item computes the spectrum, one kKindMulSpec item per piece
// reads it, one CRT entry covers the whole product).
struct Job {
    std::vector<ProductDesc> items;  // transforms, sorted by (kind, log_n)
    std::vector<ProductDesc> crt;    // CRT + carry units, request order
    std::vector<u64> res_off;        // residue offset of each item (words)
    std::vector<u32> item_owner;     // crt entry of each item
    std::vector<u8> item_spec;       // 1: the item computes a split product's spectrum
    std::vector<u64> crt_res_off;    // residue offset of each crt entry (words)
    std::vector<u32> chunk_item;     // crt entry of each carry chunk
    std::vector<std::pair<u64*, u64>> zeros;  // outputs of products with an empty operand
    u64 res_words = 0;
    u32 chunks = 0;
    int max_log_n = 0;
    // regions of the workspace, set by job_bind
    u8* fmap = nullptr;
    u32* chunk_map = nullptr;
    u32* chunk_cin = nullptr;
    void* d_desc = nullptr;          // device copy of the descriptors (job_pack layout)
    ProductDesc* d_items = nullptr;  // inside d_desc; null unless uploaded
    ProductDesc* d_crt = nullptr;
    u32* d_chunk_item = nullptr;
};

// One-time kernel attributes and constant tables (call before graph capture).
void prepare_kernels();

// Fills the job; returns 0, -1 if a product is longer than 2^kMaxLogN, or -2
// if a product does not fit the transform length of its resident spectrum.
// allow_split = false keeps every product one transform item (no pieces
// sharing an internal spectrum), so that a single product's job needs no
// descriptors in device memory.
int job_build(Job& j, const ProductIn* in, size_t count, bool allow_split = true);
// False if a non-empty operand or output of the products is a null pointer
// (which the kernels would dereference).
bool products_valid(const ProductIn* in, size_t count);
size_t job_ws_bytes(const Job& j);
void job_bind(Job& j, void* ws);
// Descriptors that kernels read from device memory (jobs of more than one
// transform or product): job_desc_bytes() bytes, to be copied from a
// job_pack()ed host buffer to j.d_desc before job_launch.  0 bytes: nothing
// to copy (the single descriptor travels as a kernel argument).
size_t job_desc_bytes(const Job& j);
void job_pack(const Job& j, void* host);
// Copies the descriptors with a synchronous-source hipMemcpyAsync.
hipError_t job_upload(const Job& j, hipStream_t s);
// Enqueues the kernels (no copies, no allocation: capturable in a graph).
hipError_t job_launch(const Job& j, const Twiddles& tw, hipStream_t s);
// True for a job of one multiply, not split into pieces, whose transform has
// a column pass: a host-pointer call may then compute the first operand's
// spectrum while the second operand uploads.
bool job_one_multiply_with_col(const Job& j);

hipError_t enqueue_zero(u64* out, u64 n, hipStream_t s);
// Plain device copy of n words (HBM calibration kernel for profilers).
hipError_t enqueue_copy(u64* dst, const u64* src, u64 n, hipStream_t s);
// Forward spectrum of one operand for one prime, canonical Montgomery form,
// reference (bit-reversed) order.  ws holds spectrum_ws_bytes(log_n) bytes.
size_t spectrum_ws_bytes(int log_n);
hipError_t enqueue_spectrum(int prime, const u64* a, u64 na, int log_n, u64* out, void* ws,
                            const Twiddles& tw, hipStream_t s);
// Resident spectrum of an operand: the 3 x 2^log_n words at spec (no other
// workspace).
hipError_t enqueue_resident_spectrum(const u64* a, u64 na, int log_n, u64* spec, const Twiddles& tw,
                                     hipStream_t s);

// Twiddle tables of the current device covering transforms of 2^log_n points
// (built on first use; tables are never freed, so a copy stays valid).
// Returns 0, or -1 with a message in *err.
int device_twiddles(int log_n, Twiddles* out, std::string* err);

// Optional per-kernel timing with HIP events on the launch stream.
// timing_collect() waits for the recorded launches and returns one
// "name count total_ms" line per kernel name.
std::string timing_collect();
void timing_enable(bool on);
bool timing_enabled();
void timing_record(const char* name, hipEvent_t start, hipEvent_t stop);
// Restricts timing to kernels of one name (NULL or "" = all).
void timing_only(const char* name);
hipError_t timing_begin(const char* name, hipStream_t s, hipEvent_t* start);
hipError_t timing_end(const char* name, hipStream_t s, hipEvent_t start);

}  // namespace fmh
