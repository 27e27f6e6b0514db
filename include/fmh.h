/*
 * fmh.h -- C-ABI of libfmh.so, the MI355X (gfx950) 3-prime NTT + CRT
 * big-integer multiply.
 *
 * Numbers are little-endian arrays of 64-bit limbs.  Outputs are exactly the
 * limbs of the product (len1 + len2 limbs, 2 * len1 for a square); the top
 * limb may be zero.  Results are bit-identical to the reference's src/
 * path: the exact product is unique.
 *
 * Every call is thread-safe.  Functions returning int give 0 on success and
 * a nonzero code on failure (fmh_last_error() describes the last failure of
 * the calling thread).
 */
#ifndef FMH_H
#define FMH_H

#ifdef __cplusplus
extern "C" {
#endif

typedef unsigned long long u64;

/* ---------------------------------------------------------------- drop-in
 * Replace src/fast_mul.h:39 and src/fast_mul.h:45 of HJimmyK/fast_mul_ntt
 * (implemented in src/fast_mul.c:308-396 and 398-460): same names, same
 * signatures, host pointers, synchronous.  `out` may alias an input.  Like
 * the reference, a device failure prints a message to stderr and calls
 * abort().  A zero-length operand gives the all-zero product (the reference
 * aborts or overruns its buffers there).                                    */
void abs_mul64(const u64* in1, u64 len1, const u64* in2, u64 len2, u64* out);
void abs_sqr64(const u64* in1, u64 len1, u64* out);

/* ------------------------------------------------------ device pointers
 * No counterpart in the reference (its API is host-only).  Inputs and output
 * live in device memory of the current HIP device; the work is enqueued on
 * `stream` (a hipStream_t; NULL is the null stream) and returns before it
 * completes.  Scratch memory is owned by the library, one workspace per
 * stream: calls on one stream run in order, calls on different streams do
 * not share scratch.  `out` must not alias an input.  A null pointer for a
 * non-empty operand or output is an error returned before anything runs.   */
typedef void* fmh_stream_t;

int fmh_mul_device(const u64* d_in1, u64 len1, const u64* d_in2, u64 len2, u64* d_out,
                   fmh_stream_t stream);
int fmh_sqr_device(const u64* d_in1, u64 len1, u64* d_out, fmh_stream_t stream);

/* Allocation-free variant of fmh_mul_device.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * It calls no hipMalloc / hipFree: its workspace is a fixed device array of
 * the library (about 49 MiB per GPU, held until the process ends), so it
 * takes products of len1 + len2 - 1 <= FMH_NOALLOC_MAX_NTT_LEN coefficients
 * (larger ones are refused with an error) and needs fmh_init(device) first.
 * A square is d_in2 == d_in1, len2 == len1.  Products are never split into
 * pieces here; results equal fmh_mul_device's.  Every call on one GPU uses
 * the same array: calls on one stream run in order, calls that overlap in
 * time (other streams, other threads) race and must not be made.          */
#define FMH_NOALLOC_MAX_NTT_LEN (1ull << 20)
int fmh_mul_device_noalloc(const u64* d_in1, u64 len1, const u64* d_in2, u64 len2, u64* d_out,
                           fmh_stream_t stream);

/* Forward transform of one operand for prime index `prime` (0, 1, 2 = the
 * reference's moduli 1, 2, 3), zero-padded to ntt_len (a power of two >=
 * len): ntt_len canonical residues X * 2^64 mod p in the reference's order,
 * position j = X[bitrev(j)] -- the values src/core.h:443-474 dif_i (and the
 * split-radix step of src/fast_mul.c:74-95 above 2^17 points) produce, reduced
 * to [0, p).  For parity checks.                                          */
int fmh_forward_spectrum_device(int prime, const u64* d_in, u64 len, u64 ntt_len, u64* d_out,
                                fmh_stream_t stream);

/* ------------------------------------------------------------- batches
 * No counterpart in the reference (one product per call there).  Many
 * products of any sizes -- multiplies and squares mixed -- run as one set of
 * launches: products are grouped by transform length, small transforms are
 * packed several to a workgroup, and the CRT/carry kernels cover the whole
 * batch.  All pointers are device pointers; in2 == NULL (or in2 == in1 with
 * len2 == len1) is a square.  Outputs must not overlap each other or any
 * input. */
typedef struct {
    const u64* in1;
    u64 len1;
    const u64* in2;
    u64 len2;
    u64* out; /* len1 + len2 limbs (2 * len1 for a square) */
} fmh_product;

/* One-shot: enqueue the batch on `stream` (scratch is the stream's workspace). */
int fmh_mul_batch_device(const fmh_product* items, u64 count, fmh_stream_t stream);

/* Planned batch: the workspace and the device copy of the descriptors are
 * allocated once and the launch sequence is captured in a HIP graph, so a
 * batch whose operand and output pointers stay the same is re-run with one
 * graph launch.  fmh_batch_run may be called any number of times; runs on
 * one batch must not overlap in time (they share its workspace). */
typedef struct fmh_batch fmh_batch;
int fmh_batch_create(const fmh_product* items, u64 count, fmh_batch** batch);
int fmh_batch_run(fmh_batch* batch, fmh_stream_t stream);
int fmh_batch_destroy(fmh_batch* batch);

/* ------------------------------------------------- resident spectra
 * The forward transform of an operand (3 primes x ntt_len words of device
 * memory, owned by the object) kept for reuse: each product with it runs one
 * forward transform instead of two, a square of it none.  This is the
 * semantics of src/fast_mul.c:135-212 conv_single_i (first operand already
 * transformed, never called in the reference).
 *
 * fmh_spectrum_create enqueues the transform of d_in (len limbs) on `stream`;
 * ntt_len = 0 picks next_pow2(2 len - 1), enough for a square and for products
 * with operands of up to len limbs.  The operand buffer is not referenced
 * afterwards.  Products with the spectrum must run after the transform
 * (same stream, or synchronised) and need len + len2 - 1 <= ntt_len.
 * Products on different streams may share one spectrum (it is read-only). */
typedef struct fmh_spectrum fmh_spectrum;
int fmh_spectrum_create(const u64* d_in, u64 len, u64 ntt_len, fmh_stream_t stream, fmh_spectrum** spec);
u64 fmh_spectrum_ntt_len(const fmh_spectrum* spec);
/* d_out (len + len2 limbs) = operand of spec * d_in2 */
int fmh_mul_spectrum_device(const fmh_spectrum* spec, const u64* d_in2, u64 len2, u64* d_out,
                            fmh_stream_t stream);
/* d_out (2 len limbs) = (operand of spec)^2 */
int fmh_sqr_spectrum_device(const fmh_spectrum* spec, u64* d_out, fmh_stream_t stream);
/* Waits for the device, then frees the spectrum. */
int fmh_spectrum_destroy(fmh_spectrum* spec);
/* The spectrum's device words (3 x ntt_len), operand length and log2(ntt_len). */
int fmh_spectrum_describe(const fmh_spectrum* spec, const u64** words, u64* len, int* log_n);

/* ------------------------------------------------------------- serving
 * No counterpart in the reference.  A request loop with continuous batching:
 * requests (device operands, or a resident spectrum and a device operand)
 * are submitted as they arrive; a scheduler thread admits each one when its
 * footprint -- output limbs * 8 + 3 primes * NTT length * 8 B * operands
 * transformed -- fits in the server's device pool, groups waiting requests of
 * one size class (transform length, kind) into one batched launch, and keeps
 * up to `streams` groups in flight on separate streams.  Outputs live in the
 * pool until fmh_server_release.  Operands must stay valid and unchanged
 * until their request completes.  Latency is submit -> completion observed
 * by the scheduler (host clock). */
typedef struct {
    u64 pool_bytes;      /* device pool size; 0: 64 GiB */
    unsigned streams;    /* launch groups in flight; 0: 4 (at most 16) */
    unsigned max_group;  /* requests per launch group; 0: 1024 */
    u64 max_group_bytes; /* residue bytes per launch group; 0: pool / 8 */
} fmh_server_config;

typedef struct {
    u64 submitted, completed, failed;
    u64 groups;              /* launch groups run */
    double mean_group;       /* requests per group */
    double p50_us, p99_us, max_us, mean_us;  /* submit -> completion */
    double products_per_s;   /* completed / span_s */
    double span_s;           /* first submit -> last completion */
    u64 pool_bytes, pool_peak_bytes;
    double device_busy_s;    /* time with at least one launch group executing
                                (HIP events around each group) */
} fmh_server_info;

typedef struct fmh_server fmh_server;
int fmh_server_create(const fmh_server_config* cfg, fmh_server** server);
/* d_b == NULL: square of d_a.  *id names the request. */
int fmh_server_submit(fmh_server* server, const u64* d_a, u64 na, const u64* d_b, u64 nb, u64* id);
/* Product of the spectrum's operand with d_b (nb limbs); d_b == NULL: its square. */
int fmh_server_submit_spectrum(fmh_server* server, const fmh_spectrum* spec, const u64* d_b, u64 nb,
                               u64* id);
/* Ids of requests completed (or failed) since the last poll, at most max. */
int fmh_server_poll(fmh_server* server, u64* ids, u64 max, u64* n);
int fmh_server_wait(fmh_server* server, u64 id);
/* Device pointer and length of a completed request's output (valid until release). */
int fmh_server_result(fmh_server* server, u64 id, u64** d_out, u64* len);
int fmh_server_release(fmh_server* server, u64 id);
/* Waits until every submitted request has completed or failed.  If no work
 * is in flight and the oldest waiting request cannot be admitted because
 * unreleased outputs hold the pool, fmh_server_drain and fmh_server_wait
 * return an error instead of blocking: release completed requests (poll)
 * while submitting. */
int fmh_server_drain(fmh_server* server);
int fmh_server_stats(fmh_server* server, fmh_server_info* stats);
/* Device time per size class since the last reset: one line per class,
 * "kind log_n groups requests device_ms" (kind 0 multiply, 1 square, 2 with a
 * spectrum, 3 square of a spectrum, 4 an unbalanced multiply split into
 * pieces of 2^log_n points; device_ms sums each group's execution on
 * its stream, so groups that overlap count in full).  Writes at most buflen
 * bytes and returns the bytes needed, terminating NUL included. */
u64 fmh_server_class_report(fmh_server* server, char* buf, u64 buflen);
/* Zeroes the counters, latencies and device times (no request in flight). */
int fmh_server_reset_stats(fmh_server* server);
/* Drains, stops the scheduler and frees the pool. */
int fmh_server_destroy(fmh_server* server);
/* Message of the calling thread's last failed fmh_server_* call. */
const char* fmh_server_last_error(void);

/* Largest supported transform length (products with len1 + len2 - 1 up to
 * this many coefficients). */
u64 fmh_max_ntt_len(void);

/* -------------------------------------------------------------- plumbing
 * Thin wrappers over the HIP runtime so that C, Go (cgo) or Python (ctypes)
 * callers can drive the device entry points without HIP headers.          */
int fmh_init(int device);
int fmh_dev_alloc(void** ptr, u64 bytes);
int fmh_dev_free(void* ptr);
int fmh_memcpy_h2d(void* dst, const void* src, u64 bytes, fmh_stream_t stream);
int fmh_memcpy_d2h(void* dst, const void* src, u64 bytes, fmh_stream_t stream);
int fmh_stream_create(fmh_stream_t* stream);
int fmh_stream_destroy(fmh_stream_t stream);
int fmh_stream_synchronize(fmh_stream_t stream);
/* Waits for the current device, then frees the library's cached scratch
 * (per-stream workspaces, staging of the host-pointer calls).  They are
 * otherwise kept, grown as needed, until fmh_stream_destroy or process exit.
 * Twiddle tables stay (planned batches may refer to them). */
int fmh_trim(void);
/* Plain word copy on the device, launched as kernel "copy": the known-size
 * reference kernel against which profiler byte counters are calibrated. */
int fmh_copy_device(u64* d_dst, const u64* d_src, u64 n, fmh_stream_t stream);
const char* fmh_last_error(void);

/* Pass geometry of transforms of 2^log_n points (log_n = -1: every size),
 * for single products (batched = 0) or for launches carrying several
 * products of the class (batched = 1): spec "col_logc=3,col_maxl=10,
 * row_maxl=11,row_tile=12,col_threads=512,row_threads=512,row_radix=4"
 * (missing keys keep their value; "" restores the defaults).  Defaults come
 * from tuning_gfx950.txt beside the library (written by
 * fast_mul_ntt_amd/autotune.py) or $FMH_TUNE_TABLE; $FMH_TUNE overrides
 * every size.  Results never depend on the geometry. */
int fmh_set_tuning(int log_n, const char* spec, int batched);
int fmh_get_tuning(int log_n, int batched, char* buf, u64 buflen);

/* Per-kernel timing with HIP events recorded on each kernel's launch stream
 * (for benchmarks).  fmh_timing_collect writes "name count total_ms" lines
 * for the launches recorded since the last collect and returns the number of
 * bytes needed (including the terminating NUL). */
int fmh_timing_enable(int on);
/* Time only the kernels called `kernel` (NULL: all), so that the events add
 * no gaps between the other kernels of a timed region. */
int fmh_timing_only(const char* kernel);
u64 fmh_timing_collect(char* buf, u64 buflen);

#ifdef __cplusplus
}
#endif

#endif /* FMH_H */
