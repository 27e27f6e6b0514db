// fmh_api.hip -- host layer and C-ABI of libfmh.so (declared in include/fmh.h).
//
// State per HIP device: the twiddle tables (built once on the device, grown
// when a longer transform arrives), one scratch workspace per stream, and
// slots of the host-pointer drop-in (a stream, device staging and a pinned
// ring each).  One mutex guards the shared state.  A drop-in call holds a
// slot of its own and takes the mutex only to enqueue its kernels, so
// concurrent callers overlap their transfers and kernels, as the reference's
// reentrant functions run concurrently on the CPU.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <map>
#include <mutex>
#include <set>
#include <thread>
#include <string>
#include <vector>

#include "../../include/fmh.h"
#include "fmh_internal.h"

namespace {

using namespace fmh;

struct Buffer {
    void* ptr = nullptr;
    size_t bytes = 0;
};

// Memory copies spread over a few persistent host threads (the calling
// thread takes one part), for moving large pageable buffers into and out of
// pinned staging memory.  A staged transfer makes one such copy per chunk,
// tens of microseconds apart; a worker that went to sleep would take about as
// long to wake, so workers spin on the generation counter for a while after
// each copy before they block.
#ifndef FMH_COPY_THREADS
#define FMH_COPY_THREADS 11
#endif
class CopyPool {
   public:
    CopyPool() {
        const unsigned hw = std::thread::hardware_concurrency();
        const int workers = (int)std::min((unsigned)FMH_COPY_THREADS, hw > 1 ? hw - 1 : 0u);
        for (int i = 0; i < workers; i++) th_.emplace_back([this, i] { work(i + 1); });
    }
    ~CopyPool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
            gen_.fetch_add(1);
        }
        cv_.notify_all();
        for (auto& t : th_) t.join();
    }
    void copy(void* dst, const void* src, size_t n) {
        if (th_.empty() || n < (1u << 20)) {
            memcpy(dst, src, n);
            return;
        }
        run(dst, src, n);
    }
    // Writes a zero byte into every 4 KiB page of [dst, dst + n), so that a
    // fresh buffer's page faults are taken here rather than during a later
    // copy into it.
    void touch(void* dst, size_t n) {
        if (th_.empty() || n < (1u << 20)) {
            touch_range((char*)dst, n);
            return;
        }
        run(dst, nullptr, n);
    }

   private:
    static void touch_range(char* p, size_t n) {
        for (size_t o = 0; o < n; o += 4096) ((volatile char*)p)[o] = 0;
    }
    void run(void* dst, const void* src, size_t n) {
        std::lock_guard<std::mutex> one(call_mu_);  // one split copy at a time
        dst_ = (char*)dst;
        src_ = (const char*)src;
        n_ = n;
        parts_ = (int)th_.size() + 1;
        pending_.store((int)th_.size());
        {
            std::lock_guard<std::mutex> lk(mu_);  // orders with a worker about to sleep
            gen_.fetch_add(1);
        }
        cv_.notify_all();
        part(0);
        while (pending_.load() != 0) cpu_relax();
    }

    static void cpu_relax() {
#if defined(__x86_64__)
        __builtin_ia32_pause();
#endif
    }
    void part(int id) {
        const size_t per = ((n_ + parts_ - 1) / parts_ + 63) & ~(size_t)63;  // parts_ x per >= n_
        const size_t b = std::min(n_, per * (size_t)id), e = std::min(n_, b + per);
        if (e <= b) return;
        if (src_)
            memcpy(dst_ + b, src_ + b, e - b);
        else
            touch_range(dst_ + b, e - b);
    }
    void work(int id) {
        unsigned long long seen = 0;
        while (true) {
            // spin about a millisecond for the next copy, then sleep
            for (int i = 0; i < 20000 && gen_.load() == seen; i++) cpu_relax();
            if (gen_.load() == seen) {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return gen_.load() != seen; });
            }
            seen = gen_.load();
            if (stop_) return;
            part(id);
            pending_.fetch_sub(1);
        }
    }
    std::vector<std::thread> th_;
    std::mutex call_mu_, mu_;
    std::condition_variable cv_;
    char* dst_ = nullptr;
    const char* src_ = nullptr;
    size_t n_ = 0;
    int parts_ = 1;
    std::atomic<int> pending_{0};
    std::atomic<unsigned long long> gen_{0};
    bool stop_ = false;
};

CopyPool& copy_pool() {
    static CopyPool pool;
    return pool;
}

// Host-pointer calls move data through a ring of kRing pinned buffers of
// kPinChunk bytes; event i marks the last transfer that used buffer i.
#ifndef FMH_PIN_CHUNK
#define FMH_PIN_CHUNK (8u << 20)  // the CPU emulation builds with a small chunk to reach the staged path
#endif
constexpr size_t kPinChunk = FMH_PIN_CHUNK;
#ifndef FMH_OVERLAP_BYTES
#define FMH_OVERLAP_BYTES (4u << 20)  // the CPU emulation builds with a small value to reach both paths
#endif
constexpr size_t kOverlapBytes = FMH_OVERLAP_BYTES;
#ifndef FMH_PIN_RING
#define FMH_PIN_RING 3
#endif
constexpr int kRing = FMH_PIN_RING;  // DMA from pinned memory: 8 MiB chunks run at 54 GB/s, 4 MiB at 51, 2 MiB at 45

// Device scratch of fmh_mul_device_noalloc: the workspace (residues of both
// operands for the three primes, per-limb carry maps, per-chunk maps and
// carry-ins) of one product of at most 2^kNoAllocLogN points, laid out as
// job_bind lays out any workspace.  Every GPU the library runs on holds its
// own copy until the process ends.  Host code never uses the array's name as
// an address: fmh_init looks up each device's copy with hipGetSymbolAddress
// and keeps it in DevCtx::scratch.
constexpr int kNoAllocLogN = 20;
constexpr size_t align256c(size_t x) { return (x + 255) & ~(size_t)255; }
constexpr size_t kNoAllocChunks = ((1ull << kNoAllocLogN) + 1 + kCarryChunk - 1) / kCarryChunk;
constexpr size_t kNoAllocBytes = align256c((size_t)6 * sizeof(u64) << kNoAllocLogN) +
                                 align256c(kNoAllocChunks * kCarryChunk) + 2 * align256c(kNoAllocChunks * sizeof(u32));
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];
static_assert((1ull << kNoAllocLogN) == FMH_NOALLOC_MAX_NTT_LEN, "fmh.h states the limit");

// What one host-pointer call uses on its own: a stream, device copies of
// the operands and the output, and the pinned ring.  Calls take a free slot
// of their device (or make one), so concurrent callers run their transfers
// and kernels side by side instead of one after the other.
// Transfers run on their own stream, so the second operand's upload overlaps
// the first operand's forward column pass on the compute stream.
struct HostSlot {
    hipStream_t stream = nullptr;  // kernels
    hipStream_t xfer = nullptr;    // uploads and downloads
    hipEvent_t up[2] = {};         // operand i is on the device
    hipEvent_t done = nullptr;     // the product is in stage_out
    Buffer stage_a, stage_b, stage_out;
    Buffer stage_spec;             // spectrum of the first operand (two-stream calls)
    void* pin[kRing] = {};
    hipEvent_t pin_ev[kRing] = {};
};

struct DevCtx {
    bool ready = false;
    hipStream_t stream = nullptr;  // private stream of set-up work (twiddles, batch descriptors)
    void* scratch = nullptr;       // this device's copy of `scratch` (set by fmh_init)
    Twiddles tw;
    std::map<hipStream_t, Buffer> ws;
    std::vector<HostSlot*> free_slots;  // host-pointer call slots not in use
    // compute streams of the slots in use: a call enqueues its job in two
    // locked steps, and fmh_trim must not free the workspace in between
    std::set<hipStream_t> slot_streams_busy;
};

constexpr int kMaxDevices = 64;
constexpr int kMinTwiddleLog = 20;

std::mutex g_mu;
DevCtx g_ctx[kMaxDevices];
thread_local std::string t_err;

int set_err(const char* what, hipError_t e) {
    t_err = std::string(what) + ": " + hipGetErrorString(e);
    return (int)e == 0 ? -1 : (int)e;
}
int set_err(const char* what) {
    t_err = what;
    return -1;
}

int ceil_log2(u64 x) {
    int l = 0;
    while ((1ull << l) < x) l++;
    return l;
}

// Callers hold g_mu.
int ctx_get(DevCtx** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return set_err("hipGetDevice", e);
    if (dev < 0 || dev >= kMaxDevices) return set_err("device index out of range");
    DevCtx& c = g_ctx[dev];
    if (!c.ready) {
        e = hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking);
        if (e != hipSuccess) return set_err("hipStreamCreate", e);
        c.ready = true;
    }
    *out = &c;
    return 0;
}

int ensure_buffer(Buffer& b, size_t bytes, hipStream_t s) {
    if (b.bytes >= bytes) return 0;
    hipError_t e;
    if (b.ptr) {
        e = hipStreamSynchronize(s);
        if (e != hipSuccess) return set_err("hipStreamSynchronize", e);
        e = hipFree(b.ptr);
        if (e != hipSuccess) return set_err("hipFree", e);
        b.ptr = nullptr;
        b.bytes = 0;
    }
    size_t want = bytes + bytes / 4;
    e = hipMalloc(&b.ptr, want);
    if (e != hipSuccess) {
        b.ptr = nullptr;
        return set_err("hipMalloc", e);
    }
    b.bytes = want;
    return 0;
}

int ensure_twiddles(DevCtx& c, int log_n) {
    if (c.tw.log_t >= log_n) return 0;
    hipError_t e;
    int lt = log_n < kMinTwiddleLog ? kMinTwiddleLog : log_n;
    if (lt > kMaxLogN) lt = kMaxLogN;
    e = build_twiddles(c.tw, lt, c.stream);
    if (e != hipSuccess) return set_err("build_twiddles", e);
    return 0;
}

// Pinned host staging of a batch's descriptors on one stream; `done` is
// recorded after the copy that reads it.
struct Staging {
    void* host = nullptr;
    size_t bytes = 0;
    hipEvent_t done = nullptr;
};
std::map<hipStream_t, Staging> g_stage;

// Builds a job of products for stream s: twiddles, the stream's workspace,
// descriptors uploaded; callers hold g_mu.
int prepare_job(DevCtx& c, Job& job, const ProductIn* in, size_t count, hipStream_t s) {
    if (!products_valid(in, count)) return set_err("NULL device pointer for a non-empty operand or output");
    const int rc = job_build(job, in, count);
    if (rc == -2) return set_err("product longer than the spectrum's transform length");
    if (rc) return set_err("product longer than fmh_max_ntt_len()");
    if (ensure_twiddles(c, job.max_log_n)) return -1;
    Buffer& w = c.ws[s];
    if (ensure_buffer(w, job_ws_bytes(job), s)) return -1;
    job_bind(job, w.ptr);
    hipError_t e;
    const size_t nd = job_desc_bytes(job);
    if (nd) {
        // descriptors go through pinned memory so the copy can be asynchronous
        Staging& st = g_stage[s];
        if (st.done) {
            e = hipEventSynchronize(st.done);
            if (e != hipSuccess) return set_err("hipEventSynchronize", e);
        } else {
            e = hipEventCreateWithFlags(&st.done, hipEventDisableTiming);
            if (e != hipSuccess) return set_err("hipEventCreate", e);
        }
        if (st.bytes < nd) {
            if (st.host) (void)hipHostFree(st.host);
            st.host = nullptr;
            st.bytes = 0;
            e = hipHostMalloc(&st.host, nd + nd / 4, hipHostMallocDefault);
            if (e != hipSuccess) return set_err("hipHostMalloc", e);
            st.bytes = nd + nd / 4;
        }
        job_pack(job, st.host);
        e = hipMemcpyAsync(job.d_desc, st.host, nd, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipEventRecord(st.done, s);
        if (e != hipSuccess) return set_err("descriptor upload", e);
    }
    return 0;
}

// Enqueue a set of products on stream s; callers hold g_mu.
int run_job(DevCtx& c, const ProductIn* in, size_t count, hipStream_t s) {
    Job job;
    if (prepare_job(c, job, in, count, s)) return -1;
    const hipError_t e = job_launch(job, c.tw, s);
    if (e != hipSuccess) return set_err("kernel launch", e);
    return 0;
}

int run_product(DevCtx& c, const u64* a, u64 na, const u64* b, u64 nb, bool square, u64* out,
                hipStream_t s) {
    const ProductIn in{a, na, square ? nullptr : b, square ? na : nb, out};
    return run_job(c, &in, 1, s);
}

int ensure_pinned(HostSlot& h) {
    for (int i = 0; i < kRing; i++) {
        if (h.pin[i]) continue;
        hipError_t e = hipHostMalloc(&h.pin[i], kPinChunk, hipHostMallocDefault);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&h.pin_ev[i], hipEventDisableTiming);
        if (e != hipSuccess) return set_err("pinned staging", e);
    }
    return 0;
}

// One transfer of a staged copy: bytes between host memory and device memory.
struct Segment {
    void* dev;
    void* host;
    size_t bytes;
};

// Chunks of a staged transfer.  Neither end of a transfer overlaps
// anything: the first H2D chunk's host copy runs before any DMA, the last D2H
// chunk's host copy after the last one.  So an upload starts with a chunk of
// at most kPinHead bytes and a download ends with one (after a few that
// shrink towards it); the rest of a transfer is cut into equal chunks of at
// most kPinChunk bytes (fewer DMAs: each costs a few microseconds of its
// own).
#ifndef FMH_PIN_HEAD
#define FMH_PIN_HEAD (FMH_PIN_CHUNK / 4)
#endif
constexpr size_t kPinHead = FMH_PIN_HEAD;

// (offset, bytes) of equal chunks of at most kPinChunk covering [off, off + n).
void equal_chunks(size_t off, size_t n, std::vector<std::pair<size_t, size_t>>& c) {
    if (!n) return;
    const size_t k = (n + kPinChunk - 1) / kPinChunk;
    const size_t per = ((n + k - 1) / k + 63) & ~(size_t)63;
    for (size_t o = 0; o < n; o += per) c.push_back({off + o, std::min(per, n - o)});
}

std::vector<std::pair<size_t, size_t>> h2d_chunks(size_t bytes, bool first) {
    std::vector<std::pair<size_t, size_t>> c;
    const size_t head = first ? std::min(bytes, kPinHead) : 0;
    if (head) c.push_back({0, head});
    equal_chunks(head, bytes - head, c);
    return c;
}

// Downloads of at least FMH_D2H_RAMP_MIN bytes end in chunks that shrink by
// 2.5x down to kPinHead, so that each one's host copy finishes while the
// next one's DMA runs and the copy after the last DMA is short (16 MiB: 4.5,
// 4.5, 5 and 2 MiB); shorter ones end in one chunk of kPinHead (4 MiB: 2 +
// 2).  With a 1 MiB kPinHead and the ramp for every size, 2^18 x 2^18 calls
// took 348 us against 317 us, and 2^20 x 2^20 calls 933 against 931 us.
#ifndef FMH_D2H_RAMP_MIN
#define FMH_D2H_RAMP_MIN (16u << 20)
#endif
std::vector<std::pair<size_t, size_t>> d2h_chunks(size_t bytes) {
    std::vector<size_t> tails;  // sizes of the last chunks, last first
    size_t used = 0, t = kPinHead;
    while (FMH_D2H_RAMP_MIN && bytes >= (size_t)FMH_D2H_RAMP_MIN && tails.size() < 3 && t < kPinChunk &&
           used + t < bytes) {
        tails.push_back(t);
        used += t;
        t = std::min(kPinChunk, (t * 5 / 2 + 63) & ~(size_t)63);
    }
    if (tails.empty() && bytes) {
        tails.push_back(std::min(bytes, kPinHead));
        used = tails[0];
    }
    std::vector<std::pair<size_t, size_t>> c;
    equal_chunks(0, bytes - used, c);
    size_t off = bytes - used;
    for (auto it = tails.rbegin(); it != tails.rend(); ++it) {
        c.push_back({off, *it});
        off += *it;
    }
    return c;
}

// Pageable host -> device through the pinned ring (a pageable hipMemcpy
// costs hundreds of microseconds even for small buffers), every segment in
// one pipeline: the host threads fill buffer k % kRing while the chunks
// before it move over DMA, so host copies and DMA overlap across segments.
// Runs on stream xs; k counts the chunks of the transfer (it continues
// across calls that upload parts of one transfer); mark (nullable) is
// recorded after the last DMA.
int h2d_staged(HostSlot& h, hipStream_t xs, const Segment* seg, int nseg, size_t& k, hipEvent_t mark) {
    if (ensure_pinned(h)) return -1;
    for (int g = 0; g < nseg; g++) {
        for (const auto& ch : h2d_chunks(seg[g].bytes, k == 0)) {
            const size_t off = ch.first, n = ch.second;
            const int i = (int)(k++ % kRing);
            hipError_t e = hipEventSynchronize(h.pin_ev[i]);  // the DMA that last read buffer i
            if (e != hipSuccess) return set_err("hipEventSynchronize", e);
            copy_pool().copy(h.pin[i], (const char*)seg[g].host + off, n);
            e = hipMemcpyAsync((char*)seg[g].dev + off, h.pin[i], n, hipMemcpyHostToDevice, xs);
            if (e == hipSuccess) e = hipEventRecord(h.pin_ev[i], xs);
            if (e != hipSuccess) return set_err("hipMemcpyAsync H2D", e);
        }
    }
    if (mark) {
        const hipError_t e = hipEventRecord(mark, xs);
        if (e != hipSuccess) return set_err("hipEventRecord", e);
    }
    return 0;
}

// Device -> pageable host: up to kRing chunks are in flight over DMA while
// the host threads copy the oldest one out of its buffer; each buffer's next
// DMA is enqueued as soon as it has been emptied.  Once the first DMAs are
// enqueued (they wait for the device's work on stream xs), the host threads
// write one byte into every page of dst, so that a freshly allocated output
// takes its page faults while the device computes and the first chunks move,
// not during the copies out (dst may alias the call's inputs: they have been
// read).  Returns after the last byte.
int d2h_staged(HostSlot& h, hipStream_t xs, void* dst, const void* src, size_t bytes) {
    if (ensure_pinned(h)) return -1;
    const std::vector<std::pair<size_t, size_t>> ch = d2h_chunks(bytes);
    const size_t chunks = ch.size();
    auto issue = [&](size_t k) {
        hipError_t r = hipMemcpyAsync(h.pin[k % kRing], (const char*)src + ch[k].first, ch[k].second,
                                      hipMemcpyDeviceToHost, xs);
        if (r == hipSuccess) r = hipEventRecord(h.pin_ev[k % kRing], xs);
        return r;
    };
    hipError_t e = hipSuccess;
    for (size_t k = 0; e == hipSuccess && k < std::min(chunks, (size_t)kRing); k++) e = issue(k);
    if (e == hipSuccess) copy_pool().touch(dst, bytes);
    for (size_t k = 0; e == hipSuccess && k < chunks; k++) {
        e = hipEventSynchronize(h.pin_ev[k % kRing]);
        if (e != hipSuccess) break;
        copy_pool().copy((char*)dst + ch[k].first, h.pin[k % kRing], ch[k].second);
        if (k + kRing < chunks) e = issue(k + kRing);
    }
    return e == hipSuccess ? 0 : set_err("staged D2H", e);
}

[[noreturn]] void die(const char* fn) {
    fprintf(stderr, "%s: %s\n", fn, t_err.c_str());
    abort();
}

void host_product(const char* fn, const u64* in1, u64 len1, const u64* in2, u64 len2, bool square,
                  u64* out) {
    const u64 out_len = square ? 2 * len1 : len1 + len2;
    if (len1 == 0 || (!square && len2 == 0)) {
        if (out_len) memset(out, 0, out_len * sizeof(u64));
        return;
    }
    // a slot of this device, taken from the free list or made
    DevCtx* c = nullptr;
    HostSlot* h = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (ctx_get(&c)) die(fn);
        if (!c->free_slots.empty()) {
            h = c->free_slots.back();
            c->free_slots.pop_back();
            c->slot_streams_busy.insert(h->stream);
        }
    }
    if (!h) {
        h = new HostSlot();
        hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->xfer, hipStreamNonBlocking);
        for (int i = 0; e == hipSuccess && i < 2; i++) e = hipEventCreateWithFlags(&h->up[i], hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&h->done, hipEventDisableTiming);
        if (e != hipSuccess) {
            set_err("host-call slot", e);
            die(fn);
        }
        std::lock_guard<std::mutex> lk(g_mu);
        c->slot_streams_busy.insert(h->stream);
    }
    // the slot's own buffers: no lock while data moves
    if (ensure_buffer(h->stage_a, len1 * sizeof(u64), h->stream)) die(fn);
    if (!square && ensure_buffer(h->stage_b, len2 * sizeof(u64), h->stream)) die(fn);
    if (ensure_buffer(h->stage_out, out_len * sizeof(u64), h->stream)) die(fn);
    u64* da = (u64*)h->stage_a.ptr;
    u64* db = square ? da : (u64*)h->stage_b.ptr;
    u64* dout = (u64*)h->stage_out.ptr;
    // both operands are read before out is written, so out may alias them
    const Segment seg[2] = {{da, (void*)in1, len1 * sizeof(u64)}, {db, (void*)in2, square ? 0 : len2 * sizeof(u64)}};
    // A second operand of at least kOverlapBytes moves on the transfer stream
    // while the device works on the first operand on the compute stream.
    // Below that size the cross-stream waits cost more than they hide
    // (2^16 x 2^16: 159 -> 170 us), and the whole call runs on the compute
    // stream.
    const ProductIn in{da, len1, square ? nullptr : db, square ? len1 : len2, dout};
    const bool two_streams = !square && len2 * sizeof(u64) >= kOverlapBytes;
    size_t k = 0;
    auto fail = [&](const char* what, hipError_t e) {
        set_err(what, e);
        die(fn);
    };
    if (two_streams) {
        Job job;
        Twiddles tw;
        // One transform with a column pass (not a product split into pieces):
        // the first operand's whole forward transform (its resident
        // spectrum) runs while the second operand uploads, and the rest is a
        // product with that spectrum, which transforms only the second
        // operand (2^20 x 2^20 limbs: the spectrum takes 88 us, under the
        // upload; the product with it 163 us, against 243 us for the whole
        // product).  Other products start once both operands are on the
        // device.
        bool spec;
        int log_n = 0;
        u64* d_spec = nullptr;
        if (h2d_staged(*h, h->xfer, &seg[0], 1, k, h->up[0])) die(fn);
        {
            std::lock_guard<std::mutex> lk(g_mu);  // the device state: twiddles, per-stream workspaces
            if (job_build(job, &in, 1)) {
                set_err("product longer than fmh_max_ntt_len()");
                die(fn);
            }
            spec = job_one_multiply_with_col(job);
            if (spec) {
                log_n = (int)job.items[0].log_n;
                if (ensure_twiddles(*c, log_n)) die(fn);
                if (ensure_buffer(h->stage_spec, (3 * sizeof(u64)) << log_n, h->stream)) die(fn);
                d_spec = (u64*)h->stage_spec.ptr;
            } else if (prepare_job(*c, job, &in, 1, h->stream)) {
                die(fn);
            }
            tw = c->tw;
            hipError_t e = hipStreamWaitEvent(h->stream, h->up[0], 0);
            if (e == hipSuccess && spec) e = enqueue_resident_spectrum(da, len1, log_n, d_spec, tw, h->stream);
            if (e != hipSuccess) fail("kernel launch", e);
        }
        if (h2d_staged(*h, h->xfer, &seg[1], 1, k, h->up[1])) die(fn);
        std::lock_guard<std::mutex> lk(g_mu);
        if (spec) {
            const ProductIn ins{db, len2, nullptr, len1, dout, d_spec, log_n};
            if (prepare_job(*c, job, &ins, 1, h->stream)) die(fn);
        }
        hipError_t e = hipStreamWaitEvent(h->stream, h->up[1], 0);
        if (e == hipSuccess) e = job_launch(job, tw, h->stream);
        if (e == hipSuccess) e = hipEventRecord(h->done, h->stream);
        if (e == hipSuccess) e = hipStreamWaitEvent(h->xfer, h->done, 0);
        if (e != hipSuccess) fail("kernel launch", e);
    } else {
        if (h2d_staged(*h, h->stream, seg, 2, k, nullptr)) die(fn);
        std::lock_guard<std::mutex> lk(g_mu);
        if (run_job(*c, &in, 1, h->stream)) die(fn);
    }
    const hipStream_t xs = two_streams ? h->xfer : h->stream;
    if (d2h_staged(*h, xs, out, dout, out_len * sizeof(u64))) die(fn);
    std::lock_guard<std::mutex> lk(g_mu);
    c->slot_streams_busy.erase(h->stream);
    c->free_slots.push_back(h);
}

}  // namespace

namespace fmh {

int device_twiddles(int log_n, Twiddles* out, std::string* err) {
    std::lock_guard<std::mutex> lk(g_mu);
    DevCtx* c = nullptr;
    if (ctx_get(&c) || ensure_twiddles(*c, log_n)) {
        if (err) *err = t_err;
        return -1;
    }
    *out = c->tw;
    return 0;
}

}  // namespace fmh

extern "C" {

void abs_mul64(const u64* in1, u64 len1, const u64* in2, u64 len2, u64* out) {
    host_product("abs_mul64", in1, len1, in2, len2, false, out);
}

void abs_sqr64(const u64* in1, u64 len1, u64* out) {
    host_product("abs_sqr64", in1, len1, in1, len1, true, out);
}

int fmh_mul_device(const u64* d_in1, u64 len1, const u64* d_in2, u64 len2, u64* d_out,
                   fmh_stream_t stream) {
    if ((len1 && !d_in1) || (len2 && !d_in2)) return set_err("NULL device pointer for a non-empty operand");
    if (len1 + len2 && !d_out) return set_err("NULL device pointer for a non-empty output");
    if (len1 == 0 || len2 == 0) {
        // the all-zero product of len1 + len2 limbs; a null operand here must
        // not be read as the square of the other one (the batch convention)
        hipError_t e = enqueue_zero(d_out, len1 + len2, (hipStream_t)stream);
        return e == hipSuccess ? 0 : set_err("kernel launch", e);
    }
    std::lock_guard<std::mutex> lk(g_mu);
    DevCtx* c = nullptr;
    if (ctx_get(&c)) return -1;
    return run_product(*c, d_in1, len1, d_in2, len2, false, d_out, (hipStream_t)stream);
}

int fmh_sqr_device(const u64* d_in1, u64 len1, u64* d_out, fmh_stream_t stream) {
    std::lock_guard<std::mutex> lk(g_mu);
    DevCtx* c = nullptr;
    if (ctx_get(&c)) return -1;
    return run_product(*c, d_in1, len1, d_in1, len1, true, d_out, (hipStream_t)stream);
}

int fmh_forward_spectrum_device(int prime, const u64* d_in, u64 len, u64 ntt_len, u64* d_out,
                                fmh_stream_t stream) {
    if (prime < 0 || prime > 2) return set_err("prime index must be 0, 1 or 2");
    if (ntt_len == 0 || (ntt_len & (ntt_len - 1)) || ntt_len < len)
        return set_err("ntt_len must be a power of two >= len");
    if ((len && !d_in) || !d_out) return set_err("NULL device pointer");
    const int log_n = ceil_log2(ntt_len);
    if (log_n > kMaxLogN) return set_err("ntt_len above fmh_max_ntt_len()");
    std::lock_guard<std::mutex> lk(g_mu);
    DevCtx* c = nullptr;
    if (ctx_get(&c)) return -1;
    if (ensure_twiddles(*c, log_n)) return -1;
    hipStream_t s = (hipStream_t)stream;
    Buffer& w = c->ws[s];
    if (ensure_buffer(w, spectrum_ws_bytes(log_n), s)) return -1;
    hipError_t e = enqueue_spectrum(prime, d_in, len, log_n, d_out, w.ptr, c->tw, s);
    if (e != hipSuccess) return set_err("kernel launch", e);
    return 0;
}

u64 fmh_max_ntt_len(void) { return 1ull << kMaxLogN; }

int fmh_mul_batch_device(const fmh_product* items, u64 count, fmh_stream_t stream) {
    if (count && !items) return set_err("items is NULL");
    std::vector<ProductIn> in(count);
    for (u64 i = 0; i < count; i++)
        in[i] = ProductIn{items[i].in1, items[i].len1, items[i].in2, items[i].len2, items[i].out};
    std::lock_guard<std::mutex> lk(g_mu);
    DevCtx* c = nullptr;
    if (ctx_get(&c)) return -1;
    return run_job(*c, in.data(), in.size(), (hipStream_t)stream);
}

}  // extern "C"

struct fmh_spectrum {
    u64* words = nullptr;  // 3 x 2^log_n
    u64 len = 0;           // limbs of the transformed operand
    int log_n = 0;
    int device = 0;
};

extern "C" {

int fmh_spectrum_create(const u64* d_in, u64 len, u64 ntt_len, fmh_stream_t stream, fmh_spectrum** out) {
    if (!out) return set_err("NULL argument");
    *out = nullptr;
    if (len == 0) return set_err("an empty operand has no spectrum");
    if (!d_in) return set_err("NULL operand");
    if (ntt_len == 0) ntt_len = 1ull << ceil_log2(2 * len - 1);
    if ((ntt_len & (ntt_len - 1)) || ntt_len < len) return set_err("ntt_len must be a power of two >= len");
    const int log_n = ceil_log2(ntt_len);
    if (log_n > kMaxLogN) return set_err("ntt_len above fmh_max_ntt_len()");
    std::lock_guard<std::mutex> lk(g_mu);
    DevCtx* c = nullptr;
    if (ctx_get(&c)) return -1;
    if (ensure_twiddles(*c, log_n)) return -1;
    fmh_spectrum* sp = new fmh_spectrum();
    sp->len = len;
    sp->log_n = log_n;
    (void)hipGetDevice(&sp->device);
    hipError_t e = hipMalloc(&sp->words, (size_t)3 * sizeof(u64) << log_n);
    if (e != hipSuccess) {
        delete sp;
        return set_err("hipMalloc", e);
    }
    e = enqueue_resident_spectrum(d_in, len, log_n, sp->words, c->tw, (hipStream_t)stream);
    if (e != hipSuccess) {
        (void)hipFree(sp->words);
        delete sp;
        return set_err("kernel launch", e);
    }
    *out = sp;
    return 0;
}

u64 fmh_spectrum_ntt_len(const fmh_spectrum* sp) { return sp ? 1ull << sp->log_n : 0; }

int fmh_spectrum_describe(const fmh_spectrum* sp, const u64** words, u64* len, int* log_n) {
    if (!sp) return set_err("NULL spectrum");
    *words = sp->words;
    *len = sp->len;
    *log_n = sp->log_n;
    return 0;
}

int fmh_mul_spectrum_device(const fmh_spectrum* sp, const u64* d_in2, u64 len2, u64* d_out,
                            fmh_stream_t stream) {
    if (!sp) return set_err("NULL spectrum");
    if (len2 && !d_in2) return set_err("NULL operand");
    if (!d_out) return set_err("NULL device pointer for a non-empty output");  // sp->len >= 1
    if (len2 && sp->len + len2 - 1 > (1ull << sp->log_n))
        return set_err("product longer than the spectrum's transform length");
    ProductIn in{d_in2, len2, nullptr, sp->len, d_out, sp->words, sp->log_n};
    std::lock_guard<std::mutex> lk(g_mu);
    DevCtx* c = nullptr;
    if (ctx_get(&c)) return -1;
    if (len2 == 0) {
        hipError_t e = enqueue_zero(d_out, sp->len, (hipStream_t)stream);
        return e == hipSuccess ? 0 : set_err("kernel launch", e);
    }
    return run_job(*c, &in, 1, (hipStream_t)stream);
}

int fmh_sqr_spectrum_device(const fmh_spectrum* sp, u64* d_out, fmh_stream_t stream) {
    if (!sp) return set_err("NULL spectrum");
    if (2 * sp->len - 1 > (1ull << sp->log_n)) return set_err("square longer than the spectrum's transform length");
    ProductIn in{nullptr, 0, nullptr, sp->len, d_out, sp->words, sp->log_n};
    std::lock_guard<std::mutex> lk(g_mu);
    DevCtx* c = nullptr;
    if (ctx_get(&c)) return -1;
    return run_job(*c, &in, 1, (hipStream_t)stream);
}

int fmh_spectrum_destroy(fmh_spectrum* sp) {
    if (!sp) return 0;
    hipError_t e = hipDeviceSynchronize();
    (void)hipFree(sp->words);
    delete sp;
    return e == hipSuccess ? 0 : set_err("hipDeviceSynchronize", e);
}

}  // extern "C"

struct fmh_batch {
    Job job;
    void* ws = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

extern "C" {

int fmh_batch_create(const fmh_product* items, u64 count, fmh_batch** out) {
    if (!out || (count && !items)) return set_err("NULL argument");
    *out = nullptr;
    std::vector<ProductIn> in(count);
    for (u64 i = 0; i < count; i++)
        in[i] = ProductIn{items[i].in1, items[i].len1, items[i].in2, items[i].len2, items[i].out};
    std::lock_guard<std::mutex> lk(g_mu);
    DevCtx* c = nullptr;
    if (ctx_get(&c)) return -1;
    fmh_batch* b = new fmh_batch();
    auto fail = [&](const char* what, hipError_t e) {
        if (b->exec) (void)hipGraphExecDestroy(b->exec);
        if (b->graph) (void)hipGraphDestroy(b->graph);
        if (b->ws) (void)hipFree(b->ws);
        delete b;
        return e == hipSuccess ? set_err(what) : set_err(what, e);
    };
    if (!products_valid(in.data(), in.size()))
        return fail("NULL device pointer for a non-empty operand or output", hipSuccess);
    if (job_build(b->job, in.data(), in.size())) return fail("product longer than fmh_max_ntt_len()", hipSuccess);
    if (ensure_twiddles(*c, b->job.max_log_n)) {
        delete b;
        return -1;
    }
    prepare_kernels();
    hipError_t e = hipMalloc(&b->ws, std::max<size_t>(job_ws_bytes(b->job), 256));
    if (e != hipSuccess) return fail("hipMalloc", e);
    job_bind(b->job, b->ws);
    e = job_upload(b->job, c->stream);
    if (e != hipSuccess) return fail("descriptor upload", e);
    // capture the kernels only (no copies, no allocation, no per-kernel events)
    const bool timing = timing_enabled();
    timing_enable(false);
    hipStream_t cs = nullptr;
    e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        hipError_t le = job_launch(b->job, c->tw, cs);
        e = hipStreamEndCapture(cs, &b->graph);
        if (le != hipSuccess) e = le;
    }
    timing_enable(timing);
    if (cs) (void)hipStreamDestroy(cs);
    if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) return fail("graph capture", e);
    *out = b;
    return 0;
}

int fmh_batch_run(fmh_batch* b, fmh_stream_t stream) {
    if (!b) return set_err("NULL batch");
    hipError_t e = hipGraphLaunch(b->exec, (hipStream_t)stream);
    return e == hipSuccess ? 0 : set_err("hipGraphLaunch", e);
}

int fmh_batch_destroy(fmh_batch* b) {
    if (!b) return 0;
    hipError_t e = hipDeviceSynchronize();
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->ws) (void)hipFree(b->ws);
    delete b;
    return e == hipSuccess ? 0 : set_err("hipDeviceSynchronize", e);
}

int fmh_copy_device(u64* d_dst, const u64* d_src, u64 n, fmh_stream_t stream) {
    hipError_t e = enqueue_copy(d_dst, d_src, n, (hipStream_t)stream);
    return e == hipSuccess ? 0 : set_err("kernel launch", e);
}

int fmh_init(int device) {
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return set_err("hipSetDevice", e);
    std::lock_guard<std::mutex> lk(g_mu);
    DevCtx* c = nullptr;
    if (ctx_get(&c)) return -1;
    if (ensure_twiddles(*c, kMinTwiddleLog)) return -1;
    prepare_kernels();
    if (!c->scratch) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return set_err("hipGetSymbolAddress(scratch)", e);
        c->scratch = p;
    }
    return 0;
}

int fmh_mul_device_noalloc(const u64* d_in1, u64 len1, const u64* d_in2, u64 len2, u64* d_out,
                           fmh_stream_t stream) {
    if ((len1 && !d_in1) || (len2 && !d_in2)) return set_err("NULL device pointer for a non-empty operand");
    if (len1 + len2 && !d_out) return set_err("NULL device pointer for a non-empty output");
    if (len1 && len2 && len1 + len2 - 1 > (1ull << kNoAllocLogN))
        return set_err("product longer than FMH_NOALLOC_MAX_NTT_LEN coefficients");
    Twiddles tw;
    void* ws = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        int dev = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e != hipSuccess) return set_err("hipGetDevice", e);
        if (dev < 0 || dev >= kMaxDevices || !g_ctx[dev].scratch)
            return set_err("fmh_mul_device_noalloc: call fmh_init on this device first");
        tw = g_ctx[dev].tw;  // fmh_init built tables of at least 2^kNoAllocLogN points
        ws = g_ctx[dev].scratch;
    }
    hipStream_t s = (hipStream_t)stream;
    if (len1 == 0 || len2 == 0) {
        hipError_t e = enqueue_zero(d_out, len1 + len2, s);
        return e == hipSuccess ? 0 : set_err("kernel launch", e);
    }
    const ProductIn in{d_in1, len1, d_in2, len2, d_out};
    Job job;
    if (job_build(job, &in, 1, /*allow_split=*/false)) return set_err("product longer than fmh_max_ntt_len()");
    // one product, one transform item: the descriptors travel as kernel
    // arguments, and the whole workspace is the scratch array
    if (job_desc_bytes(job) != 0 || job_ws_bytes(job) > kNoAllocBytes || job.max_log_n > tw.log_t)
        return set_err("fmh_mul_device_noalloc: product does not fit the scratch array");
    job_bind(job, ws);
    hipError_t e = job_launch(job, tw, s);
    return e == hipSuccess ? 0 : set_err("kernel launch", e);
}


int fmh_dev_alloc(void** ptr, u64 bytes) {
    hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
    return e == hipSuccess ? 0 : set_err("hipMalloc", e);
}

int fmh_dev_free(void* ptr) {
    hipError_t e = hipFree(ptr);
    return e == hipSuccess ? 0 : set_err("hipFree", e);
}

int fmh_memcpy_h2d(void* dst, const void* src, u64 bytes, fmh_stream_t stream) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream);
    if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
    return e == hipSuccess ? 0 : set_err("hipMemcpy H2D", e);
}

int fmh_memcpy_d2h(void* dst, const void* src, u64 bytes, fmh_stream_t stream) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream);
    if (e == hipSuccess) e = hipStreamSynchronize((hipStream_t)stream);
    return e == hipSuccess ? 0 : set_err("hipMemcpy D2H", e);
}

int fmh_stream_create(fmh_stream_t* stream) {
    hipStream_t s = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e != hipSuccess) return set_err("hipStreamCreate", e);
    *stream = (fmh_stream_t)s;
    return 0;
}

int fmh_stream_destroy(fmh_stream_t stream) {
    hipStream_t s = (hipStream_t)stream;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        int dev = 0;
        if (hipGetDevice(&dev) == hipSuccess && dev >= 0 && dev < kMaxDevices) {
            auto it = g_ctx[dev].ws.find(s);
            if (it != g_ctx[dev].ws.end()) {
                (void)hipStreamSynchronize(s);
                (void)hipFree(it->second.ptr);
                g_ctx[dev].ws.erase(it);
            }
        }
    }
    hipError_t e = hipStreamDestroy(s);
    return e == hipSuccess ? 0 : set_err("hipStreamDestroy", e);
}

int fmh_trim(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return set_err("hipGetDevice", e);
    if (dev < 0 || dev >= kMaxDevices) return set_err("device index out of range");
    DevCtx& c = g_ctx[dev];
    e = hipDeviceSynchronize();
    if (e != hipSuccess) return set_err("hipDeviceSynchronize", e);
    // workspaces of host-pointer calls in progress stay (such a call may
    // have prepared its job on one before its second operand's upload and
    // launch the job after this returns)
    for (auto it = c.ws.begin(); it != c.ws.end();) {
        if (c.slot_streams_busy.count(it->first)) {
            ++it;
            continue;
        }
        if (it->second.ptr) (void)hipFree(it->second.ptr);
        it = c.ws.erase(it);
    }
    for (HostSlot* h : c.free_slots)  // slots in use keep their buffers
        for (Buffer* b : {&h->stage_a, &h->stage_b, &h->stage_out, &h->stage_spec}) {
            if (b->ptr) (void)hipFree(b->ptr);
            *b = Buffer();
        }
    return 0;
}

int fmh_stream_synchronize(fmh_stream_t stream) {
    hipError_t e = hipStreamSynchronize((hipStream_t)stream);
    return e == hipSuccess ? 0 : set_err("hipStreamSynchronize", e);
}

const char* fmh_last_error(void) { return t_err.c_str(); }

int fmh_set_tuning(int log_n, const char* spec, int batched) {
    return set_tuning(log_n, spec, batched != 0) ? set_err("malformed tuning spec or size class") : 0;
}

int fmh_get_tuning(int log_n, int batched, char* buf, u64 buflen) {
    if (log_n < 0 || log_n > kMaxLogN || !buf || !buflen) return set_err("bad argument");
    const Tuning t = tuning(log_n, batched != 0);
    snprintf(buf, buflen,
             "col_logc=%d,col_maxl=%d,row_maxl=%d,row_tile=%d,col_threads=%d,row_threads=%d,row_radix=%d,col_dma=%d",
             t.col_logc, t.col_maxl, t.row_maxl, t.row_tile, t.col_threads, t.row_threads, t.row_radix, t.col_dma);
    return 0;
}

int fmh_timing_enable(int on) {
    timing_enable(on != 0);
    return 0;
}

int fmh_timing_only(const char* kernel) {
    timing_only(kernel);
    return 0;
}

u64 fmh_timing_collect(char* buf, u64 buflen) {
    const std::string s = timing_collect();
    if (buf && buflen) {
        const size_t n = s.size() < buflen - 1 ? s.size() : (size_t)(buflen - 1);
        memcpy(buf, s.data(), n);
        buf[n] = 0;
    }
    return (u64)s.size() + 1;
}

}  // extern "C"
