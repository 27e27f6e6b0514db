// fmh_serve.hip -- request loop: continuous batching of multiply / square
// requests over a device memory pool (C-ABI in include/fmh.h, fmh_server_*).
//
// A request names its operands by device pointer (or a resident spectrum) and
// gets its output, and the transform workspace of its product, from the
// server's pool.  Its footprint is
//     output limbs * 8 + 3 primes * NTT length * 8 B * operands transformed
// and it is admitted only when that much pool memory is free.  A scheduler
// thread forms launch groups: the oldest waiting request plus the younger
// ones of its size class (transform length and kind), up to a group size and
// byte cap.  Each group runs as one set of batched launches (the job executor
// of fmh_ntt.hip) on one of several streams, so groups overlap on the device;
// as groups complete, their workspace returns to the pool and waiting
// requests are admitted -- requests that arrive while others run join the
// next group of their class without waiting for a batch boundary.
//
// Admission is first come, first served at the head: a request that does not
// fit blocks younger requests of other classes until memory frees up, so a
// large request cannot starve.  Outputs stay in the pool until the client
// releases them.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/fmh.h"
#include "fmh_internal.h"

namespace {

using namespace fmh;

thread_local std::string t_err;

int fail(const std::string& what) {
    t_err = what;
    return -1;
}

int64_t now_ns() {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

constexpr u64 kAlign = 256;
u64 round_up(u64 x) { return (x + kAlign - 1) & ~(kAlign - 1); }

// Best-fit sub-allocator over one device allocation; free blocks coalesce.
class Pool {
   public:
    int init(u64 bytes) {
        bytes = round_up(bytes);
        hipError_t e = hipMalloc(&base_, bytes);
        if (e != hipSuccess) {
            base_ = nullptr;
            return fail(std::string("pool hipMalloc: ") + hipGetErrorString(e));
        }
        cap_ = bytes;
        free_[0] = bytes;
        return 0;
    }
    ~Pool() {
        if (base_) (void)hipFree(base_);
    }
    bool alloc(u64 bytes, u64* off) {
        bytes = round_up(std::max<u64>(bytes, 1));
        auto best = free_.end();
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= bytes && (best == free_.end() || it->second < best->second)) best = it;
        if (best == free_.end()) return false;
        *off = best->first;
        const u64 rest = best->second - bytes;
        const u64 at = best->first + bytes;
        free_.erase(best);
        if (rest) free_[at] = rest;
        used_ += bytes;
        peak_ = std::max(peak_, used_);
        return true;
    }
    void release(u64 off, u64 bytes) {
        bytes = round_up(std::max<u64>(bytes, 1));
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto next = std::next(it);
        if (next != free_.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }
    char* at(u64 off) const { return (char*)base_ + off; }
    u64 capacity() const { return cap_; }
    u64 used() const { return used_; }
    u64 peak() const { return peak_; }

   private:
    void* base_ = nullptr;
    u64 cap_ = 0, used_ = 0, peak_ = 0;
    std::map<u64, u64> free_;  // offset -> bytes
};

enum ReqState { kPending = 0, kRunning = 1, kDone = 2, kFailed = 3 };

struct Request {
    u64 id;
    ProductIn in;     // out is filled at admission
    u64 out_len;      // limbs
    int log_n;
    u32 kind;
    u64 out_off = 0;  // pool offset of the output
    u64 ws_bytes;     // residue bytes of its product
    int64_t t_submit = 0, t_done = 0;
    int state = kPending;
    bool reported = false;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t start = nullptr;  // recorded before a group's first launch
    hipEvent_t done = nullptr;   // after its last
    void* staging = nullptr;  // pinned copy of the group's descriptors
    size_t staging_bytes = 0;
    bool busy = false;
    std::vector<Request*> reqs;
    u64 ws_off = 0, ws_bytes = 0;
    u32 klass = 0;  // (kind << 8) | log_n of the group
};

// Device time of the groups of one size class.
struct ClassTime {
    u64 groups = 0, requests = 0;
    double device_ms = 0;  // sum over groups of start -> done on their stream
};

}  // namespace

struct fmh_server {
    fmh_server_config cfg{};
    int device = 0;
    Pool pool;
    std::vector<Slot> slots;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::thread loop;
    bool stop = false;
    // waiting requests per size class (kind, log_n), each in arrival order
    std::map<u32, std::deque<Request*>> pending;
    u64 n_pending = 0;
    std::unordered_map<u64, Request*> reqs;
    std::vector<u64> completed_unreported;
    u64 next_id = 1;
    u64 submitted = 0, completed = 0, failed = 0, groups = 0, grouped = 0;
    int64_t t_first_submit = 0, t_last_done = 0;
    std::vector<int64_t> latencies;  // ns, completed requests
    Twiddles tw;
    std::string loop_error;
    // Device activity: each group's interval [start, done] in ms after `ref`
    // (an event recorded at creation), and its time per size class.
    hipEvent_t ref = nullptr;
    std::vector<std::pair<float, float>> busy;
    std::map<u32, ClassTime> class_time;
    // No group in flight and the oldest request does not fit: only releasing
    // outputs can free memory.  Waits return an error instead of blocking.
    bool stalled = false;

    // -- scheduler (called with mu held) --------------------------------
    void reap() {
        for (Slot& s : slots) {
            if (!s.busy) continue;
            const hipError_t q = hipEventQuery(s.done);
            if (q == hipErrorNotReady) continue;
            const int64_t t = now_ns();
            for (Request* r : s.reqs) {
                r->state = q == hipSuccess ? kDone : kFailed;
                r->t_done = t;
                if (q == hipSuccess) {
                    completed++;
                    latencies.push_back(t - r->t_submit);
                } else {
                    failed++;
                    pool.release(r->out_off, r->out_len * sizeof(u64));
                }
                completed_unreported.push_back(r->id);
            }
            float t0 = 0, t1 = 0;
            if (q == hipSuccess && hipEventElapsedTime(&t0, ref, s.start) == hipSuccess &&
                hipEventElapsedTime(&t1, ref, s.done) == hipSuccess) {
                busy.push_back({t0, t1});
                ClassTime& c = class_time[s.klass];
                c.groups++;
                c.requests += s.reqs.size();
                c.device_ms += t1 - t0;
            }
            t_last_done = t;
            if (q != hipSuccess) loop_error = std::string("launch group failed: ") + hipGetErrorString(q);
            pool.release(s.ws_off, s.ws_bytes);
            s.reqs.clear();
            s.busy = false;
            cv_done.notify_all();
        }
    }

    bool any_busy() const {
        for (const Slot& s : slots)
            if (s.busy) return true;
        return false;
    }

    // Releases the pool memory taken for `n` outputs of `group` and a workspace.
    void undo(const std::vector<Request*>& group, size_t n, u64 ws_off, u64 ws) {
        for (size_t i = 0; i < n; i++) pool.release(group[i]->out_off, group[i]->out_len * sizeof(u64));
        if (ws) pool.release(ws_off, ws);
    }

    // Forms and launches one group on slot `si`: the oldest waiting request
    // and younger ones of its class.  If the pool cannot hold the group, it
    // is halved until it fits; returns false if not even the oldest request
    // fits now.
    bool admit(int si) {
        if (!n_pending) return false;
        // the class of the oldest waiting request
        std::deque<Request*>* q = nullptr;
        for (auto& kv : pending)
            if (!kv.second.empty() && (!q || kv.second.front()->id < q->front()->id)) q = &kv.second;
        std::vector<Request*> cand;
        u64 res_bytes = 0;
        for (Request* r : *q) {
            if (cand.size() >= cfg.max_group) break;
            if (!cand.empty() && res_bytes + r->ws_bytes > cfg.max_group_bytes) break;
            cand.push_back(r);
            res_bytes += r->ws_bytes;
        }
        for (size_t n = cand.size(); n >= 1; n /= 2) {
            size_t got = 0;
            for (; got < n; got++) {
                Request* r = cand[got];
                if (!pool.alloc(r->out_len * sizeof(u64), &r->out_off)) break;
            }
            if (got < n) {
                undo(cand, got, 0, 0);
                continue;
            }
            std::vector<ProductIn> ins(n);
            for (size_t i = 0; i < n; i++) {
                ins[i] = cand[i]->in;
                ins[i].out = (u64*)pool.at(cand[i]->out_off);
            }
            Job job;
            if (job_build(job, ins.data(), n) != 0) {  // checked at submission
                undo(cand, n, 0, 0);
                loop_error = "job_build failed";
                return false;
            }
            const u64 ws = std::max<u64>(job_ws_bytes(job), 1);
            u64 ws_off = 0;
            if (!pool.alloc(ws, &ws_off)) {
                undo(cand, n, 0, 0);
                continue;
            }
            job_bind(job, pool.at(ws_off));
            cand.resize(n);
            q->erase(q->begin(), q->begin() + (long)n);  // cand is the front of q
            n_pending -= n;
            launch(si, job, cand, ws_off, ws);
            return true;
        }
        return false;
    }

    void launch(int si, const Job& job, const std::vector<Request*>& group, u64 ws_off, u64 ws) {
        Slot& s = slots[si];
        hipError_t e = hipSuccess;
        const size_t nd = job_desc_bytes(job);
        if (nd) {
            if (s.staging_bytes < nd) {
                if (s.staging) (void)hipHostFree(s.staging);
                s.staging = nullptr;
                s.staging_bytes = 0;
                e = hipHostMalloc(&s.staging, nd, hipHostMallocDefault);
                if (e == hipSuccess) s.staging_bytes = nd;
            }
            if (e == hipSuccess) {
                // the slot is idle, so its previous upload has completed
                job_pack(job, s.staging);
                e = hipMemcpyAsync(job.d_desc, s.staging, nd, hipMemcpyHostToDevice, s.stream);
            }
        }
        if (e == hipSuccess) e = hipEventRecord(s.start, s.stream);
        if (e == hipSuccess) e = job_launch(job, tw, s.stream);
        if (e == hipSuccess) e = hipEventRecord(s.done, s.stream);
        const int64_t t = now_ns();
        for (Request* r : group) r->state = kRunning;
        if (e != hipSuccess) {
            // the group did not launch: its requests fail
            loop_error = std::string("group launch: ") + hipGetErrorString(e);
            (void)hipStreamSynchronize(s.stream);
            undo(group, group.size(), ws_off, ws);
            for (Request* r : group) {
                r->state = kFailed;
                r->t_done = t;
                failed++;
                completed_unreported.push_back(r->id);
            }
            cv_done.notify_all();
            return;
        }
        s.busy = true;
        s.klass = (group[0]->kind << 8) | (u32)group[0]->log_n;
        s.reqs = group;
        s.ws_off = ws_off;
        s.ws_bytes = ws;
        groups++;
        grouped += group.size();
    }

    void run() {
        (void)hipSetDevice(device);
        std::unique_lock<std::mutex> lk(mu);
        while (true) {
            reap();
            bool launched = false;
            for (int si = 0; si < (int)slots.size() && n_pending; si++)
                if (!slots[si].busy && admit(si)) launched = true;
            if (stop && !n_pending && !any_busy()) break;
            if (launched) continue;
            if (n_pending && !any_busy() && !stalled) {
                stalled = true;
                cv_done.notify_all();
            }
            if (any_busy())
                cv_work.wait_for(lk, std::chrono::microseconds(20));
            else
                cv_work.wait_for(lk, std::chrono::milliseconds(5));
        }
    }
};

namespace {

const char* const kStalled =
    "the pool is full of unreleased outputs: release completed requests before waiting";


int submit(fmh_server* s, const ProductIn& in, u64* id) {
    if (!s || !id) return fail("NULL argument");
    Job probe;
    const int rc = job_build(probe, &in, 1);
    if (rc == -2) return fail("product longer than the spectrum's transform length");
    if (rc) return fail("product longer than fmh_max_ntt_len()");
    auto* r = new Request();
    r->in = in;
    if (probe.items.empty()) {
        delete r;
        return fail("requests need non-empty operands");
    }
    const ProductDesc& d = probe.crt[0];
    r->out_len = d.na + d.nb;
    r->log_n = (int)probe.items[0].log_n;  // size class: the transforms'
    r->kind = probe.items[0].kind;
    r->ws_bytes = job_ws_bytes(probe);
    // alone in a group it needs its output and the job's workspace
    if (round_up(r->out_len * sizeof(u64)) + round_up(job_ws_bytes(probe)) > s->pool.capacity()) {
        delete r;
        return fail("request footprint exceeds the pool");
    }
    Twiddles tw;
    std::string err;
    if (device_twiddles(r->log_n, &tw, &err)) {
        delete r;
        return fail(err);
    }
    std::lock_guard<std::mutex> lk(s->mu);
    if (tw.log_t > s->tw.log_t) s->tw = tw;
    r->id = s->next_id++;
    r->t_submit = now_ns();
    if (!s->submitted) s->t_first_submit = r->t_submit;
    s->submitted++;
    s->reqs[r->id] = r;
    s->pending[(r->kind << 8) | (u32)r->log_n].push_back(r);
    s->n_pending++;
    s->stalled = false;
    *id = r->id;
    s->cv_work.notify_one();
    return 0;
}

}  // namespace

extern "C" {

int fmh_server_create(const fmh_server_config* cfg, fmh_server** out) {
    if (!out) return fail("NULL argument");
    *out = nullptr;
    auto* s = new fmh_server();
    if (cfg) s->cfg = *cfg;
    if (!s->cfg.pool_bytes) s->cfg.pool_bytes = 64ull << 30;
    if (!s->cfg.streams) s->cfg.streams = 4;
    if (!s->cfg.max_group) s->cfg.max_group = 1024;
    if (!s->cfg.max_group_bytes) s->cfg.max_group_bytes = s->cfg.pool_bytes / 8;
    if (s->cfg.streams > 16) s->cfg.streams = 16;
    (void)hipGetDevice(&s->device);
    std::string err;
    if (device_twiddles(1, &s->tw, &err)) {
        delete s;
        return fail(err);
    }
    prepare_kernels();
    if (s->pool.init(s->cfg.pool_bytes)) {
        delete s;
        return -1;
    }
    s->slots.resize(s->cfg.streams);
    for (Slot& sl : s->slots) {
        hipError_t e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreate(&sl.start);
        if (e == hipSuccess) e = hipEventCreate(&sl.done);
        if (e != hipSuccess) {
            for (Slot& t : s->slots) {
                if (t.start) (void)hipEventDestroy(t.start);
                if (t.done) (void)hipEventDestroy(t.done);
                if (t.stream) (void)hipStreamDestroy(t.stream);
            }
            delete s;
            return fail(std::string("stream setup: ") + hipGetErrorString(e));
        }
    }
    hipError_t e = hipEventCreate(&s->ref);
    if (e == hipSuccess) e = hipEventRecord(s->ref, s->slots[0].stream);
    if (e == hipSuccess) e = hipEventSynchronize(s->ref);
    if (e != hipSuccess) {
        for (Slot& t : s->slots) {
            (void)hipEventDestroy(t.start);
            (void)hipEventDestroy(t.done);
            (void)hipStreamDestroy(t.stream);
        }
        delete s;
        return fail(std::string("event setup: ") + hipGetErrorString(e));
    }
    s->loop = std::thread([s] { s->run(); });
    *out = s;
    return 0;
}

int fmh_server_submit(fmh_server* s, const u64* d_a, u64 na, const u64* d_b, u64 nb, u64* id) {
    if (!d_a || !na || (d_b && !nb)) return fail("requests need non-empty operands");
    ProductIn in{d_a, na, d_b, d_b ? nb : na, nullptr};
    return submit(s, in, id);
}

int fmh_server_submit_spectrum(fmh_server* s, const fmh_spectrum* spec, const u64* d_b, u64 nb, u64* id) {
    ProductIn in{};
    if (fmh_spectrum_describe(spec, &in.spec, &in.nb, &in.spec_log_n)) return fail("NULL spectrum");
    in.a = d_b;
    in.na = d_b ? nb : 0;
    if (d_b && !nb) return fail("requests need non-empty operands");
    return submit(s, in, id);
}

int fmh_server_poll(fmh_server* s, u64* ids, u64 max, u64* n) {
    if (!s || !n || (max && !ids)) return fail("NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    const u64 k = std::min<u64>(max, s->completed_unreported.size());
    for (u64 i = 0; i < k; i++) {
        ids[i] = s->completed_unreported[i];
        s->reqs[ids[i]]->reported = true;
    }
    s->completed_unreported.erase(s->completed_unreported.begin(), s->completed_unreported.begin() + k);
    *n = k;
    return 0;
}

int fmh_server_wait(fmh_server* s, u64 id) {
    if (!s) return fail("NULL argument");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail("unknown request id");
    Request* r = it->second;
    s->cv_done.wait(lk, [&] { return r->state >= kDone || (s->stalled && r->state == kPending); });
    if (r->state == kPending) return fail(kStalled);
    return r->state == kDone ? 0 : fail("request failed: " + s->loop_error);
}

int fmh_server_result(fmh_server* s, u64 id, u64** d_out, u64* len) {
    if (!s || !d_out || !len) return fail("NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail("unknown request id");
    Request* r = it->second;
    if (r->state != kDone) return fail(r->state == kFailed ? "request failed" : "request not complete");
    *d_out = (u64*)s->pool.at(r->out_off);
    *len = r->out_len;
    return 0;
}

int fmh_server_release(fmh_server* s, u64 id) {
    if (!s) return fail("NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return fail("unknown request id");
    Request* r = it->second;
    if (r->state < kDone) return fail("request not complete");
    if (r->state == kDone) s->pool.release(r->out_off, r->out_len * sizeof(u64));
    s->reqs.erase(it);
    delete r;
    s->stalled = false;
    s->cv_work.notify_one();
    return 0;
}

int fmh_server_drain(fmh_server* s) {
    if (!s) return fail("NULL argument");
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->completed + s->failed == s->submitted || s->stalled; });
    if (s->completed + s->failed != s->submitted) return fail(kStalled);
    return s->failed ? fail("some requests failed: " + s->loop_error) : 0;
}

int fmh_server_stats(fmh_server* s, fmh_server_info* st) {
    if (!s || !st) return fail("NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    memset(st, 0, sizeof(*st));
    st->submitted = s->submitted;
    st->completed = s->completed;
    st->failed = s->failed;
    st->groups = s->groups;
    st->mean_group = s->groups ? (double)s->grouped / (double)s->groups : 0.0;
    st->pool_bytes = s->pool.capacity();
    st->pool_peak_bytes = s->pool.peak();
    if (!s->latencies.empty()) {
        std::vector<int64_t> v = s->latencies;
        std::sort(v.begin(), v.end());
        auto pct = [&](double q) {
            const size_t i = std::min(v.size() - 1, (size_t)(q * (double)(v.size() - 1) + 0.5));
            return (double)v[i] * 1e-3;
        };
        double sum = 0;
        for (int64_t x : v) sum += (double)x;
        st->p50_us = pct(0.50);
        st->p99_us = pct(0.99);
        st->max_us = (double)v.back() * 1e-3;
        st->mean_us = sum / (double)v.size() * 1e-3;
        const double span = (double)(s->t_last_done - s->t_first_submit) * 1e-9;
        st->products_per_s = span > 0 ? (double)s->completed / span : 0.0;
        st->span_s = span;
    }
    // the union of the groups' device intervals
    std::vector<std::pair<float, float>> iv = s->busy;
    std::sort(iv.begin(), iv.end());
    double busy_ms = 0, end = -1e30;
    for (const auto& x : iv) {
        if (x.first > end) {
            busy_ms += x.second - x.first;
            end = x.second;
        } else if (x.second > end) {
            busy_ms += x.second - end;
            end = x.second;
        }
    }
    st->device_busy_s = busy_ms * 1e-3;
    return 0;
}

u64 fmh_server_class_report(fmh_server* s, char* buf, u64 buflen) {
    if (!s) return 0;
    std::string out;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        char line[160];
        for (const auto& kv : s->class_time) {
            snprintf(line, sizeof(line), "%u %u %llu %llu %.3f\n", kv.first >> 8, kv.first & 0xFFu,
                     (unsigned long long)kv.second.groups, (unsigned long long)kv.second.requests, kv.second.device_ms);
            out += line;
        }
    }
    if (buf && buflen) {
        const size_t n = std::min<size_t>(out.size(), (size_t)buflen - 1);
        memcpy(buf, out.data(), n);
        buf[n] = 0;
    }
    return (u64)out.size() + 1;
}

int fmh_server_reset_stats(fmh_server* s) {
    if (!s) return fail("NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->completed + s->failed != s->submitted) return fail("requests still in flight");
    s->submitted = s->completed = s->failed = s->groups = s->grouped = 0;
    s->latencies.clear();
    s->busy.clear();
    s->class_time.clear();
    return 0;
}

int fmh_server_destroy(fmh_server* s) {
    if (!s) return 0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv_work.notify_one();
    }
    s->loop.join();
    for (Slot& sl : s->slots) {
        (void)hipStreamSynchronize(sl.stream);
        (void)hipEventDestroy(sl.start);
        (void)hipEventDestroy(sl.done);
        (void)hipStreamDestroy(sl.stream);
        if (sl.staging) (void)hipHostFree(sl.staging);
    }
    for (auto& kv : s->reqs) delete kv.second;
    (void)hipEventDestroy(s->ref);
    delete s;
    return 0;
}

const char* fmh_server_last_error(void) { return t_err.c_str(); }

}  // extern "C"
