// Host implementation of tests/emu/hip/hip_runtime.h (tests only).
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <chrono>
#include <functional>
#include <memory>
#include <mutex>
#include <vector>

#include <sanitizer/common_interface_defs.h>
#include <sys/mman.h>
#include <ucontext.h>

#include "hip/hip_runtime.h"

thread_local dim3 threadIdx;
thread_local dim3 blockIdx;
dim3 blockDim;
dim3 gridDim;

// The threads of a workgroup are fibers (ucontext) on the launching host
// thread.  A fiber runs until it reaches a barrier or returns; once every
// fiber of the workgroup has done one or the other, the waiting ones resume.
// So __syncthreads has the GPU's meaning without host-thread barriers, and
// AddressSanitizer follows the stack switches through its fiber interface.
struct Fiber {
    ucontext_t ctx;
    char* stack = nullptr;
    void* fake_stack = nullptr;
    bool done = false;
};
static constexpr size_t kFiberStack = 256 << 10;
static std::mutex g_launch_mu;  // one grid at a time (the globals below)
static std::vector<Fiber> g_fibers;
static ucontext_t g_sched;
static const void* g_sched_bottom = nullptr;
static size_t g_sched_size = 0;
static unsigned g_cur = 0;
static const std::function<void()>* g_body = nullptr;
static unsigned g_slots[1024];
static void* g_dyn = nullptr;

void* emu_dyn_lds() { return g_dyn; }

// back to the scheduler: at a barrier (finishing = false) or at the end
static void to_scheduler(bool finishing) {
    Fiber& f = g_fibers[g_cur];
    __sanitizer_start_switch_fiber(finishing ? nullptr : &f.fake_stack, g_sched_bottom, g_sched_size);
    swapcontext(&f.ctx, &g_sched);
    __sanitizer_finish_switch_fiber(f.fake_stack, nullptr, nullptr);
}

static void fiber_main() {
    __sanitizer_finish_switch_fiber(nullptr, &g_sched_bottom, &g_sched_size);
    (*g_body)();
    g_fibers[g_cur].done = true;
    to_scheduler(true);
    abort();  // a finished fiber is never resumed
}

void emu_syncthreads() { to_scheduler(false); }

unsigned emu_shfl_up_u32(unsigned v, unsigned delta, int width) {
    const unsigned tid = threadIdx.x;
    g_slots[tid] = v;
    emu_syncthreads();
    const unsigned lane = tid % (unsigned)width;
    const unsigned r = lane >= delta ? g_slots[tid - delta] : v;
    emu_syncthreads();
    return r;
}

struct emu_stream {
    bool capturing = false;
    std::vector<std::function<void()>> rec;
};
struct emu_graph {
    std::vector<std::function<void()>> ops;
};
struct emu_event {
    std::chrono::steady_clock::time_point t;
};

// Runs the workgroups one after the other, each as block.x fibers.
static void run_grid(const std::function<void()>& body, dim3 grid, dim3 block, size_t lds) {
    if (block.x == 0 || block.x > 1024 || block.y != 1 || block.z != 1) {
        fprintf(stderr, "emu: bad block shape %u %u %u\n", block.x, block.y, block.z);
        abort();
    }
    std::lock_guard<std::mutex> lock(g_launch_mu);
    gridDim = grid;
    blockDim = block;
    g_body = &body;
    if (g_fibers.size() < block.x) {
        const size_t old = g_fibers.size();
        g_fibers.resize(block.x);
        for (size_t t = old; t < block.x; t++) {
            void* st = mmap(nullptr, kFiberStack, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
            if (st == MAP_FAILED) {
                fprintf(stderr, "emu: fiber stack\n");
                abort();
            }
            g_fibers[t].stack = (char*)st;
        }
    }
    const dim3 saved_t = threadIdx, saved_b = blockIdx;
    const unsigned long long nblocks = (unsigned long long)grid.x * grid.y * grid.z;
    for (unsigned long long b = 0; b < nblocks; b++) {
        // fresh dynamic LDS of exactly the requested size, filled with a pattern
        g_dyn = lds ? malloc(lds) : nullptr;
        if (lds) memset(g_dyn, 0xA5, lds);
        blockIdx = dim3((unsigned)(b % grid.x), (unsigned)((b / grid.x) % grid.y),
                        (unsigned)(b / ((unsigned long long)grid.x * grid.y)));
        for (unsigned t = 0; t < block.x; t++) {
            Fiber& f = g_fibers[t];
            getcontext(&f.ctx);
            f.ctx.uc_stack.ss_sp = f.stack;
            f.ctx.uc_stack.ss_size = kFiberStack;
            f.ctx.uc_link = nullptr;
            makecontext(&f.ctx, fiber_main, 0);
            f.done = false;
            f.fake_stack = nullptr;
        }
        unsigned live = block.x;
        while (live) {
            for (unsigned t = 0; t < block.x; t++) {
                Fiber& f = g_fibers[t];
                if (f.done) continue;
                g_cur = t;
                threadIdx = dim3(t, 0, 0);
                void* fake = nullptr;
                __sanitizer_start_switch_fiber(&fake, f.stack, kFiberStack);
                swapcontext(&g_sched, &f.ctx);
                __sanitizer_finish_switch_fiber(fake, nullptr, nullptr);
                if (f.done) live--;
            }
        }
        free(g_dyn);
        g_dyn = nullptr;
    }
    threadIdx = saved_t;
    blockIdx = saved_b;
}

void emu_launch(std::function<void()> body, dim3 grid, dim3 block, size_t lds, hipStream_t s) {
    if (s && s->capturing) {
        s->rec.push_back([=] { run_grid(body, grid, block, lds); });
        return;
    }
    run_grid(body, grid, block, lds);
}

const char* hipGetErrorString(hipError_t e) { return e == hipSuccess ? "success" : "emulated error"; }
hipError_t hipGetLastError() { return hipSuccess; }
hipError_t hipGetDevice(int* d) {
    *d = 0;
    return hipSuccess;
}
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipDeviceSynchronize() { return hipSuccess; }
hipError_t hipMalloc(void** p, size_t n) {
    *p = malloc(n ? n : 1);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipFree(void* p) {
    free(p);
    return hipSuccess;
}
hipError_t hipHostMalloc(void** p, size_t n, unsigned) { return hipMalloc(p, n); }
hipError_t hipHostFree(void* p) { return hipFree(p); }
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) {
    memmove(d, s, n);
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t st) {
    if (st && st->capturing) return hipErrorInvalidValue;  // the library never captures copies
    return hipMemcpy(d, s, n, k);
}
hipError_t hipStreamCreate(hipStream_t* s) {
    *s = new emu_stream();
    return hipSuccess;
}
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) { return hipStreamCreate(s); }
hipError_t hipStreamDestroy(hipStream_t s) {
    delete s;
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }  // streams run eagerly
hipError_t hipEventCreate(hipEvent_t* e) {
    *e = new emu_event();
    return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventRecord(hipEvent_t e, hipStream_t) {
    e->t = std::chrono::steady_clock::now();
    return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventQuery(hipEvent_t) { return hipSuccess; }  // streams run eagerly
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b) {
    *ms = std::chrono::duration<float, std::milli>(b->t - a->t).count();
    return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t e) {
    delete e;
    return hipSuccess;
}
hipError_t hipFuncSetAttribute(const void*, hipFuncAttribute, int) { return hipSuccess; }
hipError_t hipFuncGetAttributes(hipFuncAttributes* a, const void*) {
    a->sharedSizeBytes = 0;
    return hipSuccess;
}
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
    if (!s) return hipErrorInvalidValue;
    s->capturing = true;
    s->rec.clear();
    return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g) {
    s->capturing = false;
    *g = new emu_graph();
    (*g)->ops.swap(s->rec);
    return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, void*, void*, size_t) {
    *e = new emu_graph(*g);
    return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t) {
    for (auto& op : e->ops) op();
    return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t e) {
    delete e;
    return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) {
    delete g;
    return hipSuccess;
}
