// Host implementation of tests/emu/hip/hip_runtime.h (tests only).
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <atomic>
#include <barrier>
#include <chrono>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "hip/hip_runtime.h"

thread_local dim3 threadIdx;
thread_local dim3 blockIdx;
dim3 blockDim;
dim3 gridDim;

static std::barrier<>* g_bar = nullptr;
static unsigned g_slots[1024];
static void* g_dyn = nullptr;
static size_t g_dyn_bytes = 0;

void* emu_dyn_lds() { return g_dyn; }

static std::atomic<bool> g_used_barrier{false};

void emu_syncthreads() {
    g_used_barrier = true;
    g_bar->arrive_and_wait();
}

unsigned emu_shfl_up_u32(unsigned v, unsigned delta, int width) {
    g_used_barrier = true;
    const unsigned tid = threadIdx.x;
    g_slots[tid] = v;
    g_bar->arrive_and_wait();
    const unsigned lane = tid % (unsigned)width;
    const unsigned r = lane >= delta ? g_slots[tid - delta] : v;
    g_bar->arrive_and_wait();
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

// A team of block.x host threads runs every workgroup of one launch in turn.
// A kernel whose first workgroup never synchronised is barrier-free: its other
// workgroups run their threads one after the other on the calling thread.
static void run_grid(const std::function<void()>& body, dim3 grid, dim3 block, size_t lds) {
    if (block.x == 0 || block.x > 1024 || block.y != 1 || block.z != 1) {
        fprintf(stderr, "emu: bad block shape %u %u %u\n", block.x, block.y, block.z);
        abort();
    }
    gridDim = grid;
    blockDim = block;
    const unsigned long long nblocks = (unsigned long long)grid.x * grid.y * grid.z;
    auto block_of = [&](unsigned long long b) {
        return dim3((unsigned)(b % grid.x), (unsigned)((b / grid.x) % grid.y),
                    (unsigned)(b / ((unsigned long long)grid.x * grid.y)));
    };
    auto new_lds = [&] {
        // fresh dynamic LDS of exactly the requested size, filled with a pattern
        g_dyn_bytes = lds;
        g_dyn = lds ? malloc(lds) : nullptr;
        if (lds) memset(g_dyn, 0xA5, lds);
    };
    std::barrier<> bar(block.x);
    g_bar = &bar;
    g_used_barrier = false;
    // first workgroup with real threads
    new_lds();
    {
        std::vector<std::thread> th;
        for (unsigned t = 0; t < block.x; t++)
            th.emplace_back([&, t] {
                threadIdx = dim3(t, 0, 0);
                blockIdx = block_of(0);
                body();
            });
        for (auto& h : th) h.join();
    }
    free(g_dyn);
    g_dyn = nullptr;
    if (nblocks == 1) return;
    if (!g_used_barrier) {
        const dim3 saved_t = threadIdx, saved_b = blockIdx;
        for (unsigned long long b = 1; b < nblocks; b++) {
            new_lds();
            blockIdx = block_of(b);
            for (unsigned t = 0; t < block.x; t++) {
                threadIdx = dim3(t, 0, 0);
                body();
            }
            free(g_dyn);
            g_dyn = nullptr;
        }
        threadIdx = saved_t;
        blockIdx = saved_b;
        return;
    }
    std::barrier<> gate(block.x + 1);  // block start / end, with the launching thread
    std::vector<std::thread> team;
    team.reserve(block.x);
    for (unsigned t = 0; t < block.x; t++)
        team.emplace_back([&, t] {
            threadIdx = dim3(t, 0, 0);
            for (unsigned long long b = 1; b < nblocks; b++) {
                gate.arrive_and_wait();  // block b starts
                blockIdx = block_of(b);
                body();
                gate.arrive_and_wait();  // block b done
            }
        });
    for (unsigned long long b = 1; b < nblocks; b++) {
        new_lds();
        gate.arrive_and_wait();
        gate.arrive_and_wait();
        free(g_dyn);
        g_dyn = nullptr;
    }
    for (auto& h : team) h.join();
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
