// CPU execution of the library's HIP sources, for tests only.
//
// tests/emu/build_emu.sh compiles fast_mul_ntt_amd/csrc/*.hip with g++
// against this header, after one textual substitution: the dynamic LDS array
// ("extern __shared__ u64 lds[]") becomes a pointer to a host buffer of
// exactly the launch's dynamic LDS size, so its bounds are checked.  Every
// GPU thread of a workgroup is a fiber on the launching host thread;
// workgroups run one after the other; __syncthreads suspends a fiber until
// every thread of its workgroup has reached a barrier or returned; device
// memory is host memory, so AddressSanitizer checks every kernel access.
// Streams run eagerly; graph capture records launches for replay.
#pragma once

#include <stddef.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static
#define __align__(n) __attribute__((aligned(n)))
// a __device__ variable is a host object: its "device address" is its address
#define HIP_SYMBOL(x) (&(x))

struct dim3 {
    unsigned x, y, z;
    dim3(unsigned x_ = 1, unsigned y_ = 1, unsigned z_ = 1) : x(x_), y(y_), z(z_) {}
};

extern thread_local dim3 threadIdx;
extern thread_local dim3 blockIdx;
extern dim3 blockDim;
extern dim3 gridDim;

void emu_syncthreads();
void* emu_dyn_lds();
unsigned emu_shfl_up_u32(unsigned v, unsigned delta, int width);

inline void __syncthreads() { emu_syncthreads(); }
inline unsigned __shfl_up(unsigned v, unsigned delta, int width = 64) { return emu_shfl_up_u32(v, delta, width); }
inline int __clzll(long long x) { return x == 0 ? 64 : __builtin_clzll((unsigned long long)x); }

template <class T>
inline T min(T a, T b) {
    return a < b ? a : b;
}
template <class T>
inline T max(T a, T b) {
    return a < b ? b : a;
}

typedef int hipError_t;
enum {
    hipSuccess = 0,
    hipErrorInvalidValue = 1,
    hipErrorOutOfMemory = 2,
    hipErrorNotReady = 600,
    hipErrorLaunchFailure = 719,
};
typedef struct emu_stream* hipStream_t;
typedef struct emu_event* hipEvent_t;
typedef struct emu_graph* hipGraph_t;
typedef struct emu_graph* hipGraphExec_t;
typedef int hipMemcpyKind;
enum { hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2, hipMemcpyDeviceToDevice = 3 };
enum { hipStreamNonBlocking = 1, hipEventDisableTiming = 2, hipHostMallocDefault = 0 };
typedef int hipStreamCaptureMode;
enum { hipStreamCaptureModeThreadLocal = 1 };
typedef int hipFuncAttribute;
enum { hipFuncAttributeMaxDynamicSharedMemorySize = 8 };

const char* hipGetErrorString(hipError_t e);
hipError_t hipGetLastError();
hipError_t hipGetDevice(int* d);
hipError_t hipSetDevice(int d);
hipError_t hipDeviceSynchronize();
hipError_t hipMalloc(void** p, size_t n);
template <class T>
inline hipError_t hipMalloc(T** p, size_t n) {
    return hipMalloc((void**)p, n);
}
hipError_t hipFree(void* p);
hipError_t hipHostMalloc(void** p, size_t n, unsigned flags);
hipError_t hipHostFree(void* p);
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind k);
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t st);
hipError_t hipStreamCreate(hipStream_t* s);
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned flags);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned flags);
hipError_t hipEventCreate(hipEvent_t* e);
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned flags);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s = nullptr);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventQuery(hipEvent_t e);
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipFuncSetAttribute(const void* f, hipFuncAttribute a, int v);
inline hipError_t hipGetSymbolAddress(void** p, const void* sym) {
    *p = const_cast<void*>(sym);
    return hipSuccess;
}
struct hipFuncAttributes {
    size_t sharedSizeBytes;
};
hipError_t hipFuncGetAttributes(hipFuncAttributes* a, const void* f);
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode m);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g);
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, void*, void*, size_t);
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s);
hipError_t hipGraphExecDestroy(hipGraphExec_t e);
hipError_t hipGraphDestroy(hipGraph_t g);

// Launch machinery: run (or record, when the stream is capturing) a grid.
#include <functional>
void emu_launch(std::function<void()> body, dim3 grid, dim3 block, size_t lds, hipStream_t s);

#define hipLaunchKernelGGL(kern, grid, block, lds, stream, ...) \
    emu_launch([=]() { kern(__VA_ARGS__); }, dim3(grid), dim3(block), (size_t)(lds), stream)
