// Test infrastructure, never part of the library: a CPU stand-in for the
// subset of the HIP runtime and device intrinsics that the library's own
// sources use, so that walicxe3d_amd/csrc/*.hip compile as plain C++ and run
// on the host under AddressSanitizer / MemorySanitizer.  Every device array
// becomes a malloc'ed host array, so an out-of-range index in any kernel is
// reported with its source line.  Kernels run one workgroup at a time, each
// lane a fiber (tests/emu/emu_rt.cpp); graphs are recorded and replayed.
#pragma once
#include <atomic>
#include <barrier>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <thread>
#include <vector>
#include <algorithm>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __noinline__
#define __launch_bounds__(...)
#define __shared__ static
#define __align__(n) __attribute__((aligned(n)))

struct dim3 {
  unsigned x = 1, y = 1, z = 1;
  dim3() = default;
  dim3(unsigned a, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
extern thread_local dim3 threadIdx, blockIdx;
extern dim3 gridDim, blockDim;

typedef int hipError_t;
enum { hipSuccess = 0, hipErrorInvalidValue = 1 };
typedef int hipMemcpyKind;
enum { hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2, hipMemcpyDeviceToDevice = 3, hipMemcpyDefault = 4 };
enum { hipStreamNonBlocking = 1 };
enum { hipStreamCaptureModeThreadLocal = 1, hipStreamCaptureModeGlobal = 0 };
typedef int hipStreamCaptureMode;
struct EmuStream { bool capturing = false; std::vector<std::function<void()>> *rec = nullptr; };
typedef EmuStream *hipStream_t;
struct EmuEvent { int dummy; };
typedef EmuEvent *hipEvent_t;
struct EmuGraph { std::vector<std::function<void()>> ops; };
typedef EmuGraph *hipGraph_t;
typedef EmuGraph *hipGraphExec_t;
typedef void *hipGraphNode_t;

inline const char *hipGetErrorString(hipError_t) { return "emu error"; }
inline hipError_t hipGetLastError() { return hipSuccess; }
// device symbols are host arrays here: the symbol's address is the array's
#define HIP_SYMBOL(x) (&(x))
inline hipError_t hipGetSymbolAddress(void **p, const void *sym) { *p = const_cast<void *>(sym); return hipSuccess; }
inline hipError_t hipGetDeviceCount(int *n) { *n = 1; return hipSuccess; }
inline hipError_t hipGetDevice(int *d) { *d = 0; return hipSuccess; }
inline hipError_t hipSetDevice(int) { return hipSuccess; }
inline hipError_t hipDeviceSynchronize() { return hipSuccess; }
template <typename T> inline hipError_t hipMalloc(T **p, size_t n) { *p = (T *)malloc(n ? n : 1); return *p ? hipSuccess : 2; }
inline hipError_t hipFree(void *p) { free(p); return hipSuccess; }
enum { hipHostMallocDefault = 0 };
template <typename T> inline hipError_t hipHostMalloc(T **p, size_t n, unsigned) {
  *p = (T *)malloc(n ? n : 1);
  return *p ? hipSuccess : 2;
}
inline hipError_t hipHostFree(void *p) { free(p); return hipSuccess; }
inline hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind) { memcpy(d, s, n); return hipSuccess; }

void emu_record_or_run(hipStream_t st, std::function<void()> f);
inline hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind, hipStream_t st = nullptr) {
  emu_record_or_run(st, [=] { memcpy(d, s, n); });
  return hipSuccess;
}
inline hipError_t hipMemset(void *d, int v, size_t n) { memset(d, v, n); return hipSuccess; }
inline hipError_t hipMemsetAsync(void *d, int v, size_t n, hipStream_t st = nullptr) {
  emu_record_or_run(st, [=] { memset(d, v, n); });
  return hipSuccess;
}
inline hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned) { *s = new EmuStream(); return hipSuccess; }
inline hipError_t hipStreamCreate(hipStream_t *s) { *s = new EmuStream(); return hipSuccess; }
inline hipError_t hipStreamDestroy(hipStream_t s) { delete s; return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipEventCreate(hipEvent_t *e) { *e = new EmuEvent(); return hipSuccess; }
inline hipError_t hipEventDestroy(hipEvent_t e) { delete e; return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t = nullptr) { return hipSuccess; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float *ms, hipEvent_t, hipEvent_t) { *ms = 1.0f; return hipSuccess; }
inline hipError_t hipStreamBeginCapture(hipStream_t s, int) {
  s->capturing = true; s->rec = new std::vector<std::function<void()>>(); return hipSuccess;
}
inline hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g) {
  *g = new EmuGraph(); (*g)->ops = std::move(*s->rec); delete s->rec; s->rec = nullptr; s->capturing = false;
  return hipSuccess;
}
inline hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, hipGraphNode_t *, char *, size_t) {
  *e = new EmuGraph(*g); return hipSuccess;
}
inline hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t) { for (auto &f : e->ops) f(); return hipSuccess; }
inline hipError_t hipGraphDestroy(hipGraph_t g) { delete g; return hipSuccess; }
inline hipError_t hipGraphExecDestroy(hipGraphExec_t g) { delete g; return hipSuccess; }

// ---- device intrinsics -------------------------------------------------
void emu_syncthreads(const char *file, int line);
#define __syncthreads() emu_syncthreads(__FILE__, __LINE__)
double emu_shfl_xor(double v, int off);
inline double __shfl_xor(double v, int off) { return emu_shfl_xor(v, off); }
inline long long __double_as_longlong(double d) { long long r; memcpy(&r, &d, 8); return r; }
inline double __longlong_as_double(long long l) { double r; memcpy(&r, &l, 8); return r; }
inline int __float_as_int(float f) { int r; memcpy(&r, &f, 4); return r; }
inline float __int_as_float(int i) { float r; memcpy(&r, &i, 4); return r; }
inline unsigned long long atomicMin(unsigned long long *p, unsigned long long v) {
  std::atomic_ref<unsigned long long> a(*p);
  unsigned long long o = a.load();
  while (v < o && !a.compare_exchange_weak(o, v)) {}
  return o;
}
inline unsigned long long atomicAdd(unsigned long long *p, unsigned long long v) {
  return std::atomic_ref<unsigned long long>(*p).fetch_add(v);
}
inline int atomicAdd(int *p, int v) { return std::atomic_ref<int>(*p).fetch_add(v); }

// ---- launches -----------------------------------------------------------
void emu_run_grid(dim3 grid, dim3 block, std::function<void()> body);
#define hipLaunchKernelGGL(K, G, B, SH, ST, ...)                                  \
  do {                                                                            \
    dim3 g_ = (G), b_ = (B);                                                      \
    auto kfn_ = [=] { emu_run_grid(g_, b_, [=] { K(__VA_ARGS__); }); };           \
    emu_record_or_run((ST), kfn_);                                                \
  } while (0)
