// CPU stand-in for HIP (device environment + the runtime calls the library
// makes), for bounds-checking the library under AddressSanitizer.  Test
// infrastructure only (tests/cpuemu/build_emu.py).  Every GPU
// thread of a workgroup runs as a host thread; __syncthreads is a barrier
// over the workgroup, the wave barrier a barrier over its 64-thread wave.
// Device memory is host malloc (ASan redzones); every queue operation runs at
// once, in program order, except under stream capture, where launches and
// copies are recorded into the graph and replayed by hipGraphLaunch.
#pragma once
#include <pthread.h>

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

struct double2 { double x, y; };
inline double2 make_double2(double a, double b) { return double2{a, b}; }
struct dim3 {
  unsigned x, y, z;
  dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
typedef int hipError_t;
typedef void* hipStream_t;
typedef void* hipEvent_t;
struct emu_graph { std::vector<std::function<void()>> ops; };
typedef emu_graph* hipGraph_t;
typedef emu_graph* hipGraphExec_t;
enum { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorInvalidDevice = 101 };
enum hipMemcpyKind { hipMemcpyHostToHost, hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice };
enum { hipStreamNonBlocking = 1, hipEventDisableTiming = 2, hipStreamCaptureModeThreadLocal = 1 };
struct hipDeviceProp_t { char name[256]; char gcnArchName[256]; int multiProcessorCount; };

#define __device__
#define __host__
#define __forceinline__ inline
#define __global__
#define __launch_bounds__(...)
#define __shared__ static
#define __restrict__
#define __align__(n) alignas(n)

extern thread_local dim3 threadIdx;
extern thread_local dim3 blockIdx;
extern thread_local dim3 gridDim;
extern thread_local dim3 blockDim;
// Wave-scoped barrier over the wave's threads that take part in the current
// region: a thread that finishes the workgroup's body, or waits at
// __syncthreads, stops counting (as masked-off lanes on the GPU do not take
// part in a wave's instructions); __syncthreads resets every wave's count to
// its threads still running the body.
void emu_wave_arrive(unsigned wave);
void emu_syncthreads();
inline void __syncthreads() { emu_syncthreads(); }
#define __builtin_amdgcn_fence(...) ((void)0)
// the hardware reciprocal's stand-in (the kernels refine it by Newton steps)
inline double __builtin_amdgcn_rcp(double x) { return 1.0 / x; }
inline void __builtin_amdgcn_wave_barrier() { emu_wave_arrive(threadIdx.x / 64); }
extern double emu_shfl_d[1024];
extern unsigned long long emu_shfl_u[1024];
inline double __shfl_xor(double v, int off, int) {
  const unsigned t = threadIdx.x;
  emu_shfl_d[t] = v;
  __builtin_amdgcn_wave_barrier();
  const double r = emu_shfl_d[(t & ~63u) | ((t & 63u) ^ (unsigned)off)];
  __builtin_amdgcn_wave_barrier();
  return r;
}
inline unsigned long long __shfl_xor(unsigned long long v, int off, int) {
  const unsigned t = threadIdx.x;
  emu_shfl_u[t] = v;
  __builtin_amdgcn_wave_barrier();
  const unsigned long long r = emu_shfl_u[(t & ~63u) | ((t & 63u) ^ (unsigned)off)];
  __builtin_amdgcn_wave_barrier();
  return r;
}
// DPP moves: the one permutation the kernels use, quad_perm [1, 0, 3, 2]
// (lane l takes lane l ^ 1's value)
extern int emu_dpp_i[1024];
inline int __builtin_amdgcn_mov_dpp(int v, int ctrl, int, int, bool) {
  if (ctrl != 0xB1) std::abort();
  const unsigned t = threadIdx.x;
  emu_dpp_i[t] = v;
  __builtin_amdgcn_wave_barrier();
  const int r = emu_dpp_i[t ^ 1u];
  __builtin_amdgcn_wave_barrier();
  return r;
}
extern std::mutex emu_atomic_mu;
inline unsigned long long atomicMax(unsigned long long* p, unsigned long long v) {
  std::lock_guard<std::mutex> lk(emu_atomic_mu);
  const unsigned long long o = *p;
  if (v > o) *p = v;
  return o;
}
inline long long __double_as_longlong(double d) { long long r; std::memcpy(&r, &d, 8); return r; }
inline double __longlong_as_double(long long l) { double r; std::memcpy(&r, &l, 8); return r; }
using std::fabs;
using std::fmax;
using std::sqrt;

// ---- runtime ---------------------------------------------------------------
void emu_run(dim3 grid, dim3 block, std::function<void()> body);   // one launch, all workgroups
extern thread_local emu_graph* emu_capturing;
inline void emu_enqueue(std::function<void()> op) {
  if (emu_capturing) emu_capturing->ops.push_back(std::move(op));
  else op();
}
#define hipLaunchKernelGGL(K, G, B, SH, S, ...)                                       \
  do {                                                                                \
    auto emu_k_ = (K);                                                                \
    const dim3 emu_g_ = (G), emu_b_ = (B);                                            \
    emu_enqueue([=] { emu_run(emu_g_, emu_b_, [=] { emu_k_(__VA_ARGS__); }); });      \
  } while (0)
inline hipError_t hipGetLastError() { return hipSuccess; }
inline const char* hipGetErrorString(hipError_t e) { return e ? "emulated error" : "no error"; }
inline hipError_t hipGetDeviceCount(int* n) { *n = 1; return hipSuccess; }
inline hipError_t hipSetDevice(int) { return hipSuccess; }
inline hipError_t hipGetDeviceProperties(hipDeviceProp_t* p, int) {
  std::snprintf(p->name, sizeof p->name, "cpu emulation");
  std::snprintf(p->gcnArchName, sizeof p->gcnArchName, "emu");
  p->multiProcessorCount = 256;
  return hipSuccess;
}
inline hipError_t hipMemGetInfo(size_t* f, size_t* t) { *f = (size_t)8 << 30; *t = (size_t)16 << 30; return hipSuccess; }
inline hipError_t hipMalloc(void** p, size_t n) { *p = std::malloc(n ? n : 1); return *p ? hipSuccess : 2; }
inline hipError_t hipFree(void* p) { std::free(p); return hipSuccess; }
inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) { *s = (void*)new int(0); return hipSuccess; }
inline hipError_t hipStreamDestroy(hipStream_t s) { delete (int*)s; return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipEventCreate(hipEvent_t* e) { *e = (void*)new int(0); return hipSuccess; }
inline hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
inline hipError_t hipEventDestroy(hipEvent_t e) { delete (int*)e; return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t = nullptr) { return hipSuccess; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) { *ms = 1.0f; return hipSuccess; }
inline hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
inline hipError_t hipMemsetAsync(void* p, int v, size_t n, hipStream_t = nullptr) {
  emu_enqueue([=] { std::memset(p, v, n); });
  return hipSuccess;
}
inline hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) { std::memcpy(d, s, n); return hipSuccess; }
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t = nullptr) {
  emu_enqueue([=] { std::memcpy(d, s, n); });
  return hipSuccess;
}
inline hipError_t hipMemcpy2DAsync(void* d, size_t dp, const void* s, size_t sp, size_t w, size_t h, hipMemcpyKind,
                                   hipStream_t = nullptr) {
  emu_enqueue([=] {
    for (size_t r = 0; r < h; ++r) std::memcpy((char*)d + r * dp, (const char*)s + r * sp, w);
  });
  return hipSuccess;
}
inline hipError_t hipStreamBeginCapture(hipStream_t, int) { emu_capturing = new emu_graph(); return hipSuccess; }
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g) { *g = emu_capturing; emu_capturing = nullptr; return hipSuccess; }
inline hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, void*, void*, size_t) {
  *e = new emu_graph(*g);
  return hipSuccess;
}
inline hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t) {
  for (auto& op : e->ops) emu_enqueue(op);
  return hipSuccess;
}
inline hipError_t hipGraphExecDestroy(hipGraphExec_t e) { delete e; return hipSuccess; }
inline hipError_t hipGraphDestroy(hipGraph_t g) { delete g; return hipSuccess; }
#define HIP_SYMBOL(x) (&(x))
inline hipError_t hipGetSymbolAddress(void** p, const void* sym) { *p = const_cast<void*>(sym); return hipSuccess; }
