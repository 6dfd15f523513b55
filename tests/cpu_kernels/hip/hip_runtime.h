// Host stand-in for <hip/hip_runtime.h>: test infrastructure only.
//
// tests/cpu_kernels builds fdtd_amd/csrc/fdtd_hip.hip unchanged with g++ (as
// C++) against this header, which comes first on the include path. The
// library it produces has the same C ABI as libfdtd_hip.so but runs every
// kernel on the host: "device" memory is host memory, a launch runs at once
// on the calling thread, and each workgroup's work-items run as fibers that
// proceed in lockstep at __syncthreads (every work-item reaches barrier n
// before any passes it). Built under AddressSanitizer, it checks every
// kernel access for bounds on the CPU and must reproduce the GPU's results
// bit for bit (fdtd_hip.hip's arithmetic is IEEE with contraction off, and
// fma is the correctly rounded fused operation on both).
//
// Address-space and GPU-only attributes are left for g++ to ignore; the
// clang builtins the kernels call are given their host meaning below.
#pragma once

#include <ucontext.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <vector>

#if defined(__SANITIZE_ADDRESS__)
#include <sanitizer/asan_interface.h>
#endif

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __align__(n) __attribute__((aligned(n)))
// an occupancy hint whose template-id argument g++ cannot parse as an
// attribute argument: the attribute becomes a harmless one
#define amdgpu_waves_per_eu(...) unused
#define __builtin_nontemporal_load(p) (*(p))
#define __builtin_nontemporal_store(v, p) (*(p) = (v))
// used on values that are uniform over the wavefront
#define __builtin_amdgcn_readfirstlane(x) (x)

using std::fabs;
using std::fma;

template <class T>
inline T min(T a, T b) {
  return b < a ? b : a;
}
template <class T>
inline T max(T a, T b) {
  return a < b ? b : a;
}

inline int __double2loint(double d) {
  uint64_t u;
  std::memcpy(&u, &d, 8);
  return (int)(uint32_t)u;
}
inline int __double2hiint(double d) {
  uint64_t u;
  std::memcpy(&u, &d, 8);
  return (int)(uint32_t)(u >> 32);
}
inline double __hiloint2double(int hi, int lo) {
  const uint64_t u = ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
  double d;
  std::memcpy(&d, &u, 8);
  return d;
}
inline long long __double_as_longlong(double d) {
  long long v;
  std::memcpy(&v, &d, 8);
  return v;
}
inline double __longlong_as_double(long long v) {
  double d;
  std::memcpy(&d, &v, 8);
  return d;
}
inline unsigned __float_as_uint(float f) {
  unsigned v;
  std::memcpy(&v, &f, 4);
  return v;
}
inline float __uint_as_float(unsigned v) {
  float f;
  std::memcpy(&f, &v, 4);
  return f;
}

// work-items run one at a time on one host thread
inline int atomicOr(int* p, int v) {
  const int old = *p;
  *p = old | v;
  return old;
}

struct uint4 {
  unsigned x, y, z, w;
};

struct dim3 {
  unsigned x, y, z;
  dim3(unsigned x_ = 1, unsigned y_ = 1, unsigned z_ = 1) : x(x_), y(y_), z(z_) {}
};

// ------------------------------------------------------------ runtime types
typedef int hipError_t;
enum { hipSuccess = 0, hipErrorInvalidValue = 1 };
enum hipMemcpyKind { hipMemcpyHostToHost, hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice };
enum { hipStreamNonBlocking = 1, hipEventDisableTiming = 2 };
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal, hipStreamCaptureModeThreadLocal, hipStreamCaptureModeRelaxed };
enum hipDeviceAttribute_t { hipDeviceAttributeMultiprocessorCount = 0 };

namespace shim {

using Closure = std::function<void()>;

struct Stream {
  bool capturing = false;
  std::vector<Closure> recorded;
};
struct Event {
  std::chrono::steady_clock::time_point t;
};
struct Graph {
  std::vector<Closure> nodes;
};

// ---------------------------------------------------------- work-item state
struct Idx {
  unsigned x = 0, y = 0, z = 0;
};
inline Idx g_thread, g_block, g_bdim, g_gdim;

struct Fiber {
  ucontext_t ctx;
  char* stack = nullptr;
  bool done = false, at_barrier = false;
};
constexpr size_t kStack = 256 << 10;
inline ucontext_t g_sched;
inline Fiber* g_cur = nullptr;
inline const Closure* g_body = nullptr;
inline bool g_fibers = false;       // the current block runs as fibers
inline bool g_hit_barrier = false;  // a barrier was reached in this launch
inline std::vector<std::unique_ptr<Fiber>> g_pool;
inline int64_t g_launches = 0;

[[noreturn]] inline void die(const char* msg) {
  std::fprintf(stderr, "hip shim: %s\n", msg);
  std::fflush(stderr);
  std::abort();
}

inline void fiber_entry() {
  (*g_body)();
  g_cur->done = true;
  swapcontext(&g_cur->ctx, &g_sched);
  die("finished fiber resumed");
}

inline void barrier() {
  g_hit_barrier = true;
  if (!g_fibers) die("__syncthreads in a kernel whose first workgroup reached none");
  g_cur->at_barrier = true;
  swapcontext(&g_cur->ctx, &g_sched);
}

// One workgroup as nt fibers, in lockstep at barriers.
inline void run_block_fibers(unsigned nt, const Closure& body) {
  while (g_pool.size() < nt) {
    auto f = std::make_unique<Fiber>();
    f->stack = static_cast<char*>(std::malloc(kStack));
    g_pool.push_back(std::move(f));
  }
  g_body = &body;
  g_fibers = true;
  for (unsigned t = 0; t < nt; ++t) {
    Fiber& f = *g_pool[t];
#if defined(__SANITIZE_ADDRESS__)
    // the stack may hold a finished fiber's poisoned frames
    __asan_unpoison_memory_region(f.stack, kStack);
#endif
    getcontext(&f.ctx);
    f.ctx.uc_stack.ss_sp = f.stack;
    f.ctx.uc_stack.ss_size = kStack;
    f.ctx.uc_link = nullptr;
    makecontext(&f.ctx, fiber_entry, 0);
    f.done = false;
    f.at_barrier = false;
  }
  for (;;) {
    unsigned done = 0, waiting = 0;
    for (unsigned t = 0; t < nt; ++t) {
      Fiber& f = *g_pool[t];
      if (f.done) continue;
      f.at_barrier = false;
      g_cur = &f;
      g_thread.x = t;
      swapcontext(&g_sched, &f.ctx);
    }
    for (unsigned t = 0; t < nt; ++t) {
      done += g_pool[t]->done;
      waiting += g_pool[t]->at_barrier;
    }
    if (done == nt) break;
    if (done + waiting != nt || done != 0) die("work-items of one workgroup reached different barriers");
  }
  g_fibers = false;
  g_cur = nullptr;
}

// A launch: the first workgroup runs as fibers; if none of its work-items
// reached a barrier (a kernel reaches its barriers in every work-item or in
// none), the other workgroups run work-item by work-item.
inline void run_grid(dim3 grid, dim3 block, const Closure& body) {
  if (block.y != 1 || block.z != 1 || grid.y != 1 || grid.z != 1) die("only 1-D launches are modelled");
  if (block.x == 0 || block.x > 1024) die("bad workgroup size");
  ++g_launches;
  g_gdim = {grid.x, 1, 1};
  g_bdim = {block.x, 1, 1};
  g_hit_barrier = false;
  for (unsigned b = 0; b < grid.x; ++b) {
    g_block = {b, 0, 0};
    if (b == 0 || g_hit_barrier) {
      run_block_fibers(block.x, body);
    } else {
      for (unsigned t = 0; t < block.x; ++t) {
        g_thread.x = t;
        body();
      }
    }
  }
}

inline void launch(void* stream, dim3 grid, dim3 block, Closure body) {
  Stream* s = static_cast<Stream*>(stream);
  if (s && s->capturing) {
    s->recorded.push_back([grid, block, body]() { run_grid(grid, block, body); });
    return;
  }
  run_grid(grid, block, body);
}

}  // namespace shim

#define threadIdx (::shim::g_thread)
#define blockIdx (::shim::g_block)
#define blockDim (::shim::g_bdim)
#define gridDim (::shim::g_gdim)

inline void __syncthreads() { ::shim::barrier(); }

typedef shim::Stream* hipStream_t;
typedef shim::Event* hipEvent_t;
typedef shim::Graph* hipGraph_t;
typedef shim::Graph* hipGraphExec_t;

// Kernel arguments are copied at launch, as the runtime does.
#define hipLaunchKernelGGL(K, G, B, S, ST, ...) \
  ::shim::launch((void*)(ST), dim3(G), dim3(B), [=]() { K(__VA_ARGS__); })

#define HIP_SYMBOL(x) (reinterpret_cast<const void*>(&(x)))

// --------------------------------------------------------------- host calls
inline const char* hipGetErrorString(hipError_t) { return "hip shim error"; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipGetDevice(int* d) {
  *d = 0;
  return hipSuccess;
}
inline hipError_t hipSetDevice(int d) { return d == 0 ? hipSuccess : hipErrorInvalidValue; }
inline hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t, int) {
  *v = 256;
  return hipSuccess;
}
inline hipError_t hipDeviceSynchronize() { return hipSuccess; }
inline hipError_t hipMalloc(void** p, size_t n) {
  *p = std::aligned_alloc(256, (n + 255) / 256 * 256 + 256);
  return *p ? hipSuccess : hipErrorInvalidValue;
}
inline hipError_t hipFree(void* p) {
  std::free(p);
  return hipSuccess;
}
inline hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) {
  std::memcpy(d, s, n);
  return hipSuccess;
}
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t st) {
  if (st && st->capturing) return hipErrorInvalidValue;  // not modelled under capture
  std::memcpy(d, s, n);
  return hipSuccess;
}
inline hipError_t hipGetSymbolAddress(void** p, const void* sym) {
  *p = const_cast<void*>(sym);
  return hipSuccess;
}
inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) {
  *s = new shim::Stream();
  return hipSuccess;
}
inline hipError_t hipStreamDestroy(hipStream_t s) {
  delete s;
  return hipSuccess;
}
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipEventCreate(hipEvent_t* e) {
  *e = new shim::Event();
  return hipSuccess;
}
inline hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
inline hipError_t hipEventDestroy(hipEvent_t e) {
  delete e;
  return hipSuccess;
}
inline hipError_t hipEventRecord(hipEvent_t e, hipStream_t = nullptr) {
  e->t = std::chrono::steady_clock::now();
  return hipSuccess;
}
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b) {
  *ms = std::chrono::duration<float, std::milli>(b->t - a->t).count();
  return hipSuccess;
}
inline hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
inline hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
  if (!s || s->capturing) return hipErrorInvalidValue;
  s->capturing = true;
  s->recorded.clear();
  return hipSuccess;
}
inline hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g) {
  if (!s || !s->capturing) return hipErrorInvalidValue;
  s->capturing = false;
  *g = new shim::Graph{std::move(s->recorded)};
  s->recorded.clear();
  return hipSuccess;
}
inline hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, void*, void*, size_t) {
  *e = new shim::Graph{g->nodes};
  return hipSuccess;
}
inline hipError_t hipGraphDestroy(hipGraph_t g) {
  delete g;
  return hipSuccess;
}
inline hipError_t hipGraphExecDestroy(hipGraphExec_t e) {
  delete e;
  return hipSuccess;
}
inline hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s) {
  for (const auto& n : e->nodes) {
    if (s && s->capturing)
      s->recorded.push_back(n);
    else
      n();
  }
  return hipSuccess;
}
