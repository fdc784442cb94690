// hip_cpu_shim.h -- test infrastructure: just enough of HIP to compile the
// batched-fit kernels (mdff_amd/csrc/fit.hip) for the host and run a
// workgroup as std::threads meeting at a std::barrier.  Only tests/test_fit_cpu.py
// builds and loads it, to check the fit's forward and adjoint gradient on the
// CPU; the library itself always runs on the GPU.
#pragma once
#include <atomic>
#include <ucontext.h>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>
#include <algorithm>
#include <string>
#include <cstdint>
#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(x)
#define __shared__
struct dim3 { unsigned x, y, z; dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {} };
inline thread_local dim3 threadIdx, blockIdx;
inline dim3 blockDim, gridDim;
// Lanes of a workgroup run as user-level contexts (ucontext) on one OS
// thread: __syncthreads switches back to the workgroup's scheduler, which
// resumes every lane in turn, so a barrier costs a context switch per lane.
// Workgroups run in parallel on a few OS threads.
struct EmuGroup {
  ucontext_t sched;
  std::vector<ucontext_t> ctx;
  std::vector<char> done;
  int cur = 0;
};
inline thread_local EmuGroup* g_grp = nullptr;
inline void __syncthreads() { swapcontext(&g_grp->ctx[g_grp->cur], &g_grp->sched); }

inline thread_local double* g_shared = nullptr;   // the workgroup's LDS
struct int4 { int x, y, z, w; };
inline int4 make_int4(int x, int y, int z, int w) { return int4{x, y, z, w}; }
typedef int hipError_t;
constexpr int hipSuccess = 0;
typedef void* hipStream_t;
#define hipStreamNonBlocking 1
inline hipError_t hipMalloc(void** p, size_t n) { *p = calloc(1, n ? n : 1); return 0; }
inline hipError_t hipFree(void* p) { free(p); return 0; }
inline hipError_t hipMemcpy(void* d, const void* s, size_t n, int) { memcpy(d, s, n); return 0; }
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, int, hipStream_t) { memcpy(d, s, n); return 0; }
inline hipError_t hipMemset(void* d, int v, size_t n) { memset(d, v, n); return 0; }
inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, int) { *s = (void*)1; return 0; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return 0; }
inline hipError_t hipStreamDestroy(hipStream_t) { return 0; }
inline hipError_t hipGetDeviceCount(int* n) { *n = 1; return 0; }
struct hipDeviceProp_t { int multiProcessorCount; };
inline hipError_t hipGetDeviceProperties(hipDeviceProp_t* p, int) { p->multiProcessorCount = 4; return 0; }
inline hipError_t hipGetLastError() { return 0; }
inline const char* hipGetErrorString(hipError_t) { return "emu"; }
#define hipMemcpyHostToDevice 1
#define hipMemcpyDeviceToHost 2
#define hipMemcpyDeviceToDevice 3
inline void sincospi(double x, double* s, double* c) { *s = sin(M_PI * x); *c = cos(M_PI * x); }
template <typename F>
struct EmuLane {
  static void entry(unsigned lo, unsigned hi) {
    F* fn = reinterpret_cast<F*>(((uintptr_t)hi << 32) | (uintptr_t)lo);
    (*fn)();
    g_grp->done[g_grp->cur] = 1;
  }
};

// a launch: workgroups shared out over up to 8 OS threads; a workgroup's
// lanes are contexts with 128 KiB stacks, resumed round-robin until all end;
// fn is the kernel call with its arguments
template <typename F>
void emu_launch_f(F fn, dim3 g, dim3 b) {
  gridDim = g; blockDim = b;
  std::atomic<unsigned> next{0};
  auto work = [&]() {
    constexpr size_t kStack = 128 * 1024;
    std::vector<char> stacks((size_t)b.x * kStack);
    std::vector<double> lds(16384);
    EmuGroup grp;
    grp.ctx.resize(b.x);
    grp.done.resize(b.x);
    g_grp = &grp;
    g_shared = lds.data();
    const unsigned nblocks = g.x * g.y * g.z;
    for (unsigned bx; (bx = next.fetch_add(1)) < nblocks;) {
      blockIdx = dim3(bx % g.x, (bx / g.x) % g.y, bx / (g.x * g.y));
      const uintptr_t p = reinterpret_cast<uintptr_t>(&fn);
      for (unsigned t = 0; t < b.x; ++t) {
        getcontext(&grp.ctx[t]);
        grp.ctx[t].uc_stack.ss_sp = stacks.data() + (size_t)t * kStack;
        grp.ctx[t].uc_stack.ss_size = kStack;
        grp.ctx[t].uc_link = &grp.sched;
        makecontext(&grp.ctx[t], (void (*)())&EmuLane<F>::entry, 2, (unsigned)(p & 0xffffffffu),
                    (unsigned)(p >> 32));
        grp.done[t] = 0;
      }
      for (unsigned left = b.x; left > 0;) {
        left = 0;
        for (unsigned t = 0; t < b.x; ++t) {
          if (grp.done[t]) continue;
          grp.cur = (int)t;
          threadIdx = dim3(t, 0, 0);
          swapcontext(&grp.sched, &grp.ctx[t]);
          if (!grp.done[t]) ++left;
        }
      }
    }
  };
  const unsigned nth = std::min(8u, g.x);
  std::vector<std::thread> th;
  for (unsigned i = 1; i < nth; ++i) th.emplace_back(work);
  work();
  for (auto& x : th) x.join();
}
#define hipLaunchKernelGGL(K, G, B, SHM, ST, ...) emu_launch_f([&]() { K(__VA_ARGS__); }, dim3(G), dim3(B))
using std::min; using std::max;
inline double nint_d(double x) { return x >= 0.0 ? floor(x + 0.5) : -floor(0.5 - x); }
#define MDFF_PI 3.14159265358979323846264338327950
template <typename T>
inline T errfc_pos(T z) {
  T t = T(1) / (T(1) + T(0.5) * z);
  T p = t * (T(1.48851587) + t * (T(-.82215223) + t * T(.17087277)));
  p = t * (T(-1.13520398) + p); p = t * (T(.27886807) + p); p = t * (T(-.18628806) + p);
  p = t * (T(.09678418) + p); p = t * (T(.37409196) + p); p = t * (T(1.00002368) + p);
  return t * std::exp(-z * z - T(1.26551223) + p);
}
template <typename T>
inline void tt_damp(T b, T c, T r, int order, T& f, T& fd) {
  const T E[9] = {T(1), T(1), T(0.5), T(1.0 / 6.0), T(1.0 / 24.0), T(1.0 / 120.0), T(1.0 / 720.0), T(1.0 / 5040.0), T(1.0 / 40320.0)};
  T br = b * r; T ex = std::exp(-br) * c; T ff = E[order];
  for (int k = order - 1; k >= 1; --k) ff = ff * br + E[k];
  ff = ff * br + E[0]; f = T(1) - ff * ex;
  T brn = T(1), a = br;
  for (int k = order;;) { if (k & 1) brn *= a; k >>= 1; if (k == 0) break; a *= a; }
  fd = E[order] * brn * ex * b;
}
#include "mdff_gpu.h"
