// CPU emulation of the subset of HIP used by cm_kernels.hip: every GPU thread
// is a ucontext fiber; blocks run one at a time; __syncthreads and shuffles are
// block-wide cooperative barriers.  Device memory is host memory (malloc), so
// AddressSanitizer sees every out-of-bounds access.
#pragma once
#include <ucontext.h>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <vector>
#include <algorithm>
using std::min;
using std::max;
using std::isfinite;
inline int __clz(int x) { return x == 0 ? 32 : __builtin_clz((unsigned)x); }

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __align__(n) alignas(n)
#define __restrict__ __restrict
#define __shared__ static

struct alignas(8) float2 { float x, y; };
struct alignas(16) double2 { double x, y; };
struct alignas(2) uchar2 { unsigned char x, y; };
struct alignas(16) float4 { float x, y, z, w; };
inline float2 make_float2(float a, float b) { return float2{a, b}; }
inline double2 make_double2(double a, double b) { return double2{a, b}; }
inline uchar2 make_uchar2(unsigned char a, unsigned char b) { return uchar2{a, b}; }

struct dim3 {
  unsigned x, y, z;
  dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};

namespace emu {
struct Fiber {
  ucontext_t ctx;
  std::vector<char> stack;
  unsigned tid = 0;
  int state = 0;  // 0 running, 1 at barrier, 2 done
  double shfl = 0;
};
struct Block {
  std::vector<Fiber> f;
  ucontext_t sched;
  int cur = -1;
  unsigned n = 0;
  std::function<void()> body;
};
Block& blk();
extern dim3 g_grid, g_block, g_bidx;
extern unsigned char* g_dyn;
extern size_t g_dyn_bytes;
inline Fiber& me() { return blk().f[blk().cur]; }
void barrier();
template <typename V>
V shfl_xor(V v, int off, int width) {
  Fiber& m = me();
  std::memcpy(&m.shfl, &v, sizeof(V));
  barrier();
  const unsigned lane = m.tid & 63u, base = m.tid & ~63u;
  const unsigned src = base + (lane ^ (unsigned)off);
  V r;
  if (src >= blk().n) {
    const double nan = std::nan("");
    std::memcpy(&r, &nan, sizeof(V) < sizeof(double) ? sizeof(V) : sizeof(double));
    if (getenv("EMU_WARN_SHFL")) fprintf(stderr, "emu: shuffle from inactive lane %u (block of %u)\n", src, blk().n);
  } else {
    std::memcpy(&r, &blk().f[src].shfl, sizeof(V));
  }
  barrier();
  (void)width;
  return r;
}
// The value v of lane src of this thread's wavefront (every thread of the
// block calls it, as with a wave-wide instruction).
template <typename V>
V lane_read(V v, unsigned src_lane) {
  Fiber& m = me();
  std::memcpy(&m.shfl, &v, sizeof(V));
  barrier();
  const unsigned src = (m.tid & ~63u) + src_lane;
  V r;
  if (src >= blk().n) {
    const double nan = std::nan("");
    std::memcpy(&r, &nan, sizeof(V) < sizeof(double) ? sizeof(V) : sizeof(double));
  } else {
    std::memcpy(&r, &blk().f[src].shfl, sizeof(V));
  }
  barrier();
  return r;
}
struct U2 {
  int v[2];
  int operator[](int i) const { return v[i]; }
};
void run_block(std::function<void()> body, unsigned nthreads);
}  // namespace emu

// Cross-lane builtins of gfx950 used by the wave reductions, as the ISA
// documents them: DPP quad permutes and row mirrors (row_mask and bank_mask
// all set), and the permlane16/32 swaps of odd rows of the first operand
// with even rows of the second.
inline int __builtin_amdgcn_update_dpp(int, int src, int ctrl, int, int, bool) {
  const unsigned i = emu::me().tid & 63u;
  unsigned s;
  if (ctrl <= 0xFF) s = (i & ~3u) + ((ctrl >> (2 * (i & 3u))) & 3u);
  else if (ctrl == 0x140) s = (i & ~15u) + (15u - (i & 15u));
  else if (ctrl == 0x141) s = (i & ~7u) + (7u - (i & 7u));
  else { fprintf(stderr, "emu: DPP control %#x not emulated\n", ctrl); abort(); }
  return emu::lane_read(src, s);
}
inline emu::U2 __builtin_amdgcn_permlane16_swap(int old, int src, bool, bool) {
  const unsigned i = emu::me().tid & 63u;
  const int src_lo = emu::lane_read(src, i >= 16 ? i - 16 : 0);   // src of the row below
  const int old_hi = emu::lane_read(old, i + 16 < 64 ? i + 16 : 63);  // old of the row above
  const bool odd = (i & 16u) != 0;
  return emu::U2{{odd ? src_lo : old, odd ? src : old_hi}};
}
inline emu::U2 __builtin_amdgcn_permlane32_swap(int old, int src, bool, bool) {
  const unsigned i = emu::me().tid & 63u;
  const int src_lo = emu::lane_read(src, i >= 32 ? i - 32 : 0);
  const int old_hi = emu::lane_read(old, i + 32 < 64 ? i + 32 : 63);
  const bool up = (i & 32u) != 0;
  return emu::U2{{up ? src_lo : old, up ? src : old_hi}};
}
inline int __double2loint(double d) {
  long long b;
  std::memcpy(&b, &d, 8);
  return (int)(unsigned)(b & 0xffffffffLL);
}
inline int __double2hiint(double d) {
  long long b;
  std::memcpy(&b, &d, 8);
  return (int)(b >> 32);
}
inline double __hiloint2double(int hi, int lo) {
  const unsigned long long b = ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
  double d;
  std::memcpy(&d, &b, 8);
  return d;
}

struct EmuIdx {
  unsigned x, y, z;
};
#define threadIdx (EmuIdx{emu::me().tid, 0u, 0u})
#define blockIdx (emu::g_bidx)
#define blockDim (emu::g_block)
#define gridDim (emu::g_grid)

inline void __syncthreads() { emu::barrier(); }
inline void __threadfence() {}
#define __builtin_amdgcn_fence(...) ((void)0)
template <typename V>
inline V __shfl_xor(V v, int off, int width = 64) { return emu::shfl_xor(v, off, width); }
inline unsigned atomicAdd(unsigned* p, unsigned v) {
  unsigned o = *p;
  *p = o + v;
  return o;
}
inline unsigned long long atomicAdd(unsigned long long* p, unsigned long long v) {
  unsigned long long o = *p;
  *p = o + v;
  return o;
}
#define __ATOMIC_RELAXED_EMU 0
#define __HIP_MEMORY_SCOPE_AGENT 0
template <typename A, typename B>
inline void __hip_atomic_store(A* p, B v, int, int) { *p = (A)v; }

// ---- runtime API
typedef int hipError_t;
enum { hipSuccess = 0, hipErrorNotSupported = 801, hipErrorInvalidValue = 1 };
typedef struct emu_stream* hipStream_t;
typedef struct emu_event* hipEvent_t;
typedef struct emu_graph* hipGraph_t;
typedef struct emu_graphexec* hipGraphExec_t;
enum hipMemcpyKind { hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice, hipMemcpyDefault };
enum { hipStreamNonBlocking = 1 };
enum hipStreamCaptureMode { hipStreamCaptureModeThreadLocal = 1 };
enum hipFuncAttribute { hipFuncAttributeMaxDynamicSharedMemorySize = 8 };
enum { hipHostMallocDefault = 0 };

inline hipError_t hipMalloc(void** p, size_t n) { *p = malloc(n ? n : 16); return *p ? hipSuccess : 2; }
template <typename T> inline hipError_t hipMalloc(T** p, size_t n) { return hipMalloc((void**)p, n); }
inline hipError_t hipFree(void* p) { free(p); return hipSuccess; }
inline hipError_t hipHostMalloc(void** p, size_t n, unsigned = 0) { *p = malloc(n ? n : 16); return hipSuccess; }
template <typename T> inline hipError_t hipHostMalloc(T** p, size_t n, unsigned f = 0) { return hipHostMalloc((void**)p, n, f); }
inline hipError_t hipHostFree(void* p) { free(p); return hipSuccess; }
inline hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) { if (n) memmove(d, s, n); return hipSuccess; }
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t) { if (n) memmove(d, s, n); return hipSuccess; }
inline hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t = nullptr) { memset(d, v, n); return hipSuccess; }
inline hipError_t hipMemset(void* d, int v, size_t n) { memset(d, v, n); return hipSuccess; }
inline hipError_t hipStreamCreate(hipStream_t* s) { *s = (hipStream_t)malloc(8); return hipSuccess; }
inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) { return hipStreamCreate(s); }
inline hipError_t hipStreamDestroy(hipStream_t s) { free(s); return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipDeviceSynchronize() { return hipSuccess; }
inline hipError_t hipEventCreate(hipEvent_t* e) { *e = (hipEvent_t)malloc(8); return hipSuccess; }
inline hipError_t hipEventDestroy(hipEvent_t e) { free(e); return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t = nullptr) { return hipSuccess; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventQuery(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) { *ms = 0.001f; return hipSuccess; }
inline hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
inline hipError_t hipSetDevice(int) { return hipSuccess; }
inline hipError_t hipGetDevice(int* d) { *d = 0; return hipSuccess; }
inline hipError_t hipGetDeviceCount(int* n) { *n = 1; return hipSuccess; }
inline const char* hipGetErrorString(hipError_t) { return "emu error"; }
extern hipError_t emu_last_error;
inline hipError_t hipGetLastError() { hipError_t e = emu_last_error; emu_last_error = 0; return e; }
inline hipError_t hipPeekAtLastError() { return emu_last_error; }
inline hipError_t hipFuncSetAttribute(const void*, hipFuncAttribute, int) { return hipSuccess; }
#define HIP_SYMBOL(x) ((const void*)&(x))
inline hipError_t hipGetSymbolAddress(void** p, const void* sym) { *p = (void*)sym; return hipSuccess; }
inline hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) { return hipErrorNotSupported; }
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t*) { return hipErrorNotSupported; }
inline hipError_t hipGraphInstantiate(hipGraphExec_t*, hipGraph_t, void*, void*, size_t) { return hipErrorNotSupported; }
inline hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return hipErrorNotSupported; }
inline hipError_t hipGraphDestroy(hipGraph_t) { return hipSuccess; }
inline hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }

namespace emu {
template <typename K, typename... A>
void launch(K kern, dim3 grid, dim3 block, size_t shmem, hipStream_t, A... args) {
  const unsigned nt = block.x * block.y * block.z;
  if (nt == 0 || nt > 1024 || shmem > 160 * 1024 || grid.x == 0 || grid.y == 0) {
    fprintf(stderr, "emu: invalid launch grid (%u,%u) block %u lds %zu\n", grid.x, grid.y, nt, shmem);
    emu_last_error = hipErrorInvalidValue;
    abort();
  }
  g_grid = grid;
  g_block = block;
  for (unsigned by = 0; by < grid.y; ++by)
    for (unsigned bx = 0; bx < grid.x; ++bx) {
      g_bidx = dim3(bx, by, 0);
      g_dyn = shmem ? (unsigned char*)malloc(shmem) : nullptr;
      g_dyn_bytes = shmem;
      if (g_dyn) memset(g_dyn, 0xA5, shmem);
      run_block([&]() { kern(args...); }, nt);
      free(g_dyn);
      g_dyn = nullptr;
    }
}
}  // namespace emu
#define hipLaunchKernelGGL(kern, grid, block, shmem, stream, ...) \
  emu::launch(kern, dim3(grid), dim3(block), (size_t)(shmem), stream, __VA_ARGS__)
