// gfx950 kernels and C ABI of the contact solver (libcmhip.so).
//
// Convolution u = G (*) f on an (Mx, My) real grid, My even, runs as three
// passes over HBM:
//   1. row pass  (k_r2c): real rows -> half-length complex FFT in LDS ->
//      real-to-complex split -> spectrum S in a column-blocked layout;
//   2. column pass (k_col): forward FFT along x, multiply by the real Green's
//      function G(kx, ky), inverse FFT, all inside LDS;
//   3. row pass  (k_c2r): complex-to-real merge -> inverse FFT -> real rows.
// Row passes take a "producer"/"consumer" functor so the CCG element-wise work
// and its FP64 reductions are fused into the loads/stores of the transforms.
//
// Spectrum layout: S holds Mx x (My/2) complex values.  Column ky = 0 packs the
// real DC (ky = 0) and Nyquist (ky = My/2) columns as one complex column.
// Element (x, ky) lives at ((ky / Cb) * Mx + x) * Cb + ky % Cb: each column
// block of Cb columns is contiguous, so the column pass reads and writes one
// contiguous slab and the row pass writes R*Cb-element contiguous runs.
//
// Reference: ContactMechanics/FFTElasticHalfSpace.py (operators) and
// ContactMechanics/Optimization/ConstrainedConjugateGradients.py (solver);
// each CCG phase below cites the lines it implements.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/cmhip.h"
#include "cm_fft.hpp"

namespace cm {

// ---------------------------------------------------------------------------
// Geometry and per-instance state

struct Geom {
  int nx, ny;   // topography grid
  int Mx, My;   // computational grid
  int N2;       // My / 2 (complex row length)
  int Cb;       // spectrum column-block width
  int R;        // rows per row-pass block
};

enum { DS_RESET = 0, DS_SD = 1, DS_CG = 2, DS_MIX = 3 };
enum { MODE_PK = 0, MODE_MIX = 1 };

// Scalar state of one CCG solve (device resident).  Doubles first.
struct Scal {
  // parameters
  double F, H, pentol, forcetol, thermotol, mixfac, mixdecfac;
  // constants from set-up
  double max_h, nb_surface_pts, n_pad;
  // iteration state
  double offset, shift, G, G_old, beta, tau, mix, last_tf, last_mh;
  // per-iteration values
  double A_contact, A_cg, A_fl, Scgt, gmax, gmin, max_pres, pad_pres, scale, total_force;
  double nc_count, maxdu, max_pen, rms_pen, fun;
  int load_ctl, has_H, H_truthy, minmixsteps, maxiter;
  int it, nfev, delta, delta_str, mode, mixsteps, done, status;
  int has_last_tf, has_last_mh, use_beta, do_nc, uniform, pad_;
  // batched bisection across launches: bracket end xa, step dm, f(xa), steps
  // taken, state (1 running, 2 finished)
  double bis_xa, bis_dm, bis_fa;
  int bis_steps, bis_state;
};

// One instance of a solve (or of a plain operator evaluation).  T is the
// precision of the plan: the caller's displacement and force buffers, the
// conjugate direction and the transforms of the iteration.  During a solve
// the forces and the gap are kept in FP64 (fm, w); FP64 plans keep them in
// the caller's buffers themselves.
template <typename T>
struct Inst {
  T* u;            // displacements (M grid), the caller's buffer
  double* w;       // during a solve: the gap u - h on the computational
                   // points, u elsewhere (see ElemPhaseA); FP64 plans: == u
  T* f;            // forces (M grid), the caller's buffer
  double* fm;      // during a solve: the forces in FP64; FP64 plans: == f
  T* t;            // conjugate direction (M grid)
  uint8_t* cm;     // bit0 contact set c, bit1 gap<0 at start of iteration
  const double* h; // heights (nx*ny)
  const uint8_t* surf;  // valid-pixel mask (nx*ny) or null
  cplx<T>* S;      // spectrum work buffer
  const T* G;      // evaluate_disp kernel, (N2+1) x Mx
  const T* K;      // evaluate_force kernel, (N2+1) x Mx
  // FP64 transforms of an FP32 plan (set-up and the periodic displacement
  // refresh); FP64 plans: the same buffers as S, G, K
  cplx<double>* S64;
  const double* G64;
  const double* K64;
  const T* src;    // operator input (plain evaluate)
  T* dst;          // operator output (plain evaluate)
  Scal* sc;
  double* part;    // per-block partial reductions
  unsigned* ctr;   // last-block tickets (8 slots of kCtrSlot counters)
};

// Spectrum buffer and Green's function for transforms in precision TT of an
// instance with fields in precision TF.
template <typename TT, typename TF>
__device__ __forceinline__ cplx<TT>* spec_of(const Inst<TF>& in) {
  if constexpr (std::is_same<TT, TF>::value)
    return in.S;
  else
    return in.S64;
}
template <typename TT, typename TF>
__device__ __forceinline__ const TT* green_of(const Inst<TF>& in, int useK) {
  if constexpr (std::is_same<TT, TF>::value)
    return useK ? in.K : in.G;
  else
    return useK ? in.K64 : in.G64;
}

template <typename T>
struct InstArr {
  const Inst<T>* arr;  // device array (batched) or null
  Inst<T> one;         // single instance by value
  __device__ __forceinline__ const Inst<T>& get(int b) const { return arr ? arr[b] : one; }
};

__device__ __forceinline__ bool is_comp(const Geom& g, const uint8_t* surf, int i, int j) {
  return i < g.nx && j < g.ny && (surf == nullptr || surf[i * g.ny + j] != 0);
}

// ---------------------------------------------------------------------------
// Wait until this wave's outstanding memory operations (here: the
// write-through partial stores, issued as flat stores, which count in both
// counters) have completed.
__device__ __forceinline__ void wait_stores_done() { asm volatile("s_waitcnt vmcnt(0) lgkmcnt(0)" ::: "memory"); }
// Agent-scope acquire: invalidates this CU's view of L2 so later plain loads
// see data written through by other XCDs.
__device__ __forceinline__ void acquire_agent() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }

// Tickets are taken in kTicketGroups groups (block b in group b % groups),
// each counter on its own 64-byte line: the block completing a group takes a
// ticket on the top counter.  Same-address atomics serialise (about 6 ns each
// on gfx950), so one counter for 1-2k blocks would add 6-12 us per kernel.
constexpr int kTicketGroups = 16;
constexpr int kCtrLine = 16;                             // unsigned per 64-byte line
constexpr int kCtrSlot = kCtrLine * (kTicketGroups + 1); // counters of one reduction slot

// Records of the last kLogRing iterations kept per instance behind its scalar
// record (the log ring): a chunk of up to kLogRing iterations runs with no
// host check in between, and the host still reports every iteration.
constexpr int kLogRing = 64;

// FP32 plans: default number of iterations per FP64 displacement refresh.
constexpr int kRefreshEvery = 8;

// A NaN that a thread's kmax/kmin remembered in a max or min slot's
// correction term becomes the slot's value (red_apply's result), and the
// term goes back to zero (the cross-thread reduction carries no correction
// for such slots).
template <class P>
__device__ __forceinline__ void fold_nan(double* acc) {
  constexpr int NS = P::NS;
#pragma unroll
  for (int i = 0; i < NS; ++i)
    if (P::op(i) == RMAX || P::op(i) == RMIN) {
      if (acc[NS + i] != acc[NS + i]) acc[i] = acc[NS + i];
      acc[NS + i] = 0.0;
    }
}

// Ticket of block bid of nblocks, taken by one thread once the block's
// partials are stored and complete: true in the block that takes the last
// ticket (that block resets the top counter when it is done with the
// partials; the group counters reset themselves).
__device__ __forceinline__ bool take_ticket(unsigned* ctr, int nblocks, int bid) {
  const int g = bid % kTicketGroups;
  const int gsize = nblocks / kTicketGroups + (g < nblocks % kTicketGroups ? 1 : 0);
  const int ngroups = nblocks < kTicketGroups ? nblocks : kTicketGroups;
  unsigned* grp = ctr + kCtrLine * (1 + g);
  bool last = false;
  if (atomicAdd(grp, 1u) == (unsigned)(gsize - 1)) {  // this block completes its group
    __hip_atomic_store(grp, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    last = atomicAdd(ctr, 1u) == (unsigned)(ngroups - 1);
  }
  return last;
}

// Cross-block reduction: every block reduces NS slots, writes its partials,
// takes a ticket; the block drawing the last ticket reduces all partials in a
// fixed order (deterministic result) and returns true.
//
// No device-scope release fence per block: on gfx950 such a fence writes back
// and invalidates the XCD's whole L2, once per block.  Instead the partials are
// stored with agent-scope (L2 write-through) atomic stores, the wave waits for
// them to complete (s_waitcnt vmcnt(0)) before it takes its ticket, and only
// the last block invalidates its L2 (agent-scope acquire) before reading them.
//
// acc holds 2 NS doubles: the slot values, then the correction terms of the
// compensated sums (RCSUM); a compensated slot's partial is the pair.
template <class P>
__device__ __forceinline__ bool reduce_blocks(double* acc, double* part, unsigned* ctr, int nblocks, int bid) {
  constexpr int NS = P::NS;
  __shared__ double scr[2 * 8 * 16];
  __shared__ int s_last;
  int ops[NS > 0 ? NS : 1];
#pragma unroll
  for (int i = 0; i < NS; ++i) ops[i] = P::op(i);
  block_reduce<NS>(acc, acc + NS, ops, scr);
  // partials are slot-major, part[i * nblocks + b] (corrections of slot i at
  // part[(NS + i) * nblocks + b]), so the last block reads each slot with
  // coalesced loads
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NS; ++i) {
      __hip_atomic_store(&part[(size_t)i * nblocks + bid], acc[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (ops[i] == RCSUM)
        __hip_atomic_store(&part[(size_t)(NS + i) * nblocks + bid], acc[NS + i], __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
    wait_stores_done();
    s_last = take_ticket(ctr, nblocks, bid);
  }
  __syncthreads();
  if (!s_last) return false;
  acquire_agent();
  double v[NS > 0 ? NS : 1], vc[NS > 0 ? NS : 1];
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    v[i] = red_identity(ops[i]);
    vc[i] = 0.0;
  }
  // every slot of kB blocks per thread loaded together (one round trip for
  // up to kB * blockDim blocks), then combined in block order; plain loads,
  // the acquire above makes every block's partials visible.  kB * NS stays
  // at most 8 values (16 with corrections) in registers.
  constexpr int kB = NS <= 2 ? 4 : (NS <= 4 ? 2 : 1);
  for (int b0 = threadIdx.x; b0 < nblocks; b0 += kB * blockDim.x) {
    double w[kB][NS > 0 ? NS : 1], wc[kB][NS > 0 ? NS : 1];
#pragma unroll
    for (int k = 0; k < kB; ++k) {
      const int b = b0 + k * (int)blockDim.x;
#pragma unroll
      for (int i = 0; i < NS; ++i) {
        const bool cs = ops[i] == RCSUM;
        w[k][i] = b < nblocks ? part[(size_t)i * nblocks + b] : red_identity(ops[i]);
        wc[k][i] = (cs && b < nblocks) ? part[(size_t)(NS + i) * nblocks + b] : 0.0;
      }
    }
#pragma unroll
    for (int k = 0; k < kB; ++k) {
#pragma unroll
      for (int i = 0; i < NS; ++i) {
        if (ops[i] == RCSUM)
          kcombine(v[i], vc[i], w[k][i], wc[k][i]);
        else
          v[i] = red_apply(ops[i], v[i], w[k][i]);
      }
    }
  }
  block_reduce<NS>(v, vc, ops, scr);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NS; ++i) {
      acc[i] = v[i];
      acc[NS + i] = vc[i];
    }
    __hip_atomic_store(ctr, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  return true;
}

// ---------------------------------------------------------------------------
// Per-point functors work on pairs of horizontally adjacent points (col even):
//   Rec fetch(g, row, col)                 loads only;
//   make(g, row, col, rec, a, b, acc)      producer: outputs + FFT input pair;
//   take(g, row, col, rec, a, b, acc, cx)  consumer of an FFT output pair
//                                          (cx = ctx(): scalars read once per thread);
//   emit(g, row, col, rec, acc)            element-wise kernel body.
// Kernels fetch a batch of pairs per thread (kPairs in the element-wise
// kernel, a functor's kBatch in the row passes) before computing any, so the
// loads are in flight together instead of waiting behind stores.
constexpr int kPairs = 4;

// Field pairs are accessed as aligned 2-vectors in the global address space
// (instance-table pointers would otherwise compile to flat accesses, which the
// compiler also counts against LDS traffic).
template <typename T>
using Gv2 = T __attribute__((ext_vector_type(2)));

template <typename T>
struct P2 {
  T a, b;
};
template <typename T>
__device__ __forceinline__ P2<T> ld2(const T* p) {
  const Gv2<T> v = *(const __attribute__((address_space(1))) Gv2<T>*)p;
  return P2<T>{v[0], v[1]};
}
template <typename T>
__device__ __forceinline__ void st2(T* p, T a, T b) {
  Gv2<T> v;
  v[0] = a;
  v[1] = b;
  *(__attribute__((address_space(1))) Gv2<T>*)p = v;
}
// Contact-mask bytes of a pair, kept as the loaded 16-bit word until used:
// splitting it right after the load makes the compiler wait for that load,
// and with in-order memory counters that drains every load still in flight,
// so a batch of point pairs would go to memory one pair at a time.
struct M2 {
  uint16_t v;
  __device__ __forceinline__ uint8_t a() const { return (uint8_t)(v & 0xff); }
  __device__ __forceinline__ uint8_t b() const { return (uint8_t)(v >> 8); }
};
__device__ __forceinline__ M2 ldm(const uint8_t* p) {
  return M2{*(const __attribute__((address_space(1))) uint16_t*)p};
}
__device__ __forceinline__ void stm(uint8_t* p, uint8_t a, uint8_t b) {
  *(__attribute__((address_space(1))) uint16_t*)p = (uint16_t)(a | (b << 8));
}

// Heights and validity of the pair (row, col), (row, col+1).
struct HPair {
  double h0, h1;
  bool c0, c1;  // in the computational region
};
// Validity of the pair (row, col), (row, col+1): in the topography block and
// not masked out.  Masked grids read the mask at a clamped, always valid
// index, so no lane-dependent branch guards the load.
struct CPair {
  bool c0, c1;
};
__device__ __forceinline__ CPair ldc(const Geom& g, const uint8_t* surf, int row, int col) {
  CPair r;
  r.c0 = row < g.nx && col < g.ny;
  r.c1 = row < g.nx && col + 1 < g.ny;
  if (surf) {
    const auto m = (const __attribute__((address_space(1))) uint8_t*)surf;
    const size_t q = (size_t)row * g.ny + col;
    const uint8_t m0 = m[r.c0 ? q : 0], m1 = m[r.c1 ? q + 1 : 0];
    r.c0 = r.c0 && m0;
    r.c1 = r.c1 && m1;
  }
  return r;
}

// Heights and validity of the pair (row, col), (row, col+1); heights read at
// clamped indices (no divergent branch) and zero off the valid points.
template <typename T>
__device__ __forceinline__ HPair ldh(const Geom& g, const T* h, const uint8_t* surf, int row, int col) {
  const CPair c = ldc(g, surf, row, col);
  const auto hg = (const __attribute__((address_space(1))) T*)h;
  const size_t q = (size_t)row * g.ny + col;
  const double h0 = (double)hg[c.c0 ? q : 0], h1 = (double)hg[c.c1 ? q + 1 : 0];
  HPair r;
  r.c0 = c.c0;
  r.c1 = c.c1;
  r.h0 = c.c0 ? h0 : 0.0;
  r.h1 = c.c1 ? h1 : 0.0;
  return r;
}

// Global-memory access to complex arrays that reach kernels through instance
// tables: plain (generic) pointers compile to flat loads, which the compiler
// must also count against LDS traffic, so every LDS access would wait for
// them.  The views address the real and imaginary parts as scalars.
template <typename T>
struct GView {
  const __attribute__((address_space(1))) T* p;
  __device__ explicit GView(const cplx<T>* q) : p((const __attribute__((address_space(1))) T*)q) {}
  __device__ __forceinline__ cplx<T> operator[](size_t i) const { return mk(p[2 * i], p[2 * i + 1]); }
};
template <typename T>
struct GViewW {
  __attribute__((address_space(1))) T* p;
  __device__ explicit GViewW(cplx<T>* q) : p((__attribute__((address_space(1))) T*)q) {}
  __device__ __forceinline__ void put(size_t i, cplx<T> v) const {
    p[2 * i] = v.x;
    p[2 * i + 1] = v.y;
  }
};

// Slab copies between HBM and LDS with kCopyU loads (or stores) in flight per
// thread: a plain strided loop waits for each load before the next.
constexpr int kCopyU = 8;

// Contiguous slab: lds[pad(i)] <- S[base + i], i < n.
template <typename T>
__device__ __forceinline__ void slab_load(cplx<T>* lds, const cplx<T>* Sp, size_t base, int n) {
  const GView<T> S(Sp);
  for (int i0 = threadIdx.x; i0 < n; i0 += kCopyU * (int)blockDim.x) {
    // out-of-range slots repeat the last element (an identical, harmless
    // second write): no branches, so the values stay in registers (as
    // separate real/imaginary arrays, which the compiler keeps out of
    // private memory) and all loads are in flight together
    T tx[kCopyU], ty[kCopyU];
#pragma unroll
    for (int k = 0; k < kCopyU; ++k) {
      const cplx<T> z = S[base + min(i0 + k * (int)blockDim.x, n - 1)];
      tx[k] = z.x;
      ty[k] = z.y;
    }
#pragma unroll
    for (int k = 0; k < kCopyU; ++k) lds[lpad<T>(min(i0 + k * (int)blockDim.x, n - 1))] = mk(tx[k], ty[k]);
  }
}
template <typename T>
__device__ __forceinline__ void slab_store(cplx<T>* Sp, const cplx<T>* lds, size_t base, int n) {
  const GViewW<T> S(Sp);
  for (int i0 = threadIdx.x; i0 < n; i0 += kCopyU * (int)blockDim.x) {
#pragma unroll
    for (int k = 0; k < kCopyU; ++k) {
      const int i = i0 + k * (int)blockDim.x;
      if (i < n) S.put(base + i, lds[lpad<T>(i)]);
    }
  }
}

// Row-block <-> column-blocked spectrum.  Element i of the block's R rows x N2
// values: kk = i % Cb, rr = (i / Cb) % R, cb = i / (R Cb) (Cb, R powers of two);
// LDS holds row rr at rr * N2, the spectrum element (x = row, ky = cb Cb + kk).
struct RowBlockMap {
  int cbs, rs, Cb, R, N2, Mx, row0;
  __device__ RowBlockMap(const Geom& g, int blk)
      : cbs(__builtin_ctz(g.Cb)), rs(__builtin_ctz(g.R)), Cb(g.Cb), R(g.R), N2(g.N2), Mx(g.Mx), row0(blk * g.R) {}
  __device__ __forceinline__ size_t spec(int i) const {
    const int kk = i & (Cb - 1), rr = (i >> cbs) & (R - 1), cb = i >> (cbs + rs);
    return ((size_t)cb * Mx + row0 + rr) * Cb + kk;
  }
  __device__ __forceinline__ int lds(int i) const {
    const int kk = i & (Cb - 1), rr = (i >> cbs) & (R - 1), cb = i >> (cbs + rs);
    return rr * N2 + cb * Cb + kk;
  }
};

// Row block of a spectrum-reading row pass, XCD-contiguous: the hardware hands
// the workgroups of a launch to its 8 XCDs in turn (workgroup b to XCD b % 8),
// each XCD with its own L2.  Neighbouring row blocks read the two halves of
// the same 128-byte lines of the column-blocked spectrum (runs of Cb x R
// values), so this map gives each XCD one contiguous range of row blocks: the
// partner of a line then reads it from the same L2 instead of fetching the
// whole line again.  A bijection of [0, n) for any n.
__device__ __forceinline__ int xcd_row_block(int b, int n) {
  constexpr int kXcd = 8;
  const int q = n / kXcd, rem = n - q * kXcd, x = b % kXcd;
  return x * q + min(x, rem) + b / kXcd;
}

// The last block's scalar logic (thread 0) on a copy of the scalar record in
// LDS: the block loads the record with one load per thread and stores it
// back the same way, so the logic's dependent reads and writes of many fields
// cost LDS latency instead of one memory round trip each.  Called by every
// thread of the last block; when it runs, every other block of the kernel has
// read the record already (each reads it when it starts, before its ticket)
// and no kernel behind this one has started.  With RING, the record as the
// logic left it is also stored in the instance's log ring (kLogRing records
// behind the scalar record, slot it % kLogRing), so the host can report every
// iteration of a chunk.
template <bool RING = false, typename F>
__device__ __forceinline__ void with_scal_in_lds(Scal* gs, F&& logic) {
  __shared__ Scal ls;
  constexpr int nw = (int)(sizeof(Scal) / sizeof(unsigned));
  static_assert(sizeof(Scal) % sizeof(unsigned) == 0, "the record copies as 32-bit words");
  const auto src = (const __attribute__((address_space(1))) unsigned*)gs;
  unsigned* dst = reinterpret_cast<unsigned*>(&ls);
  for (int i = threadIdx.x; i < nw; i += blockDim.x) dst[i] = src[i];
  __syncthreads();
  if (threadIdx.x == 0) logic(&ls);
  __syncthreads();
  const auto out = (__attribute__((address_space(1))) unsigned*)gs;
  for (int i = threadIdx.x; i < nw; i += blockDim.x) out[i] = dst[i];
  if constexpr (RING) {
    const int slot = (ls.it % kLogRing + kLogRing) % kLogRing;
    const auto rg = (__attribute__((address_space(1))) unsigned*)(gs + 1 + slot);
    for (int i = threadIdx.x; i < nw; i += blockDim.x) rg[i] = dst[i];
  }
}
// Functors whose finalize runs the long scalar logic (P::kScalInLds) use the
// LDS copy; the others write a field or two directly.
template <class P, class = void>
struct scal_in_lds : std::false_type {};
template <class P>
struct scal_in_lds<P, std::void_t<decltype(P::kScalInLds)>> : std::bool_constant<P::kScalInLds> {};
// Functors that end an iteration (P::kLogsRing) also store the record in the log ring.
template <class P, class = void>
struct logs_ring : std::false_type {};
template <class P>
struct logs_ring<P, std::void_t<decltype(P::kLogsRing)>> : std::bool_constant<P::kLogsRing> {};

// The last block's finalize on the reduced slots (a copy, so the accumulator
// array itself never has its address taken and stays in registers: the
// functors' finalize is not necessarily inlined).  Every thread of the last
// block calls it.
template <class P>
__device__ __forceinline__ void finalize_block(const P& p, const double* acc, Scal* gs) {
  double fin[P::NS > 0 ? P::NS : 1];
#pragma unroll
  for (int i = 0; i < P::NS; ++i) fin[i] = P::op(i) == RCSUM ? kvalue(acc[i], acc[P::NS + i]) : acc[i];
  if constexpr (scal_in_lds<P>::value)
    with_scal_in_lds<logs_ring<P>::value>(gs, [&](Scal* s) { p.finalize(fin, s); });
  else if (threadIdx.x == 0)
    p.finalize(fin, gs);
}

// Largest block of an FFT kernel: 512 threads.  The lean 8-element builds
// (about 120 VGPRs) fit two such blocks per CU at four waves per SIMD; the
// 16/32-element builds (up to 256 VGPRs) one, and they serve the long lines
// (up to 8192 points, e.g. 8000 = 16*4*5^3) that need more than 256 threads.
__host__ __device__ constexpr int fft_block_max(int) { return 512; }
// Waves per SIMD the compiler must leave room for: four for the 8-element
// builds (two 512-thread blocks per CU at <= 128 VGPRs; without the bound the
// column pass took 153 VGPRs and three waves), one for the long-line builds.
__host__ __device__ constexpr int fft_min_waves(int emax) { return emax <= 8 ? 4 : 1; }

// Row pass: real rows -> packed half spectrum (forward, unnormalised).
template <typename T, int EMAX, class P, bool GEN = false>
__global__ void __launch_bounds__(fft_block_max(EMAX), fft_min_waves(EMAX)) k_r2c(Geom g, LinePlan lp, const cplx<T>* __restrict__ tw,
                                             const cplx<T>* __restrict__ twp, InstArr<typename P::F> ia,
                                             int ctr_slot) {
  extern __shared__ __align__(16) unsigned char smem_r2c[];
  cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_r2c);
  const Inst<typename P::F>& in = ia.get(blockIdx.y);
  P p(in);
  if (p.skip()) return;
  const int N2 = g.N2;
  const int r = threadIdx.x / lp.tpl, tl = threadIdx.x - r * lp.tpl;
  const int row = blockIdx.x * g.R + r;
  double acc[P::NS > 0 ? 2 * P::NS : 1];  // slot values, then corrections of compensated sums
#pragma unroll
  for (int i = 0; i < P::NS; ++i) {
    acc[i] = red_identity(P::op(i));
    acc[P::NS + i] = 0.0;
  }
  const int rb = r * N2;
  auto L = [&](int n) -> cplx<T>& { return lds[lpad<T>(rb + n)]; };
  // the producer feeds the first FFT pass directly (first-pass order)
  cplx<T> v[EMAX];
  {
    const int R0 = lp.radix(0);
    const int nv0 = nbutterflies(N2, R0, tl, lp.tpl) * R0;  // slots holding a pair
    const SlotMap sm(R0, N2, lp.tpl);
    constexpr int KB = P::kBatch;
    if (nv0 >= EMAX) {
      // every register slot holds a point pair: no per-slot branch between
      // the loads of a batch, so they all issue before the first is used
#pragma unroll
      for (int e0 = 0; e0 < EMAX; e0 += KB) {
        typename P::Rec rec[KB];
#pragma unroll
        for (int k = 0; k < KB; ++k) rec[k] = p.fetch(g, row, 2 * (tl + sm.off(e0 + k)));
#pragma unroll
        for (int k = 0; k < KB; ++k) {
          T a, b;
          p.make(g, row, 2 * (tl + sm.off(e0 + k)), rec[k], a, b, acc);
          v[e0 + k] = mk(a, b);
        }
      }
    } else {
#pragma unroll
      for (int e0 = 0; e0 < EMAX; e0 += KB) {
        typename P::Rec rec[KB];
#pragma unroll
        for (int k = 0; k < KB; ++k) {
          const int e = e0 + k;
          if (e < nv0) rec[k] = p.fetch(g, row, 2 * (tl + sm.off(e)));
        }
#pragma unroll
        for (int k = 0; k < KB; ++k) {
          const int e = e0 + k;
          if (e < nv0) {
            T a, b;
            p.make(g, row, 2 * (tl + sm.off(e)), rec[k], a, b, acc);
            v[e] = mk(a, b);
          }
        }
      }
    }
  }
  fft_line_x<T, false, EMAX, true, false, true, GEN>(v, lds, rb, 1, lp, tl, tw);
  // split the half-length transform into the real-input spectrum X[k],
  // k = 0..N2, packing X[0] (real) and X[N2] (real) into slot 0.
  for (int k = tl; k <= N2 / 2; k += lp.tpl) {
    if (k == 0) {
      cplx<T> z = L(0);
      L(0) = mk(z.x + z.y, z.x - z.y);
    } else {
      const int km = N2 - k;
      cplx<T> zk = L(k), zm = L(km);
      cplx<T> E = (zk + conj(zm)) * T(0.5);
      cplx<T> D = zk - conj(zm);
      cplx<T> O = mk(D.y * T(0.5), -D.x * T(0.5));
      cplx<T> B = cmul(twp[k], O);
      L(k) = E + B;
      if (km != k) L(km) = conj(E - B);
    }
  }
  __syncthreads();
  cplx<T>* S = spec_of<T>(in);
  const int tot = g.R * N2;
  const RowBlockMap map(g, blockIdx.x);
  const GViewW<T> Sg(S);
  for (int i0 = threadIdx.x; i0 < tot; i0 += kCopyU * (int)blockDim.x) {
#pragma unroll
    for (int k = 0; k < kCopyU; ++k) {
      const int i = i0 + k * (int)blockDim.x;
      if (i < tot) Sg.put(map.spec(i), lds[lpad<T>(map.lds(i))]);
    }
  }
  if constexpr (P::NS > 0) {
    fold_nan<P>(acc);
    if (reduce_blocks<P>(acc, in.part, in.ctr + ctr_slot * kCtrSlot, gridDim.x, blockIdx.x))
      finalize_block(p, acc, in.sc);
  }
}

// A column-pass thread's first-pass register slots straight from the
// column-blocked spectrum (column c of the block at base): every slot loads,
// slots past the thread's butterflies from an in-range element whose value
// nothing reads, so the loads carry no branch and all issue before the first
// is waited for.
template <typename T, int EMAX>
__device__ __forceinline__ void col_load_slots(cplx<T>* v, const cplx<T>* S, size_t base, int Cb, int c, int Mx,
                                               int tl, const SlotMap& sm) {
  const GView<T> Sg(S);
#pragma unroll
  for (int e = 0; e < EMAX; ++e) v[e] = Sg[base + (size_t)min(tl + sm.off(e), Mx - 1) * Cb + c];
}

// Column kernels load twiddles early (before each pass's barrier) only in
// FP32: in FP64 the extra registers would cost the 8-element column kernel
// its fourth wave per SIMD.
template <typename T> constexpr bool kColEarlyTw = sizeof(T) == 4;

__device__ void tau_finalize(Scal* s, double sum_rt);

// Reduction slots of the spectral tau pass (one sum).
struct TauRed {
  static constexpr int NS = 1;
  __device__ static constexpr int op(int) { return RCSUM; }
};

// Column pass.  MODE 0: forward * G * inverse; 1: forward only; 2: inverse
// only; 3: forward only, then sum_k G(k) |T(k)|^2 over the whole spectrum and
// the step length tau from it (no output).
// skip: 0 never, 1 when the solve is done, 2 when done or in a mixing step.
// T: transform precision; TF: field precision of the instances.
template <typename T, int EMAX, int MODE, typename TF, bool GEN = false>
__global__ void __launch_bounds__(fft_block_max(EMAX), fft_min_waves(EMAX)) k_col(Geom g, LinePlan lp, LinePlan rv, const cplx<T>* __restrict__ tw,
                                              InstArr<TF> ia, int useK, int skip) {
  extern __shared__ __align__(16) unsigned char smem_col[];
  cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_col);
  const Inst<TF>& in = ia.get(blockIdx.y);
  if (skip && in.sc) {
    if (in.sc->done) return;
    if (skip == 2 && in.sc->mode == MODE_MIX) return;
  }
  cplx<T>* S = spec_of<T>(in);
  const T* G = green_of<T>(in, useK);
  const int cb = blockIdx.x;
  const int Mx = g.Mx, Cb = g.Cb;
  const size_t base = (size_t)cb * Mx * Cb;
  const int n = Mx * Cb;
  if (MODE == 3) {
    // r = evaluate_disp(t) is only needed for sum_c r t (:287-292), and t
    // vanishes off the contact set, so by Parseval
    //   sum_x r(x) t(x) = sum_k G(k) |T(k)|^2
    // over the full spectrum: columns 0 < ky < My/2 count twice (Hermitian
    // half), the packed DC and Nyquist columns once each.
    double acc[2] = {0.0, 0.0};  // compensated sum
    const int c = threadIdx.x % Cb, tl = threadIdx.x / Cb;
    if (cb != 0) {
      const int nvl = nbutterflies(Mx, lp.rlast, tl, lp.tpl) * lp.rlast;
      const SlotMap sm0(lp.radix(0), Mx, lp.tpl), sml(lp.rlast, Mx, lp.tpl);
      cplx<T> v[EMAX];
      col_load_slots<T, EMAX>(v, S, base, Cb, c, Mx, tl, sm0);
      fft_line_x<T, false, EMAX, true, true, kColEarlyTw<T>, GEN>(v, lds, c, Cb, lp, tl, tw);
      const auto Gk = (const __attribute__((address_space(1))) T*)(G + (size_t)(cb * Cb + c) * Mx);
      T gk[EMAX];
#pragma unroll
      for (int e = 0; e < EMAX; ++e) gk[e] = Gk[min(tl + sml.off(e), Mx - 1)];
#pragma unroll
      for (int e = 0; e < EMAX; ++e)
        if (e < nvl) {
          const double a = (double)v[e].x, b = (double)v[e].y;
          kadd(acc[0], acc[1], 2.0 * (double)gk[e] * (a * a + b * b));
        }
    } else {
      slab_load(lds, S, base, n);
      __syncthreads();
      fft_line<T, false, EMAX, kColEarlyTw<T>, GEN>(lds, c, Cb, lp, tl, true, tw);
      if (c != 0) {
        const T* Gk = G + (size_t)c * Mx;
        for (int x = tl; x < Mx; x += lp.tpl) {
          const cplx<T> q = lds[lpad<T>(x * Cb + c)];
          kadd(acc[0], acc[1], 2.0 * (double)Gk[x] * ((double)q.x * q.x + (double)q.y * q.y));
        }
      } else {
        // Q = A + iB: A(kx) = (Q(kx) + conj Q(-kx)) / 2 (ky = 0),
        // B(kx) = (Q(kx) - conj Q(-kx)) / 2i (ky = My/2)
        const T* G0 = G;
        const T* GN = G + (size_t)g.N2 * Mx;
        for (int x = tl; x < Mx; x += lp.tpl) {
          const int xm = (Mx - x) % Mx;
          const cplx<T> q = lds[lpad<T>(x * Cb)], qm = lds[lpad<T>(xm * Cb)];
          const double ar = 0.5 * ((double)q.x + qm.x), ai = 0.5 * ((double)q.y - qm.y);
          const double br = 0.5 * ((double)q.x - qm.x), bi = 0.5 * ((double)q.y + qm.y);
          kadd(acc[0], acc[1], (double)G0[x] * (ar * ar + ai * ai) + (double)GN[x] * (br * br + bi * bi));
        }
      }
    }
    if (reduce_blocks<TauRed>(acc, in.part, in.ctr + 4 * kCtrSlot, gridDim.x, blockIdx.x) && threadIdx.x == 0)
      tau_finalize(in.sc, kvalue(acc[0], acc[1]));
    return;
  }
  if (MODE == 0 && cb != 0) {
    // every column is an ordinary one: HBM -> first forward pass, last
    // forward pass -> x G in registers -> first inverse pass (same radix, same
    // slots), last inverse pass -> HBM.  Two LDS round trips fewer each way.
    const int c = threadIdx.x % Cb, tl = threadIdx.x / Cb;
    const int R0 = lp.radix(0);
    const int nv0 = nbutterflies(Mx, R0, tl, lp.tpl) * R0;
    const SlotMap sm0(R0, Mx, lp.tpl), sml(lp.rlast, Mx, lp.tpl);
    cplx<T> v[EMAX];
    col_load_slots<T, EMAX>(v, S, base, Cb, c, Mx, tl, sm0);
    fft_line_x<T, false, EMAX, true, true, kColEarlyTw<T>, GEN>(v, lds, c, Cb, lp, tl, tw);
    __syncthreads();
    const auto Gk = (const __attribute__((address_space(1))) T*)(G + (size_t)(cb * Cb + c) * Mx);
    {
      // (slots past the last pass's butterflies read an in-range G and
      // carry values nothing reads)
      T gk[EMAX];
#pragma unroll
      for (int e = 0; e < EMAX; ++e) gk[e] = Gk[min(tl + sml.off(e), Mx - 1)];
#pragma unroll
      for (int e = 0; e < EMAX; ++e) v[e] = v[e] * gk[e];
    }
    // rv: the plan with its radices reversed, so the inverse transform's first
    // pass takes the forward transform's last-pass register slots as they are
    fft_line_x<T, true, EMAX, true, true, kColEarlyTw<T>, GEN>(v, lds, c, Cb, rv, tl, tw);
    const GViewW<T> Sw(S);
    // the store addresses are those of the loads: an opaque copy of the lane
    // index makes the compiler form them again here instead of keeping the
    // load addresses live (spilled) through both transforms
    int tls = tl;
    asm volatile("" : "+v"(tls));
    if (nv0 >= EMAX) {
#pragma unroll
      for (int e = 0; e < EMAX; ++e) Sw.put(base + (size_t)(tls + sm0.off(e)) * Cb + c, v[e]);
    } else {
#pragma unroll
      for (int e = 0; e < EMAX; ++e)
        if (e < nv0) Sw.put(base + (size_t)(tls + sm0.off(e)) * Cb + c, v[e]);
    }
    return;
  }
  slab_load(lds, S, base, n);
  __syncthreads();
  const int c = threadIdx.x % Cb, tl = threadIdx.x / Cb;
  if (MODE != 2) fft_line<T, false, EMAX, kColEarlyTw<T>, GEN>(lds, c, Cb, lp, tl, true, tw);
  if (MODE == 0) {
    const int ky = cb * Cb + c;
    if (ky != 0) {
      const T* Gk = G + (size_t)ky * Mx;
      for (int x = tl; x < Mx; x += lp.tpl) {
        cplx<T>& e = lds[lpad<T>(x * Cb + c)];
        e = e * Gk[x];
      }
    } else {
      // packed DC/Nyquist column: Q = A + iB with A, B spectra of real
      // columns; A *= G(kx, 0), B *= G(kx, My/2).
      const T* G0 = G;
      const T* GN = G + (size_t)g.N2 * Mx;
      for (int x = tl; x <= Mx / 2; x += lp.tpl) {
        const int xm = (Mx - x) % Mx;
        cplx<T> q = lds[lpad<T>(x * Cb)], qm = lds[lpad<T>(xm * Cb)];
        const T a0 = G0[x], aN = GN[x], b0 = G0[xm], bN = GN[xm];
        cplx<T> nq = q * (T(0.5) * (a0 + aN)) + conj(qm) * (T(0.5) * (a0 - aN));
        cplx<T> nqm = qm * (T(0.5) * (b0 + bN)) + conj(q) * (T(0.5) * (b0 - bN));
        lds[lpad<T>(x * Cb)] = nq;
        if (xm != x) lds[lpad<T>(xm * Cb)] = nqm;
      }
    }
    __syncthreads();
  }
  if (MODE != 1) fft_line<T, true, EMAX, kColEarlyTw<T>, GEN>(lds, c, Cb, lp, tl, true, tw);
  slab_store(S, lds, base, n);
}

// Last inverse pass of a row (radix R) into registers, then the consumer on
// batches of P::kBatch output pairs.
template <typename T, int EMAX, int R, class P>
__device__ __forceinline__ void c2r_finish(const P& p, const Geom& g, int row, int tl, const LinePlan& lp,
                                           const cplx<T>* lds, int rb, const cplx<T>* __restrict__ tw, double* acc) {
  const int N2 = g.N2;
  cplx<T> v[EMAX];
  const typename P::Ctx cx = p.ctx();
  last_pass_regs<T, R, true, EMAX>(v, lds, rb, 1, N2, tl, lp.tpl, tw);
  const int nbl = nbutterflies(N2, R, tl, lp.tpl);
  if (nbl * R >= EMAX) {
    // every slot holds a pair: the batch's loads issue with no branch between
#pragma unroll
    for (int e0 = 0; e0 < EMAX; e0 += P::kBatch) {
      typename P::Rec rec[P::kBatch];
#pragma unroll
      for (int k = 0; k < P::kBatch; ++k) rec[k] = p.fetch(g, row, 2 * slot_element(e0 + k, R, N2, tl, lp.tpl));
#pragma unroll
      for (int k = 0; k < P::kBatch; ++k)
        p.take(g, row, 2 * slot_element(e0 + k, R, N2, tl, lp.tpl), rec[k], v[e0 + k].x, v[e0 + k].y, acc, cx);
    }
    return;
  }
#pragma unroll
  for (int e0 = 0; e0 < EMAX; e0 += P::kBatch) {
    typename P::Rec rec[P::kBatch];
#pragma unroll
    for (int k = 0; k < P::kBatch; ++k) {
      const int e = e0 + k;
      if (e / R < nbl) rec[k] = p.fetch(g, row, 2 * slot_element(e, R, N2, tl, lp.tpl));
    }
#pragma unroll
    for (int k = 0; k < P::kBatch; ++k) {
      const int e = e0 + k;
      if (e / R < nbl) p.take(g, row, 2 * slot_element(e, R, N2, tl, lp.tpl), rec[k], v[e].x, v[e].y, acc, cx);
    }
  }
}

// Row pass: packed half spectrum -> real rows (inverse, unnormalised; the
// round trip r2c -> c2r scales by Mx*My, which the kernels G/K absorb).
template <typename T, int EMAX, class P, bool GEN = false>
__global__ void __launch_bounds__(fft_block_max(EMAX), fft_min_waves(EMAX)) k_c2r(Geom g, LinePlan lp, const cplx<T>* __restrict__ tw,
                                             const cplx<T>* __restrict__ twp, InstArr<typename P::F> ia,
                                             int ctr_slot) {
  extern __shared__ __align__(16) unsigned char smem_c2r[];
  cplx<T>* lds = reinterpret_cast<cplx<T>*>(smem_c2r);
  const Inst<typename P::F>& in = ia.get(blockIdx.y);
  P p(in);
  if (p.skip()) return;
  const int N2 = g.N2;
  const cplx<T>* S = spec_of<T>(in);
  const int tot = g.R * N2;
  const int blk = xcd_row_block(blockIdx.x, gridDim.x);
  const RowBlockMap map(g, blk);
  const int r = threadIdx.x / lp.tpl, tl = threadIdx.x - r * lp.tpl;
  // the merge pass's first twiddles, loaded with the slab: the merge runs
  // after the slab's barrier, so their latency overlaps the slab's
  constexpr int kTwPre = 4;
  cplx<T> twk[kTwPre];
#pragma unroll
  for (int j = 0; j < kTwPre; ++j) twk[j] = twp[min(tl + j * lp.tpl, N2 / 2)];
  const GView<T> Sg(S);
  for (int i0 = threadIdx.x; i0 < tot; i0 += kCopyU * (int)blockDim.x) {
    // (out-of-range slots repeat the last element, as in slab_load)
    T tx[kCopyU], ty[kCopyU];
#pragma unroll
    for (int k = 0; k < kCopyU; ++k) {
      const cplx<T> z = Sg[map.spec(min(i0 + k * (int)blockDim.x, tot - 1))];
      tx[k] = z.x;
      ty[k] = z.y;
    }
#pragma unroll
    for (int k = 0; k < kCopyU; ++k) lds[lpad<T>(map.lds(min(i0 + k * (int)blockDim.x, tot - 1)))] = mk(tx[k], ty[k]);
  }
  __syncthreads();
  const int row = blk * g.R + r;
  const int rb = r * N2;
  auto L = [&](int n) -> cplx<T>& { return lds[lpad<T>(rb + n)]; };
  auto merge = [&](int k, const cplx<T>& w) {
    if (k == 0) {
      cplx<T> z = L(0);  // (X[0], X[N2])
      L(0) = mk(z.x + z.y, z.x - z.y);
    } else {
      const int km = N2 - k;
      cplx<T> xk = L(k), xm = L(km);
      cplx<T> E = xk + conj(xm);
      cplx<T> O = cmulc(xk - conj(xm), w);
      L(k) = mk(E.x - O.y, E.y + O.x);            // E + iO
      if (km != k) L(km) = mk(E.x + O.y, O.x - E.y);  // conj(E) + i conj(O)
    }
  };
#pragma unroll
  for (int j = 0; j < kTwPre; ++j) {
    const int k = tl + j * lp.tpl;
    if (k <= N2 / 2) merge(k, twk[j]);
  }
  for (int k = tl + kTwPre * lp.tpl; k <= N2 / 2; k += lp.tpl) merge(k, twp[k]);
  __syncthreads();
  double acc[P::NS > 0 ? 2 * P::NS : 1];  // slot values, then corrections of compensated sums
#pragma unroll
  for (int i = 0; i < P::NS; ++i) {
    acc[i] = red_identity(P::op(i));
    acc[P::NS + i] = 0.0;
  }
  if constexpr (EMAX == 8) {
    // last inverse pass into registers and the consumer, specialised on the
    // last radix so every register slot index is a compile-time constant
    fft_line_to_last<T, true, EMAX>(lds, rb, 1, lp, tl, tw);
    switch (lp.rlast) {
      case 2: c2r_finish<T, EMAX, 2>(p, g, row, tl, lp, lds, rb, tw, acc); break;
      case 3: c2r_finish<T, EMAX, 3>(p, g, row, tl, lp, lds, rb, tw, acc); break;
      case 4: c2r_finish<T, EMAX, 4>(p, g, row, tl, lp, lds, rb, tw, acc); break;
      case 5: c2r_finish<T, EMAX, 5>(p, g, row, tl, lp, lds, rb, tw, acc); break;
      default: c2r_finish<T, EMAX, 8>(p, g, row, tl, lp, lds, rb, tw, acc); break;
    }
  } else {
    // long lines (16/32-element builds): one generic consumer loop
    cplx<T> v[EMAX];
    const typename P::Ctx cx = p.ctx();
    fft_line_x<T, true, EMAX, false, true, true, GEN>(v, lds, rb, 1, lp, tl, tw);
    const int nvl = nbutterflies(N2, lp.rlast, tl, lp.tpl) * lp.rlast;
    const SlotMap sml(lp.rlast, N2, lp.tpl);
    constexpr int KB = P::kBatch;
#pragma unroll
    for (int e0 = 0; e0 < EMAX; e0 += KB) {
      typename P::Rec rec[KB];
#pragma unroll
      for (int k = 0; k < KB; ++k) {
        const int e = e0 + k;
        if (e < nvl) rec[k] = p.fetch(g, row, 2 * (tl + sml.off(e)));
      }
#pragma unroll
      for (int k = 0; k < KB; ++k) {
        const int e = e0 + k;
        if (e < nvl) p.take(g, row, 2 * (tl + sml.off(e)), rec[k], v[e].x, v[e].y, acc, cx);
      }
    }
  }
  if constexpr (P::NS > 0) {
    fold_nan<P>(acc);
    if (reduce_blocks<P>(acc, in.part, in.ctr + ctr_slot * kCtrSlot, gridDim.x, blk))
      finalize_block(p, acc, in.sc);
  }
}

// Element-wise kernel with fused reductions: blocks stride over rows, threads
// over point pairs of a row (no index division, coalesced pair loads).
template <typename T, class P>
__global__ void __launch_bounds__(256, 4) k_elem(Geom g, InstArr<T> ia, int ctr_slot) {
  const Inst<T>& in = ia.get(blockIdx.y);
  P p(in);
  if (p.skip()) return;
  double acc[P::NS > 0 ? 2 * P::NS : 1];  // slot values, then corrections of compensated sums
#pragma unroll
  for (int i = 0; i < P::NS; ++i) {
    acc[i] = red_identity(P::op(i));
    acc[P::NS + i] = 0.0;
  }
  const int npair = g.My / 2;
  for (int row = blockIdx.x; row < g.Mx; row += gridDim.x) {
    for (int n0 = threadIdx.x; n0 < npair; n0 += kPairs * blockDim.x) {
      typename P::Rec rec[kPairs];
      if (n0 + (kPairs - 1) * (int)blockDim.x < npair) {
        // a full batch: no per-pair branch between the loads, so every load
        // of the batch issues before the first value is used
#pragma unroll
        for (int k = 0; k < kPairs; ++k) rec[k] = p.fetch(g, row, 2 * (n0 + k * (int)blockDim.x));
#pragma unroll
        for (int k = 0; k < kPairs; ++k) p.emit(g, row, 2 * (n0 + k * (int)blockDim.x), rec[k], acc);
        continue;
      }
#pragma unroll
      for (int k = 0; k < kPairs; ++k) {
        const int n = n0 + k * blockDim.x;
        if (n < npair) rec[k] = p.fetch(g, row, 2 * n);
      }
#pragma unroll
      for (int k = 0; k < kPairs; ++k) {
        const int n = n0 + k * blockDim.x;
        if (n < npair) p.emit(g, row, 2 * n, rec[k], acc);
      }
    }
  }
  if constexpr (P::NS > 0) {
    fold_nan<P>(acc);
    if (reduce_blocks<P>(acc, in.part, in.ctr + ctr_slot * kCtrSlot, gridDim.x, blockIdx.x))
      finalize_block(p, acc, in.sc);
  }
}

// ---------------------------------------------------------------------------
// Functors for plain operator evaluation.  TF: field precision of the
// instance, TT: precision of the transform they feed or consume.

template <typename TF, typename TT = TF>
struct ProdCopy {  // input row from src
  using F = TF;
  static constexpr int kBatch = kPairs;
  static constexpr int NS = 0;
  __device__ static constexpr int op(int) { return RSUM; }
  struct Rec {
    P2<TF> v;
  };
  const Inst<TF>& in;
  __device__ ProdCopy(const Inst<TF>& i) : in(i) {}
  __device__ bool skip() const { return false; }
  __device__ Rec fetch(const Geom& g, int row, int col) const { return Rec{ld2(in.src + (size_t)row * g.My + col)}; }
  __device__ void make(const Geom&, int, int, const Rec& r, TT& a, TT& b, double*) const {
    a = (TT)r.v.a;
    b = (TT)r.v.b;
  }
  __device__ void finalize(double*, Scal*) const {}
};

template <typename TF, typename TT = TF>
struct ConsCopy {  // output row to dst
  using F = TF;
  static constexpr int NS = 0;
  static constexpr int kBatch = kPairs;
  __device__ static constexpr int op(int) { return RSUM; }
  struct Rec {};
  const Inst<TF>& in;
  __device__ ConsCopy(const Inst<TF>& i) : in(i) {}
  __device__ bool skip() const { return false; }
  __device__ Rec fetch(const Geom&, int, int) const { return Rec{}; }
  struct Ctx {};
  __device__ Ctx ctx() const { return Ctx{}; }
  __device__ void take(const Geom& g, int row, int col, const Rec&, TT a, TT b, double*, const Ctx&) const {
    st2(in.dst + (size_t)row * g.My + col, (TF)a, (TF)b);
  }
  __device__ void finalize(double*, Scal*) const {}
};

// ---------------------------------------------------------------------------
// CCG set-up (ConstrainedConjugateGradients.py:179-219)

// u[comp] = max(u, h + offset) (:201); max over valid heights (:196-197).
template <typename T>
struct ElemInit {
  using F = T;
  static constexpr int NS = 2;  // max_h, n_comp
  __device__ static constexpr int op(int i) { return i == 0 ? RMAX : RSUM; }
  struct Rec {
    P2<T> u;
    HPair h;
  };
  const Inst<T>& in;
  __device__ ElemInit(const Inst<T>& i) : in(i) {}
  __device__ bool skip() const { return false; }
  __device__ Rec fetch(const Geom& g, int row, int col) const {
    return Rec{ld2(in.u + (size_t)row * g.My + col), ldh(g, in.h, in.surf, row, col)};
  }
  __device__ void emit(const Geom& g, int row, int col, const Rec& r, double* acc) const {
    const double off = in.sc->offset;
    T u0 = r.u.a, u1 = r.u.b;
    if (r.h.c0) {
      const double lim = r.h.h0 + off;
      if ((double)u0 < lim) u0 = (T)lim;
      acc[0] = red_apply(RMAX, acc[0], r.h.h0);
      acc[1] += 1.0;
    }
    if (r.h.c1) {
      const double lim = r.h.h1 + off;
      if ((double)u1 < lim) u1 = (T)lim;
      acc[0] = red_apply(RMAX, acc[0], r.h.h1);
      acc[1] += 1.0;
    }
    if (r.h.c0 || r.h.c1) st2(in.u + (size_t)row * g.My + col, u0, u1);
  }
  // the host stores the full grid size M in n_pad; subtract the comp count
  __device__ void finalize(double* acc, Scal* s) const {
    s->max_h = acc[0];
    s->n_pad = s->n_pad - acc[1];
  }
};

// f = evaluate_force(u) output with f[pad] = 0 (:212, :219), into the FP64
// forces of the solve.
template <typename TF, typename TT>
struct ConsInitForce {
  using F = TF;
  static constexpr int NS = 0;
  static constexpr int kBatch = kPairs;
  __device__ static constexpr int op(int) { return RSUM; }
  struct Rec {
    CPair c;
  };
  const Inst<TF>& in;
  __device__ ConsInitForce(const Inst<TF>& i) : in(i) {}
  __device__ bool skip() const { return false; }
  __device__ Rec fetch(const Geom& g, int row, int col) const { return Rec{ldc(g, in.surf, row, col)}; }
  struct Ctx {};
  __device__ Ctx ctx() const { return Ctx{}; }
  __device__ void take(const Geom& g, int row, int col, const Rec& r, TT a, TT b, double*, const Ctx&) const {
    st2(in.fm + (size_t)row * g.My + col, r.c.c0 ? (double)a : 0.0, r.c.c1 ? (double)b : 0.0);
  }
  __device__ void finalize(double*, Scal*) const {}
};

// The initial forces of the solve with f[pad] = 0 after u =
// evaluate_disp(initial_forces) (:213-219): FP64 plans zero the padding in
// place, FP32 plans copy the forces into their FP64 forces.
template <typename T>
struct ElemZeroPad {
  using F = T;
  static constexpr int NS = 0;
  __device__ static constexpr int op(int) { return RSUM; }
  struct Rec {
    CPair c;
    P2<T> f;
  };
  const Inst<T>& in;
  __device__ ElemZeroPad(const Inst<T>& i) : in(i) {}
  __device__ bool skip() const { return false; }
  __device__ Rec fetch(const Geom& g, int row, int col) const {
    Rec r;
    r.c = ldc(g, in.surf, row, col);
    if (!std::is_same<T, double>::value) r.f = ld2(in.f + (size_t)row * g.My + col);
    return r;
  }
  __device__ void emit(const Geom& g, int row, int col, const Rec& r, double*) const {
    double* fp = in.fm + (size_t)row * g.My + col;
    if (std::is_same<T, double>::value) {
      if (!r.c.c0) fp[0] = 0.0;
      if (!r.c.c1) fp[1] = 0.0;
    } else {
      st2(fp, r.c.c0 ? (double)r.f.a : 0.0, r.c.c1 ? (double)r.f.b : 0.0);
    }
  }
  __device__ void finalize(double*, Scal*) const {}
};

// ---------------------------------------------------------------------------
// CCG iteration phases

// Start of an iteration (:235-273, :300-309): counts, PK/mixing switch,
// offset and G from the shifted gap sums a = {A_contact, A_cg, Sd, Sdd, gmax,
// gmin} accumulated with shift s->shift.
__device__ void phase_a_logic(Scal* s, const double* a) {
  s->A_contact = a[0];
  s->A_cg = a[1];
  s->gmax = a[4];
  s->gmin = a[5];
  const bool pk = !s->has_H || (a[0] > 0 && a[1] / a[0] > 0.01 && s->mixsteps <= 0);  // :260
  if (pk) {
    s->mode = MODE_PK;
    if (s->delta_str == DS_MIX) s->delta_str = DS_SD;  // :261-262
    if (s->load_ctl) s->offset = a[1] > 0 ? s->shift + a[2] / a[1] : 0.0;  // :266-269
    const double m = s->offset - s->shift;
    double G = a[3] - 2.0 * m * a[2] + a[1] * m * m;  // sum_c (g - offset)^2 (:273)
    s->G = G < 0.0 ? 0.0 : G;  // clamp round-off below zero, keep NaN
    s->use_beta = (s->delta > 0 && s->G_old > 0) ? 1 : 0;  // :276
    s->beta = s->use_beta ? s->delta * (s->G / s->G_old) : 0.0;
  } else {
    s->mode = MODE_MIX;
    if (s->delta_str != DS_MIX) {  // :307-309
      s->mixsteps = s->minmixsteps;
      s->delta_str = DS_MIX;
    }
  }
}

// Per-point part of the start of an iteration: contact set from the current
// forces (:239-249) and the sums of the gap g0 = u - h shifted by `shift` (the
// offset of the previous iteration, which keeps sum (g - offset)^2 free of
// cancellation); returns c.
// acc: the six slots {A_contact, A_cg, Sd, Sdd, gmax, gmin}.  Sd and Sdd are
// sums of the shifted gap d = g - shift, centred on the previous offset, so
// they carry no cancellation of large terms; they are plain FP64 sums (a
// compensated pair in the displacement pass, where they are fused, took that
// kernel from 122 VGPRs to 237, or 380 spills at four waves per SIMD).
__device__ __forceinline__ bool phase_a_point(bool has_H, double H, double shift, double fv, double g0, bool comp,
                                              double* acc) {
  const bool c0 = fv < 0.0;
  const bool c = c0 && (!has_H || fv > -H);
  acc[0] += c0 ? 1.0 : 0.0;
  acc[1] += c ? 1.0 : 0.0;
  if (comp) {
    const double d = g0 - shift;
    if (c) acc[2] += d;                  // sum over the contact set only (:269)
    acc[3] += (c ? 1.0 : 0.0) * d * d;   // c * g * g: a NaN gap poisons G as in :273
    acc[4] = red_apply(RMAX, acc[4], g0);
    acc[5] = red_apply(RMIN, acc[5], g0);
  }
  return c;
}

// Reduction ops of the phase-A slots (shared by ElemPhaseA and ConsDisp).
__device__ constexpr int phase_a_op(int i) { return i == 4 ? RMAX : (i == 5 ? RMIN : RSUM); }

// Contact-mask bits: bit 0 = contact set c of the last finished iteration
// (result.active_set), bit 1 = gap < 0 at the start of that iteration
// (result.plastic), bit 2 = contact set of the iteration in progress, bit 3
// = final force < 0 (written with bit 2 by the force pass for the contact
// counts of the next iteration).
constexpr uint8_t kCur = 1, kGneg = 2, kNext = 4, kNeg = 8;

// Phase A of the first iteration (later iterations fuse it into ConsDisp):
// contact set and gap statistics (:238-273, :312-318 brackets).  It also
// replaces u by the gap u - h (:265) on the computational points, in FP64 in
// the gap buffer w: every later pass of the iteration uses only the gap, so
// the heights are read at most once per iteration; ElemFinish restores u.
// FP64 plans form the gap the oracle forms, fl(u - h), in place.  It runs
// also for an instance that is already done (maxiter < 1), so that
// ElemFinish always finds the gap.
template <typename T>
struct ElemPhaseA {
  using F = T;
  static constexpr int NS = 6;  // A_contact, A_cg, Sd, Sdd, gmax, gmin
  __device__ static constexpr int op(int i) { return phase_a_op(i); }
  struct Rec {
    P2<double> f;
    P2<T> u;
    HPair h;
  };
  const Inst<T>& in;
  __device__ ElemPhaseA(const Inst<T>& i) : in(i) {}
  __device__ bool skip() const { return false; }
  __device__ Rec fetch(const Geom& g, int row, int col) const {
    const size_t p = (size_t)row * g.My + col;
    return Rec{ld2(in.fm + p), ld2(in.u + p), ldh(g, in.h, in.surf, row, col)};
  }
  __device__ void emit(const Geom& g, int row, int col, const Rec& r, double* acc) const {
    const Scal* s = in.sc;
    const double w0 = r.h.c0 ? (double)r.u.a - r.h.h0 : (double)r.u.a;
    const double w1 = r.h.c1 ? (double)r.u.b - r.h.h1 : (double)r.u.b;
    const bool c0 = phase_a_point(s->has_H, s->H, s->shift, r.f.a, w0, r.h.c0, acc);
    const bool c1 = phase_a_point(s->has_H, s->H, s->shift, r.f.b, w1, r.h.c1, acc);
    const size_t p = (size_t)row * g.My + col;
    stm(in.cm + p, c0 ? kNext : 0, c1 ? kNext : 0);
    // FP64 plans write only where the gap differs from u (w aliases u)
    if (!std::is_same<T, double>::value || r.h.c0 || r.h.c1) st2(in.w + p, w0, w1);
  }
  __device__ void finalize(double* a, Scal* s) const {
    if (!s->done) phase_a_logic(s, a);
  }
};

// scipy.optimize.bisect(f, max g, min g, xtol=2e-12, rtol=4*eps, maxiter=100)
// of f(x) = count(g - x < 0) * H - F over the computational region
// (:316-318), with its steps and stopping rule:
//   dm = xb - xa; repeat: dm *= 0.5; xm = xa + dm; fm = f(xm);
//   if fm * fa >= 0: xa = xm;  if fm == 0 or |dm| < xtol + rtol |xm|: xm.
// f is evaluated as numpy does, fl(fl(cnt * H) - F), with no fused multiply
// -add; count(g - x < 0) equals count(g < x) for every pair of doubles
// (a difference of distinct finite doubles is never rounded to zero; NaN
// compares false either way).
//
// Batched rounds: the points that the next k steps can visit form a binary
// tree fixed by (xa, dm) alone (each step halves dm and moves xa to xm or
// not), so one pass over the gaps counts the gaps below all 2^k - 1 of them
// (in round 0 also below the brackets), and the block that finishes the pass
// replays the steps with those counts: the same points, counts and decisions
// as the serial bisection, in a few passes over the gaps by the whole grid
// instead of one pass per step by one block.  Round 0 takes 8 steps (its
// gaps spread over all bins: per-block bin counts, summed by the last block);
// later rounds take 11 (the gaps left inside the bracket are few: their bins
// go to a per-instance array with 64-bit atomics).
constexpr int kBisDepth0 = 8;                         // steps of round 0
constexpr int kBisDepth = 11;                         // steps of later rounds
constexpr int kBisPts0 = (1 << kBisDepth0) + 1;       // round 0: tree and brackets
constexpr int kBisSearch = 1 << kBisDepth;            // sorted points, NaN-padded
constexpr int kBisStride = 264;                       // round-0 bins per block partial, padded
constexpr int kBisRounds = 4;                         // batched steps: 8 + 3 * 11
constexpr int kBisBlockMax = 128;
#ifndef CM_BIS_PER_BLOCK
#define CM_BIS_PER_BLOCK 8192  // the CPU emulation builds with fewer, so small grids span several blocks
#endif
constexpr int kBisPerBlock = CM_BIS_PER_BLOCK;        // gaps per block (32 per thread)
static_assert(kBisPts0 + 1 <= kBisStride && kBisPts0 < kBisSearch, "bisection tables");

int bisect_blocks(const Geom& g) {
  const size_t Mt = (size_t)g.nx * g.ny;
  return (int)std::max<size_t>(1, std::min<size_t>(kBisBlockMax, (Mt + kBisPerBlock - 1) / kBisPerBlock));
}

// The per-instance bin counts of the rounds after the first, behind the
// eight ticket slots (zero between rounds: the last block clears them).
template <typename T>
__device__ __forceinline__ unsigned long long* bis_bins(const Inst<T>& in) {
  return (unsigned long long*)(in.ctr + 8 * kCtrSlot);
}

__device__ __forceinline__ double bis_f(double cnt, double H, double F) {
#pragma clang fp contract(off)
  return cnt * H - F;
}
__device__ __forceinline__ bool bis_stop(double fm, double dm, double xm) {
#pragma clang fp contract(off)
  return fm == 0.0 || fabs(dm) < 2e-12 + 8.881784197001252e-16 * fabs(xm);
}

// Point t (level order) of the tree of the next steps from (xa, dm): the bits
// of its path are the decisions "xa moved to xm".
__device__ __forceinline__ double bis_tree_point(double xa, double dm, int t) {
  const int l = 31 - __clz(t + 1);
  const int p = t + 1 - (1 << l);
  double x = xa, d = dm;
  for (int lev = 0; lev < l; ++lev) {
    d *= 0.5;
    const double xm = x + d;
    if ((p >> (l - 1 - lev)) & 1) x = xm;
  }
  d *= 0.5;
  return x + d;
}

// The np points of a round (the tree of `depth` steps, and with `brackets`
// gmin and gmax) sorted ascending in pts[0, np), NaN in pts[np, kBisSearch).
// In exact arithmetic in-order position along dm is the sorted order; rounding
// can tie points (harmless) or, with dm near an ulp of xa, invert neighbours:
// then they are sorted by rank.  Whole block; ends synchronised.
__device__ void bis_place(double xa, double dm, int depth, bool brackets, double gmax, double gmin, double* pts,
                          double* raw, int* flag) {
  const int P = (1 << depth) - 1, np = P + (brackets ? 2 : 0), off = brackets ? 1 : 0;
  if (threadIdx.x == 0) *flag = 0;
  for (int t = threadIdx.x; t < kBisSearch; t += blockDim.x) {
    if (t < np) {
      double v;
      int si;
      if (t < P) {
        v = bis_tree_point(xa, dm, t);
        const int l = 31 - __clz(t + 1);
        const int p = t + 1 - (1 << l);
        const int io = ((2 * p + 1) << (depth - 1 - l)) - 1;
        si = (dm < 0.0 ? P - 1 - io : io) + off;
      } else {
        v = t == P ? gmax : gmin;
        si = t == P ? np - 1 : 0;
      }
      raw[t] = v;
      pts[si] = v;
    } else {
      pts[t] = __builtin_nan("");
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t + 1 < np; t += blockDim.x)
    if (!(pts[t] <= pts[t + 1])) *flag = 1;
  __syncthreads();
  if (*flag) {
    for (int t = threadIdx.x; t < np; t += blockDim.x) {
      const double v = raw[t];
      int r = 0;
      for (int k = 0; k < np; ++k) {
        const double u = raw[k];
        r += (u < v || (u == v && k < t)) ? 1 : 0;
      }
      pts[r] = v;
    }
    __syncthreads();
  }
}

// Bins the gaps of the computational points [q0, q1) (row-major) into hist:
// bin k = number of points <= g; NaN gaps and masked points are not counted.
// Whole block; ends synchronised.
template <typename T>
__device__ void bis_bin_gaps(const Geom& g, const Inst<T>& in, size_t q0, size_t q1, const double* pts, int np,
                             unsigned* hist) {
  const double lo = pts[0], hi = pts[np - 1];
  const unsigned bd = blockDim.x;
  const int di = (int)(bd / (unsigned)g.ny), dj = (int)(bd - (unsigned)di * (unsigned)g.ny);
  size_t q = q0 + threadIdx.x;
  int i = (int)(q / (size_t)g.ny), j = (int)(q - (size_t)i * g.ny);
  unsigned below = 0, above = 0;
  while (q < q1) {
    double v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const bool ok = q < q1 && (!in.surf || in.surf[q]);
      v[k] = ok ? in.w[(size_t)i * g.My + j] : __builtin_nan("");
      q += bd;
      i += di;
      j += dj;
      if (j >= g.ny) {
        j -= g.ny;
        ++i;
      }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const double x = v[k];
      if (x < lo) {
        ++below;
      } else if (x >= hi) {
        ++above;
      } else if (x == x) {
        int pos = 0;
#pragma unroll
        for (int st = kBisSearch / 2; st > 0; st >>= 1)
          if (pts[pos + st - 1] <= x) pos += st;
        atomicAdd(&hist[pos], 1u);
      }
    }
  }
  if (below) atomicAdd(&hist[0], below);
  if (above) atomicAdd(&hist[np], above);
  __syncthreads();
}

// Inclusive prefix of cum[0, n) in place.  Whole block; ends synchronised.
__device__ void bis_scan(unsigned long long* cum, int n) {
  constexpr int kPer = kBisSearch / 256;
  for (int off = 1; off < n; off <<= 1) {
    unsigned long long a[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int t = threadIdx.x + k * (int)blockDim.x;
      a[k] = (t < n && t >= off) ? cum[t - off] : 0ull;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int t = threadIdx.x + k * (int)blockDim.x;
      if (t < n) cum[t] += a[k];
    }
    __syncthreads();
  }
}

// Bisection state between rounds and passes.
struct BisState {
  double xa, dm, fa, res;
  int it, state;  // 1 running, 2 found res, 3 failed
};

// Replays up to `depth` steps with the prefix counts cum of the sorted points
// (thread 0).  Round 0 (`first`) starts with the brackets' checks.
__device__ void bis_walk(BisState& b, bool first, int depth, const double* pts, int np,
                         const unsigned long long* cum, double H, double F, double gmax, double gmin) {
  // count(g < x) for x one of the points (-1: x is not one of them)
  auto cnt_at = [&](double x) -> double {
    int lb = 0;
    for (int st = kBisSearch / 2; st > 0; st >>= 1)
      if (pts[lb + st - 1] < x) lb += st;
    return (lb < np && pts[lb] == x) ? (double)cum[lb] : -1.0;
  };
  if (first) {
    b.xa = gmax;
    b.dm = gmin - gmax;
    b.it = 0;
    b.state = 1;
    b.fa = bis_f(cnt_at(gmax), H, F);
    const double fb = bis_f(cnt_at(gmin), H, F);
    if (b.fa * fb > 0) {
      b.state = 3;
    } else if (b.fa == 0) {
      b.res = gmax;
      b.state = 2;
    } else if (fb == 0) {
      b.res = gmin;
      b.state = 2;
    }
  }
  for (int lev = 0; b.state == 1 && lev < depth && b.it < 100; ++lev) {
    const double dn = b.dm * 0.5;
    const double xm = b.xa + dn;
    const double c = cnt_at(xm);
    if (c < 0.0) break;  // not reached (the tree holds every point of these steps): no step taken
    b.dm = dn;
    const double fm = bis_f(c, H, F);
    if (fm * b.fa >= 0) b.xa = xm;
    ++b.it;
    if (bis_stop(fm, b.dm, xm)) {
      b.res = xm;
      b.state = 2;
    }
  }
  if (b.state == 1 && b.it >= 100) b.state = 3;
}

__device__ __forceinline__ void bis_commit(Scal* s, const BisState& b) {
  if (b.state == 2) {
    s->offset = b.res;
  } else if (b.state == 3) {
    s->status = CM_E_BISECT;
    s->done = 1;
    s->it += 1;  // the failed iteration counts
  }
  s->bis_xa = b.xa;
  s->bis_dm = b.dm;
  s->bis_fa = b.fa;
  s->bis_steps = b.it;
  s->bis_state = b.state == 1 ? 1 : 2;
}

// count(g < x) over the computational points by one whole block (every
// thread gets the count).
template <typename T>
__device__ double bis_count_block(const Geom& g, const Inst<T>& in, double x, double* scr) {
  __shared__ double s_cnt;
  const size_t Mt = (size_t)g.nx * g.ny;
  const int i0 = (int)(threadIdx.x / (unsigned)g.ny), j0 = (int)threadIdx.x - i0 * g.ny;
  const int di = (int)(blockDim.x / (unsigned)g.ny), dj = (int)blockDim.x - di * g.ny;
  double v = 0.0;
  int i = i0, j = j0;
  for (size_t q = threadIdx.x; q < Mt; q += blockDim.x) {
    if (!in.surf || in.surf[q]) v += in.w[(size_t)i * g.My + j] < x ? 1.0 : 0.0;
    i += di;
    j += dj;
    if (j >= g.ny) {
      j -= g.ny;
      ++i;
    }
  }
  int ops[1] = {RSUM};
  double unused = 0.0;
  block_reduce<1>(&v, &unused, ops, scr);
  if (threadIdx.x == 0) s_cnt = v;
  __syncthreads();
  const double cnt = s_cnt;
  __syncthreads();
  return cnt;
}

// Whole bisection by one block, one count per step (brackets that are not
// finite: scipy's checks and steps on whatever the counts give).
template <typename T>
__device__ void bis_whole_serial(const Geom& g, const Inst<T>& in, Scal* s, double* scr) {
  const double H = s->H, F = s->F;
  BisState b{};
  b.xa = s->gmax;
  const double xb = s->gmin;
  b.fa = bis_f(bis_count_block(g, in, b.xa, scr), H, F);
  const double fb = bis_f(bis_count_block(g, in, xb, scr), H, F);
  b.state = 1;
  if (b.fa * fb > 0) {
    b.state = 3;
  } else if (b.fa == 0) {
    b.res = b.xa;
    b.state = 2;
  } else if (fb == 0) {
    b.res = xb;
    b.state = 2;
  }
  b.dm = xb - b.xa;
  for (; b.state == 1 && b.it < 100;) {
    b.dm *= 0.5;
    const double xm = b.xa + b.dm;
    const double fm = bis_f(bis_count_block(g, in, xm, scr), H, F);
    if (fm * b.fa >= 0) b.xa = xm;
    ++b.it;
    if (bis_stop(fm, b.dm, xm)) {
      b.res = xm;
      b.state = 2;
    }
  }
  if (b.state == 1) b.state = 3;
  if (threadIdx.x == 0) bis_commit(s, b);
}

// One round of the batched bisection; launched kBisRounds times per
// iteration on a grid (nblk, instances); scb + b * sc_stride is instance b's
// scalar record.  Round 0 starts from the brackets,
// later rounds continue while the record's bis_state is 1.  The last round's
// last block finishes an unfinished bisection alone, kBisDepth steps per pass
// over the gaps (more than 8 + 3 * 11 steps: brackets about 2^41 xtol apart
// or more).
template <typename T>
__global__ void __launch_bounds__(256) k_bisect_round(Geom g, InstArr<T> ia, char* scb, size_t sc_stride, int round,
                                                      int nblk) {
  // the scalar record first, addressed from the arguments (instance b's at
  // scb + b * sc_stride, as in the instance record): a round with nothing to
  // do makes one dependent load, not two
  Scal* s = (Scal*)(scb + (size_t)blockIdx.y * sc_stride);
  if (s->done || s->mode != MODE_MIX || !s->load_ctl) return;
  if (round > 0 && s->bis_state != 1) return;
  const Inst<T>& in = ia.get(blockIdx.y);
  __shared__ double pts[kBisSearch];
  __shared__ union {
    double raw[kBisSearch];               // the points in tree order
    unsigned long long cum[kBisSearch];  // then the prefix counts
  } ub;
  __shared__ unsigned hist[kBisSearch];  // gaps per bin: bin k = #points <= g
  __shared__ double scr[2 * 8 * 16];
  __shared__ int s_last, s_flag;
  __shared__ BisState s_b;
  const double gmax = s->gmax, gmin = s->gmin, H = s->H, F = s->F;
  const bool first = round == 0;
  const double xa0 = first ? gmax : s->bis_xa;
  const double dm0 = first ? gmin - gmax : s->bis_dm;
  if (first && !(isfinite(gmax) && isfinite(gmin) && isfinite(dm0))) {
    if (blockIdx.x == 0) bis_whole_serial(g, in, s, scr);
    return;
  }
  const int depth = first ? kBisDepth0 : kBisDepth;
  const int np = (1 << depth) - 1 + (first ? 2 : 0);
  for (int t = threadIdx.x; t < kBisSearch; t += blockDim.x) hist[t] = 0u;
  bis_place(xa0, dm0, depth, first, gmax, gmin, pts, ub.raw, &s_flag);
  // this block's share of the computational points, in row-major order
  const size_t Mt = (size_t)g.nx * g.ny;
  const size_t per = (Mt + nblk - 1) / nblk;
  const size_t q0 = (size_t)blockIdx.x * per, q1 = q0 + per < Mt ? q0 + per : Mt;
  bis_bin_gaps(g, in, q0 < Mt ? q0 : Mt, q1, pts, np, hist);
  unsigned* part = (unsigned*)in.part;
  unsigned long long* gb = bis_bins(in);
  unsigned* ctr = in.ctr;  // ticket slot 0 (set-up passes use it too, never at the same time)
  if (first) {
    for (int t = threadIdx.x; t <= np; t += blockDim.x)
      __hip_atomic_store(&part[(size_t)blockIdx.x * kBisStride + t], hist[t], __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
  } else {
    for (int t = threadIdx.x; t <= np; t += blockDim.x)
      if (hist[t]) atomicAdd(&gb[t], (unsigned long long)hist[t]);
  }
  wait_stores_done();
  __syncthreads();
  if (threadIdx.x == 0) s_last = take_ticket(ctr, nblk, blockIdx.x);
  __syncthreads();
  if (!s_last) return;
  acquire_agent();
  if (first) {
    // (16 independent loads in flight per thread: one round trip per 16 blocks)
    for (int t = threadIdx.x; t <= np; t += blockDim.x) {
      unsigned long long c = 0;
      int b = 0;
      for (; b + 16 <= nblk; b += 16) {
        unsigned v[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) v[k] = part[(size_t)(b + k) * kBisStride + t];
#pragma unroll
        for (int k = 0; k < 16; ++k) c += v[k];
      }
      for (; b < nblk; ++b) c += part[(size_t)b * kBisStride + t];
      ub.cum[t] = c;
    }
  } else {
    for (int t = threadIdx.x; t <= np; t += blockDim.x) {
      ub.cum[t] = gb[t];
      __hip_atomic_store(&gb[t], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  __syncthreads();
  bis_scan(ub.cum, np + 1);
  if (threadIdx.x == 0) {
    __hip_atomic_store(ctr, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    BisState b{};
    if (!first) {
      b.xa = xa0;
      b.dm = dm0;
      b.fa = s->bis_fa;
      b.it = s->bis_steps;
      b.state = 1;
    }
    bis_walk(b, first, depth, pts, np, ub.cum, H, F, gmax, gmin);
    s_b = b;
  }
  __syncthreads();
  // the last round: this block goes on alone, a tree of kBisDepth steps per
  // pass over all the gaps
  while (s_b.state == 1 && round + 1 >= kBisRounds) {
    const BisState b0 = s_b;
    for (int t = threadIdx.x; t < kBisSearch; t += blockDim.x) hist[t] = 0u;
    bis_place(b0.xa, b0.dm, kBisDepth, false, gmax, gmin, pts, ub.raw, &s_flag);
    const int n1 = (1 << kBisDepth) - 1;
    bis_bin_gaps(g, in, 0, Mt, pts, n1, hist);
    for (int t = threadIdx.x; t <= n1; t += blockDim.x) ub.cum[t] = hist[t];
    __syncthreads();
    bis_scan(ub.cum, n1 + 1);
    if (threadIdx.x == 0) {
      BisState b = b0;
      bis_walk(b, false, kBisDepth, pts, n1, ub.cum, H, F, gmax, gmin);
      s_b = b;
    }
    __syncthreads();
    if (s_b.state == 1 && s_b.it == b0.it) {
      // no step taken (not reached): one step with its own count, so every
      // pass of this loop takes at least one of the at most 100 steps
      BisState b = b0;
      b.dm *= 0.5;
      const double xm = b.xa + b.dm;
      const double fm = bis_f(bis_count_block(g, in, xm, scr), H, F);
      if (fm * b.fa >= 0) b.xa = xm;
      ++b.it;
      if (bis_stop(fm, b.dm, xm)) {
        b.res = xm;
        b.state = 2;
      } else if (b.it >= 100) {
        b.state = 3;
      }
      if (threadIdx.x == 0) s_b = b;
      __syncthreads();
    }
  }
  if (threadIdx.x == 0) bis_commit(s, s_b);
}

// Conjugate direction t = c (g + beta t) on comp (:265-279) fused into the
// first forward transform; partial sum_c g t for tau (:292).
template <typename T>
struct ProdDir {
  using F = T;
  static constexpr int kBatch = kPairs;
  static constexpr int NS = 1;
  __device__ static constexpr int op(int) { return RCSUM; }
  struct Rec {
    P2<double> w;
    P2<T> t;
    M2 cm;
    CPair cp;
  };
  const Inst<T> in;  // by value: the field pointers stay in registers (stores cannot alias them)
  const Scal sv;  // scalar record as the kernel began: only the last block's finalize writes it
  __device__ ProdDir(const Inst<T>& i) : in(i), sv(*i.sc) {}
  __device__ bool skip() const { return sv.done || sv.mode == MODE_MIX; }
  __device__ Rec fetch(const Geom& g, int row, int col) const {
    const size_t p = (size_t)row * g.My + col;
    return Rec{ld2(in.w + p), ld2(in.t + p), ldm(in.cm + p), ldc(g, in.surf, row, col)};
  }
  // wv: the stored gap u - h
  // Masks multiply, as in the reference (t = c (g + ...), sum c g t), so a
  // NaN gap outside the contact set still reaches t and the sums.
  __device__ T one(const Scal* s, bool comp, bool c, double wv, T told, double* acc) const {
    if (!comp) return told;  // t is only ever set on comp (zero elsewhere)
    const double gap = wv - s->offset;
    const double cm = c ? 1.0 : 0.0;
    const T tt = (T)(cm * (s->use_beta ? gap + s->beta * (double)told : gap));
    kadd(acc[0], acc[NS], cm * gap * (double)tt);
    return tt;
  }
  __device__ void make(const Geom& g, int row, int col, const Rec& r, T& a, T& b, double* acc) const {
    const Scal* s = &sv;
    a = one(s, r.cp.c0, r.cm.a() & kNext, r.w.a, r.t.a, acc);
    b = one(s, r.cp.c1, r.cm.b() & kNext, r.w.b, r.t.b, acc);
    st2(in.t + (size_t)row * g.My + col, a, b);
  }
  __device__ void finalize(double* a, Scal* s) const { s->Scgt = a[0]; }
};

// Step length from sum_c r t (:284-294); r = evaluate_disp(t).
__device__ void tau_finalize(Scal* s, double sum_rt) {
  s->nfev += 1;
  s->tau = 0.0;
  if (s->A_cg > 0) {
    const double x = -sum_rt;
    if (x > 0.0)
      s->tau = s->Scgt / x;
    else
      s->G = 0.0;
  }
}

// Per-pixel force after the PK or mixing update, before projection
// (:296 and :322-323).  Shared by phase B and the second transform.
// f + (tau c) t multiplies by the mask as the reference does (a NaN in t
// reaches f even where c = 0).  HARD = false: no instance of the launch has a
// hardness, so every step is a PK step (mixing needs a hardness, :260) and
// the hardness branches compile away.
template <bool HARD>
__device__ __forceinline__ double force_update(const Scal* s, double fv, double tv, bool c) {
  if (!HARD || s->mode == MODE_PK) return fv + (c ? s->tau : 0.0) * tv;
  return (1.0 - s->mix) * fv - (s->mix * s->H) * (c ? 1.0 : 0.0);
}

// Phase B: projection statistics (:332-389).  HARD as in force_update.
template <typename T, bool HARD = true>
struct ElemPhaseB {
  using F = T;
  static constexpr int NS = 7;  // n_tensile, max_tensile, pad_max, A_fl, min_fl, sum_f2, nc
  static constexpr bool kScalInLds = true;  // projection and step-kind logic
  __device__ static constexpr int op(int i) {
    return (i == 1 || i == 2) ? RMAX : (i == 4 ? RMIN : (i == 5 ? RCSUM : RSUM));
  }
  struct Rec {
    P2<double> f;
    P2<T> t;
    P2<double> w;
    M2 cm;
    CPair cp;
  };
  const Inst<T> in;  // by value: the field pointers stay in registers (stores cannot alias them)
  const Scal sv;  // scalar record as the kernel began: only the last block's finalize writes it
  __device__ ElemPhaseB(const Inst<T>& i) : in(i), sv(*i.sc) {}
  __device__ bool skip() const { return sv.done != 0; }
  __device__ Rec fetch(const Geom& g, int row, int col) const {
    const size_t p = (size_t)row * g.My + col;
    return Rec{ld2(in.fm + p), ld2(in.t + p), ld2(in.w + p), ldm(in.cm + p), ldc(g, in.surf, row, col)};
  }
  // wv: the stored gap u - h
  __device__ uint8_t one(const Scal* s, bool comp, uint8_t cmv, double fv, double tv, double wv,
                         double* acc) const {
    double gap = 0.0;
    if (comp) gap = wv - s->offset;
    bool c = (cmv & kNext) != 0;
    if (HARD && s->mode == MODE_MIX && comp) c = gap < 0.0;  // c[comp] = g < 0 (:322)
    const double f1 = force_update<HARD>(s, fv, tv, c);
    const bool tensile = f1 >= 0.0;
    const bool flowing = HARD && s->has_H && f1 <= -s->H;
    if (tensile) {
      acc[0] += 1.0;
      kmax(acc[1], acc[NS + 1], f1);
    }
    if (!comp) kmax(acc[2], acc[NS + 2], fabs(f1));
    if (HARD && s->H_truthy && flowing) {
      acc[3] += 1.0;
      kmin(acc[4], acc[NS + 4], f1 + s->H);
    }
    if (comp) {
      kadd(acc[5], acc[NS + 5], tensile ? 0.0 : f1);
      const bool nc = (tensile && gap < 0.0) || (flowing && gap > 0.0);
      acc[6] += nc ? 1.0 : 0.0;
    }
    return (uint8_t)((cmv & ~kNext) | (c ? kNext : 0));
  }
  __device__ void emit(const Geom& g, int row, int col, const Rec& r, double* acc) const {
    const Scal* s = &sv;
    const uint8_t c0 = one(s, r.cp.c0, r.cm.a(), r.f.a, (double)r.t.a, r.w.a, acc);
    const uint8_t c1 = one(s, r.cp.c1, r.cm.b(), r.f.b, (double)r.t.b, r.w.b, acc);
    if (HARD && s->mode == MODE_MIX) stm(in.cm + (size_t)row * g.My + col, c0, c1);
  }
  __device__ void finalize(double* a, Scal* s) const {
    s->pad_pres = s->n_pad > 0 ? a[2] : 0.0;             // :344-346
    s->max_pres = a[0] > 0 ? a[1] : 0.0;                 // :350-352
    s->A_fl = a[3];
    if (s->H_truthy && a[3] > 0) {                       // :353-356
      const double v = -a[4];
      s->max_pres = (v > s->max_pres) ? v : s->max_pres;
    }
    s->uniform = 0;
    s->scale = 1.0;
    if (s->load_ctl) {                                   // :362-368
      const double tf = -a[5];
      if (tf != 0)
        s->scale = s->F / tf;
      else
        s->uniform = 1;
    }
    s->nc_count = a[6];
    s->do_nc = 0;
    if (s->mode == MODE_MIX) {                           // :375-377
      s->delta = 0;
      s->G = 0;
    } else if (a[6] > 0) {                               // :379-386
      s->do_nc = 1;
      s->delta = 0;
      s->delta_str = DS_SD;
    } else {                                             // :387-389
      s->delta = 1;
      s->delta_str = DS_CG;
    }
  }
};

// Final projected force (:359-384) fused into the second forward transform.
// The transform (precision TT) takes the new force (INC = false) or, for the
// incremental displacement update of FP32 plans, the change of the force
// f_new - f_old (INC = true), formed in FP64.
template <typename T, typename TT, bool INC, bool HARD = true>
struct ProdForce {
  using F = T;
  static constexpr int kBatch = kPairs;
  static constexpr int NS = 1;  // sum f over comp
  __device__ static constexpr int op(int) { return RCSUM; }
  struct Rec {
    P2<double> f;
    P2<T> t;
    P2<double> w;
    M2 cm;
    CPair cp;
  };
  const Inst<T> in;  // by value: the field pointers stay in registers (stores cannot alias them)
  const Scal sv;  // scalar record as the kernel began: only the last block's finalize writes it
  __device__ ProdForce(const Inst<T>& i) : in(i), sv(*i.sc) {}
  __device__ bool skip() const { return sv.done != 0; }
  __device__ Rec fetch(const Geom& g, int row, int col) const {
    const size_t p = (size_t)row * g.My + col;
    return Rec{ld2(in.fm + p), ld2(in.t + p), ld2(in.w + p), ldm(in.cm + p), ldc(g, in.surf, row, col)};
  }
  // wv: the stored gap u - h
  __device__ double one(const Scal* s, bool comp, uint8_t& cmv, double fv, double tv, double wv,
                        double* acc) const {
    double gap = 0.0;
    if (comp) gap = wv - s->offset;
    const bool c = (cmv & kNext) != 0;
    const double f1 = force_update<HARD>(s, fv, tv, c);
    const bool tensile = f1 >= 0.0;
    const bool flowing = HARD && s->has_H && f1 <= -s->H;
    double fo = tensile ? 0.0 : f1;                       // :359
    if (s->load_ctl) {
      if (s->uniform)
        fo = comp ? -s->F / s->nb_surface_pts : 0.0;      // :367-368
      else
        fo = fo * s->scale;                               // :365
    }
    if (flowing) fo = -s->H;                              // :372-373
    if (comp) {
      const bool nc = (tensile && gap < 0.0) || (flowing && gap > 0.0);
      if (s->do_nc && nc) fo = fo + s->tau * gap;         // :384
      // contact set of the next iteration from the final force (:239-249),
      // bits 2 and 3 for the displacement pass and the next iteration
      const bool cn0 = fo < 0.0;
      const bool cn = cn0 && (!HARD || !s->has_H || fo > -s->H);
      cmv = (uint8_t)((c ? kCur : 0) | (gap < 0.0 ? kGneg : 0) | (cn ? kNext : 0) | (cn0 ? kNeg : 0));
      kadd(acc[0], acc[NS], fo);
    }
    return fo;
  }
  __device__ void make(const Geom& g, int row, int col, const Rec& r, TT& a, TT& b, double* acc) const {
    const Scal* s = &sv;
    uint8_t m0 = r.cm.a(), m1 = r.cm.b();
    const double f0 = one(s, r.cp.c0, m0, r.f.a, (double)r.t.a, r.w.a, acc);
    const double f1 = one(s, r.cp.c1, m1, r.f.b, (double)r.t.b, r.w.b, acc);
    const size_t p = (size_t)row * g.My + col;
    st2(in.fm + p, f0, f1);
    if (r.cp.c0 || r.cp.c1) stm(in.cm + p, m0, m1);
    if (INC) {
      a = (TT)(f0 - r.f.a);
      b = (TT)(f1 - r.f.b);
    } else {
      a = (TT)f0;
      b = (TT)f1;
    }
  }
  __device__ void finalize(double* a, Scal* s) const { s->total_force = -a[0]; }
};

// u_new = evaluate_disp(f) consumed on the fly: maxdu, max_pen and the
// convergence tests (:391-434), NaN check (:491-492), maxiter (:494-509);
// fused with the start of the next iteration (contact set from the final
// forces, shifted gap sums of u_new), so no separate phase-A pass is needed.
// The gap buffer keeps u - h on the computational points (FP64): the old
// gap is read and the new one stored; maxdu and max_pen are formed from the
// gaps (u_new - u_old = gap_new - gap_old, h + offset - u_new = offset -
// gap_new).  INC = false: the transform gives u_new and the heights are read
// (gap_new = u_new - h).  INC = true (FP32 plans between FP64 refreshes):
// the transform gives evaluate_disp(f_new - f_old) and gap_new = gap_old +
// that change, so FP32 round-off enters only through the change of the
// displacements, not through the displacements themselves.
template <typename T, typename TT, bool INC>
struct ConsDisp {
  using F = T;
  // maxdu, max_pen | A_contact, A_cg, Sd, Sdd, gmax, gmin of the next
  // iteration (its contact set comes from the mask bits the force pass wrote
  // from the final forces)
  static constexpr int NS = 8;
  // output pairs per batch: two with FP64 transforms (the record by value
  // and the transform's registers leave no room for four under 128 VGPRs,
  // four waves per SIMD), four with FP32 transforms (88 VGPRs with two)
  static constexpr int kBatch = sizeof(TT) == 4 ? 4 : 2;
  static constexpr bool kScalInLds = true;  // convergence tests and the next iteration's start
  static constexpr bool kLogsRing = true;   // the end of an iteration
  __device__ static constexpr int op(int i) { return i < 2 ? RMAX : phase_a_op(i - 2); }
  struct Rec {
    P2<double> w;
    M2 cm;
    HPair h;
  };
  const Inst<T> in;  // by value: the field pointers stay in registers (stores cannot alias them)
  __device__ ConsDisp(const Inst<T>& i) : in(i) {}
  __device__ bool skip() const { return in.sc->done != 0; }
  __device__ Rec fetch(const Geom& g, int row, int col) const {
    const size_t p = (size_t)row * g.My + col;
    Rec r;
    r.w = ld2(in.w + p);
    r.cm = ldm(in.cm + p);
    if (INC) {
      const CPair c = ldc(g, in.surf, row, col);
      r.h = HPair{0.0, 0.0, c.c0, c.c1};
    } else {
      r.h = ldh(g, in.h, in.surf, row, col);
    }
    return r;
  }
  // The scalars the consumer needs, read once per thread just before the
  // last transform pass (per point pair they would be dependent loads after
  // every store; across the whole transform they would cost registers).
  struct Ctx {
    double offset, H;
    bool has_H;
  };
  __device__ Ctx ctx() const { return Ctx{in.sc->offset, in.sc->H, in.sc->has_H != 0}; }
  // wo: stored value (gap on comp, displacement elsewhere); v: transform
  // output; wn: the new stored value; the next iteration's contact set is
  // bit 2 of the mask (written by the force pass)
  __device__ void one(const Ctx& cx, bool comp, uint8_t cmv, double wo, double hv, TT v, double& wn,
                      double* acc) const {
    if (INC)
      wn = wo + (double)v;
    else
      wn = comp ? (double)v - hv : (double)v;
    kmax(acc[0], acc[NS + 0], fabs(wn - wo));
    if (comp) {
      const double pen = (cmv & kCur) ? (cx.offset - wn) : 0.0;
      kmax(acc[1], acc[NS + 1], pen);
      // contact counts and shifted gap sums of the next iteration, shift =
      // this offset (:239-273)
      const bool cn = (cmv & kNext) != 0;
      acc[2] += (cmv & kNeg) ? 1.0 : 0.0;
      acc[3] += cn ? 1.0 : 0.0;
      const double d = wn - cx.offset;
      if (cn) acc[4] += d;                  // sum over the contact set only (:269)
      acc[5] += (cn ? 1.0 : 0.0) * d * d;   // c * g * g: a NaN gap poisons G as in :273
      kmax(acc[6], acc[NS + 6], wn);
      kmin(acc[7], acc[NS + 7], wn);
    }
  }
  __device__ void take(const Geom& g, int row, int col, const Rec& r, TT a, TT b, double* acc, const Ctx& cx) const {
    double w0, w1;
    one(cx, r.h.c0, r.cm.a(), r.w.a, r.h.h0, a, w0, acc);
    one(cx, r.h.c1, r.cm.b(), r.w.b, r.h.h1, b, w1, acc);
    st2(in.w + (size_t)row * g.My + col, w0, w1);
  }
  __device__ void finalize(double* a, Scal* s) const {
    s->it += 1;  // iterations completed (:235-236)
    s->nfev += 1;
    s->maxdu = a[0];
    s->G_old = s->G;
    s->rms_pen = s->A_cg > 0 ? sqrt(s->G / s->A_cg) : sqrt(s->G);
    s->max_pen = a[1] > 0.0 ? a[1] : 0.0;
    bool conv = true;
    const double tf = s->total_force;
    if (s->load_ctl)
      conv = conv && fabs((tf - s->F) / tf) < s->thermotol;
    else if (s->has_last_tf)
      conv = conv && fabs((tf - s->last_tf) / tf) < s->thermotol;
    s->last_tf = tf;
    s->has_last_tf = 1;
    const double mh = s->max_h + s->offset;
    if (s->load_ctl && s->has_last_mh) conv = conv && fabs((mh - s->last_mh) / mh) < s->thermotol;
    s->last_mh = mh;
    s->has_last_mh = 1;
    if (s->mode == MODE_MIX)
      conv = conv && s->maxdu < s->pentol && s->max_pres < s->forcetol && s->pad_pres < s->forcetol;
    else
      conv = conv && s->rms_pen < s->pentol && s->max_pen < s->pentol && s->maxdu < s->pentol &&
             s->max_pres < s->forcetol && s->pad_pres < s->forcetol;
    if (s->mode == MODE_MIX) {  // :326-330
      s->mix *= s->mixdecfac;
      if (s->mixsteps > 0) s->mixsteps -= 1;
    } else {
      s->mix = s->mixfac;  // :299
    }
    if (conv) {
      s->status = CM_CONVERGED;
      s->done = 1;
    } else if (s->G != s->G || s->rms_pen != s->rms_pen) {
      s->status = CM_E_NAN;
      s->done = 1;
    } else if (s->it >= s->maxiter) {
      s->status = CM_MAXITER;
      s->done = 1;
    } else {
      s->shift = s->offset;     // the sums a[2..7] were taken with this shift
      phase_a_logic(s, a + 2);  // start of the next iteration
    }
  }
};

// End of a solve: the displacements u = gap + h back on the computational
// points (undoing ElemPhaseA's substitution; FP32 plans: everywhere, from the
// FP64 gap, and the forces from the FP64 forces) and the elastic energy
// -sum(f u)/2 over the topography block (:471, :505-506), from the FP64
// values.
template <typename T>
struct ElemFinish {
  using F = T;
  static constexpr int NS = 1;
  __device__ static constexpr int op(int) { return RCSUM; }
  struct Rec {
    P2<double> f, w;
    HPair h;
  };
  const Inst<T>& in;
  __device__ ElemFinish(const Inst<T>& i) : in(i) {}
  __device__ bool skip() const { return false; }
  __device__ Rec fetch(const Geom& g, int row, int col) const {
    const size_t p = (size_t)row * g.My + col;
    return Rec{ld2(in.fm + p), ld2(in.w + p), ldh(g, in.h, in.surf, row, col)};
  }
  __device__ void emit(const Geom& g, int row, int col, const Rec& r, double* acc) const {
    const double u0 = r.h.c0 ? r.w.a + r.h.h0 : r.w.a;
    const double u1 = r.h.c1 ? r.w.b + r.h.h1 : r.w.b;
    const size_t p = (size_t)row * g.My + col;
    if (std::is_same<T, double>::value) {
      if (r.h.c0 || r.h.c1) st2(in.u + p, (T)u0, (T)u1);
    } else {
      st2(in.u + p, (T)u0, (T)u1);
      st2(in.f + p, (T)r.f.a, (T)r.f.b);
    }
    if (row < g.nx && col < g.ny) kadd(acc[0], acc[NS], r.f.a * u0);
    if (row < g.nx && col + 1 < g.ny) kadd(acc[0], acc[NS], r.f.b * u1);
  }
  __device__ void finalize(double* a, Scal* s) const { s->fun = -a[0] / 2; }
};

// ---------------------------------------------------------------------------
// Green's function set-up kernels

// Love's displacement at (x, y) under unit pressure on one pixel of
// half-widths a, b centred at the origin (Johnson 1985, p. 54).
__device__ double love_patch(double x, double y, double a, double b, double young) {
  const double xp = x + a, xm = x - a, yp = y + b, ym = y - b;
  auto term = [](double p, double q, double r, double s) {
    return p * log((q + sqrt(q * q + p * p)) / (r + sqrt(r * r + s * s)));
  };
  return (term(xp, yp, ym, xp) + term(yp, xp, xm, yp) + term(xm, ym, yp, xm) + term(ym, xm, xp, ym)) /
         (M_PI * young);
}

// Free half-space: the Love kernel on the wrapped 2N grid, coordinate i*dx
// for i <= nx and (i - 2nx)*dx above (FFTElasticHalfSpace.py:1041-1137).
__global__ void k_love_kernel(int nx, int ny, double dx, double dy, double young, double* out) {
  const int Mx = 2 * nx, My = 2 * ny;
  const size_t M = (size_t)Mx * My;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < M; idx += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(idx / My), j = (int)(idx - (size_t)i * My);
    out[idx] = love_patch((i <= nx ? i : i - Mx) * dx, (j <= ny ? j : j - My) * dy, dx * 0.5, dy * 0.5, young);
  }
}

// Semi-periodic half-space (FFTElasticHalfSpace.py:1365-1497): the Love
// kernel at numpy-fftfreq-wrapped coordinates (i*dx for i <= (Mx-1)/2, else
// (i - Mx)*dx), summed over the images shifted by ix*sx, |ix| <= nimx, and
// iy*sy, |iy| <= nimy (images in the x-outer, y-inner order of the
// reference).  The grid holds `tile` periods of My1 points along y; each holds
// the kernel of one period divided by tile, so that a convolution over the
// tiled grid of a field tiled the same way is the one-period convolution,
// tiled.
__global__ void k_semi_kernel(int Mx, int My1, int tile, double dx, double dy, double sx, double sy, int nimx,
                              int nimy, double young, double* out) {
  const int My = My1 * tile;
  const size_t M = (size_t)Mx * My;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < M; idx += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(idx / My), j = (int)(idx - (size_t)i * My) % My1;
    const double x0 = (i <= (Mx - 1) / 2 ? i : i - Mx) * dx;
    const double y0 = (j <= (My1 - 1) / 2 ? j : j - My1) * dy;
    double acc = 0.0;
    for (int ix = -nimx; ix <= nimx; ++ix)
      for (int iy = -nimy; iy <= nimy; ++iy) acc += love_patch(x0 + ix * sx, y0 + iy * sy, dx * 0.5, dy * 0.5, young);
    out[idx] = acc / tile;
  }
}

// Real part of the packed forward spectrum in the (N2+1) x Mx Green layout.
__global__ void k_unpack_green(Geom g, const cplx<double>* S, double* out) {
  const size_t n = (size_t)(g.N2 + 1) * g.Mx;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += (size_t)gridDim.x * blockDim.x) {
    const int ky = (int)(idx / g.Mx), x = (int)(idx - (size_t)ky * g.Mx);
    double v;
    if (ky > 0 && ky < g.N2) {
      v = S[((size_t)(ky / g.Cb) * g.Mx + x) * g.Cb + ky % g.Cb].x;
    } else {
      const int xm = (g.Mx - x) % g.Mx;
      cplx<double> q = S[(size_t)x * g.Cb], qm = S[(size_t)xm * g.Cb];
      v = ky == 0 ? 0.5 * (q.x + qm.x) : 0.5 * (q.y + qm.y);
    }
    out[idx] = v;
  }
}

// ---------------------------------------------------------------------------
// k-space helpers (FFTElasticHalfSpace.py:613-885) in the reference's fourier
// layout: (Mx/2 + 1) x My complex, C order, r2c along x (:119-128).

// Spectrum value X[kx, ky] of a real field from the column-blocked half
// spectrum (ky along the r2c axis, column 0 packing the DC and Nyquist
// columns); ky > My/2 by Hermitian symmetry X[kx, ky] = conj X[-kx, -ky].
template <typename T>
__device__ __forceinline__ cplx<T> spectrum_at(const Geom& g, const cplx<T>* S, int kx, int ky) {
  bool cj = false;
  if (ky > g.N2) {
    ky = g.My - ky;
    kx = (g.Mx - kx) % g.Mx;
    cj = true;
  }
  cplx<T> v;
  if (ky > 0 && ky < g.N2) {
    v = S[((size_t)(ky / g.Cb) * g.Mx + kx) * g.Cb + ky % g.Cb];
  } else {
    const int km = (g.Mx - kx) % g.Mx;
    const cplx<T> q = S[(size_t)kx * g.Cb], qm = S[(size_t)km * g.Cb];
    v = ky == 0 ? mk(T(0.5) * (q.x + qm.x), T(0.5) * (q.y - qm.y))     // (Q + conj Q(-k)) / 2
                : mk(T(0.5) * (q.y + qm.y), T(0.5) * (qm.x - q.x));    // (Q - conj Q(-k)) / 2i
  }
  return cj ? conj(v) : v;
}

// out[kx, ky] = scale * W(kx, ky) * X[kx, ky] (W = 1 for kind 0, else the
// real kernel W, (N2+1) x Mx rows by |ky|).
template <typename T>
__global__ void k_unpack_fourier(Geom g, const cplx<T>* S, const T* W, T scale, cplx<T>* out) {
  const int Rx = g.Mx / 2 + 1;
  const size_t n = (size_t)Rx * g.My;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += (size_t)gridDim.x * blockDim.x) {
    const int kx = (int)(idx / g.My), ky = (int)(idx - (size_t)kx * g.My);
    cplx<T> v = spectrum_at(g, S, kx, ky);
    if (W) v = v * (scale * W[(size_t)(ky <= g.N2 ? ky : g.My - ky) * g.Mx + kx]);
    out[idx] = v;
  }
}

// evaluate_k_force_k (:650-666) on the fourier layout: out = scale * K * in.
template <typename T>
__global__ void k_kforce_k(Geom g, const T* K, T scale, const cplx<T>* in, cplx<T>* out) {
  const size_t n = (size_t)(g.Mx / 2 + 1) * g.My;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += (size_t)gridDim.x * blockDim.x) {
    const int kx = (int)(idx / g.My), ky = (int)(idx - (size_t)kx * g.My);
    out[idx] = in[idx] * (scale * K[(size_t)(ky <= g.N2 ? ky : g.My - ky) * g.Mx + kx]);
  }
}

// One compensated FP64 sum across the blocks of a kernel, result to *res.
struct SumRed {
  static constexpr int NS = 1;
  __device__ static constexpr int op(int) { return RCSUM; }
};

// evaluate_scalar_product_k_space (:686-805): sum over the fourier layout of
// w(kx) Re(conj(a) b) / (Mx My), w = 1 for kx = 0, 2 inside, and for the last
// row kx = Mx/2 1 if Mx is even (it holds its own mirror) or 2 if odd.
template <typename T>
__global__ void __launch_bounds__(256) k_sum_k(Geom g, const cplx<T>* a, const cplx<T>* b, double* part,
                                               unsigned* ctr, double* res) {
  const int Rx = g.Mx / 2 + 1;
  const size_t n = (size_t)Rx * g.My;
  double acc[2] = {0.0, 0.0};
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += (size_t)gridDim.x * blockDim.x) {
    const int kx = (int)(idx / g.My);
    const double w = kx == 0 ? 1.0 : (kx == Rx - 1 ? (g.Mx % 2 == 0 ? 1.0 : 2.0) : 2.0);
    const cplx<T> u = a[idx], v = b[idx];
    kadd(acc[0], acc[1], w * ((double)u.x * (double)v.x + (double)u.y * (double)v.y));
  }
  if (reduce_blocks<SumRed>(acc, part, ctr, gridDim.x, blockIdx.x) && threadIdx.x == 0)
    *res = kvalue(acc[0], acc[1]) / ((double)g.Mx * g.My);
}

// sum_i a_i b_i over the (Mx, My) grid (evaluate_elastic_energy, :676-684).
template <typename T>
__global__ void __launch_bounds__(256) k_dot(size_t n, const T* a, const T* b, double* part, unsigned* ctr,
                                             double* res) {
  double acc[2] = {0.0, 0.0};
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    kadd(acc[0], acc[1], (double)a[i] * (double)b[i]);
  if (reduce_blocks<SumRed>(acc, part, ctr, gridDim.x, blockIdx.x) && threadIdx.x == 0) *res = kvalue(acc[0], acc[1]);
}

}  // namespace cm

// ===========================================================================
// Host side

using namespace cm;

// Device scratch of the allocation-free solve (cm_ccg_solve_noalloc): the
// whole workspace of one solve -- conjugate direction, spectra, FP64 gap and
// forces of FP32 plans, scalar record, partial reductions, tickets, instance
// tables -- lives in this file-scope array.  Every GPU the library runs on
// holds its own copy until the process ends; host code reaches it only through
// the address hipGetSymbolAddress returns for the current device (cm_ctx).
constexpr size_t kScratchBytes = (size_t)136 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

struct cm_ctx {
  int device = 0;
  size_t budget = 0;
  size_t used = 0;
  hipStream_t stream = nullptr;
  void* scratch = nullptr;  // device address of `scratch` on this context's device
  std::string err;
  std::mutex mu;
};

namespace {

thread_local std::string g_err_noctx;

int fail(cm_ctx* ctx, int code, const std::string& msg) {
  if (ctx) {
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->err = msg;
  } else {
    g_err_noctx = msg;
  }
  return code;
}

#define HIPCHK(ctx, expr)                                                                       \
  do {                                                                                          \
    hipError_t e_ = (expr);                                                                     \
    if (e_ != hipSuccess)                                                                       \
      return fail((ctx), CM_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));         \
  } while (0)

int dev_alloc(cm_ctx* ctx, void** p, size_t bytes) {
  if (ctx->budget && ctx->used + bytes > ctx->budget)
    return fail(ctx, CM_E_OOM, "device memory budget exceeded");
  hipError_t e = hipMalloc(p, bytes ? bytes : 16);
  if (e != hipSuccess) return fail(ctx, CM_E_OOM, std::string("hipMalloc: ") + hipGetErrorString(e));
  ctx->used += bytes;
  return 0;
}
void dev_free(cm_ctx* ctx, void* p, size_t bytes) {
  if (p) {
    (void)hipFree(p);
    ctx->used -= std::min(ctx->used, bytes);
  }
}

// Radix sequence and threads-per-line for a line of length L.
// The factors 2, 3 and 5 go to unrolled radices 16, 8, 4, 5, 3, 2 (largest
// first).  A pass of radix R has L/R butterflies dealt round-robin to tpl
// threads; a thread holds at most ceil(L/R/tpl)*R elements, which must fit
// the kernel's register budget emax.  Any other prime factor p gets a generic
// pass (one thread per output element, ceil(L/tpl) <= emax; the 16-element
// GEN builds only); such lines begin and end with an identity pass, so the kernels' register-ended first
// and last passes see radix 1 (element tl + e*tpl in register slot e).
// Prefer tpl dividing every L/R (balanced passes), then larger tpl.
bool make_line_plan(int L, int max_tpl, int emax, LinePlan& best) {
  if (L < 2) return false;
  LinePlan lp{};
  lp.L = L;
  std::vector<int> unrolled, generic;
  int rem = L;
  while (rem > 1) {
    int pick = 0;
    for (int R : {16, 8, 4, 5, 3, 2})
      if (R <= emax && rem % R == 0) {
        pick = R;
        break;
      }
    if (!pick) break;
    unrolled.push_back(pick);
    rem /= pick;
  }
  for (int p = 7; rem > 1 && (long)p * p <= rem; p += 2)
    while (rem % p == 0) {
      generic.push_back(p);
      rem /= p;
    }
  if (rem > 1) {
    if (rem % 2 == 0 || rem % 3 == 0 || rem % 5 == 0) return false;  // a factor the budget emax left over
    generic.push_back(rem);
  }
  if (!generic.empty() && emax != 16) return false;  // generic passes exist in the 16-element GEN builds
  std::vector<int> seq;
  if (generic.empty()) {
    seq = unrolled;
  } else {
    seq.push_back(1);
    seq.insert(seq.end(), unrolled.begin(), unrolled.end());
    seq.insert(seq.end(), generic.begin(), generic.end());
    seq.push_back(1);
  }
  if ((int)seq.size() > LinePlan::kMaxPasses) return false;
  int np = 0;
  for (int R : seq) lp.set_radix(np++, R);
  lp.npass = np;
  lp.rlast = lp.radix(np - 1);
  int need = 1;  // smallest tpl meeting the register budget
  for (int s = 0; s < np; ++s) {
    const int R = lp.radix(s);
    if (!unrolled_radix(R)) {
      need = std::max(need, (L + emax - 1) / emax);
      continue;
    }
    const int nbmax = emax / R;
    if (nbmax < 1) return false;
    need = std::max(need, (L / R + nbmax - 1) / nbmax);
  }
  if (need > max_tpl) return false;
  int tpl = need;
  for (int t = need; t <= max_tpl && t <= L; ++t) {
    bool balanced = true;
    for (int s = 0; s < np; ++s) {
      const int R = lp.radix(s);
      balanced = balanced && (unrolled_radix(R) ? L / R : L) % t == 0;
    }
    if (balanced) {
      tpl = t;
      break;
    }
  }
  lp.tpl = tpl;
  int E = 0;
  for (int s = 0; s < np; ++s) {
    const int R = lp.radix(s);
    E = std::max(E, unrolled_radix(R) ? ((L / R + tpl - 1) / tpl) * R : (L + tpl - 1) / tpl);
  }
  lp.E = E;
  best = lp;
  return true;
}

template <typename T>
struct TwTables {
  cplx<T>* row = nullptr;   // N2 entries
  cplx<T>* col = nullptr;   // Mx entries
  cplx<T>* post = nullptr;  // N2/2+1 entries: exp(-2 pi i k / My)
};

// tw[k] = exp(-2 pi i k / denom), k < n, rounded from long double.
template <typename T>
std::vector<cplx<T>> host_twiddles(int n, int denom) {
  std::vector<cplx<T>> v(n);
  const long double twopi = 6.283185307179586476925286766559005768L;
  for (int k = 0; k < n; ++k) {
    const long double ang = twopi * (long double)k / (long double)denom;
    v[k].x = (T)cosl(ang);
    v[k].y = (T)(-sinl(ang));
  }
  return v;
}

}  // namespace

// One FFT executor per precision.
struct FftExec {
  Geom g{};
  LinePlan row{}, col{}, col_rev{};  // col_rev: col with the radix order reversed
  int emax_row = 16, emax_col = 16;
  bool gen_row = false, gen_col = false;  // a generic radix: the 16-element GEN builds run the line
  void* tw_row = nullptr;
  void* tw_col = nullptr;
  void* tw_post = nullptr;
  size_t lds_row = 0, lds_col = 0;
};

struct cm_plan {
  cm_ctx* ctx = nullptr;
  cm_substrate_desc d{};
  int Mx = 0, My = 0;
  double area_per_pt = 0;
  int prec = CM_FP64;
  FftExec ex;            // executor in plan precision
  void* G = nullptr;     // (N2+1) x Mx, plan precision, scaled for evaluate_disp
  void* K = nullptr;     // (N2+1) x Mx, plan precision, scaled for evaluate_force
  // FP64 executor and kernels: FP32 plans use them for the set-up transforms
  // and the periodic FP64 displacement refresh of a solve; FP64 plans alias
  // ex, G and K
  FftExec ex64;
  void* G64 = nullptr;
  void* K64 = nullptr;
  std::vector<double> G_host;  // FP64 copy (scaled) for inspection
  void* S = nullptr;     // spectrum work buffer for plain operator calls
  // reduction workspace of the k-space helpers: partials, tickets, result
  double* red_part = nullptr;
  unsigned* red_ctr = nullptr;
  double* red_res = nullptr;
  size_t bytes_owned = 0;
  int nblk_elem = 1;
};

namespace {

// Launch shape of the FFT passes; 0 = choose automatically.
struct Shape {
  int Cb = 0;    // columns per column-pass block
  int R = 0;     // rows per row-pass block
  int erow = 0;  // register elements per thread, row transforms (16 or 32)
  int ecol = 0;  // register elements per thread, column transforms (16 or 32)
};

int emax_bucket(int E) { return E <= 8 ? 8 : (E <= 16 ? 16 : 32); }
bool has_generic(const LinePlan& lp) {
  for (int s = 0; s < lp.npass; ++s)
    if (!unrolled_radix(lp.radix(s))) return true;
  return false;
}
// Register build of a line plan: the smallest holding E elements; lines with
// a generic radix run the 16-element GEN builds.
int plan_emax(const LinePlan& lp) { return has_generic(lp) ? 16 : emax_bucket(lp.E); }

// Solves with 32-element row transforms are disabled: a 2048 x 2048 FP64
// solve at the explicit shape (Cb 2, R 2, 32-element rows, 8-element columns)
// ended in a GPU memory fault whose cause was not found (the host build of the
// same kernels runs that case with no out-of-bounds access).  Explicit
// 32-element shapes are refused, the autotuner does not try them, and a solve
// on a plan whose automatic shape needs 32-element rows (rows over 16384
// points) returns CM_E_UNSUPPORTED.  32-element columns stay: the FP64
// set-up and refresh passes of FP32 solves on 8192-point columns run them.
bool uses_e32(const FftExec& ex) { return plan_emax(ex.row) == 32; }
const char* const kE32Refused =
    "32-element row transforms are disabled for solves (rows over 16384 points, or explicit 32-element launch "
    "shapes): a solve with them ended in a GPU memory fault whose cause was not found";

template <typename T>
bool choose_shape(FftExec& ex, const Shape& sh) {
  Geom& g = ex.g;
  const int Mx = g.Mx;
  const size_t cs = sizeof(cplx<T>);
  constexpr size_t kLdsAuto = 80 * 1024;   // automatic shapes: two blocks per CU (160 KiB LDS)
  constexpr size_t kLdsMax = 160 * 1024;   // largest block allocation on gfx950
  // register elements per thread: 8 (lean kernels, four or more waves per
  // SIMD), more when the line does not fit a block otherwise
  auto plan_len = [&](int L, int e, LinePlan& lp) {
    if (e) return make_line_plan(L, fft_block_max(e), e, lp);
    return make_line_plan(L, fft_block_max(8), 8, lp) || make_line_plan(L, fft_block_max(16), 16, lp) ||
           make_line_plan(L, fft_block_max(32), 32, lp);
  };
  if (!plan_len(g.N2, sh.erow, ex.row) || !plan_len(Mx, sh.ecol, ex.col)) return false;
  const int row_max = fft_block_max(plan_emax(ex.row)), col_max = fft_block_max(plan_emax(ex.col));
  // Row blocks write runs of R*Cb spectrum values: aim at >= 128-byte runs
  // with blocks that leave room for two per CU.
  int Cb = 1;
  if (sh.Cb) {
    Cb = sh.Cb;
    if ((Cb & (Cb - 1)) || g.N2 % Cb || ex.col.tpl * Cb > col_max) return false;
  } else {
    const int cb_target = (int)(64 / cs) > 0 ? (int)(64 / cs) : 1;
    while (Cb * 2 <= cb_target && g.N2 % (Cb * 2) == 0 && lds_bytes<T>((size_t)Mx * Cb * 2) <= kLdsAuto &&
           ex.col.tpl * Cb * 2 <= col_max)
      Cb *= 2;
  }
  if (lds_bytes<T>((size_t)Mx * Cb) > kLdsMax) return false;
  int R = 1;
  if (sh.R) {
    R = sh.R;
    if ((R & (R - 1)) || Mx % R || ex.row.tpl * R > row_max) return false;
  } else {
    // FP64 row blocks of at most 256 threads: on 2048^2 the CCG iteration
    // runs 2 % faster with R = 2 (256 threads, 64-byte runs) than with R = 4
    // (512 threads, 128-byte runs), the r2c/c2r passes gaining most
    const int row_auto = cs == 16 ? std::min(row_max, 256) : row_max;
    while (Mx % (R * 2) == 0 && ex.row.tpl * R * 2 <= row_auto &&
           lds_bytes<T>((size_t)R * 2 * g.N2) <= kLdsAuto && (size_t)R * Cb * cs < 128)
      R *= 2;
  }
  if (lds_bytes<T>((size_t)R * g.N2) > kLdsMax) return false;
  g.Cb = Cb;
  g.R = R;
  ex.lds_row = lds_bytes<T>((size_t)R * g.N2);
  ex.lds_col = lds_bytes<T>((size_t)Mx * Cb);
  ex.emax_row = plan_emax(ex.row);
  ex.emax_col = plan_emax(ex.col);
  ex.gen_row = has_generic(ex.row);
  ex.gen_col = has_generic(ex.col);
  ex.col_rev = ex.col;
  for (int q = 0; q < ex.col.npass; ++q) ex.col_rev.set_radix(q, ex.col.radix(ex.col.npass - 1 - q));
  ex.col_rev.rlast = ex.col_rev.radix(ex.col.npass - 1);
  return true;
}

template <typename T>
int exec_setup(cm_ctx* ctx, FftExec& ex, int nx, int ny, int Mx, int My, size_t& owned) {
  Geom& g = ex.g;
  g.nx = nx;
  g.ny = ny;
  g.Mx = Mx;
  g.My = My;
  g.N2 = My / 2;
  const size_t cs = sizeof(cplx<T>);
  if (!choose_shape<T>(ex, Shape{}))
    return fail(ctx, CM_E_UNSUPPORTED,
                "grid " + std::to_string(Mx) + "x" + std::to_string(My) +
                    " not supported: a row of My/2 or a column of Mx points must fit one block (512 threads x 32 "
                    "elements) and a column must fit in LDS");
  auto twr = host_twiddles<T>(g.N2, g.N2);
  auto twc = host_twiddles<T>(Mx, Mx);
  auto twp = host_twiddles<T>(g.N2 / 2 + 1, My);
  int rc;
  if ((rc = dev_alloc(ctx, &ex.tw_row, twr.size() * cs))) return rc;
  if ((rc = dev_alloc(ctx, &ex.tw_col, twc.size() * cs))) return rc;
  if ((rc = dev_alloc(ctx, &ex.tw_post, twp.size() * cs))) return rc;
  owned += (twr.size() + twc.size() + twp.size()) * cs;
  HIPCHK(ctx, hipMemcpy(ex.tw_row, twr.data(), twr.size() * cs, hipMemcpyHostToDevice));
  HIPCHK(ctx, hipMemcpy(ex.tw_col, twc.data(), twc.size() * cs, hipMemcpyHostToDevice));
  HIPCHK(ctx, hipMemcpy(ex.tw_post, twp.data(), twp.size() * cs, hipMemcpyHostToDevice));
  return 0;
}

// Allow more than 64 KB of dynamic LDS for a kernel.  The limit set for each
// kernel is remembered and raised again whenever a launch needs more (a shape
// sweep asks for growing sizes); returns false if the runtime refuses.
bool allow_big_lds(const void* kern, size_t lds) {
  if (lds <= 65536) return true;
  static std::mutex mu;
  static std::vector<std::pair<const void*, size_t>> set;
  std::lock_guard<std::mutex> lk(mu);
  for (auto& e : set)
    if (e.first == kern) {
      if (e.second >= lds) return true;
      if (hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess) {
        (void)hipGetLastError();
        return false;
      }
      e.second = lds;
      return true;
    }
  if (hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  set.emplace_back(kern, lds);
  return true;
}

// A launch whose dynamic LDS the runtime refused is skipped and remembered
// here; the entry point that enqueued it reports CM_E_UNSUPPORTED.
thread_local bool t_lds_refused = false;
bool lds_ok(const void* kern, size_t lds) {
  if (allow_big_lds(kern, lds)) return true;
  t_lds_refused = true;
  return false;
}
int launch_status(cm_ctx* ctx) {
  if (t_lds_refused) {
    t_lds_refused = false;
    return fail(ctx, CM_E_UNSUPPORTED, "dynamic LDS size of a launch shape refused by the runtime");
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(ctx, CM_E_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
  return 0;
}

// Launchers: T is the transform precision of ex; the instance table holds
// instances with the functor's field precision P::F.
template <typename T, class P>
void launch_r2c(const FftExec& ex, InstArr<typename P::F> ia, int B, int slot, hipStream_t st) {
  dim3 grid(ex.g.Mx / ex.g.R, B), block(ex.row.tpl * ex.g.R);
  const cplx<T>* tw = (const cplx<T>*)ex.tw_row;
  const cplx<T>* twp = (const cplx<T>*)ex.tw_post;
  if (ex.gen_row) {
    if (lds_ok((const void*)k_r2c<T, 16, P, true>, ex.lds_row))
      hipLaunchKernelGGL((k_r2c<T, 16, P, true>), grid, block, ex.lds_row, st, ex.g, ex.row, tw, twp, ia, slot);
  } else if (ex.emax_row == 8) {
    if (lds_ok((const void*)k_r2c<T, 8, P>, ex.lds_row))
      hipLaunchKernelGGL((k_r2c<T, 8, P>), grid, block, ex.lds_row, st, ex.g, ex.row, tw, twp, ia, slot);
  } else if (ex.emax_row == 16) {
    if (lds_ok((const void*)k_r2c<T, 16, P>, ex.lds_row))
      hipLaunchKernelGGL((k_r2c<T, 16, P>), grid, block, ex.lds_row, st, ex.g, ex.row, tw, twp, ia, slot);
  } else {
    if (lds_ok((const void*)k_r2c<T, 32, P>, ex.lds_row))
      hipLaunchKernelGGL((k_r2c<T, 32, P>), grid, block, ex.lds_row, st, ex.g, ex.row, tw, twp, ia, slot);
  }
}

template <typename T, class P>
void launch_c2r(const FftExec& ex, InstArr<typename P::F> ia, int B, int slot, hipStream_t st) {
  dim3 grid(ex.g.Mx / ex.g.R, B), block(ex.row.tpl * ex.g.R);
  const cplx<T>* tw = (const cplx<T>*)ex.tw_row;
  const cplx<T>* twp = (const cplx<T>*)ex.tw_post;
  if (ex.gen_row) {
    if (lds_ok((const void*)k_c2r<T, 16, P, true>, ex.lds_row))
      hipLaunchKernelGGL((k_c2r<T, 16, P, true>), grid, block, ex.lds_row, st, ex.g, ex.row, tw, twp, ia, slot);
  } else if (ex.emax_row == 8) {
    if (lds_ok((const void*)k_c2r<T, 8, P>, ex.lds_row))
      hipLaunchKernelGGL((k_c2r<T, 8, P>), grid, block, ex.lds_row, st, ex.g, ex.row, tw, twp, ia, slot);
  } else if (ex.emax_row == 16) {
    if (lds_ok((const void*)k_c2r<T, 16, P>, ex.lds_row))
      hipLaunchKernelGGL((k_c2r<T, 16, P>), grid, block, ex.lds_row, st, ex.g, ex.row, tw, twp, ia, slot);
  } else {
    if (lds_ok((const void*)k_c2r<T, 32, P>, ex.lds_row))
      hipLaunchKernelGGL((k_c2r<T, 32, P>), grid, block, ex.lds_row, st, ex.g, ex.row, tw, twp, ia, slot);
  }
}

template <typename T, int MODE, typename TF = T>
void launch_col(const FftExec& ex, InstArr<TF> ia, int B, int useK, int skip, hipStream_t st) {
  dim3 grid(ex.g.N2 / ex.g.Cb, B), block(ex.col.tpl * ex.g.Cb);
  const cplx<T>* tw = (const cplx<T>*)ex.tw_col;
  if (ex.gen_col) {
    if (lds_ok((const void*)k_col<T, 16, MODE, TF, true>, ex.lds_col))
      hipLaunchKernelGGL((k_col<T, 16, MODE, TF, true>), grid, block, ex.lds_col, st, ex.g, ex.col, ex.col_rev, tw, ia,
                         useK, skip);
  } else if (ex.emax_col == 8) {
    if (lds_ok((const void*)k_col<T, 8, MODE, TF>, ex.lds_col))
      hipLaunchKernelGGL((k_col<T, 8, MODE, TF>), grid, block, ex.lds_col, st, ex.g, ex.col, ex.col_rev, tw, ia, useK,
                         skip);
  } else if (ex.emax_col == 16) {
    if (lds_ok((const void*)k_col<T, 16, MODE, TF>, ex.lds_col))
      hipLaunchKernelGGL((k_col<T, 16, MODE, TF>), grid, block, ex.lds_col, st, ex.g, ex.col, ex.col_rev, tw, ia,
                         useK, skip);
  } else {
    if (lds_ok((const void*)k_col<T, 32, MODE, TF>, ex.lds_col))
      hipLaunchKernelGGL((k_col<T, 32, MODE, TF>), grid, block, ex.lds_col, st, ex.g, ex.col, ex.col_rev, tw, ia,
                         useK, skip);
  }
}

template <typename T, class P>
void launch_elem(const FftExec& ex, int nblk, InstArr<T> ia, int B, int slot, hipStream_t st) {
  dim3 grid(nblk, B), block(256);
  hipLaunchKernelGGL((k_elem<T, P>), grid, block, 0, st, ex.g, ia, slot);
}

int plan_free(cm_plan* p) {
  if (!p) return 0;
  cm_ctx* ctx = p->ctx;
  dev_free(ctx, p->ex.tw_row, 0);
  dev_free(ctx, p->ex.tw_col, 0);
  dev_free(ctx, p->ex.tw_post, 0);
  dev_free(ctx, p->G, 0);
  dev_free(ctx, p->K, 0);
  dev_free(ctx, p->S, 0);
  dev_free(ctx, p->red_part, 0);
  dev_free(ctx, p->red_ctr, 0);
  dev_free(ctx, p->red_res, 0);
  if (p->prec != CM_FP64) {
    dev_free(ctx, p->ex64.tw_row, 0);
    dev_free(ctx, p->ex64.tw_col, 0);
    dev_free(ctx, p->ex64.tw_post, 0);
    dev_free(ctx, p->G64, 0);
    dev_free(ctx, p->K64, 0);
  }
  ctx->used -= std::min(ctx->used, p->bytes_owned);
  delete p;
  return 0;
}

// Periodic Green's function on the host in FP64 (FFTElasticHalfSpace.py:494-561)
// in the (N2+1) x Mx layout: ky along the r2c axis, x along the full axis.
void periodic_green_host(const cm_substrate_desc& d, int Mx, int My, std::vector<double>& G) {
  const int N2 = My / 2;
  G.assign((size_t)(N2 + 1) * Mx, 0.0);
  const double Es = d.young / (1 - d.poisson * d.poisson);
  const bool thick = d.thickness > 0 && std::isfinite(d.thickness);
  auto qcomp = [](int k, int n, double s) { return (k <= n / 2 ? k : n - k) / s; };
  for (int ky = 0; ky <= N2; ++ky) {
    const double qy = qcomp(ky, My, d.sy);
    for (int x = 0; x < Mx; ++x) {
      const double qx = qcomp(x, Mx, d.sx);
      double q = std::sqrt(qx * qx + qy * qy);
      double stiff;
      if (x == 0 && ky == 0) {
        if (thick) {
          stiff = d.young / d.thickness * (1 - d.poisson) / ((1 - 2 * d.poisson) * (1 + d.poisson));
        } else if (d.q0_mode == 0) {
          stiff = (M_PI * Es * (1.0 / d.sx) + M_PI * Es * (1.0 / d.sy)) / 2;
        } else if (d.q0_mode == 1) {
          stiff = 1.0;
        } else {
          stiff = d.stiffness_q0;
        }
      } else {
        stiff = M_PI * Es * q;
        if (thick) {
          const double qt = q * 2 * M_PI * d.thickness;
          const double fac = 3 - 4 * d.poisson;
          const double off = 4 * d.poisson * (2 * d.poisson - 3) + 5;
          double corr = (fac * std::cosh(2 * qt) + 2 * qt * qt + off) / (fac * std::sinh(2 * qt) - 2 * qt);
          if (std::isnan(corr)) corr = 1.0;
          stiff *= corr;
        }
      }
      double gv = 1.0 / stiff;
      if (x == 0 && ky == 0 && !thick && d.q0_mode == 1) gv = 0.0;
      G[(size_t)ky * Mx + x] = gv;
    }
  }
}

// Free and semi-periodic Green's functions: the real-space kernel on the
// computational grid, FP64 forward transform on the device, real part
// unpacked (FFTElasticHalfSpace.py:1033-1139, 1353-1398; both kernels are even
// in x and y, so their spectra are real).
int free_green_device(cm_plan* P, std::vector<double>& G) {
  cm_ctx* ctx = P->ctx;
  FftExec ex;
  size_t owned = 0;
  int rc = exec_setup<double>(ctx, ex, P->d.nx, P->d.ny, P->Mx, P->My, owned);
  if (rc) return rc;
  const size_t M = (size_t)P->Mx * P->My;
  double* field = nullptr;
  cplx<double>* S = nullptr;
  double* Gd = nullptr;
  const size_t gn = (size_t)(ex.g.N2 + 1) * P->Mx;
  if ((rc = dev_alloc(ctx, (void**)&field, M * sizeof(double)))) return rc;
  if ((rc = dev_alloc(ctx, (void**)&S, (M / 2) * sizeof(cplx<double>)))) return rc;
  if ((rc = dev_alloc(ctx, (void**)&Gd, gn * sizeof(double)))) return rc;
  hipStream_t st = ctx->stream;
  const double dx = P->d.sx / P->d.nx, dy = P->d.sy / P->d.ny;
  if (P->d.bc == CM_FREE) {
    hipLaunchKernelGGL(k_love_kernel, dim3(1024), dim3(256), 0, st, P->d.nx, P->d.ny, dx, dy, P->d.young, field);
  } else {
    // semi-periodic: images along the periodic directions; a tiled plan
    // (two periods along y) sums the images of one period
    const int tile = std::max(1, P->d.tile_y);
    const int nim = std::max(0, P->d.n_images);
    hipLaunchKernelGGL(k_semi_kernel, dim3(1024), dim3(256), 0, st, P->Mx, P->My / tile, tile, dx, dy,
                       P->d.sx, P->d.sy / tile, P->d.periodic_x ? nim : 0, P->d.periodic_y ? nim : 0, P->d.young,
                       field);
  }
  InstArr<double> ia{};
  ia.arr = nullptr;
  ia.one.src = field;
  ia.one.S = S;
  launch_r2c<double, ProdCopy<double>>(ex, ia, 1, 0, st);
  launch_col<double, 1>(ex, ia, 1, 0, 0, st);
  if ((rc = launch_status(ctx))) return rc;
  hipLaunchKernelGGL(k_unpack_green, dim3(1024), dim3(256), 0, st, ex.g, S, Gd);
  G.resize(gn);
  HIPCHK(ctx, hipMemcpyAsync(G.data(), Gd, gn * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(ctx, hipStreamSynchronize(st));
  dev_free(ctx, field, M * sizeof(double));
  dev_free(ctx, S, (M / 2) * sizeof(cplx<double>));
  dev_free(ctx, Gd, gn * sizeof(double));
  dev_free(ctx, ex.tw_row, 0);
  dev_free(ctx, ex.tw_col, 0);
  dev_free(ctx, ex.tw_post, 0);
  ctx->used -= std::min(ctx->used, owned);
  return 0;
}

// Scaled evaluate_disp / evaluate_force kernels in precision T into Gd, Kd.
template <typename T>
int upload_kernels(cm_plan* P, const std::vector<double>& Graw, void*& Gd, void*& Kd) {
  cm_ctx* ctx = P->ctx;
  const size_t n = Graw.size();
  const double M = (double)P->Mx * P->My;
  std::vector<T> Gs(n), Ks(n);
  P->G_host.resize(n);
  for (size_t i = 0; i < n; ++i) {
    const double gv = Graw[i];
    const double kv = gv != 0.0 ? 1.0 / gv : 0.0;  // FFTElasticHalfSpace.py:563-571
    const double gs = gv * (-1.0 / (P->area_per_pt * M));
    P->G_host[i] = gs;
    Gs[i] = (T)gs;
    Ks[i] = (T)(kv * (-P->area_per_pt / M));
  }
  int rc;
  if ((rc = dev_alloc(ctx, &Gd, n * sizeof(T)))) return rc;
  if ((rc = dev_alloc(ctx, &Kd, n * sizeof(T)))) return rc;
  P->bytes_owned += 2 * n * sizeof(T);
  HIPCHK(ctx, hipMemcpy(Gd, Gs.data(), n * sizeof(T), hipMemcpyHostToDevice));
  HIPCHK(ctx, hipMemcpy(Kd, Ks.data(), n * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

constexpr int kRedBlocks = 1024;  // blocks of the k-space helper reductions

template <typename T>
int plan_build(cm_plan* P) {
  cm_ctx* ctx = P->ctx;
  int rc = exec_setup<T>(ctx, P->ex, P->d.nx, P->d.ny, P->Mx, P->My, P->bytes_owned);
  if (rc) return rc;
  std::vector<double> Graw;
  if (P->d.bc == CM_PERIODIC) {
    const int tile = std::max(1, P->d.tile_y);
    if (tile == 1) {
      periodic_green_host(P->d, P->Mx, P->My, Graw);
    } else {
      // the grid holds `tile` periods along y: the kernel of one period at
      // the frequencies a tiled field has (ky a multiple of tile), zero at
      // the others
      cm_substrate_desc d1 = P->d;
      d1.ny /= tile;
      d1.sy /= tile;
      std::vector<double> G1;
      periodic_green_host(d1, P->Mx, P->My / tile, G1);
      const int N2 = P->My / 2;
      Graw.assign((size_t)(N2 + 1) * P->Mx, 0.0);
      for (int ky = 0; ky <= N2; ky += tile)
        std::copy(G1.begin() + (size_t)(ky / tile) * P->Mx, G1.begin() + (size_t)(ky / tile + 1) * P->Mx,
                  Graw.begin() + (size_t)ky * P->Mx);
    }
  } else {
    if ((rc = free_green_device(P, Graw))) return rc;
    // a tiled semi-periodic grid: the transform of the tiled kernel vanishes
    // at odd ky up to round-off; exact zeros there keep 1/G (evaluate_force)
    // from amplifying that round-off
    const int tile = std::max(1, P->d.tile_y);
    for (int ky = 0; tile > 1 && ky <= P->My / 2; ++ky)
      if (ky % tile) std::fill(Graw.begin() + (size_t)ky * P->Mx, Graw.begin() + (size_t)(ky + 1) * P->Mx, 0.0);
  }
  if ((rc = upload_kernels<T>(P, Graw, P->G, P->K))) return rc;
  if (std::is_same<T, double>::value) {
    P->ex64 = P->ex;
    P->G64 = P->G;
    P->K64 = P->K;
  } else {
    if ((rc = exec_setup<double>(ctx, P->ex64, P->d.nx, P->d.ny, P->Mx, P->My, P->bytes_owned))) return rc;
    if ((rc = upload_kernels<double>(P, Graw, P->G64, P->K64))) return rc;
  }
  const size_t Sb = (size_t)P->Mx * (P->My / 2) * sizeof(cplx<T>);
  if ((rc = dev_alloc(ctx, &P->S, Sb))) return rc;
  P->bytes_owned += Sb;
  // reduction workspace of the k-space helpers (tickets start at zero; the
  // last block of every reduction resets them)
  if ((rc = dev_alloc(ctx, (void**)&P->red_part, kRedBlocks * 2 * sizeof(double)))) return rc;
  if ((rc = dev_alloc(ctx, (void**)&P->red_ctr, kCtrSlot * sizeof(unsigned)))) return rc;
  if ((rc = dev_alloc(ctx, (void**)&P->red_res, sizeof(double)))) return rc;
  P->bytes_owned += kRedBlocks * 2 * sizeof(double) + kCtrSlot * sizeof(unsigned) + sizeof(double);
  HIPCHK(ctx, hipMemset(P->red_ctr, 0, kCtrSlot * sizeof(unsigned)));
  const size_t M = (size_t)P->Mx * P->My;
  // element-wise kernels: blocks stride over rows; few enough blocks that the
  // last block's pass over the partials stays short
  // element-wise passes: up to 1024 blocks (four per CU); blocks loop over
  // rows, so waves drift out of step and loads of one overlap the arithmetic
  // and reductions of another (measured at 2048^2: 1024 blocks beat 512 and
  // 2048 by 4-7 us per pass)
  P->nblk_elem = (int)std::max<size_t>(1, std::min<size_t>(std::min<size_t>(1024, (size_t)P->Mx), (M + 1023) / 1024));
  return 0;
}

template <typename T>
void evaluate_ex(cm_plan* P, const FftExec& ex, const void* in, void* out, int useK, hipStream_t st) {
  InstArr<T> ia{};
  ia.arr = nullptr;
  ia.one.src = (const T*)in;
  ia.one.dst = (T*)out;
  ia.one.S = (cplx<T>*)P->S;
  ia.one.G = (const T*)P->G;
  ia.one.K = (const T*)P->K;
  launch_r2c<T, ProdCopy<T>>(ex, ia, 1, 0, st);
  launch_col<T, 0>(ex, ia, 1, useK, 0, st);
  launch_c2r<T, ConsCopy<T>>(ex, ia, 1, 0, st);
}

// Time the three passes of evaluate_disp over candidate launch shapes and keep
// the fastest (columns per block, rows per block, register elements).
template <typename T>
int autotune(cm_plan* P, int reps, std::string& desc) {
  cm_ctx* ctx = P->ctx;
  hipStream_t st = ctx->stream;
  const size_t M = (size_t)P->Mx * P->My;
  void* a = nullptr;
  void* b = nullptr;
  int rc;
  if ((rc = dev_alloc(ctx, &a, M * sizeof(T)))) return rc;
  if ((rc = dev_alloc(ctx, &b, M * sizeof(T)))) {
    dev_free(ctx, a, M * sizeof(T));
    return rc;
  }
  HIPCHK(ctx, hipMemsetAsync(a, 0, M * sizeof(T), st));
  hipEvent_t e0, e1;
  HIPCHK(ctx, hipEventCreate(&e0));
  HIPCHK(ctx, hipEventCreate(&e1));
  FftExec best = P->ex;
  float best_us = 1e30f;
  int tried = 0;
  for (int Cb : {1, 2, 4, 8})
    for (int R : {1, 2, 4, 8})
      for (int er : {8, 16})  // (32-element shapes are disabled for solves)
        for (int ec : {8, 16}) {
          FftExec trial = P->ex;
          if (!choose_shape<T>(trial, Shape{Cb, R, er, ec})) continue;
          evaluate_ex<T>(P, trial, a, b, 0, st);  // warm-up (first launch of a kernel)
          if (t_lds_refused) {  // the runtime refused this shape's LDS: not a candidate
            t_lds_refused = false;
            continue;
          }
          if ((rc = launch_status(ctx))) return rc;
          HIPCHK(ctx, hipEventRecord(e0, st));
          for (int k = 0; k < reps; ++k) evaluate_ex<T>(P, trial, a, b, 0, st);
          HIPCHK(ctx, hipEventRecord(e1, st));
          HIPCHK(ctx, hipEventSynchronize(e1));
          if ((rc = launch_status(ctx))) return rc;
          float ms = 0;
          HIPCHK(ctx, hipEventElapsedTime(&ms, e0, e1));
          const float us = ms * 1000.f / reps;
          ++tried;
          if (us < best_us) {
            best_us = us;
            best = trial;
          }
        }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  dev_free(ctx, a, M * sizeof(T));
  dev_free(ctx, b, M * sizeof(T));
  P->ex = best;
  char buf[512];
  snprintf(buf, sizeof buf,
           "{\"Cb\": %d, \"R\": %d, \"erow\": %d, \"ecol\": %d, \"row_tpl\": %d, \"col_tpl\": %d, "
           "\"evaluate_us\": %.3f, \"candidates\": %d}",
           best.g.Cb, best.g.R, plan_emax(best.row), plan_emax(best.col), best.row.tpl, best.col.tpl,
           best_us, tried);
  desc = buf;
  return 0;
}

template <typename T>
int evaluate(cm_plan* P, const void* in, void* out, int useK, void* stream) {
  hipStream_t st = stream ? (hipStream_t)stream : P->ctx->stream;
  InstArr<T> ia{};
  ia.arr = nullptr;
  ia.one.src = (const T*)in;
  ia.one.dst = (T*)out;
  ia.one.S = (cplx<T>*)P->S;
  ia.one.G = (const T*)P->G;
  ia.one.K = (const T*)P->K;
  launch_r2c<T, ProdCopy<T>>(P->ex, ia, 1, 0, st);
  launch_col<T, 0>(P->ex, ia, 1, useK, 0, st);
  launch_c2r<T, ConsCopy<T>>(P->ex, ia, 1, 0, st);
  return launch_status(P->ctx);
}

// ---------------------------------------------------------------------------
// k-space helpers

// Forward transform of a real (Mx, My) field into the fourier layout, times
// M G (kind 1: G DFT(-f) / area_per_pt, since G holds -1/(area M)) or M K
// (kind 2: -K DFT(u) area_per_pt) or nothing (kind 0).
template <typename T>
int fourier_out(cm_plan* P, const void* field, void* out, int kind, void* stream) {
  hipStream_t st = stream ? (hipStream_t)stream : P->ctx->stream;
  InstArr<T> ia{};
  ia.arr = nullptr;
  ia.one.src = (const T*)field;
  ia.one.S = (cplx<T>*)P->S;
  launch_r2c<T, ProdCopy<T>>(P->ex, ia, 1, 0, st);
  launch_col<T, 1>(P->ex, ia, 1, 0, 0, st);
  const T* W = kind == 1 ? (const T*)P->G : (kind == 2 ? (const T*)P->K : nullptr);
  const T scale = (T)((double)P->Mx * P->My);
  hipLaunchKernelGGL(k_unpack_fourier<T>, dim3(1024), dim3(256), 0, st, P->ex.g, (const cplx<T>*)P->S, W, scale,
                     (cplx<T>*)out);
  return launch_status(P->ctx);
}

template <typename T>
int kforce_k(cm_plan* P, const void* in, void* out, void* stream) {
  hipStream_t st = stream ? (hipStream_t)stream : P->ctx->stream;
  const T scale = (T)((double)P->Mx * P->My);
  hipLaunchKernelGGL(k_kforce_k<T>, dim3(1024), dim3(256), 0, st, P->ex.g, (const T*)P->K, scale,
                     (const cplx<T>*)in, (cplx<T>*)out);
  return launch_status(P->ctx);
}

int reduction_blocks(size_t n) { return (int)std::max<size_t>(1, std::min<size_t>(kRedBlocks, (n + 255) / 256)); }

template <typename T>
int scalar_product_k(cm_plan* P, const void* a, const void* b, double* out, void* stream) {
  hipStream_t st = stream ? (hipStream_t)stream : P->ctx->stream;
  const size_t n = (size_t)(P->Mx / 2 + 1) * P->My;
  hipLaunchKernelGGL(k_sum_k<T>, dim3(reduction_blocks(n)), dim3(256), 0, st, P->ex.g, (const cplx<T>*)a,
                     (const cplx<T>*)b, P->red_part, P->red_ctr, P->red_res);
  int rc = launch_status(P->ctx);
  if (rc) return rc;
  HIPCHK(P->ctx, hipMemcpyAsync(out, P->red_res, sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(P->ctx, hipStreamSynchronize(st));
  return 0;
}

template <typename T>
int dot(cm_plan* P, const void* a, const void* b, double* out, void* stream) {
  hipStream_t st = stream ? (hipStream_t)stream : P->ctx->stream;
  const size_t n = (size_t)P->Mx * P->My;
  hipLaunchKernelGGL(k_dot<T>, dim3(reduction_blocks(n)), dim3(256), 0, st, n, (const T*)a, (const T*)b,
                     P->red_part, P->red_ctr, P->red_res);
  int rc = launch_status(P->ctx);
  if (rc) return rc;
  HIPCHK(P->ctx, hipMemcpyAsync(out, P->red_res, sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(P->ctx, hipStreamSynchronize(st));
  return 0;
}

// ---------------------------------------------------------------------------
// CCG driver

struct WsLayout {
  size_t t, S, sc, part, ctr, w, fm, S64, inst, inst0, total;
  size_t st_t, st_S, st_sc, st_part, st_ctr, st_w, st_fm, st_S64;  // per-instance strides
};

size_t al256(size_t v) { return (v + 255) & ~(size_t)255; }

WsLayout ws_layout(const cm_plan* P, int B) {
  const size_t w = P->prec == CM_FP64 ? 8 : 4;
  const size_t M = (size_t)P->Mx * P->My;
  // partials: up to 8 slots and their 8 correction terms per block of any
  // reducing launch (element-wise, row passes, column pass) of either executor
  const size_t nblk_max = std::max<size_t>({(size_t)P->nblk_elem, (size_t)(P->Mx / P->ex.g.R),
                                            (size_t)(P->ex.g.N2 / P->ex.g.Cb), (size_t)(P->Mx / P->ex64.g.R),
                                            (size_t)(P->ex64.g.N2 / P->ex64.g.Cb)});
  const size_t nparts = nblk_max * 16;
  WsLayout L{};
  L.st_t = al256(M * w);
  L.st_S = al256((M / 2) * 2 * w);
  L.st_sc = al256(sizeof(Scal) * (1 + kLogRing));  // the record and its log ring
  // (the bisection rounds keep their bin counts there too)
  L.st_part = al256(std::max(nparts * sizeof(double), (size_t)bisect_blocks(P->ex.g) * kBisStride * sizeof(unsigned)));
  // the tickets, then the bisection's bin counts
  L.st_ctr = al256(8 * kCtrSlot * sizeof(unsigned) + kBisSearch * sizeof(unsigned long long));
  // FP32 plans: FP64 gap and forces, FP64 spectrum of the refresh transforms
  L.st_w = w == 4 ? al256(M * 8) : 0;
  L.st_fm = w == 4 ? al256(M * 8) : 0;
  L.st_S64 = w == 4 ? al256((M / 2) * 16) : 0;
  const size_t inst_bytes = al256(std::max(sizeof(Inst<double>), sizeof(Inst<float>)) * B);
  size_t o = 0;
  L.t = o; o += L.st_t * B;
  L.S = o; o += L.st_S * B;
  L.sc = o; o += L.st_sc * B;
  L.part = o; o += L.st_part * B;
  L.ctr = o; o += L.st_ctr * B;
  L.w = o; o += L.st_w * B;
  L.fm = o; o += L.st_fm * B;
  L.S64 = o; o += L.st_S64 * B;
  L.inst = o; o += inst_bytes;
  L.inst0 = o; o += inst_bytes;
  L.total = o;
  return L;
}

void fill_stats(const Scal& s, cm_ccg_stats* o) {
  o->status = s.done ? s.status : CM_RUNNING;
  o->nit = s.it;
  o->nfev = s.nfev;
  o->last_mode = s.delta_str;
  o->offset = s.offset;
  o->fun = s.fun;
  o->area = s.A_contact;
  o->cg_area = s.A_cg;
  o->plastic_area = s.A_fl;
  o->total_force = s.total_force;
  o->rms_pen = s.rms_pen;
  o->max_pen = s.max_pen;
  o->max_pres = s.max_pres;
  o->pad_pres = s.pad_pres;
  o->maxdu = s.maxdu;
  o->tau = s.tau;
  o->mixfac = s.mix;
  o->nc_count = s.nc_count;
}

}  // namespace

// A solve in flight: B instances of one plan, their workspace (conjugate
// directions, spectra, scalar state, partial reductions) resident on the
// device between calls, and the captured iteration graphs.
struct cm_session {
  cm_plan* P = nullptr;
  FftExec ex;  // the plan's launch shapes when the session began (the workspace is sized for them)
  FftExec ex64;  // FP64 executor (FP32 plans: set-up and refresh transforms; FP64 plans: == ex)
  int refresh_every = 0;  // FP32 plans: every k-th iteration's displacement update is an FP64 refresh
  int B = 0;
  hipStream_t st = nullptr;
  void* ws = nullptr;
  size_t ws_bytes = 0;
  WsLayout L{};
  bool any_bisect = false;
  bool any_hard = false;  // an instance has a hardness (else the PK-only passes run)
  int use_graph = 0;
  long long issued = 0;
  int maxiter = 0;
  bool finished = false;  // ElemFinish has restored the displacements
  Scal* host = nullptr;  // pinned copy of the B scalar records
  void* hin = nullptr;   // pinned instance tables (main, set-up)
  void* hin0 = nullptr;
  std::vector<std::pair<long long, hipGraphExec_t>> graphs;  // key: chunk length, refresh phase
  std::vector<hipGraph_t> graph_defs;
};

namespace {

// Whether iteration number k (0-based, counted over the session) updates the
// displacements with an FP64 refresh: always for FP64 plans; for FP32 plans
// every refresh_every-th iteration, starting with the first.
bool refresh_iteration(const cm_session* S, long long k) {
  return S->P->prec == CM_FP64 || S->refresh_every <= 1 || k % S->refresh_every == 0;
}

// The second convolution of an iteration and the displacement pass: FP64 on
// FP64 plans; FP32 plans take the FP64 transforms of the new forces on
// refresh iterations and the FP32 transforms of the force change otherwise.
// The rounds of the plastic step's bisection (every round after the one that
// finishes it returns at once).
template <typename T>
void launch_bisect(const cm_session* S, const FftExec& ex, InstArr<T> ia, int B, hipStream_t st) {
  const int nb = bisect_blocks(ex.g);
  char* scb = (char*)S->ws + S->L.sc;
  for (int r = 0; r < kBisRounds; ++r)
    hipLaunchKernelGGL((k_bisect_round<T>), dim3(nb, B), dim3(256), 0, st, ex.g, ia, scb, S->L.st_sc, r, nb);
}

template <typename T, bool HARD>
void enqueue_displacement_update(cm_session* S, InstArr<T> ia, bool refresh) {
  const int B = S->B;
  hipStream_t st = S->st;
  if (refresh) {
    const FftExec& e64 = S->ex64;
    launch_r2c<double, ProdForce<T, double, false, HARD>>(e64, ia, B, 6, st);
    launch_col<double, 0, T>(e64, ia, B, 0, 1, st);
    launch_c2r<double, ConsDisp<T, double, false>>(e64, ia, B, 7, st);
  } else if constexpr (!std::is_same<T, double>::value) {
    const FftExec& ex = S->ex;
    launch_r2c<T, ProdForce<T, T, true, HARD>>(ex, ia, B, 6, st);
    launch_col<T, 0>(ex, ia, B, 0, 1, st);
    launch_c2r<T, ConsDisp<T, T, true>>(ex, ia, B, 7, st);
  }
}

template <typename T>
void enqueue_iteration(cm_session* S, bool refresh) {
  const FftExec& ex = S->ex;
  InstArr<T> ia{};
  ia.arr = (const Inst<T>*)((char*)S->ws + S->L.inst);
  const int B = S->B, nbe = S->P->nblk_elem;
  hipStream_t st = S->st;
  // phase A of iteration k+1 is fused into ConsDisp of iteration k; the first
  // one runs at set-up
  if (S->any_bisect) launch_bisect<T>(S, ex, ia, B, st);
  launch_r2c<T, ProdDir<T>>(ex, ia, B, 3, st);
  launch_col<T, 3>(ex, ia, B, 0, 2, st);  // tau from the spectrum: no inverse pass
  if (S->any_hard) {
    launch_elem<T, ElemPhaseB<T, true>>(ex, nbe, ia, B, 5, st);
    enqueue_displacement_update<T, true>(S, ia, refresh);
  } else {
    launch_elem<T, ElemPhaseB<T, false>>(ex, nbe, ia, B, 5, st);
    enqueue_displacement_update<T, false>(S, ia, refresh);
  }
}

void session_free(cm_session* S);

// Launch shapes, workspace layout and stream of a session of B instances
// (no allocation).
template <typename T>
void session_configure(cm_session* S, cm_plan* P, int B, const cm_ccg_params* prm, void* stream) {
  S->P = P;
  S->ex = P->ex;
  S->ex64 = std::is_same<T, double>::value ? P->ex : P->ex64;
  S->refresh_every = prm[0].refresh_every > 0 ? prm[0].refresh_every : kRefreshEvery;
  S->B = B;
  S->st = stream ? (hipStream_t)stream : P->ctx->stream;
  S->L = ws_layout(P, B);
  S->use_graph = prm[0].use_graph;
}

// Instance tables and scalar records (host buffers S->hin, S->hin0, S->host)
// and the set-up launches on the workspace S->ws; allocates nothing.
template <typename T>
int session_init(cm_session* S, const cm_ccg_params* prm, const double* const* hs, const uint8_t* const* surfs,
                 void* const* us, void* const* fs, uint8_t* const* cms);

template <typename T>
int session_begin(cm_plan* P, int B, const cm_ccg_params* prm, const double* const* hs,
                  const uint8_t* const* surfs, void* const* us, void* const* fs, uint8_t* const* cms,
                  void* stream, cm_session** out) {
  cm_ctx* ctx = P->ctx;
  if (uses_e32(P->ex) || (P->prec != CM_FP64 && uses_e32(P->ex64))) return fail(ctx, CM_E_UNSUPPORTED, kE32Refused);
  for (int b = 1; b < B; ++b)
    if (prm[b].init_mode != prm[0].init_mode)
      return fail(ctx, CM_E_ARG, "all instances of a batch must share init_mode");
  cm_session* S = new cm_session();
  session_configure<T>(S, P, B, prm, stream);
  int rc = dev_alloc(ctx, &S->ws, S->L.total);
  if (rc) {
    delete S;
    return rc;
  }
  S->ws_bytes = S->L.total;
  if (hipHostMalloc((void**)&S->host, sizeof(Scal) * B, 0) != hipSuccess ||
      hipHostMalloc(&S->hin, sizeof(Inst<T>) * B, 0) != hipSuccess ||
      hipHostMalloc(&S->hin0, sizeof(Inst<T>) * B, 0) != hipSuccess) {
    session_free(S);
    return fail(ctx, CM_E_OOM, "pinned host allocation failed");
  }
  if ((rc = session_init<T>(S, prm, hs, surfs, us, fs, cms))) {
    session_free(S);
    return rc;
  }
  *out = S;
  return 0;
}

template <typename T>
int session_init(cm_session* S, const cm_ccg_params* prm, const double* const* hs, const uint8_t* const* surfs,
                 void* const* us, void* const* fs, uint8_t* const* cms) {
  cm_plan* P = S->P;
  cm_ctx* ctx = P->ctx;
  const int B = S->B;
  int rc = 0;
  const WsLayout& L = S->L;
  char* base = (char*)S->ws;
  const size_t M = (size_t)P->Mx * P->My;
  if (P->d.tile_y > 1 && P->d.bc == CM_SEMI && !P->d.periodic_y)
    return fail(ctx, CM_E_UNSUPPORTED,
                "a plan repeating a non-periodic y has no solve (the contact set is one top-left block)");
  Inst<T>* hin = (Inst<T>*)S->hin;
  Inst<T>* hin0 = (Inst<T>*)S->hin0;
  Scal* hsc = S->host;
  for (int b = 0; b < B; ++b) {
    const cm_ccg_params& q = prm[b];
    Inst<T>& I = hin[b];
    std::memset(&I, 0, sizeof(I));
    I.u = (T*)us[b];
    I.w = std::is_same<T, double>::value ? (double*)us[b] : (double*)(base + L.w + b * L.st_w);
    I.f = (T*)fs[b];
    I.fm = std::is_same<T, double>::value ? (double*)fs[b] : (double*)(base + L.fm + b * L.st_fm);
    I.t = (T*)(base + L.t + b * L.st_t);
    I.cm = cms[b];
    I.h = hs[b];
    I.surf = surfs ? surfs[b] : nullptr;
    I.S = (cplx<T>*)(base + L.S + b * L.st_S);
    I.G = (const T*)P->G;
    I.K = (const T*)P->K;
    I.S64 = std::is_same<T, double>::value ? (cplx<double>*)I.S : (cplx<double>*)(base + L.S64 + b * L.st_S64);
    I.G64 = (const double*)P->G64;
    I.K64 = (const double*)P->K64;
    I.sc = (Scal*)(base + L.sc + b * L.st_sc);
    I.part = (double*)(base + L.part + b * L.st_part);
    I.ctr = (unsigned*)(base + L.ctr + b * L.st_ctr);
    // set-up table: the initial transform reads u (forces from displacements)
    // or f (displacements from initial forces) and writes the other field
    hin0[b] = I;
    if (q.init_mode == 0) {
      hin0[b].src = I.u;
    } else {
      hin0[b].src = I.f;
      hin0[b].dst = I.u;
    }
    Scal& s = hsc[b];
    std::memset(&s, 0, sizeof(Scal));
    s.F = q.external_force;
    s.H = q.hardness;
    s.pentol = q.pentol;
    s.forcetol = q.forcetol;
    s.thermotol = q.thermotol;
    s.mixfac = q.mixfac;
    s.mixdecfac = q.mixdecfac;
    s.nb_surface_pts = (double)P->d.nx * P->d.ny;
    s.n_pad = (double)M;  // ElemInit subtracts the number of computational points
    s.offset = std::isnan(q.offset) ? 0.0 : q.offset;  // offset=None -> 0 (:167-168)
    s.shift = s.offset;
    s.G_old = 1.0;  // :224
    s.mix = q.mixfac;
    s.load_ctl = q.load_control;
    s.has_H = q.has_hardness;
    s.H_truthy = q.has_hardness && q.hardness != 0.0;
    s.minmixsteps = q.minmixsteps;
    s.maxiter = q.maxiter;
    s.it = 0;
    s.nfev = 1;  // the initial evaluation (:217)
    s.delta_str = DS_RESET;
    if (q.maxiter < 1) {
      s.done = 1;
      s.status = CM_MAXITER;
    }
    S->any_bisect = S->any_bisect || (q.has_hardness && q.load_control);
    S->any_hard = S->any_hard || q.has_hardness;
    S->maxiter = std::max(S->maxiter, q.maxiter);
  }
  hipStream_t st = S->st;
  Inst<T>* dinst = (Inst<T>*)(base + L.inst);
  Inst<T>* dinst0 = (Inst<T>*)(base + L.inst0);
  for (int b = 0; b < B; ++b)
    HIPCHK(ctx, hipMemcpyAsync(hin[b].sc, &hsc[b], sizeof(Scal), hipMemcpyHostToDevice, st));
  HIPCHK(ctx, hipMemcpyAsync(dinst, hin, sizeof(Inst<T>) * B, hipMemcpyHostToDevice, st));
  HIPCHK(ctx, hipMemcpyAsync(dinst0, hin0, sizeof(Inst<T>) * B, hipMemcpyHostToDevice, st));
  HIPCHK(ctx, hipMemsetAsync(base + L.ctr, 0, L.st_ctr * B, st));
  HIPCHK(ctx, hipMemsetAsync(base + L.t, 0, L.st_t * B, st));
  InstArr<T> ia{};
  ia.arr = dinst;
  InstArr<T> ia0{};
  ia0.arr = dinst0;
  const FftExec& ex = S->ex;
  const FftExec& e64 = S->ex64;
  const int nbe = P->nblk_elem;
  // ---- set-up (:179-219); the initial transform is FP64 in both precisions
  launch_elem<T, ElemInit<T>>(ex, nbe, ia, B, 0, st);
  if (prm[0].init_mode == 0) {
    launch_r2c<double, ProdCopy<T, double>>(e64, ia0, B, 1, st);  // f = evaluate_force(u), f[pad] = 0
    launch_col<double, 0, T>(e64, ia0, B, 1, 0, st);
    launch_c2r<double, ConsInitForce<T, double>>(e64, ia0, B, 1, st);
  } else {
    launch_r2c<double, ProdCopy<T, double>>(e64, ia0, B, 1, st);  // u = evaluate_disp(f)
    launch_col<double, 0, T>(e64, ia0, B, 0, 0, st);
    launch_c2r<double, ConsCopy<T, double>>(e64, ia0, B, 1, st);
    launch_elem<T, ElemZeroPad<T>>(ex, nbe, ia, B, 1, st);
  }
  launch_elem<T, ElemPhaseA<T>>(ex, nbe, ia, B, 2, st);  // start of iteration 1
  return launch_status(ctx);
}

// Enqueue n iterations; instances that have finished turn the kernels into
// no-ops.  With use_graph, chunks of iterations replay a captured hipGraph.
template <typename T>
int session_step(cm_session* S, int n) {
  cm_ctx* ctx = S->P->ctx;
  if (n <= 0) return 0;
  if (S->use_graph && n >= 8) {
    // a captured chunk is valid for its length and for where it starts in
    // the refresh cycle of an FP32 plan
    const long long phase = S->refresh_every > 1 ? S->issued % S->refresh_every : 0;
    const long long key = (long long)n * 4096 + phase;
    hipGraphExec_t gx = nullptr;
    for (auto& g : S->graphs)
      if (g.first == key) gx = g.second;
    if (!gx) {
      hipGraph_t gdef = nullptr;
      hipError_t e = hipStreamBeginCapture(S->st, hipStreamCaptureModeThreadLocal);
      if (e == hipSuccess) {
        for (int k = 0; k < n; ++k) enqueue_iteration<T>(S, refresh_iteration(S, S->issued + k));
        e = hipStreamEndCapture(S->st, &gdef);
      }
      if (e == hipSuccess) e = hipGraphInstantiate(&gx, gdef, nullptr, nullptr, 0);
      if (e != hipSuccess) {
        (void)hipGetLastError();
        if (gdef) (void)hipGraphDestroy(gdef);
        S->use_graph = 0;  // capture unavailable on this stream: launch directly
        gx = nullptr;
      } else {
        S->graphs.emplace_back(key, gx);
        S->graph_defs.push_back(gdef);
      }
    }
    if (gx) {
      HIPCHK(ctx, hipGraphLaunch(gx, S->st));
      S->issued += n;
      return 0;
    }
  }
  for (int k = 0; k < n; ++k) enqueue_iteration<T>(S, refresh_iteration(S, S->issued + k));
  S->issued += n;
  return launch_status(ctx);
}

template <typename T>
int session_finish(cm_session* S) {
  const FftExec& ex = S->ex;
  InstArr<T> ia{};
  ia.arr = (const Inst<T>*)((char*)S->ws + S->L.inst);
  S->finished = true;
  launch_elem<T, ElemFinish<T>>(ex, S->P->nblk_elem, ia, S->B, 0, S->st);
  return 0;
}

// Per-kernel durations of one iteration (HIP events between launches).
template <typename T>
int session_profile(cm_session* S, int n, float* us) {
  cm_ctx* ctx = S->P->ctx;
  const FftExec& ex = S->ex;
  InstArr<T> ia{};
  ia.arr = (const Inst<T>*)((char*)S->ws + S->L.inst);
  const int B = S->B, nbe = S->P->nblk_elem;
  hipStream_t st = S->st;
  constexpr int NK = 9;
  hipEvent_t ev[NK + 1];
  for (auto& e : ev) HIPCHK(ctx, hipEventCreate(&e));
  std::vector<double> acc(NK, 0.0);
  for (int it = 0; it < n; ++it) {
    int k = 0;
    HIPCHK(ctx, hipEventRecord(ev[k++], st));
    // (phase A is fused into the displacement pass: nothing to launch)
    HIPCHK(ctx, hipEventRecord(ev[k++], st));
    if (S->any_bisect) launch_bisect<T>(S, ex, ia, B, st);
    HIPCHK(ctx, hipEventRecord(ev[k++], st));
    launch_r2c<T, ProdDir<T>>(ex, ia, B, 3, st);
    HIPCHK(ctx, hipEventRecord(ev[k++], st));
    launch_col<T, 3>(ex, ia, B, 0, 2, st);
    HIPCHK(ctx, hipEventRecord(ev[k++], st));
    // (tau comes from the spectrum in the column pass: no c2r pass)
    HIPCHK(ctx, hipEventRecord(ev[k++], st));
    if (S->any_hard)
      launch_elem<T, ElemPhaseB<T, true>>(ex, nbe, ia, B, 5, st);
    else
      launch_elem<T, ElemPhaseB<T, false>>(ex, nbe, ia, B, 5, st);
    // the displacement update as between refreshes: FP64 transforms of the
    // new forces (FP64 plans), FP32 transforms of the force change (FP32)
    HIPCHK(ctx, hipEventRecord(ev[k++], st));
    if constexpr (std::is_same<T, double>::value) {
      if (S->any_hard)
        launch_r2c<T, ProdForce<T, T, false, true>>(ex, ia, B, 6, st);
      else
        launch_r2c<T, ProdForce<T, T, false, false>>(ex, ia, B, 6, st);
      HIPCHK(ctx, hipEventRecord(ev[k++], st));
      launch_col<T, 0>(ex, ia, B, 0, 1, st);
      HIPCHK(ctx, hipEventRecord(ev[k++], st));
      launch_c2r<T, ConsDisp<T, T, false>>(ex, ia, B, 7, st);
    } else {
      if (S->any_hard)
        launch_r2c<T, ProdForce<T, T, true, true>>(ex, ia, B, 6, st);
      else
        launch_r2c<T, ProdForce<T, T, true, false>>(ex, ia, B, 6, st);
      HIPCHK(ctx, hipEventRecord(ev[k++], st));
      launch_col<T, 0>(ex, ia, B, 0, 1, st);
      HIPCHK(ctx, hipEventRecord(ev[k++], st));
      launch_c2r<T, ConsDisp<T, T, true>>(ex, ia, B, 7, st);
    }
    HIPCHK(ctx, hipEventRecord(ev[k++], st));
    HIPCHK(ctx, hipEventSynchronize(ev[NK]));
    for (int i = 0; i < NK; ++i) {
      float ms = 0;
      HIPCHK(ctx, hipEventElapsedTime(&ms, ev[i], ev[i + 1]));
      acc[i] += ms * 1000.0;
    }
  }
  S->issued += n;
  for (auto& e : ev) (void)hipEventDestroy(e);
  for (int i = 0; i < NK; ++i) us[i] = (float)(acc[i] / std::max(1, n));
  return 0;
}

int session_poll(cm_session* S, cm_ccg_stats* stats, int* all_done) {
  cm_ctx* ctx = S->P->ctx;
  const char* base = (const char*)S->ws;
  for (int b = 0; b < S->B; ++b)
    HIPCHK(ctx, hipMemcpyAsync(&S->host[b], base + S->L.sc + b * S->L.st_sc, sizeof(Scal),
                               hipMemcpyDeviceToHost, S->st));
  HIPCHK(ctx, hipStreamSynchronize(S->st));
  bool done = true;
  for (int b = 0; b < S->B; ++b) {
    if (stats) fill_stats(S->host[b], &stats[b]);
    done = done && S->host[b].done;
  }
  if (all_done) *all_done = done ? 1 : 0;
  return 0;
}

// The log ring of instance b (after a poll: the stream is idle).
int session_ring(cm_session* S, int b, std::vector<Scal>& out) {
  cm_ctx* ctx = S->P->ctx;
  out.resize(kLogRing);
  const char* base = (const char*)S->ws + S->L.sc + (size_t)b * S->L.st_sc + sizeof(Scal);
  HIPCHK(ctx, hipMemcpy(out.data(), base, sizeof(Scal) * kLogRing, hipMemcpyDeviceToHost));
  return 0;
}

void session_free(cm_session* S) {
  if (!S) return;
  cm_ctx* ctx = S->P->ctx;
  (void)hipStreamSynchronize(S->st);
  for (auto& g : S->graphs) (void)hipGraphExecDestroy(g.second);
  for (auto& g : S->graph_defs) (void)hipGraphDestroy(g);
  dev_free(ctx, S->ws, S->ws_bytes);
  if (S->host) (void)hipHostFree(S->host);
  if (S->hin) (void)hipHostFree(S->hin);
  if (S->hin0) (void)hipHostFree(S->hin0);
  delete S;
}

// Whole solve: set-up, chunks of iterations with a host check of the done
// flags after each chunk (chunks double up to 64), energy, statistics.
int solve_driver(cm_plan* P, int B, const cm_ccg_params* prm, const double* const* hs,
                 const uint8_t* const* surfs, void* const* us, void* const* fs, uint8_t* const* cms,
                 cm_ccg_stats* stats, cm_progress_fn progress, void* user, void* stream) {
  cm_session* S = nullptr;
  int rc = P->prec == CM_FP64 ? session_begin<double>(P, B, prm, hs, surfs, us, fs, cms, stream, &S)
                              : session_begin<float>(P, B, prm, hs, surfs, us, fs, cms, stream, &S);
  if (rc) return rc;
  const int ips = prm[0].iters_per_sync;
  int chunk = ips > 0 ? ips : 4;
  // adaptive chunks with a progress function: every iteration of a chunk is
  // reported from the log ring (the chunks never exceed the ring)
  const bool from_ring = progress && ips <= 0;
  std::vector<cm_ccg_stats> tmp(B);
  std::vector<Scal> ring;
  long long reported = 0;
  int result = 0;
  while (true) {
    rc = P->prec == CM_FP64 ? session_step<double>(S, chunk) : session_step<float>(S, chunk);
    if (rc) {
      result = rc;
      break;
    }
    int done = 0;
    if ((rc = session_poll(S, tmp.data(), &done))) {
      result = rc;
      break;
    }
    if (from_ring) {
      if ((rc = session_ring(S, 0, ring))) {
        result = rc;
        break;
      }
      for (long long it = reported + 1; it <= tmp[0].nit; ++it) {
        const Scal& r = ring[it % kLogRing];
        if (it == tmp[0].nit) {
          progress(user, &tmp[0]);  // the latest record (also an iteration ended outside the ring)
        } else if (r.it == it) {
          cm_ccg_stats st;
          fill_stats(r, &st);
          progress(user, &st);
        }
      }
      reported = tmp[0].nit;
    } else if (progress) {
      progress(user, &tmp[0]);
    }
    if (done) break;
    if (S->issued > (long long)S->maxiter + 8) {
      result = fail(P->ctx, CM_E_ARG, "solver did not terminate within maxiter");
      break;
    }
    if (ips <= 0) chunk = std::min(chunk * 2, kLogRing);
  }
  if (result == 0) {
    P->prec == CM_FP64 ? session_finish<double>(S) : session_finish<float>(S);
    if ((rc = session_poll(S, stats, nullptr))) result = rc;
  }
  if (result == 0) {
    result = stats[0].status;
    for (int b = 1; b < B; ++b)
      if (stats[b].status < 0) result = stats[b].status;
  }
  session_free(S);
  return result;
}

// Allocation-free solve of one instance: the workspace is this device's
// static scratch, the host-side records live on the stack, and the solve runs
// without graphs (instantiating one allocates).  Same kernels, same order as
// solve_driver, hence the same results.
template <typename T>
int solve_noalloc(cm_plan* P, const cm_ccg_params* prm, const double* h, const uint8_t* surf, void* u, void* f,
                  uint8_t* cm, cm_ccg_stats* stats, void* stream) {
  cm_ctx* ctx = P->ctx;
  if (!ctx->scratch) return fail(ctx, CM_E_HIP, "static scratch has no device address");
  if (uses_e32(P->ex) || (P->prec != CM_FP64 && uses_e32(P->ex64))) return fail(ctx, CM_E_UNSUPPORTED, kE32Refused);
  cm_ccg_params q = *prm;
  q.use_graph = 0;
  cm_session S;
  session_configure<T>(&S, P, 1, &q, stream);
  if (S.L.total > kScratchBytes)
    return fail(ctx, CM_E_OOM,
                "solve workspace of " + std::to_string(S.L.total) + " bytes exceeds the static scratch of " +
                    std::to_string(kScratchBytes) + " bytes of cm_ccg_solve_noalloc");
  Scal host{};
  Inst<T> hin{}, hin0{};
  S.ws = ctx->scratch;
  S.ws_bytes = 0;  // not owned
  S.host = &host;
  S.hin = &hin;
  S.hin0 = &hin0;
  const double* hs[1] = {h};
  const uint8_t* ms[1] = {surf};
  void* us[1] = {u};
  void* fs[1] = {f};
  uint8_t* cs[1] = {cm};
  int rc = session_init<T>(&S, &q, hs, ms, us, fs, cs);
  int result = rc;
  const int ips = q.iters_per_sync;
  int chunk = ips > 0 ? ips : 4;
  cm_ccg_stats st{};
  while (result == 0) {
    if ((rc = session_step<T>(&S, chunk))) {
      result = rc;
      break;
    }
    int done = 0;
    if ((rc = session_poll(&S, &st, &done))) {
      result = rc;
      break;
    }
    if (done) break;
    if (S.issued > (long long)S.maxiter + 8) {
      result = fail(ctx, CM_E_ARG, "solver did not terminate within maxiter");
      break;
    }
    if (ips <= 0) chunk = std::min(chunk * 2, kLogRing);
  }
  if (result == 0) {
    session_finish<T>(&S);
    if ((rc = session_poll(&S, stats, nullptr))) result = rc;
  }
  (void)hipStreamSynchronize(S.st);  // the stack records must outlive every copy
  if (result == 0) result = stats->status;
  return result;
}

}  // namespace

// ===========================================================================
// C ABI

extern "C" {

const char* cm_version(void) { return "cmhip 0.1 (gfx950)"; }

const char* cm_last_error(cm_ctx* ctx) {
  if (!ctx) return g_err_noctx.c_str();
  return ctx->err.c_str();
}

int cm_ctx_create(int device, size_t mem_budget, cm_ctx** out) {
  if (!out) return fail(nullptr, CM_E_ARG, "out is null");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(nullptr, CM_E_HIP, "no HIP device available");
  if (device < 0 || device >= n) return fail(nullptr, CM_E_ARG, "bad device index");
  if (hipSetDevice(device) != hipSuccess) return fail(nullptr, CM_E_HIP, "hipSetDevice failed");
  cm_ctx* c = new cm_ctx();
  c->device = device;
  c->budget = mem_budget;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return fail(nullptr, CM_E_HIP, "hipStreamCreate failed");
  }
  // this device's copy of the static scratch (the array's name is only a
  // host-side stand-in: kernels and copies take this address)
  hipError_t e = hipGetSymbolAddress(&c->scratch, HIP_SYMBOL(scratch));
  if (e != hipSuccess) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return fail(nullptr, CM_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  }
  *out = c;
  return 0;
}

int cm_ctx_destroy(cm_ctx* ctx) {
  if (!ctx) return 0;
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return 0;
}

size_t cm_ctx_mem_used(cm_ctx* ctx) { return ctx ? ctx->used : 0; }

int cm_ctx_stream(cm_ctx* ctx, void** stream) {
  if (!ctx || !stream) return fail(ctx, CM_E_ARG, "null argument");
  *stream = (void*)ctx->stream;
  return 0;
}

int cm_plan_create(cm_ctx* ctx, const cm_substrate_desc* d, cm_plan** out) {
  if (!ctx || !d || !out) return fail(ctx, CM_E_ARG, "null argument");
  if (d->nx < 1 || d->ny < 1) return fail(ctx, CM_E_SHAPE, "grid must be at least 1x1");
  if (!(d->sx > 0) || !(d->sy > 0)) return fail(ctx, CM_E_ARG, "physical sizes must be positive");
  if (d->bc != CM_PERIODIC && d->bc != CM_FREE && d->bc != CM_SEMI)
    return fail(ctx, CM_E_ARG, "bc must be CM_PERIODIC, CM_FREE or CM_SEMI");
  const int tile = std::max(1, d->tile_y);
  if (tile > 2 || (tile == 2 && (d->ny % 2 != 0 || d->bc == CM_FREE)))
    return fail(ctx, CM_E_ARG, "tile_y = 2 needs a periodic or semi-periodic plan and an even ny");
  if (d->precision != CM_FP64 && d->precision != CM_FP32) return fail(ctx, CM_E_ARG, "bad precision");
  (void)hipSetDevice(ctx->device);
  cm_plan* P = new cm_plan();
  P->ctx = ctx;
  P->d = *d;
  P->prec = d->precision;
  if (d->bc == CM_SEMI) {
    // Hockney padding to 2n - 1 along the non-periodic directions
    // (FFTElasticHalfSpace.py:1344-1352); with tile_y, ny counts the points
    // of all periods and My is tile_y computational periods
    P->Mx = d->periodic_x ? d->nx : 2 * d->nx - 1;
    const int ny1 = d->ny / tile;
    P->My = tile * (d->periodic_y ? ny1 : 2 * ny1 - 1);
  } else {
    P->Mx = d->bc == CM_FREE ? 2 * d->nx : d->nx;
    P->My = d->bc == CM_FREE ? 2 * d->ny : d->ny;
  }
  P->area_per_pt = d->sx * d->sy / ((double)d->nx * d->ny);
  if (P->My % 2 != 0) {
    delete P;
    return fail(ctx, CM_E_UNSUPPORTED,
                "the contiguous grid dimension My must be even (the row transforms pack pairs of points); "
                "an odd My is computed on the transposed grid or, along a periodic y, with tile_y = 2");
  }
  int rc = P->prec == CM_FP64 ? plan_build<double>(P) : plan_build<float>(P);
  if (rc) {
    plan_free(P);
    return rc;
  }
  *out = P;
  return 0;
}

int cm_plan_destroy(cm_plan* plan) {
  if (!plan) return 0;
  (void)hipStreamSynchronize(plan->ctx->stream);
  return plan_free(plan);
}

int cm_plan_grid(cm_plan* plan, int* Mx, int* My) {
  if (!plan) return fail(nullptr, CM_E_ARG, "null plan");
  if (Mx) *Mx = plan->Mx;
  if (My) *My = plan->My;
  return 0;
}

int cm_plan_green_host(cm_plan* plan, double* out) {
  if (!plan || !out) return fail(nullptr, CM_E_ARG, "null argument");
  std::memcpy(out, plan->G_host.data(), plan->G_host.size() * sizeof(double));
  return 0;
}

// Explicit launch shapes are only offered for FP64 plans: an FP32 2048^2
// solve with non-automatic shapes ended in a GPU fault whose cause is not
// established, so FP32 plans keep the automatic shape.
int refuse_fp32_shapes(cm_plan* plan) {
  return fail(plan->ctx, CM_E_UNSUPPORTED, "FP32 plans use the automatic launch shape");
}


int cm_plan_autotune(cm_plan* plan, int reps, char* out, size_t out_len) {
  if (!plan || reps < 1) return fail(plan ? plan->ctx : nullptr, CM_E_ARG, "bad argument");
  if (plan->prec != CM_FP64) return refuse_fp32_shapes(plan);
  std::string d;
  int rc = plan->prec == CM_FP64 ? autotune<double>(plan, reps, d) : autotune<float>(plan, reps, d);
  if (rc) return rc;
  if (out && out_len) {
    std::strncpy(out, d.c_str(), out_len - 1);
    out[out_len - 1] = 0;
  }
  return 0;
}

int cm_plan_set_shape(cm_plan* plan, int Cb, int R, int erow, int ecol) {
  if (!plan) return fail(nullptr, CM_E_ARG, "null plan");
  if (plan->prec != CM_FP64) return refuse_fp32_shapes(plan);
  if (erow == 32 || ecol == 32) return fail(plan->ctx, CM_E_UNSUPPORTED, kE32Refused);
  FftExec trial = plan->ex;
  const bool ok = plan->prec == CM_FP64 ? choose_shape<double>(trial, Shape{Cb, R, erow, ecol})
                                        : choose_shape<float>(trial, Shape{Cb, R, erow, ecol});
  if (!ok) return fail(plan->ctx, CM_E_ARG, "launch shape not valid for this grid");
  plan->ex = trial;
  return 0;
}

int cm_plan_shape(cm_plan* plan, int* Cb, int* R, int* erow, int* ecol) {
  if (!plan) return fail(nullptr, CM_E_ARG, "null plan");
  if (Cb) *Cb = plan->ex.g.Cb;
  if (R) *R = plan->ex.g.R;
  if (erow) *erow = plan_emax(plan->ex.row);
  if (ecol) *ecol = plan_emax(plan->ex.col);
  return 0;
}

int cm_evaluate_disp(cm_plan* plan, const void* f, void* u, void* stream) {
  if (!plan || !f || !u) return fail(plan ? plan->ctx : nullptr, CM_E_ARG, "null argument");
  return plan->prec == CM_FP64 ? evaluate<double>(plan, f, u, 0, stream) : evaluate<float>(plan, f, u, 0, stream);
}

int cm_evaluate_force(cm_plan* plan, const void* u, void* f, void* stream) {
  if (!plan || !f || !u) return fail(plan ? plan->ctx : nullptr, CM_E_ARG, "null argument");
  return plan->prec == CM_FP64 ? evaluate<double>(plan, u, f, 1, stream) : evaluate<float>(plan, u, f, 1, stream);
}

int cm_fourier(cm_plan* plan, const void* field, void* out_k, int kind, void* stream) {
  if (!plan || !field || !out_k) return fail(plan ? plan->ctx : nullptr, CM_E_ARG, "null argument");
  if (kind < 0 || kind > 2) return fail(plan->ctx, CM_E_ARG, "kind must be 0, 1 or 2");
  return plan->prec == CM_FP64 ? fourier_out<double>(plan, field, out_k, kind, stream)
                               : fourier_out<float>(plan, field, out_k, kind, stream);
}

int cm_k_force_k(cm_plan* plan, const void* disp_k, void* out_k, void* stream) {
  if (!plan || !disp_k || !out_k) return fail(plan ? plan->ctx : nullptr, CM_E_ARG, "null argument");
  return plan->prec == CM_FP64 ? kforce_k<double>(plan, disp_k, out_k, stream)
                               : kforce_k<float>(plan, disp_k, out_k, stream);
}

int cm_scalar_product_k(cm_plan* plan, const void* ka, const void* kb, double* out, void* stream) {
  if (!plan || !ka || !kb || !out) return fail(plan ? plan->ctx : nullptr, CM_E_ARG, "null argument");
  return plan->prec == CM_FP64 ? scalar_product_k<double>(plan, ka, kb, out, stream)
                               : scalar_product_k<float>(plan, ka, kb, out, stream);
}

int cm_dot(cm_plan* plan, const void* a, const void* b, double* out, void* stream) {
  if (!plan || !a || !b || !out) return fail(plan ? plan->ctx : nullptr, CM_E_ARG, "null argument");
  return plan->prec == CM_FP64 ? dot<double>(plan, a, b, out, stream) : dot<float>(plan, a, b, out, stream);
}

int cm_ccg_solve(cm_plan* plan, const cm_ccg_params* params, const double* heights, const uint8_t* surf_mask,
                 void* u, void* f, uint8_t* cmask, cm_ccg_stats* stats, cm_progress_fn progress, void* user,
                 void* stream) {
  if (!plan || !params || !heights || !u || !f || !cmask || !stats)
    return fail(plan ? plan->ctx : nullptr, CM_E_ARG, "null argument");
  const double* hs[1] = {heights};
  const uint8_t* ms[1] = {surf_mask};
  void* us[1] = {u};
  void* fs[1] = {f};
  uint8_t* cs[1] = {cmask};
  return solve_driver(plan, 1, params, hs, ms, us, fs, cs, stats, progress, user, stream);
}

int cm_ccg_solve_noalloc(cm_plan* plan, const cm_ccg_params* params, const double* heights, const uint8_t* surf_mask,
                         void* u, void* f, uint8_t* cmask, cm_ccg_stats* stats, void* stream) {
  if (!plan || !params || !heights || !u || !f || !cmask || !stats)
    return fail(plan ? plan->ctx : nullptr, CM_E_ARG, "null argument");
  return plan->prec == CM_FP64 ? solve_noalloc<double>(plan, params, heights, surf_mask, u, f, cmask, stats, stream)
                               : solve_noalloc<float>(plan, params, heights, surf_mask, u, f, cmask, stats, stream);
}

int cm_ccg_solve_batch(cm_plan* plan, int B, const cm_ccg_params* params, const double* const* heights,
                       const uint8_t* const* surf_masks, void* const* u, void* const* f, uint8_t* const* cmask,
                       cm_ccg_stats* stats, void* stream) {
  if (!plan || B < 1 || !params || !heights || !u || !f || !cmask || !stats)
    return fail(plan ? plan->ctx : nullptr, CM_E_ARG, "bad argument");
  return solve_driver(plan, B, params, heights, surf_masks, u, f, cmask, stats, nullptr, nullptr, stream);
}

size_t cm_ccg_session_bytes(cm_plan* plan, int B) {
  if (!plan || B < 1) return 0;
  return ws_layout(plan, B).total;
}

int cm_ccg_begin(cm_plan* plan, int B, const cm_ccg_params* params, const double* const* heights,
                 const uint8_t* const* surf_masks, void* const* u, void* const* f, uint8_t* const* cmask,
                 void* stream, cm_session** out) {
  if (!plan || B < 1 || !params || !heights || !u || !f || !cmask || !out)
    return fail(plan ? plan->ctx : nullptr, CM_E_ARG, "bad argument");
  return plan->prec == CM_FP64
             ? session_begin<double>(plan, B, params, heights, surf_masks, u, f, cmask, stream, out)
             : session_begin<float>(plan, B, params, heights, surf_masks, u, f, cmask, stream, out);
}

int cm_ccg_step(cm_session* s, int n_iters) {
  if (!s) return fail(nullptr, CM_E_ARG, "null session");
  if (s->finished) return fail(s->P->ctx, CM_E_ARG, "session already finished");
  return s->P->prec == CM_FP64 ? session_step<double>(s, n_iters) : session_step<float>(s, n_iters);
}

int cm_ccg_poll(cm_session* s, cm_ccg_stats* stats, int* all_done) {
  if (!s) return fail(nullptr, CM_E_ARG, "null session");
  return session_poll(s, stats, all_done);
}

int cm_ccg_finish(cm_session* s, cm_ccg_stats* stats) {
  if (!s || !stats) return fail(nullptr, CM_E_ARG, "null argument");
  // a second call only reports: the displacements are restored once
  if (!s->finished) s->P->prec == CM_FP64 ? session_finish<double>(s) : session_finish<float>(s);
  return session_poll(s, stats, nullptr);
}

int cm_ccg_end(cm_session* s) {
  session_free(s);
  return 0;
}

int cm_ccg_profile(cm_session* s, int n_iters, float* kernel_us) {
  if (!s || !kernel_us || n_iters < 1) return fail(nullptr, CM_E_ARG, "bad argument");
  if (s->finished) return fail(s->P->ctx, CM_E_ARG, "session already finished");
  return s->P->prec == CM_FP64 ? session_profile<double>(s, n_iters, kernel_us)
                               : session_profile<float>(s, n_iters, kernel_us);
}

}  // extern "C"
