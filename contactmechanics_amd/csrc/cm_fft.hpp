// Building blocks for the gfx950 contact solver: complex arithmetic, radix-R
// DFT butterflies, an in-LDS Stockham autosort FFT engine and wavefront/block
// reductions in FP64.
//
// The FFT engine transforms one "line" (a row or a column of the computational
// grid) held in LDS by `tpl` cooperating threads.  A radix-R pass has L/R
// butterflies dealt round-robin to the threads.  A pass loads all of a
// thread's butterfly inputs into registers, synchronises, computes and writes
// the outputs back in place (Stockham index map), then synchronises again, so
// one LDS buffer of L complex values per line suffices.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cm {

template <typename T> struct cplx { T x, y; };

template <typename T> __device__ __forceinline__ cplx<T> mk(T a, T b) { cplx<T> r; r.x = a; r.y = b; return r; }
template <typename T> __device__ __forceinline__ cplx<T> operator+(cplx<T> a, cplx<T> b) { return mk(a.x + b.x, a.y + b.y); }
template <typename T> __device__ __forceinline__ cplx<T> operator-(cplx<T> a, cplx<T> b) { return mk(a.x - b.x, a.y - b.y); }
template <typename T> __device__ __forceinline__ cplx<T> operator*(cplx<T> a, T s) { return mk(a.x * s, a.y * s); }
template <typename T> __device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
template <typename T> __device__ __forceinline__ cplx<T> cmulc(cplx<T> a, cplx<T> b) {
  return mk(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
template <typename T> __device__ __forceinline__ cplx<T> conj(cplx<T> a) { return mk(a.x, -a.y); }
// multiply by -i (forward) or +i (inverse)
template <typename T, bool INV> __device__ __forceinline__ cplx<T> rot90(cplx<T> a) {
  return INV ? mk(-a.y, a.x) : mk(a.y, -a.x);
}

// ---------------------------------------------------------------------------
// Radix-R DFT of v[0..R-1] in registers.  Forward uses exp(-2 pi i/R),
// inverse exp(+2 pi i/R).  Constants are the exact trigonometric values.
template <typename T, int R, bool INV> struct Dft;

// Radix 1: the identity.  Lines whose length has a prime factor above 5 start
// and end with an identity pass, so the register-ended first and last passes
// of the kernels stay compile-time radices (see make_line_plan).
template <typename T, bool INV> struct Dft<T, 1, INV> {
  __device__ __forceinline__ static void run(cplx<T>*) {}
};

template <typename T, bool INV> struct Dft<T, 2, INV> {
  __device__ __forceinline__ static void run(cplx<T>* v) {
    cplx<T> a = v[0], b = v[1];
    v[0] = a + b; v[1] = a - b;
  }
};

template <typename T, bool INV> struct Dft<T, 3, INV> {
  __device__ __forceinline__ static void run(cplx<T>* v) {
    const T c1 = T(-0.5);
    const T s1 = T(0.86602540378443864676372317075294);  // sin(2pi/3)
    cplx<T> a = v[0], b = v[1], c = v[2];
    cplx<T> s = b + c, d = b - c;
    cplx<T> t = a + s * c1;
    cplx<T> e = rot90<T, INV>(d) * s1;  // -i*s1*(b-c) forward
    v[0] = a + s;
    v[1] = t + e;
    v[2] = t - e;
  }
};

template <typename T, bool INV> struct Dft<T, 4, INV> {
  __device__ __forceinline__ static void run(cplx<T>* v) {
    cplx<T> a = v[0], b = v[1], c = v[2], d = v[3];
    cplx<T> s0 = a + c, s1 = a - c, s2 = b + d, s3 = rot90<T, INV>(b - d);
    v[0] = s0 + s2; v[2] = s0 - s2;
    v[1] = s1 + s3; v[3] = s1 - s3;
  }
};

template <typename T, bool INV> struct Dft<T, 5, INV> {
  __device__ __forceinline__ static void run(cplx<T>* v) {
    const T c1 = T(0.30901699437494742410229341718282);   // cos(2pi/5)
    const T c2 = T(-0.80901699437494742410229341718282);  // cos(4pi/5)
    const T s1 = T(0.95105651629515357211643933337938);   // sin(2pi/5)
    const T s2 = T(0.58778525229247312916870595463907);   // sin(4pi/5)
    cplx<T> a = v[0];
    cplx<T> p1 = v[1] + v[4], m1 = v[1] - v[4];
    cplx<T> p2 = v[2] + v[3], m2 = v[2] - v[3];
    cplx<T> t1 = a + p1 * c1 + p2 * c2;
    cplx<T> t2 = a + p1 * c2 + p2 * c1;
    // forward: X1 = t1 - i(s1 m1 + s2 m2), X4 = t1 + i(...)
    //          X2 = t2 - i(s2 m1 - s1 m2), X3 = t2 + i(...)
    cplx<T> e1 = rot90<T, INV>(m1 * s1 + m2 * s2);
    cplx<T> e2 = rot90<T, INV>(m1 * s2 - m2 * s1);
    v[0] = a + p1 + p2;
    v[1] = t1 + e1; v[4] = t1 - e1;
    v[2] = t2 + e2; v[3] = t2 - e2;
  }
};

template <typename T, bool INV> struct Dft<T, 8, INV> {
  __device__ __forceinline__ static void run(cplx<T>* v) {
    const T h = T(0.70710678118654752440084436210485);
    // radix-2 across stride 4, then twiddle, then two radix-4
    cplx<T> a0 = v[0] + v[4], a4 = v[0] - v[4];
    cplx<T> a1 = v[1] + v[5], a5 = v[1] - v[5];
    cplx<T> a2 = v[2] + v[6], a6 = v[2] - v[6];
    cplx<T> a3 = v[3] + v[7], a7 = v[3] - v[7];
    // twiddles w8^k for the odd half: k=1: (h, -h) fwd, k=2: -i, k=3: (-h,-h) fwd
    a5 = INV ? mk(h * (a5.x - a5.y), h * (a5.x + a5.y)) : mk(h * (a5.x + a5.y), h * (a5.y - a5.x));
    a6 = rot90<T, INV>(a6);
    a7 = INV ? mk(-h * (a7.x + a7.y), h * (a7.x - a7.y)) : mk(h * (a7.y - a7.x), -h * (a7.x + a7.y));
    cplx<T> e[4] = {a0, a1, a2, a3};
    cplx<T> o[4] = {a4, a5, a6, a7};
    Dft<T, 4, INV>::run(e);
    Dft<T, 4, INV>::run(o);
    v[0] = e[0]; v[2] = e[1]; v[4] = e[2]; v[6] = e[3];
    v[1] = o[0]; v[3] = o[1]; v[5] = o[2]; v[7] = o[3];
  }
};

// Radix 16 as 4 x 4: n = 4 n1 + n2, k = k1 + 4 k2; twiddles W16^(n2 k1).
template <typename T, bool INV>
__device__ __forceinline__ cplx<T> w16(int m) {
  const T c = T(0.92387953251128675612818318939679);  // cos(pi/8)
  const T s = T(0.38268343236508977172845998403040);  // sin(pi/8)
  const T h = T(0.70710678118654752440084436210485);
  cplx<T> w;
  switch (m) {
    case 1: w = mk(c, -s); break;
    case 2: w = mk(h, -h); break;
    case 3: w = mk(s, -c); break;
    case 4: w = mk(T(0), T(-1)); break;
    case 6: w = mk(-h, -h); break;
    default: w = mk(-c, s); break;  // m = 9
  }
  return INV ? conj(w) : w;
}

template <typename T, bool INV> struct Dft<T, 16, INV> {
  __device__ __forceinline__ static void run(cplx<T>* v) {
    cplx<T> y[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
      cplx<T> a[4] = {v[n2], v[n2 + 4], v[n2 + 8], v[n2 + 12]};
      Dft<T, 4, INV>::run(a);
#pragma unroll
      for (int k1 = 0; k1 < 4; ++k1) y[n2][k1] = a[k1];
    }
#pragma unroll
    for (int n2 = 1; n2 < 4; ++n2)
#pragma unroll
      for (int k1 = 1; k1 < 4; ++k1) y[n2][k1] = cmul(y[n2][k1], w16<T, INV>(n2 * k1));
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      cplx<T> b[4] = {y[0][k1], y[1][k1], y[2][k1], y[3][k1]};
      Dft<T, 4, INV>::run(b);
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = b[k2];
    }
  }
};

// LDS keeps complex values at padded positions: one slot is skipped every
// 128 bytes, so the stride-R Stockham writes of the early passes spread over
// all banks instead of piling onto a few.
template <typename T> constexpr int kPadShift = sizeof(cplx<T>) == 16 ? 3 : 4;
template <typename T> __device__ __forceinline__ int lpad(int i) { return i + (i >> kPadShift<T>); }
template <typename T> inline size_t lds_bytes(size_t n) {
  return (n + (n >> kPadShift<T>)) * sizeof(cplx<T>);
}

// ---------------------------------------------------------------------------
// FFT plan for one line length, passed by value to kernels.
struct LinePlan {
  int L;          // transform length
  int tpl;        // threads per line
  int E;          // most elements a thread holds in any pass
  int npass;      // number of radix passes
  int rlast;      // radix of the last pass
  // radix of pass s in 16 bits, four passes per word (at most kMaxPasses):
  // shifts and selects instead of an indexed array, which a kernel would
  // have to copy to private memory to index at run time.  Radix 1 is the
  // identity; 2, 3, 4, 5, 8 and 16 have unrolled butterflies; any other
  // value is a generic (prime) radix.
  uint64_t rc0, rc1, rc2;
  static constexpr int kMaxPasses = 12;
  __host__ __device__ __forceinline__ int radix(int s) const {
    const uint64_t w = s < 4 ? rc0 : (s < 8 ? rc1 : rc2);
    return (int)((w >> (16 * (s & 3))) & 0xffffu);
  }
  __host__ void set_radix(int s, int R) {
    uint64_t& w = s < 4 ? rc0 : (s < 8 ? rc1 : rc2);
    const int sh = 16 * (s & 3);
    w = (w & ~((uint64_t)0xffffu << sh)) | ((uint64_t)R << sh);
  }
};

// Radices with unrolled butterflies (and the identity).
__host__ __device__ __forceinline__ bool unrolled_radix(int R) {
  return R == 1 || R == 2 || R == 3 || R == 4 || R == 5 || R == 8 || R == 16;
}

// Twiddle table: tw[k] = exp(-2 pi i k / L), k in [0, L).
// One Stockham pass (radix R, sub-transform length Ns before the pass).
// Line element n lives at lds[base + n*stride].
template <typename T, int R, bool INV, int EMAX, bool EARLY = true>
__device__ __forceinline__ void stockham_pass(cplx<T>* lds, int base, int stride, int L, int Ns,
                                              int tl, int tpl, int nb, const cplx<T>* __restrict__ tw) {
  constexpr int NBMAX = EMAX / R > 0 ? EMAX / R : 1;
  cplx<T> v[NBMAX][R];
  const int LR = L / R;
  const int twstep = L / (Ns * R);
  // j = q * Ns + k; Ns is uniform, and a power of two for 2^n lengths
  const bool pow2 = (Ns & (Ns - 1)) == 0;
  const int nsh = __builtin_ctz(Ns);
  auto split = [&](int j, int& q, int& k) {
    if (pow2) {
      q = j >> nsh;
      k = j & (Ns - 1);
    } else {
      q = j / Ns;
      k = j - q * Ns;
    }
  };
  if constexpr (EARLY) {
    // twiddles first: they are global (cached) loads independent of the LDS
    // contents, so their latency overlaps the LDS reads and the barrier (at
    // the price of holding them in registers across the barrier)
    int qq[NBMAX], kk[NBMAX];
    cplx<T> w[NBMAX][R - 1];
#pragma unroll
    for (int b = 0; b < NBMAX; ++b) {
      if (b < nb) {
        split(tl + b * tpl, qq[b], kk[b]);
        if (Ns > 1) {
          const int kt = kk[b] * twstep;
#pragma unroll
          for (int r = 1; r < R; ++r) w[b][r - 1] = tw[r * kt];
        }
      }
    }
#pragma unroll
    for (int b = 0; b < NBMAX; ++b) {
      if (b < nb) {
        const int j = tl + b * tpl;
#pragma unroll
        for (int r = 0; r < R; ++r) v[b][r] = lds[lpad<T>(base + (j + r * LR) * stride)];
      }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < NBMAX; ++b) {
      if (b < nb) {
        if (Ns > 1) {
#pragma unroll
          for (int r = 1; r < R; ++r) v[b][r] = INV ? cmulc(v[b][r], w[b][r - 1]) : cmul(v[b][r], w[b][r - 1]);
        }
        Dft<T, R, INV>::run(v[b]);
        const int o = qq[b] * Ns * R + kk[b];
#pragma unroll
        for (int r = 0; r < R; ++r) lds[lpad<T>(base + (o + r * Ns) * stride)] = v[b][r];
      }
    }
  } else {
#pragma unroll
    for (int b = 0; b < NBMAX; ++b) {
      if (b < nb) {
        const int j = tl + b * tpl;
#pragma unroll
        for (int r = 0; r < R; ++r) v[b][r] = lds[lpad<T>(base + (j + r * LR) * stride)];
      }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < NBMAX; ++b) {
      if (b < nb) {
        int q, k;
        split(tl + b * tpl, q, k);
        if (Ns > 1) {
          const int kt = k * twstep;
#pragma unroll
          for (int r = 1; r < R; ++r) {
            const cplx<T> t = tw[r * kt];
            v[b][r] = INV ? cmulc(v[b][r], t) : cmul(v[b][r], t);
          }
        }
        Dft<T, R, INV>::run(v[b]);
        const int o = q * Ns * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) lds[lpad<T>(base + (o + r * Ns) * stride)] = v[b][r];
      }
    }
  }
  __syncthreads();
}

// Butterflies of a pass (L/R of them, dealt round-robin to tpl threads) that
// thread tl of the line owns.
__device__ __forceinline__ int nbutterflies(int L, int R, int tl, int tpl) {
  const int nbf = L / R;
  const int full = nbf / tpl;  // uniform: no per-thread division
  return full + (tl < nbf - full * tpl ? 1 : 0);
}

// Generic radix-R pass (R any value, in practice a prime above 5), computed
// per output element instead of per butterfly: with o = q*Ns*R + r*Ns + k
// (k < Ns, r < R) and j = q*Ns + k, the Stockham output is
//   X[o] = sum_m x[j + m*L/R] * w^(m*(k + r*Ns)*L/(Ns*R)),  w = exp(-2 pi i/L)
// (the pass twiddle w^(m*k*L/(Ns*R)) and the radix-R DFT factor
// exp(-2 pi i r m/R) in one table entry; the inverse conjugates it).  Thread
// tl < tpl owns elements tl + i*tpl, i < EMAX (the plan makes ceil(L/tpl) <=
// EMAX); each costs R complex multiply-adds and R LDS reads, which is slow
// for large primes but exact to the twiddle table's rounding.
template <typename T, bool INV, int EMAX>
__device__ void generic_pass(cplx<T>* lds, int base, int stride, int L, int Ns, int R, int tl, int tpl,
                             const cplx<T>* __restrict__ tw) {
  cplx<T> out[EMAX];
  const int twstep = L / (Ns * R), LR = L / R;
  const bool act = tl < tpl;
#pragma unroll
  for (int i = 0; i < EMAX; ++i) {
    const int o = tl + i * tpl;
    if (act && o < L) {
      const int k = o % Ns, r = (o / Ns) % R, q = o / (Ns * R);
      const int j = q * Ns + k;
      const int a = (k + r * Ns) * twstep;  // < L
      T sx = T(0), sy = T(0);
      int idx = 0;
      for (int m = 0; m < R; ++m) {
        const cplx<T> x = lds[lpad<T>(base + (j + m * LR) * stride)];
        const cplx<T> w = tw[idx];
        const cplx<T> p = INV ? cmulc(x, w) : cmul(x, w);
        sx += p.x;
        sy += p.y;
        idx += a;
        if (idx >= L) idx -= L;
      }
      out[i] = mk(sx, sy);
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < EMAX; ++i) {
    const int o = tl + i * tpl;
    if (act && o < L) lds[lpad<T>(base + o * stride)] = out[i];
  }
  __syncthreads();
}

// One Stockham pass of radix R (unrolled radices R <= EMAX; radix 16 only
// exists in kernels built for 16 or more register elements, so smaller
// builds stay lean).  Radix 1 (identity) leaves LDS as it is.  Generic
// radices exist only in the builds with GEN (lines with a prime factor above
// 5 launch those), so the other builds keep their register budget.
template <typename T, bool INV, int EMAX, bool EARLY = true, bool GEN = false>
__device__ __forceinline__ void pass_dispatch(int R, cplx<T>* lds, int base, int stride, int L, int Ns, int tl,
                                              int tpl, int nb, const cplx<T>* __restrict__ tw) {
  switch (R) {
    case 1: break;
    case 2: stockham_pass<T, 2, INV, EMAX, EARLY>(lds, base, stride, L, Ns, tl, tpl, nb, tw); break;
    case 3: stockham_pass<T, 3, INV, EMAX, EARLY>(lds, base, stride, L, Ns, tl, tpl, nb, tw); break;
    case 4: stockham_pass<T, 4, INV, EMAX, EARLY>(lds, base, stride, L, Ns, tl, tpl, nb, tw); break;
    case 5: stockham_pass<T, 5, INV, EMAX, EARLY>(lds, base, stride, L, Ns, tl, tpl, nb, tw); break;
    case 8: stockham_pass<T, 8, INV, EMAX, EARLY>(lds, base, stride, L, Ns, tl, tpl, nb, tw); break;
    case 16:
      if constexpr (EMAX >= 16) stockham_pass<T, 16, INV, EMAX, EARLY>(lds, base, stride, L, Ns, tl, tpl, nb, tw);
      break;
    default:
      if constexpr (GEN) generic_pass<T, INV, EMAX>(lds, base, stride, L, Ns, R, tl, tpl, tw);
      break;
  }
}

// Full in-place LDS FFT of one line (all threads of the block must call it;
// threads whose tl >= tpl must pass active=false and still hit the barriers).
template <typename T, bool INV, int EMAX, bool EARLY = true, bool GEN = false>
__device__ void fft_line(cplx<T>* lds, int base, int stride, const LinePlan& p, int tl, bool active,
                         const cplx<T>* __restrict__ tw) {
  int Ns = 1;
  for (int s = 0; s < p.npass; ++s) {
    const int R = p.radix(s);
    const int nb = active ? nbutterflies(p.L, R, tl, p.tpl) : 0;
    pass_dispatch<T, INV, EMAX, EARLY, GEN>(R, lds, base, stride, p.L, Ns, tl, p.tpl, nb, tw);
    Ns *= R;
  }
}

// ---------------------------------------------------------------------------
// Register-ended passes.  "First-pass order" for radix R: a thread's register
// slot e = b*R + r holds element tl + b*tpl + r*(L/R) (b < nb butterflies);
// "last-pass order" is the same map with the last pass's radix.  Feeding the
// first pass from registers and leaving the last pass in registers lets
// kernels read inputs from / write outputs to HBM without an LDS round trip.


// Element index held by register slot e for a pass of radix R.
__device__ __forceinline__ int slot_element(int e, int R, int L, int tl, int tpl) {
  return tl + (e / R) * tpl + (e % R) * (L / R);
}

// slot_element for a run-time radix R of a first or last pass (1, 2, 3, 4, 5,
// 8 or 16) and compile-time slots e < 32: e / R as a multiply and shift with
// m = ceil(2^16 / R), exact while e < 2^16 / R, so the offsets are a few
// scalar operations on uniform values instead of a division per slot.
struct SlotMap {
  int R, m, LR, tpl;
  __device__ __forceinline__ SlotMap(int R_, int L, int tpl_)
      : R(R_), m((65536 + R_ - 1) / R_), LR(L / R_), tpl(tpl_) {}
  // line element of slot e, less the thread's lane
  __device__ __forceinline__ int off(int e) const {
    const int q = (e * m) >> 16;
    return q * tpl + (e - q * R) * LR;
  }
};

template <typename T, int R, bool INV, int EMAX>
__device__ __forceinline__ void first_pass_regs(const cplx<T>* v, cplx<T>* lds, int base, int stride, int L,
                                                int tl, int tpl) {
  constexpr int NBMAX = EMAX / R > 0 ? EMAX / R : 1;
  const int nb = nbutterflies(L, R, tl, tpl);
#pragma unroll
  for (int b = 0; b < NBMAX; ++b) {
    if (b < nb) {
      cplx<T> w[R];
#pragma unroll
      for (int r = 0; r < R; ++r) w[r] = v[b * R + r];
      Dft<T, R, INV>::run(w);
      const int o = (tl + b * tpl) * R;  // Ns = 1
#pragma unroll
      for (int r = 0; r < R; ++r) lds[lpad<T>(base + (o + r) * stride)] = w[r];
    }
  }
}

template <typename T, int R, bool INV, int EMAX>
__device__ __forceinline__ void last_pass_regs(cplx<T>* v, const cplx<T>* lds, int base, int stride, int L,
                                               int tl, int tpl, const cplx<T>* __restrict__ tw) {
  constexpr int NBMAX = EMAX / R > 0 ? EMAX / R : 1;
  const int nb = nbutterflies(L, R, tl, tpl);
  const int LR = L / R;  // = Ns of the last pass
#pragma unroll
  for (int b = 0; b < NBMAX; ++b) {
    if (b < nb) {
      const int j = tl + b * tpl;
      cplx<T> w[R];
#pragma unroll
      for (int r = 0; r < R; ++r) w[r] = lds[lpad<T>(base + (j + r * LR) * stride)];
      if (LR > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          cplx<T> t = tw[r * j];
          w[r] = INV ? cmulc(w[r], t) : cmul(w[r], t);
        }
      }
      Dft<T, R, INV>::run(w);
#pragma unroll
      for (int r = 0; r < R; ++r) v[b * R + r] = w[r];
    }
  }
}

// Line transform with the input in registers (IN_REGS, first-pass order) or
// in LDS (natural order), and the output left in registers (OUT_REGS,
// last-pass order) or in LDS (natural order).  With OUT_REGS the caller must
// synchronise before writing LDS again.
template <typename T, bool INV, int EMAX, bool IN_REGS, bool OUT_REGS, bool EARLY = true, bool GEN = false>
__device__ void fft_line_x(cplx<T>* v, cplx<T>* lds, int base, int stride, const LinePlan& p, int tl,
                           const cplx<T>* __restrict__ tw) {
  int s0 = 0, s1 = p.npass, Ns = 1;
  if (IN_REGS) {
    switch (p.radix(0)) {
      case 1:
        if constexpr (GEN) first_pass_regs<T, 1, INV, EMAX>(v, lds, base, stride, p.L, tl, p.tpl);
        break;
      case 2: first_pass_regs<T, 2, INV, EMAX>(v, lds, base, stride, p.L, tl, p.tpl); break;
      case 3: first_pass_regs<T, 3, INV, EMAX>(v, lds, base, stride, p.L, tl, p.tpl); break;
      case 4: first_pass_regs<T, 4, INV, EMAX>(v, lds, base, stride, p.L, tl, p.tpl); break;
      case 5: first_pass_regs<T, 5, INV, EMAX>(v, lds, base, stride, p.L, tl, p.tpl); break;
      case 16:
        if constexpr (EMAX >= 16) first_pass_regs<T, 16, INV, EMAX>(v, lds, base, stride, p.L, tl, p.tpl);
        break;
      default: first_pass_regs<T, 8, INV, EMAX>(v, lds, base, stride, p.L, tl, p.tpl); break;
    }
    __syncthreads();
    Ns = p.radix(0);
    s0 = 1;
  }
  if (OUT_REGS) s1 = p.npass - 1;
  for (int s = s0; s < s1; ++s) {
    const int R = p.radix(s);
    const int nb = nbutterflies(p.L, R, tl, p.tpl);
    pass_dispatch<T, INV, EMAX, EARLY, GEN>(R, lds, base, stride, p.L, Ns, tl, p.tpl, nb, tw);
    Ns *= R;
  }
  if (OUT_REGS) {
    if (IN_REGS && p.npass == 1) {
      // the single pass already wrote natural order to LDS
      const int R = p.radix(0);
      const int nb = nbutterflies(p.L, R, tl, p.tpl);
#pragma unroll
      for (int e = 0; e < EMAX; ++e)
        if (e / R < nb) v[e] = lds[lpad<T>(base + slot_element(e, R, p.L, tl, p.tpl) * stride)];
      return;
    }
    switch (p.rlast) {
      case 1:
        if constexpr (GEN) last_pass_regs<T, 1, INV, EMAX>(v, lds, base, stride, p.L, tl, p.tpl, tw);
        break;
      case 2: last_pass_regs<T, 2, INV, EMAX>(v, lds, base, stride, p.L, tl, p.tpl, tw); break;
      case 3: last_pass_regs<T, 3, INV, EMAX>(v, lds, base, stride, p.L, tl, p.tpl, tw); break;
      case 4: last_pass_regs<T, 4, INV, EMAX>(v, lds, base, stride, p.L, tl, p.tpl, tw); break;
      case 5: last_pass_regs<T, 5, INV, EMAX>(v, lds, base, stride, p.L, tl, p.tpl, tw); break;
      case 16:
        if constexpr (EMAX >= 16) last_pass_regs<T, 16, INV, EMAX>(v, lds, base, stride, p.L, tl, p.tpl, tw);
        break;
      default: last_pass_regs<T, 8, INV, EMAX>(v, lds, base, stride, p.L, tl, p.tpl, tw); break;
    }
  }
}

// All passes but the last, from LDS (natural order) to LDS; the caller runs
// the last pass itself (last_pass_regs with its radix as a template argument,
// so the register slots it fills are compile-time indices).
template <typename T, bool INV, int EMAX, bool EARLY = true, bool GEN = false>
__device__ __forceinline__ void fft_line_to_last(cplx<T>* lds, int base, int stride, const LinePlan& p, int tl,
                                 const cplx<T>* __restrict__ tw) {
  int Ns = 1;
  for (int s = 0; s + 1 < p.npass; ++s) {
    const int R = p.radix(s);
    pass_dispatch<T, INV, EMAX, EARLY, GEN>(R, lds, base, stride, p.L, Ns, tl, p.tpl, nbutterflies(p.L, R, tl, p.tpl), tw);
    Ns *= R;
  }
}

// ---------------------------------------------------------------------------
// Reductions in FP64.  A reduction record holds NS slots; OPS[i] is the op.
// RCSUM is a compensated sum: each thread accumulates with Neumaier's
// correction term (kadd) and the tree combines (sum, correction) pairs with
// Knuth's TwoSum (kcombine), so the result is the exactly rounded sum up to
// O(n eps^2) of the sum of magnitudes.  RSUM is a plain sum, used for counts
// (exact in FP64).
enum RedOp : int { RSUM = 0, RMAX = 1, RMIN = 2, RCSUM = 3 };

__device__ __forceinline__ double red_identity(int op) {
  return (op == RSUM || op == RCSUM) ? 0.0 : (op == RMAX ? -__builtin_huge_val() : __builtin_huge_val());
}
__device__ __forceinline__ double red_apply(int op, double a, double b) {
  // NaN-propagating max/min so that a NaN anywhere reaches the scalar logic
  if (op == RSUM || op == RCSUM) return a + b;
  if (a != a || b != b) return a + b;
  return op == RMAX ? (a > b ? a : b) : (a < b ? a : b);
}

// Running max (min) over a thread's points with red_apply's NaN propagation
// at a third of its instructions: the hardware max (which returns the other
// operand when one is NaN) plus the last NaN seen, kept in nan (a max or min
// slot's otherwise unused correction term) and folded into the value once per
// thread (fold_nan) before the cross-thread reduction.
__device__ __forceinline__ void kmax(double& a, double& nan, double x) {
  a = __builtin_fmax(a, x);
  nan = (x != x) ? x : nan;
}
__device__ __forceinline__ void kmin(double& a, double& nan, double x) {
  a = __builtin_fmin(a, x);
  nan = (x != x) ? x : nan;
}

// s + x with the rounding error of the addition collected in c (Neumaier).
__device__ __forceinline__ void kadd(double& s, double& c, double x) {
  const double t = s + x;
  c += (fabs(s) >= fabs(x)) ? (s - t) + x : (x - t) + s;
  s = t;
}
// (s1, c1) += (s2, c2); the rounding error of s1 + s2 is exact (TwoSum), and
// the result does not depend on the order of the two operands.
__device__ __forceinline__ void kcombine(double& s1, double& c1, double s2, double c2) {
  const double t = s1 + s2;
  const double bp = t - s1;
  const double e = (s1 - (t - bp)) + (s2 - bp);
  s1 = t;
  c1 = (c1 + c2) + e;
}
// Value of a compensated pair; an infinite or NaN sum keeps its value.
__device__ __forceinline__ double kvalue(double s, double c) { return (s - s == 0.0) ? s + c : s; }

// Value of the partner lane of step OFF of an ascending butterfly (OFF = 1,
// 2, 4, 8, 16, 32) over a full wavefront, in the vector ALU instead of the
// LDS crossbar: DPP quad permutes for 1 and 2 (lane ^ OFF), DPP row
// half-mirror and mirror for 4 and 8 (lane 7 - i, 15 - i of its group: a lane
// of the partner group, which holds one value once the smaller steps are
// done), permlane16/32 swaps for 16 and 32 (lane ^ OFF).
template <int OFF>
__device__ __forceinline__ double lane_partner(double v) {
  const int lo = __double2loint(v), hi = __double2hiint(v);
  int nlo, nhi;
  if constexpr (OFF == 1 || OFF == 2 || OFF == 4 || OFF == 8) {
    constexpr int ctrl = OFF == 1 ? 0xB1 : (OFF == 2 ? 0x4E : (OFF == 4 ? 0x141 : 0x140));
    nlo = __builtin_amdgcn_update_dpp(0, lo, ctrl, 0xF, 0xF, false);
    nhi = __builtin_amdgcn_update_dpp(0, hi, ctrl, 0xF, 0xF, false);
  } else if constexpr (OFF == 16) {
    // odd rows of the first result hold the even rows' values, even rows of
    // the second the odd rows'
    const auto a = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
    const auto b = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
    const bool odd = (threadIdx.x & 16u) != 0;
    nlo = odd ? a[0] : a[1];
    nhi = odd ? b[0] : b[1];
  } else {
    static_assert(OFF == 32, "butterfly steps are 1, 2, 4, 8, 16, 32");
    const auto a = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
    const auto b = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
    const bool up = (threadIdx.x & 32u) != 0;
    nlo = up ? a[0] : a[1];
    nhi = up ? b[0] : b[1];
  }
  return __hiloint2double(nhi, nlo);
}

// One step of the full-wave butterfly for the slots [G0, G0 + kG) of NS.
template <int NS, int OFF, int G0, int kG>
__device__ __forceinline__ void wave_step(double* v, double* c, const int* ops) {
  double o[kG], oc[kG];
#pragma unroll
  for (int k = 0; k < kG; ++k) {
    const int i = G0 + k;
    if (i < NS) {
      o[k] = lane_partner<OFF>(v[i]);
      oc[k] = ops[i] == RCSUM ? lane_partner<OFF>(c[i]) : 0.0;
    }
  }
#pragma unroll
  for (int k = 0; k < kG; ++k) {
    const int i = G0 + k;
    if (i < NS) {
      if (ops[i] == RCSUM)
        kcombine(v[i], c[i], o[k], oc[k]);
      else
        v[i] = red_apply(ops[i], v[i], o[k]);
    }
  }
}
template <int NS, int G0, int kG>
__device__ __forceinline__ void wave_steps(double* v, double* c, const int* ops) {
  wave_step<NS, 1, G0, kG>(v, c, ops);
  wave_step<NS, 2, G0, kG>(v, c, ops);
  wave_step<NS, 4, G0, kG>(v, c, ops);
  wave_step<NS, 8, G0, kG>(v, c, ops);
  wave_step<NS, 16, G0, kG>(v, c, ops);
  wave_step<NS, 32, G0, kG>(v, c, ops);
  if constexpr (G0 + kG < NS) wave_steps<NS, G0 + kG, kG>(v, c, ops);
}

// Butterfly reduction of NS slots over the lanes of one wavefront (slot i
// with op ops[i], compile-time after inlining); every lane ends with the
// result.  Slots go in groups of up to four, interleaved step by step, so
// their exchanges overlap without raising the register peak of the kernels
// that reduce seven or eight slots.  The combinations are commutative
// (TwoSum's error term is exact in either order), so both lanes of a pair
// form the same value.  A block whose size is not a multiple of 64 has a
// partial last wave: lanes past the block end do not exist, so that wave
// exchanges through the LDS crossbar and skips them.
template <int NS>
__device__ __forceinline__ void wave_reduce(double* v, double* c, const int* ops) {
  const int lane = threadIdx.x & 63;
  const int nlanes = min(64, (int)blockDim.x - (int)(threadIdx.x & ~63u));
  constexpr int kG = 4;
  if (nlanes == 64) {
    wave_steps<NS, 0, kG>(v, c, ops);
    return;
  }
#pragma unroll
  for (int g0 = 0; g0 < NS; g0 += kG) {
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      double o[kG], oc[kG];
#pragma unroll
      for (int k = 0; k < kG; ++k) {
        const int i = g0 + k;
        if (i < NS) {
          o[k] = __shfl_xor(v[i], off, 64);
          oc[k] = ops[i] == RCSUM ? __shfl_xor(c[i], off, 64) : 0.0;
        }
      }
      if ((lane ^ off) < nlanes) {
#pragma unroll
        for (int k = 0; k < kG; ++k) {
          const int i = g0 + k;
          if (i < NS) {
            if (ops[i] == RCSUM)
              kcombine(v[i], c[i], o[k], oc[k]);
            else
              v[i] = red_apply(ops[i], v[i], o[k]);
          }
        }
      }
    }
  }
}

// Block-wide reduction of NS slots (vals) and the correction terms of the
// compensated ones (comp); result valid in thread 0.  `scratch` must hold
// 2 * NS * 16 doubles of shared memory (up to 16 waves).  Slot i is combined
// across the waves, in wave order, by thread i (all slots at once).
template <int NS>
__device__ __forceinline__ void block_reduce(double* vals, double* comp, const int* ops, double* scratch) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  double v[NS], c[NS];
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    v[i] = vals[i];
    c[i] = comp[i];
  }
  wave_reduce<NS>(v, c, ops);
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < NS; ++i) {
      scratch[i * 16 + wid] = v[i];
      scratch[(NS + i) * 16 + wid] = ops[i] == RCSUM ? c[i] : 0.0;
    }
  }
  __syncthreads();
  if (threadIdx.x < NS) {
    const int i = threadIdx.x;
    // ops[i] from bit masks of the (compile-time) ops, so no private copy of
    // the array is indexed at run time
    unsigned m_cs = 0, m_max = 0, m_min = 0;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      m_cs |= (ops[k] == RCSUM ? 1u : 0u) << k;
      m_max |= (ops[k] == RMAX ? 1u : 0u) << k;
      m_min |= (ops[k] == RMIN ? 1u : 0u) << k;
    }
    const int op = ((m_cs >> i) & 1u) ? RCSUM : ((m_max >> i) & 1u) ? RMAX : ((m_min >> i) & 1u) ? RMIN : RSUM;
    double v = scratch[i * 16], c = scratch[(NS + i) * 16];
    for (int w = 1; w < nw; ++w) {
      if (op == RCSUM)
        kcombine(v, c, scratch[i * 16 + w], scratch[(NS + i) * 16 + w]);
      else
        v = red_apply(op, v, scratch[i * 16 + w]);
    }
    scratch[i * 16] = v;
    scratch[(NS + i) * 16] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NS; ++i) {
      vals[i] = scratch[i * 16];
      comp[i] = scratch[(NS + i) * 16];
    }
  }
  __syncthreads();
}

}  // namespace cm
