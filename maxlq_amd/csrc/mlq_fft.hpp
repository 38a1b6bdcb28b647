// Device FFT building blocks for gfx950 (64-wide wavefronts).
//
// One 1-D transform of length N (a power of two, 8..1024) is computed by Tf =
// N/E cooperating threads, each owning E elements: element t + m*Tf lives in
// register slot m of thread t on entry and on exit.  The transform is a
// Stockham auto-sort sequence of radix-8/4/2 passes; the first pass works on
// the registers loaded straight from HBM, later passes exchange through LDS,
// and the last pass leaves natural-order results in the same register slots,
// so global loads and stores of every pass kernel use the same
// (t + m*Tf) pattern and stay coalesced across the batch/column dimension.
#pragma once
#include <type_traits>

#include "mlq_common.hpp"

namespace mlq {

template <typename T> __device__ __forceinline__ c2<T> cmk(T x, T y) {
  c2<T> r; r.x = x; r.y = y; return r;
}
#define MLQ_CPLX_OPS(V, S)                                                              \
  __device__ __forceinline__ V cadd(V a, V b) { return cmk<S>(a.x + b.x, a.y + b.y); } \
  __device__ __forceinline__ V csub(V a, V b) { return cmk<S>(a.x - b.x, a.y - b.y); } \
  __device__ __forceinline__ V cmul(V a, V b) {                                         \
    return cmk<S>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);                        \
  }                                                                                     \
  __device__ __forceinline__ V cscale(V a, S s) { return cmk<S>(a.x * s, a.y * s); }
MLQ_CPLX_OPS(float2, float)
MLQ_CPLX_OPS(double2, double)
#undef MLQ_CPLX_OPS
// multiply by (DIR * i): DIR = -1 forward, +1 inverse
template <typename T, int DIR> __device__ __forceinline__ c2<T> cmul_i(c2<T> a) {
  return DIR > 0 ? cmk<T>(-a.y, a.x) : cmk<T>(a.y, -a.x);
}

// first radix pass of FP32 transforms in FP32 arithmetic (0: FP64 as the rest)
#ifndef MLQ_P0_F32
#define MLQ_P0_F32 1
#endif

// ---- compile-time radix plan ------------------------------------------------
constexpr int radix_for(int rem, int E) {
  return (rem % 8 == 0 && E % 8 == 0) ? 8 : ((rem % 4 == 0 && E % 4 == 0) ? 4 : 2);
}
constexpr int radix_at(int N, int E, int p) {
  int rem = N;
  for (int i = 0; i < p; ++i) rem /= radix_for(rem, E);
  return radix_for(rem, E);
}
constexpr int ns_at(int N, int E, int p) {
  int ns = 1, rem = N;
  for (int i = 0; i < p; ++i) { int r = radix_for(rem, E); ns *= r; rem /= r; }
  return ns;
}
constexpr int npasses(int N, int E) {
  int p = 0, rem = N;
  while (rem > 1) { rem /= radix_for(rem, E); ++p; }
  return p;
}
template <int N> struct FftShape {
  static constexpr int E = (N >= 1024) ? 16 : 8;
  static constexpr int Tf = N / E;
};

// ---- small DFTs, natural order in and out ------------------------------------
// x * sqrt(1/2); in FP32 with the constant split into two floats (hh + hl
// within 3e-16 of sqrt(1/2)): a single FP32 constant would scale every such
// product by the same factor 1 - 1.7e-8
template <typename T> __device__ __forceinline__ T mul_h(T x);
template <> __device__ __forceinline__ double mul_h<double>(double x) {
  return 0.70710678118654752440084436210485 * x;
}
template <> __device__ __forceinline__ float mul_h<float>(float x) {
  return __builtin_fmaf(0.70710676908493042f, x, 1.2101617485882343e-08f * x);
}
template <typename T, int DIR> __device__ __forceinline__ void dft2(c2<T>* x) {
  c2<T> a = x[0], b = x[1];
  x[0] = cadd(a, b);
  x[1] = csub(a, b);
}
template <typename T, int DIR> __device__ __forceinline__ void dft4(c2<T>* x) {
  c2<T> s02 = cadd(x[0], x[2]), d02 = csub(x[0], x[2]);
  c2<T> s13 = cadd(x[1], x[3]), d13 = cmul_i<T, DIR>(csub(x[1], x[3]));
  x[0] = cadd(s02, s13);
  x[2] = csub(s02, s13);
  x[1] = cadd(d02, d13);
  x[3] = csub(d02, d13);
}
template <typename T, int DIR> __device__ __forceinline__ void dft8(c2<T>* x) {
  c2<T> e[4] = {x[0], x[2], x[4], x[6]};
  c2<T> o[4] = {x[1], x[3], x[5], x[7]};
  dft4<T, DIR>(e);
  dft4<T, DIR>(o);
  // o[s] *= w8^s, w8 = exp(DIR i pi/4)
  c2<T> o1 = cmk<T>(mul_h<T>(o[1].x - DIR * o[1].y), mul_h<T>(o[1].y + DIR * o[1].x));
  c2<T> o2 = cmul_i<T, DIR>(o[2]);
  c2<T> o3 = cmk<T>(mul_h<T>(-o[3].x - DIR * o[3].y), mul_h<T>(-o[3].y + DIR * o[3].x));
  x[0] = cadd(e[0], o[0]);
  x[4] = csub(e[0], o[0]);
  x[1] = cadd(e[1], o1);
  x[5] = csub(e[1], o1);
  x[2] = cadd(e[2], o2);
  x[6] = csub(e[2], o2);
  x[3] = cadd(e[3], o3);
  x[7] = csub(e[3], o3);
}
// dft8 of an input whose x[3] and x[4] are zero (the first pass of an
// inverse transform of a band-limited input: elements 3N/8 .. 5N/8 - 1 lie
// outside |m| < 3N/8); the same results as dft8 without the terms in zero
template <typename T, int DIR> __device__ __forceinline__ void dft8_z34(c2<T>* x) {
  // even half {x0, x2, 0, x6}
  const c2<T> es13 = cadd(x[2], x[6]), ed13 = cmul_i<T, DIR>(csub(x[2], x[6]));
  c2<T> e[4] = {cadd(x[0], es13), cadd(x[0], ed13), csub(x[0], es13), csub(x[0], ed13)};
  // odd half {x1, 0, x5, x7}
  const c2<T> os02 = cadd(x[1], x[5]), od02 = csub(x[1], x[5]);
  const c2<T> od13 = cmul_i<T, DIR>(cmk<T>(-x[7].x, -x[7].y));
  c2<T> o[4] = {cadd(os02, x[7]), cadd(od02, od13), csub(os02, x[7]), csub(od02, od13)};
  c2<T> o1 = cmk<T>(mul_h<T>(o[1].x - DIR * o[1].y), mul_h<T>(o[1].y + DIR * o[1].x));
  c2<T> o2 = cmul_i<T, DIR>(o[2]);
  c2<T> o3 = cmk<T>(mul_h<T>(-o[3].x - DIR * o[3].y), mul_h<T>(-o[3].y + DIR * o[3].x));
  x[0] = cadd(e[0], o[0]);
  x[4] = csub(e[0], o[0]);
  x[1] = cadd(e[1], o1);
  x[5] = csub(e[1], o1);
  x[2] = cadd(e[2], o2);
  x[6] = csub(e[2], o2);
  x[3] = cadd(e[3], o3);
  x[7] = csub(e[3], o3);
}
template <typename T, int R, int DIR> __device__ __forceinline__ void dftR(c2<T>* x) {
  if constexpr (R == 8) dft8<T, DIR>(x);
  else if constexpr (R == 4) dft4<T, DIR>(x);
  else dft2<T, DIR>(x);
}

// ---- LDS views -------------------------------------------------------------------
// Row view: one transform per contiguous LDS row, one pad slot every 8
// elements so that the stride-R writes of the first passes hit distinct banks.
// Storage type S is the field precision; ld returns it unchanged and st rounds
// an FP64 butterfly result to it.
template <typename S> struct LdsRow {
  using Store = S;
  // the Tf threads of a row transform are consecutive (Tf divides 64): one wave
  static constexpr bool wave_local = true;
  c2<S>* base;
  static __host__ __device__ constexpr int pitch(int N) { return N + N / 8; }
  __device__ __forceinline__ c2<S> ld(int n) const { return base[n + (n >> 3)]; }
  __device__ __forceinline__ void st(int n, c2<double> v) const {
    base[n + (n >> 3)] = cmk<S>((S)v.x, (S)v.y);
  }
  __device__ __forceinline__ void st_s(int n, c2<S> v) const { base[n + (n >> 3)] = v; }
};
// Column view: C transforms interleaved (element n of column c at n*C + c), so
// consecutive threads (consecutive columns) touch consecutive LDS words.
template <typename S, int C> struct LdsCol {
  using Store = S;
  static constexpr bool wave_local = false;   // a column's threads are C apart
  c2<S>* base;
  int c;
  __device__ __forceinline__ c2<S> ld(int n) const { return base[n * C + c]; }
  __device__ __forceinline__ void st(int n, c2<double> v) const {
    base[n * C + c] = cmk<S>((S)v.x, (S)v.y);
  }
  __device__ __forceinline__ void st_s(int n, c2<S> v) const { base[n * C + c] = v; }
};

// precision conversion of a complex value (c2<S> is not deducible: one
// overload per source type)
template <typename D> __device__ __forceinline__ c2<D> ccvt(float2 a) {
  return cmk<D>((D)a.x, (D)a.y);
}
template <typename D> __device__ __forceinline__ c2<D> ccvt(double2 a) {
  return cmk<D>((D)a.x, (D)a.y);
}

// Synchronisation between the LDS exchanges of one transform: a workgroup
// barrier when the transform's threads span waves (column view), a
// wave-level one when each transform lies within one wave (row view).  LDS
// instructions of one wave execute in order, so the wave-level form only keeps
// the compiler from moving LDS accesses across it.
template <class L> __device__ __forceinline__ void lds_sync() {
  if constexpr (L::wave_local) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// ---- twiddle sources ---------------------------------------------------------------
// Both multiply x[r], r = 1 .. R-1, of a butterfly with local index k (within
// its Ns-point sub-transforms) by w^(r k), w = exp(-+2 pi i / (Ns R)), from
// the FP64 table tw[m] = exp(-2 pi i m / N).
// TwPow: one table load per butterfly (w^k); the other twiddles follow as
// FP64 products (within a few FP64 ulps of the table values).  For the row
// view, whose 64 lanes of a wave need 64 distinct twiddles: measured at
// 512^3 FP32, one global table load per twiddle cost the passes as much as
// their whole butterfly arithmetic.
struct TwPow {
  const double2* __restrict__ tw;
  template <int N, int R, int Ns, int DIR>
  __device__ __forceinline__ void apply(c2<double> (&x)[R], int k) const {
    const double2 t1 = tw[k * (N / (Ns * R))];
    const c2<double> w1 = cmk<double>(t1.x, DIR > 0 ? -t1.y : t1.y);
    c2<double> w = w1;
    x[1] = cmul(x[1], w);
#pragma unroll
    for (int r = 2; r < R; ++r) {
      w = cmul(w, w1);
      x[r] = cmul(x[r], w);
    }
  }
};
// TwTab: every twiddle read from a copy of the table in LDS (the caller fills
// it and orders the fill before the first pass with twiddles).  For the
// column view, whose lanes share their twiddles (C adjacent columns per row
// index: a few distinct addresses per wave, broadcast reads).
struct TwTab {
  const double2* tw;
  template <int N, int R, int Ns, int DIR>
  __device__ __forceinline__ void apply(c2<double> (&x)[R], int k) const {
#pragma unroll
    for (int r = 1; r < R; ++r) {
      const double2 w = tw[r * k * (N / (Ns * R))];
      x[r] = cmul(x[r], cmk<double>(w.x, DIR > 0 ? -w.y : w.y));
    }
  }
};
// Copy of the twiddle table into LDS by the NT threads of a workgroup, in two
// steps: load() at the kernel's start (the table loads overlap the data
// loads), store() before the transform's first LDS barrier.
template <int N, int NT> struct TwStage {
  static constexpr int PT = (N + NT - 1) / NT;   // entries per thread
  double2 r[PT];
  __device__ __forceinline__ void load(const double2* __restrict__ tw) {
#pragma unroll
    for (int u = 0; u < PT; ++u) {
      const int i = threadIdx.x + u * NT;
      r[u] = tw[i < N ? i : 0];
    }
  }
  __device__ __forceinline__ void store(double2* sm) const {
#pragma unroll
    for (int u = 0; u < PT; ++u) {
      const int i = threadIdx.x + u * NT;
      if (i < N) sm[i] = r[u];
    }
  }
};

// ---- one Stockham pass -----------------------------------------------------------
// Precision: every twiddle product and butterfly is evaluated in FP64.  The
// per-thread register array v has type RT: the storage precision in the
// strided pass kernels (an FP32 field then holds E complex floats between
// passes, the same rounding as the LDS exchange, and the kernel stays within
// 64 VGPRs), FP64 where the caller keeps more registers to spare.
template <typename RT, int N, int E, int DIR, int P, class L, bool ZB = false, class TW>
__device__ __forceinline__ void fft_pass(c2<RT> (&v)[E], int t, const TW& tw, const L& lds) {
  constexpr int Tf = N / E;
  constexpr int R = radix_at(N, E, P);
  constexpr int Ns = ns_at(N, E, P);
  constexpr int NB = E / R;  // butterflies per thread
  constexpr bool last = (P == npasses(N, E) - 1);
  static_assert(!L::wave_local || (Tf <= 64 && 64 % Tf == 0), "a row transform within one wave");
  if constexpr (P > 0) {
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int j = t + b * Tf;
#pragma unroll
      for (int r = 0; r < R; ++r) v[b + r * NB] = ccvt<RT>(lds.ld(j + r * (N / R)));
    }
    lds_sync<L>();
  }
  // The first pass of an FP32 transform (no twiddles, its outputs go to FP32
  // LDS rows) runs in FP32: its adds round as the LDS store would, and the
  // sqrt(1/2) products are compensated.  Every later pass is FP64.
  if constexpr (P == 0 && !last && std::is_same<RT, float>::value &&
                std::is_same<typename L::Store, float>::value && MLQ_P0_F32) {
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int j = t + b * Tf;
      c2<float> x[R];
#pragma unroll
      for (int r = 0; r < R; ++r) x[r] = v[b + r * NB];
      if constexpr (ZB && DIR > 0 && R == 8) dft8_z34<float, DIR>(x);
      else dftR<float, R, DIR>(x);
#pragma unroll
      for (int r = 0; r < R; ++r) lds.st_s(j * R + r, x[r]);   // Ns = 1: idxD = j R
    }
    lds_sync<L>();
    fft_pass<RT, N, E, DIR, P + 1, L, ZB, TW>(v, t, tw, lds);
    return;
  }
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = t + b * Tf;
    c2<double> x[R];
#pragma unroll
    for (int r = 0; r < R; ++r) x[r] = ccvt<double>(v[b + r * NB]);
    if constexpr (Ns > 1) tw.template apply<N, R, Ns, DIR>(x, j & (Ns - 1));
    if constexpr (ZB && DIR > 0 && P == 0 && R == 8) {
      dft8_z34<double, DIR>(x);   // inverse: inputs 3N/8 .. 5N/8 - 1 are zero
    } else {
      dftR<double, R, DIR>(x);
    }
    if constexpr (ZB && DIR < 0 && last && R == 8) {
      // forward: outputs 3N/8 .. 5N/8 - 1 are never stored (their adds drop out)
      x[3] = cmk<double>(0.0, 0.0);
      x[4] = cmk<double>(0.0, 0.0);
    }
    if constexpr (!last) {
      const int idxD = (j / Ns) * Ns * R + (j & (Ns - 1));
#pragma unroll
      for (int r = 0; r < R; ++r) lds.st(idxD + r * Ns, x[r]);
    } else {
#pragma unroll
      for (int r = 0; r < R; ++r) v[b + r * NB] = ccvt<RT>(x[r]);
    }
  }
  if constexpr (!last) {
    lds_sync<L>();
    fft_pass<RT, N, E, DIR, P + 1, L, ZB, TW>(v, t, tw, lds);
  }
}

// Full length-N transform; DIR = -1 forward (exp(-i...)), +1 inverse
// (unnormalised), twiddles from tw (TwPow or TwTab over the table
// exp(-2 pi i k / N) in FP64).
// ZB, for a band-limited launch with keep < 3N/8: an inverse transform's
// elements 3N/8 .. 5N/8 - 1 are zero on entry (the first radix-8 pass skips
// them); a forward transform's outputs there are never stored (the last pass
// leaves them zero).
template <typename RT, int N, int E, int DIR, class L, bool ZB = false, class TW>
__device__ __forceinline__ void fft_core(c2<RT> (&v)[E], int t, const TW& tw, const L& lds) {
  fft_pass<RT, N, E, DIR, 0, L, ZB, TW>(v, t, tw, lds);
}

}  // namespace mlq
