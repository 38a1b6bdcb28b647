// Elementwise functors (physical space, x fastest) and spectral-space kernels.
#pragma once
#include "mlq_kernels.hpp"

namespace mlq {

constexpr int MAP_THREADS = 256;
#ifndef MLQ_MAP_CAP
#define MLQ_MAP_CAP ((1LL << 24) - 1)
#endif

template <class F> __global__ __launch_bounds__(MAP_THREADS) void k_map(F f, long long n) {
  for (long long i = (long long)blockIdx.x * MAP_THREADS + threadIdx.x; i < n;
       i += (long long)gridDim.x * MAP_THREADS)
    f(i);
}

// ---- physical-space functors -------------------------------------------------
template <typename T> struct F_fill { T* o; T v; __device__ void operator()(long long i) const { o[i] = v; } };
// exact widening of FP32 values to FP64 (4 points per call)
struct F_widen {
  static constexpr int vec = 4;
  const float* a; double* o;
  double abytes() const { return 12.0; }
  __device__ void operator()(long long i) const {
    float v[4];
    ld16(a, i, v);
    const double w0[2] = {(double)v[0], (double)v[1]}, w1[2] = {(double)v[2], (double)v[3]};
    st16(o, 2 * i, w0);
    st16(o, 2 * i + 1, w1);
  }
};

template <typename T> struct F_copy {   // 16-byte vectors
  static constexpr int vec = V16<T>::w;
  const T* a; T* o;
  double abytes() const { return 2.0 * sizeof(T); }
  __device__ void operator()(long long i) const {
    T v[vec];
    ld16(a, i, v);
    st16(o, i, v);
  }
};
template <typename T> struct F_scale {   // 16-byte vectors
  static constexpr int vec = V16<T>::w;
  T* o; T s;
  double abytes() const { return 2.0 * sizeof(T); }
  __device__ void operator()(long long i) const {
    T v[vec];
    ld16(o, i, v);
#pragma unroll
    for (int l = 0; l < vec; ++l) v[l] *= s;
    st16(o, i, v);
  }
};
template <typename T> struct F_divide {
  T* o; double s;
  double abytes() const { return 2.0 * sizeof(T); }
  __device__ void operator()(long long i) const { o[i] = (T)((double)o[i] / s); }
};
// o = a*x + b*y
template <typename T> struct F_axpby {   // 16-byte vectors; o may alias x or y
  static constexpr int vec = V16<T>::w;
  const T* x; const T* y; T* o; T a, b;
  double abytes() const { return (x == y ? 2.0 : 3.0) * sizeof(T); }
  __device__ void operator()(long long i) const {
    T u[vec], v[vec];
    ld16(x, i, u);
    ld16(y, i, v);
#pragma unroll
    for (int l = 0; l < vec; ++l) u[l] = a * u[l] + b * v[l];
    st16(o, i, u);
  }
};
template <typename T> struct F_mul {
  const T* a; const T* b; T* o;
  double abytes() const { return 3.0 * sizeof(T); }
  __device__ void operator()(long long i) const { o[i] = a[i] * b[i]; }
};
// o = u0^2 + u1^2 + u2^2  (calc_uk, function_ops.f90:2332)
template <typename T> struct F_dot3self {
  const T* u; long long nr; T* o;
  double abytes() const { return 4.0 * sizeof(T); }
  __device__ void operator()(long long i) const {
    const T a = u[i], b = u[nr + i], c = u[2 * nr + i];
    o[i] = a * a + b * b + c * c;
  }
};
// g = sqrt(s) where s > 0 else 0 (function_ops.f90:2335-2339)
template <typename T> struct F_sqrtclip {
  const T* s; T* o;
  __device__ void operator()(long long i) const { const T v = s[i]; o[i] = v <= (T)0 ? (T)0 : sqrt(v); }
};
// g^k for 0 <= k < 2 (function_ops.f90:2303) and the guarded negative power (:2297-2301)
template <typename T> struct F_pow {
  const T* g; T* o; double k; int neg;
  __device__ void operator()(long long i) const {
    const T v = g[i];
    if (neg) { o[i] = (fabs((double)v) < 2.0e-16) ? (T)0 : (T)pow((double)v, k); return; }
    if (k == 0.0) { o[i] = (T)1; return; }
    if (k == 1.0) { o[i] = v; return; }
    // x^(1/2) correctly rounded (what a correctly rounded pow returns; g >= 0
    // here), at a fraction of pow's cost
    if (k == 0.5) { o[i] = (T)sqrt((double)v); return; }
    o[i] = (T)pow((double)v, k);
  }
};
// |grad u|^2 = sum_ij (d_j u_i)^2   (function_ops.f90:2378-2386); G[i*3+j] = d_j u_i
template <typename T> struct F_gradsq {
  const T* G; long long nr; T* o;
  __device__ void operator()(long long i) const {
    T s = 0;
#pragma unroll
    for (int a = 0; a < 9; ++a) { const T g = G[a * nr + i]; s += g * g; }
    o[i] = s;
  }
};
// sum_ij d_i u_j d_j u_i   (function_ops.f90:1933-1938)
template <typename T> struct F_gradgradT {
  const T* G; long long nr; T* o;
  __device__ void operator()(long long i) const {
    T s = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) s += G[(b * 3 + a) * nr + i] * G[(a * 3 + b) * nr + i];
    o[i] = s;
  }
};
// w_j = sum_i v_i d_j u_i, j = 0..2 (function_ops.f90:2189-2194)
template <typename T> struct F_vgradu {
  const T* v; const T* G; long long nr; T* w0; T* w1; T* w2;
  __device__ void operator()(long long i) const {
    const T v0 = v[i], v1 = v[nr + i], v2 = v[2 * nr + i];
    w0[i] = v0 * G[0 * nr + i] + v1 * G[3 * nr + i] + v2 * G[6 * nr + i];
    w1[i] = v0 * G[1 * nr + i] + v1 * G[4 * nr + i] + v2 * G[7 * nr + i];
    w2[i] = v0 * G[2 * nr + i] + v1 * G[5 * nr + i] + v2 * G[8 * nr + i];
  }
};
// o = sum_j w_j^2
template <typename T> struct F_sumsq3 {
  const T* w0; const T* w1; const T* w2; T* o;
  __device__ void operator()(long long i) const {
    const T a = w0[i], b = w1[i], c = w2[i];
    o[i] = a * a + b * b + c * c;
  }
};
// o = sum_j v_j w_j  (v a vector with stride nr)
template <typename T> struct F_dot3 {
  const T* v; long long nr; const T* w0; const T* w1; const T* w2; T* o;
  __device__ void operator()(long long i) const {
    o[i] = v[i] * w0[i] + v[nr + i] * w1[i] + v[2 * nr + i] * w2[i];
  }
};
// e_u = u/|u| with |u| = sqrt(|s|), 0 where |u| < eps (function_ops.f90:1833-1841)
template <typename T> struct F_unit {
  const T* u; const T* s; long long nr; T* e;
  __device__ void operator()(long long i) const {
    const T nrm = sqrt(fabs(s[i]));
    const bool small = (double)nrm < 2.0e-16;
#pragma unroll
    for (int c = 0; c < 3; ++c) e[c * nr + i] = small ? (T)0 : u[c * nr + i] / nrm;
  }
};
// aux2 = sum_i v_i (vnu * Lap u_i - pc * grad_i p)   (function_ops.f90:1119, 1132)
// u . (vnu Lap u - pc grad p); with P null, L holds the force itself
template <typename T> struct F_udotforce {
  const T* v; const T* L; const T* P; long long nr; T vnu, pc; T* o;
  __device__ T at(long long i) const {
    T s = 0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const T f = P ? vnu * gld<NT_PROD>(L + c * nr + i) - pc * gld<NT_PROD>(P + c * nr + i)
                    : gld<NT_PROD>(L + c * nr + i);
      s += gld<NT_PROD>(v + c * nr + i) * f;
    }
    return s;
  }
  __device__ void operator()(long long i) const { o[i] = at(i); }
};
// producer form for the fused forward x pass
template <typename T> struct P_udotforce {
  static constexpr int NB = 1;
  F_udotforce<T> f;   // o unused
  double inbytes() const { return (f.P ? 9.0 : 6.0) * sizeof(T); }
  __device__ void operator()(long long i, T (&r)[1]) const { r[0] = f.at(i); }
};
// o = sum_j u_j g_j (one component of (u . grad) u from g_j = d_j u_i,
// function_ops.f90:1456-1461)
template <typename T> struct F_udotgrad {
  const T* u; const T* g; long long nr; T* o;
  __device__ void operator()(long long i) const {
    o[i] = u[i] * g[i] + u[nr + i] * g[nr + i] + u[2 * nr + i] * g[2 * nr + i];
  }
};
// aux5_i = sum_j d_i u_j * d_j phi   (function_ops.f90:1165-1169)
template <typename T> struct F_gradT_dot {
  const T* G; const T* gphi; long long nr; T* o;  // o: 3 components
  __device__ void operator()(long long i) const {
    const T g0 = gphi[i], g1 = gphi[nr + i], g2 = gphi[2 * nr + i];
#pragma unroll
    for (int c = 0; c < 3; ++c)
      o[c * nr + i] = G[(0 * 3 + c) * nr + i] * g0 + G[(1 * 3 + c) * nr + i] * g1 +
                      G[(2 * 3 + c) * nr + i] * g2;
  }
};
// V = a3*q2 - pc*a4 + 2 pc*a5 + vnu*a6 + vnu*a7  (function_ops.f90:1186-1190)
template <typename T> struct F_gradcombine {
  const T* a3; const T* a4; const T* a5; const T* a6; const T* a7; T* o;
  T q2, pc, vnu;
  __device__ void operator()(long long i) const {
    o[i] = q2 * a3[i] - pc * a4[i] + pc * (T)2 * a5[i] + vnu * a6[i] + vnu * a7[i];
  }
};
// vector transport final combination (function_ops.f90:1746-1747)
template <typename T> struct F_transport {
  const T* xi; const T* upe; T* o; T c1, sc;
  __device__ void operator()(long long i) const { o[i] = sc * (xi[i] - c1 * upe[i]); }
};

// seeded white noise: SplitMix64 finaliser + Box-Muller, FP64
__device__ __forceinline__ unsigned long long splitmix64(unsigned long long x) {
  unsigned long long z = x + 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
template <typename T> struct F_noise {
  T* o; unsigned long long base; long long npts; int comp;
  __device__ void operator()(long long i) const {
    const unsigned long long lin = ((unsigned long long)comp * (unsigned long long)npts +
                                    (unsigned long long)i) * 2ULL;
    const unsigned long long h1 = splitmix64(base + lin), h2 = splitmix64(base + lin + 1ULL);
    const double u1 = ((double)(h1 >> 11) + 1.0) * 0x1.0p-53;
    const double u2 = (double)(h2 >> 11) * 0x1.0p-53;
    o[i] = (T)(sqrt(-2.0 * log(u1)) * cos(TWO_PI * u2));
  }
};

// ---- reduction functors (accumulate in FP64) --------------------------------------
template <typename T> struct R_dot {
  const T* a; const T* b;
  double abytes() const { return (a == b ? 1.0 : 2.0) * sizeof(T); }
  __device__ void operator()(long long i, double (&acc)[1]) const { acc[0] += (double)a[i] * (double)b[i]; }
};
template <typename T> struct R_sum {
  const T* a;
  __device__ void operator()(long long i, double (&acc)[1]) const { acc[0] += (double)a[i]; }
};
template <typename T> struct R_nan {
  const T* a;
  __device__ void operator()(long long i, double (&acc)[1]) const { if (a[i] != a[i]) acc[0] += 1.0; }
};

// Weighted Parseval sums over the half spectrum: sum_k m(kx) w(k)^2 Re(a conj b)
// with m = 1 on the kx = 0 and kx = N/2 planes and 2 elsewhere.
// kind 0: w = 1; 1: w = 1 + (lam|K|)^s (H_l); 2: w = (1 + |K|)^s (H).
template <typename T> struct R_specdot {
  static constexpr int unroll = 1;
  const c2<T>* a; const c2<T>* b; int n, px, kind; double lam, s;
  __device__ void operator()(long long i, double (&acc)[1]) const {
    const RowIdx ri = row_split(i, px);   // px: row pitch
    const int kx = ri.kx;
    if (kx > (n >> 1)) return;             // row padding
    const int ky = row_ky(ri.row, n), kz = row_kz(ri.row, n);
    const int h = n >> 1;
    const int my = ky <= h ? ky : ky - n, mz = kz <= h ? kz : kz - n;
    double w = 1.0;
    if (kind != 0) {
      const double K = TWO_PI * sqrt((double)kx * kx + (double)my * my + (double)mz * mz);
      w = kind == 1 ? 1.0 + pow(lam * K, s) : pow(1.0 + K, s);
    }
    const double m = (kx == 0 || kx == h) ? 1.0 : 2.0;
    const c2<T> x = a[i], y = b[i];
    acc[0] += m * w * w * ((double)x.x * (double)y.x + (double)x.y * (double)y.y);
  }
};

// ---- spectral-space kernels ------------------------------------------------------
// Leray projection + mean removal on three half spectra in place
// (divAvg_free, function_ops.f90:2867-2909): u_j - K'_j (K'.u)/|K|^2 with K'
// the Nyquist-zeroed wave vector; k = 0 -> 0.
template <typename T> struct S_leray {
  c2<T>* u0; c2<T>* u1; c2<T>* u2; int n, px;
  __device__ void operator()(long long i) const {
    const RowIdx ri = row_split(i, px);   // px: row pitch
    const int kx = ri.kx;
    if (kx > (n >> 1)) return;             // row padding
    const int ky = row_ky(ri.row, n), kz = row_kz(ri.row, n);
    const int h = n >> 1;
    const int mx = kx, my = ky <= h ? ky : ky - n, mz = kz <= h ? kz : kz - n;
    const double K1 = TWO_PI * mx, K2 = TWO_PI * my, K3 = TWO_PI * mz;
    const double ksq = K1 * K1 + K2 * K2 + K3 * K3;
    if (ksq <= 2.0e-16) {
      u0[i] = cmk<T>(0, 0); u1[i] = cmk<T>(0, 0); u2[i] = cmk<T>(0, 0);
      return;
    }
    const double e1 = mx == h ? 0.0 : K1, e2 = my == h ? 0.0 : K2, e3 = mz == h ? 0.0 : K3;
    const c2<T> a = u0[i], b = u1[i], c = u2[i];
    const double dr = (e1 * a.x + e2 * b.x + e3 * c.x) / ksq;
    const double di = (e1 * a.y + e2 * b.y + e3 * c.y) / ksq;
    u0[i] = cmk<T>((T)(a.x - e1 * dr), (T)(a.y - e1 * di));
    u1[i] = cmk<T>((T)(b.x - e2 * dr), (T)(b.y - e2 * di));
    u2[i] = cmk<T>((T)(c.x - e3 * dr), (T)(c.y - e3 * di));
  }
};

// Sobolev weight of HilbertGradient (function_ops.f90:3605-3615) on the three
// spectra of a vector: u_j / (1 + (lam |K|)^e); the k = 0 mode keeps weight 1.
template <typename T> struct S_hilbert3 {
  c2<T>* u0; c2<T>* u1; c2<T>* u2; int n, px; double lam, e;
  __device__ void operator()(long long i) const {
    const RowIdx ri = row_split(i, px);   // px: row pitch
    const int kx = ri.kx;
    if (kx > (n >> 1)) return;             // row padding
    const int ky = row_ky(ri.row, n), kz = row_kz(ri.row, n);
    const int h = n >> 1;
    const int mx = kx, my = ky <= h ? ky : ky - n, mz = kz <= h ? kz : kz - n;
    const double K1 = TWO_PI * mx, K2 = TWO_PI * my, K3 = TWO_PI * mz;
    const double w = 1.0 / (1.0 + pow(lam * sqrt(K1 * K1 + K2 * K2 + K3 * K3), e));
    const c2<T> a = u0[i], b = u1[i], c = u2[i];
    u0[i] = cmk<T>((T)(w * a.x), (T)(w * a.y));
    u1[i] = cmk<T>((T)(w * b.x), (T)(w * b.y));
    u2[i] = cmk<T>((T)(w * c.x), (T)(w * c.y));
  }
};
// divergence spectrum: sum_l i K'_l w_l  (function_ops.f90:3231)
template <typename T> struct S_div {
  const c2<T>* w0; const c2<T>* w1; const c2<T>* w2; c2<T>* o; int n, px;
  __device__ void operator()(long long i) const {
    const RowIdx ri = row_split(i, px);   // px: row pitch
    const int kx = ri.kx;
    if (kx > (n >> 1)) return;             // row padding
    const int ky = row_ky(ri.row, n), kz = row_kz(ri.row, n);
    const int h = n >> 1;
    const int mx = kx, my = ky <= h ? ky : ky - n, mz = kz <= h ? kz : kz - n;
    const double e1 = mx == h ? 0.0 : TWO_PI * mx, e2 = my == h ? 0.0 : TWO_PI * my,
                 e3 = mz == h ? 0.0 : TWO_PI * mz;
    const c2<T> a = w0[i], b = w1[i], c = w2[i];
    const double re = e1 * a.x + e2 * b.x + e3 * c.x;
    const double im = e1 * a.y + e2 * b.y + e3 * c.y;
    o[i] = cmk<T>((T)(-im), (T)re);
  }
};
// o = a*x + b*y on spectra
template <typename T> struct S_axpby {
  const c2<T>* x; const c2<T>* y; c2<T>* o; double a, b;
  __device__ void operator()(long long i) const {
    const c2<T> p = x[i], q = y[i];
    o[i] = cmk<T>((T)(a * p.x + b * q.x), (T)(a * p.y + b * q.y));
  }
};
// seeded envelope + k=0 removal + Leray in place (mlq_seeded_field)
template <typename T> struct S_envelope_leray {
  c2<T>* u0; c2<T>* u1; c2<T>* u2; int n, px, kind; double a;
  __device__ void operator()(long long i) const {
    const RowIdx ri = row_split(i, px);   // px: row pitch
    const int kx = ri.kx;
    if (kx > (n >> 1)) return;             // row padding
    const int ky = row_ky(ri.row, n), kz = row_kz(ri.row, n);
    const int h = n >> 1;
    const int mx = kx, my = ky <= h ? ky : ky - n, mz = kz <= h ? kz : kz - n;
    const double m2 = (double)mx * mx + (double)my * my + (double)mz * mz;
    double E = exp(-m2 / a);
    if (kind == 1) E *= m2;
    if (m2 == 0.0) E = 0.0;
    const double K1 = TWO_PI * mx, K2 = TWO_PI * my, K3 = TWO_PI * mz;
    double ksq = K1 * K1 + K2 * K2 + K3 * K3;
    if (ksq == 0.0) ksq = 1.0;
    const double e1 = mx == h ? 0.0 : K1, e2 = my == h ? 0.0 : K2, e3 = mz == h ? 0.0 : K3;
    const c2<T> p = u0[i], q = u1[i], r = u2[i];
    const double ar = E * p.x, ai = E * p.y, br = E * q.x, bi = E * q.y, cr = E * r.x, ci = E * r.y;
    const double dr = (e1 * ar + e2 * br + e3 * cr) / ksq;
    const double di = (e1 * ai + e2 * bi + e3 * ci) / ksq;
    u0[i] = cmk<T>((T)(ar - e1 * dr), (T)(ai - e1 * di));
    u1[i] = cmk<T>((T)(br - e2 * dr), (T)(bi - e2 * di));
    u2[i] = cmk<T>((T)(cr - e3 * dr), (T)(ci - e3 * di));
  }
};

// ---- energy spectrum in spherical shells (calculate_spectral_data,
// function_ops.f90:2614-2677) ----------------------------------------------
// Mode m of the full complex spectrum lies in shell i when
//   2 pi i < |K| <= 2 pi (i+1),  |K| = SQRT(K1**2 + K2**2 + K3**2),
// K_j = (2 pi) m_j (initialize.f90:170-192), i = 0 .. n-1 (:2640-2659).  |K|
// is evaluated in the reference's FP64 operation order with no contraction
// and a correctly rounded square root, and the shell bounds as the reference
// forms them, so a mode on a shell boundary (|m| an integer) lands in the
// same shell.  The half spectrum holds the modes with 0 < kx < n/2 for
// themselves and their mirror -m (same |K|, same |u_hat|^2): weight 2.
__device__ __forceinline__ double shell_bound(int i) {   // 2.0_pr*PI*REAL(i,pr)
  return TWO_PI * (double)i;
}
// K1**2 + K2**2 + K3**2, K_j = 2 pi m_j, as the reference evaluates it: each
// square rounded, summed left to right (a fused multiply-add would move modes
// with integer |m| across a shell edge)
__device__ __forceinline__ double kabs2_ref(int m1, int m2, int m3) {
#pragma clang fp contract(off)
  const double K1 = TWO_PI * (double)m1, K2 = TWO_PI * (double)m2, K3 = TWO_PI * (double)m3;
  return (K1 * K1 + K2 * K2) + K3 * K3;
}
// the shell of |K| = nk > 0 (n when beyond the last shell); both loops are
// bounded whatever nk holds
__device__ __forceinline__ int shell_of(double nk, int n) {
  int i = (int)fmin(fmax(nk / TWO_PI, 0.0), (double)n);
  while (i > 0 && !(shell_bound(i) < nk)) --i;
  while (i < n && !(nk <= shell_bound(i + 1))) ++i;
  return i;
}
// Each workgroup sums its contiguous chunk of modes into LDS bins and writes
// them to part[blockIdx.x * n + shell]; k_shell_final adds the workgroups'
// partials in workgroup order.
template <typename T>
__global__ __launch_bounds__(256) void k_shell_sum(const c2<T>* s0, const c2<T>* s1,
                                                   const c2<T>* s2, int n, int px,
                                                   long long nmodes, long long chunk,
                                                   double* part) {
  extern __shared__ double bins[];
  for (int b = threadIdx.x; b < n; b += blockDim.x) bins[b] = 0.0;
  __syncthreads();
  const long long i0 = (long long)blockIdx.x * chunk;
  const long long i1 = i0 + chunk < nmodes ? i0 + chunk : nmodes;
  for (long long i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    KIdx k;
    if (!kdecode(i, n, px, k)) continue;
    const double ksq = kabs2_ref(k.mx, k.my, k.mz);
    if (ksq == 0.0) continue;   // |K| = 0 is in no shell
    const int sh = shell_of(__dsqrt_rn(ksq), n);
    if (sh >= n) continue;
    const c2<T> a = gld<NT_PROD>(s0 + i), b = gld<NT_PROD>(s1 + i), c = gld<NT_PROD>(s2 + i);
    const double e = ((double)a.x * a.x + (double)a.y * a.y) + ((double)b.x * b.x + (double)b.y * b.y) +
                     ((double)c.x * c.x + (double)c.y * c.y);
    const double w = (k.kx == 0 || k.kx == (n >> 1)) ? 1.0 : 2.0;
    atomicAdd(&bins[sh], w * e);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < n; b += blockDim.x) part[(long long)blockIdx.x * n + b] = bins[b];
}
__global__ void k_shell_final(const double* part, int nblk, int n, double* out);

}  // namespace mlq
