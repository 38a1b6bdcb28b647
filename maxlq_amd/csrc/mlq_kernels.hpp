// Pass kernels of the batched 3-D real transforms and the elementwise /
// spectral / reduction kernels of the maxLq hot path.
//
// Data layout in HBM (per field):
//   physical  T[n3][n2][n1]            (x fastest: the reference's Fortran order)
//   spectrum  c2<T>[n3][n2][px], px = n1/2+1  (kx fastest; kx in [0, n1/2])
// A 3-D forward transform is x (R2C, contiguous rows) -> y -> z passes; the
// inverse runs z -> y -> x.  Spectral multipliers are fused into the z passes
// (epilogue of the forward, prologue of the inverse), the 1/N^3 of fftbwd
// (fftwfunction.f90:109) into the inverse x pass.
#pragma once
#include <type_traits>

#include "mlq_fft.hpp"

namespace mlq {

constexpr int MAXB = 12;
constexpr double TWO_PI = 6.283185307179586476925286766559;

// Global loads and stores of the transform passes.  Every element of a pass
// is read once and written once, so they may carry the nontemporal hint
// (streamed past the caches).  The masks select it per pass class (bit c):
// 0 forward x, 1 inverse x, 2 inverse y, 3 forward y, 4 inverse z, 5 forward z.
// Measured at 512^3 FP32 (batches of 12): nontemporal stores gain in every
// class, nontemporal loads in every class but the inverse z pass (whose
// masked rows re-read row 0 of their column, a cache hit without the hint).
#ifndef MLQ_NT_LD
#define MLQ_NT_LD 47
#endif
#ifndef MLQ_NT_ST
#define MLQ_NT_ST 63
#endif
enum PassClass { PC_XF = 0, PC_XI = 1, PC_YI = 2, PC_YF = 3, PC_ZI = 4, PC_ZF = 5 };
template <int CLS> constexpr bool nt_ld() { return (MLQ_NT_LD >> CLS) & 1; }
template <int CLS> constexpr bool nt_st() { return (MLQ_NT_ST >> CLS) & 1; }
template <int MODE, int DIR> constexpr int strided_class() {
  return MODE == 1 ? PC_ZI : MODE == 2 ? PC_ZF : DIR > 0 ? PC_YI : PC_YF;
}
// the pointwise producers of the fused forward x passes and the kept-cube
// spectral kernels read (and write) each value once per launch
#ifndef MLQ_NT_PROD
#define MLQ_NT_PROD 1
#endif
constexpr bool NT_PROD = MLQ_NT_PROD != 0;
template <typename V> struct NatV { using type = V; };
template <> struct NatV<float2> { typedef float type __attribute__((ext_vector_type(2))); };
template <> struct NatV<double2> { typedef double type __attribute__((ext_vector_type(2))); };
template <bool NT, typename V> __device__ __forceinline__ V gld(const V* p) {
  if constexpr (NT) {
    using NV = typename NatV<V>::type;
    const NV n = __builtin_nontemporal_load(reinterpret_cast<const NV*>(p));
    V v;
    __builtin_memcpy(&v, &n, sizeof(V));
    return v;
  } else {
    return *p;
  }
}
template <bool NT, typename V> __device__ __forceinline__ void gst(V* p, const V& v) {
  if constexpr (NT) {
    using NV = typename NatV<V>::type;
    NV n;
    __builtin_memcpy(&n, &v, sizeof(V));
    __builtin_nontemporal_store(n, reinterpret_cast<NV*>(p));
  } else {
    *p = v;
  }
}

// Row pitch of a half spectrum in elements: n/2+1 rounded up so that every
// row starts on a 128-byte boundary (16 complex floats, 8 complex doubles).
// The padding entries kx in [n/2+1, pitch) are never read or written.
template <typename T> constexpr int spec_pitch(int n) {
  constexpr int a = 128 / (int)sizeof(c2<T>);
  return ((n / 2 + 1) + a - 1) / a * a;
}

// Flat spectral indices are below 2^31 (one half spectrum, padding
// included, up to 1024^3), so they decode in 32-bit arithmetic; n is a
// power of two.
struct RowIdx { unsigned row; int kx; };
__device__ __forceinline__ RowIdx row_split(long long i, int pp) {
  const unsigned ii = (unsigned)i, row = ii / (unsigned)pp;
  return RowIdx{row, (int)(ii - row * (unsigned)pp)};
}
__device__ __forceinline__ int row_ky(unsigned row, int n) { return (int)(row & (unsigned)(n - 1)); }
__device__ __forceinline__ int row_kz(unsigned row, int n) { return (int)(row >> (__ffs(n) - 1)); }

// Decode a flat half-spectrum index (pitch pp) into mode indices; false on
// padding.  m_y, m_z in (-n/2, n/2] (Nyquist counted +n/2).
struct KIdx { int kx, ky, kz, mx, my, mz; };
__device__ __forceinline__ bool kdecode(long long i, int n, int pp, KIdx& k) {
  const RowIdx r = row_split(i, pp);
  k.kx = r.kx;
  if (k.kx > (n >> 1)) return false;
  k.ky = row_ky(r.row, n);
  k.kz = row_kz(r.row, n);
  const int h = n >> 1;
  k.mx = k.kx;
  k.my = k.ky <= h ? k.ky : k.ky - n;
  k.mz = k.kz <= h ? k.kz : k.kz - n;
  return true;
}

// Band-limited (pruned) transforms: with keep >= 0 every spectrum of the
// launch is zero outside the cube max_j |m_j| <= keep (the 2/3 dealias mask),
// either because the forward pass's spectral factor masks it or because the
// inverse's input is already masked.  The passes then skip the HBM traffic of
// the modes outside: the forward x pass writes only kx <= keep, the forward y
// pass transforms only those columns and writes only the kept ky rows, the
// inverse passes read only kept modes and treat the rest as zero.  keep < 0:
// full transforms.
template <typename T> struct XArgs {
  const T* rsrc[MAXB];
  T* rdst[MAXB];
  c2<T>* spec;        // batch-contiguous spectra, stride sstride
  const double2* tw;  // exp(-2 pi i k/N), FP64 for every field precision
  long long sstride;
  int nb;
  int keep = -1;
  double scale;
  // xmap: output b reads scratch spectrum xin[b] (else b) and, with xik[b],
  // multiplies it by i K_x first (the x part of a factored derivative)
  int xmap = 0;
  int xin[MAXB];
  int xik[MAXB];
  // forward x pass (k_xfwd): a non-null rmul[b] makes input b the pointwise
  // product rsrc[b] * rmul[b], formed in T as a pointwise pass would store it
  const T* rmul[MAXB] = {};
};

template <typename T> struct SArgs {
  c2<T>* spec;        // scratch spectra, batch-contiguous
  const c2<T>* src[MAXB][3];
  double coef[MAXB][3];
  c2<T>* dst[MAXB];
  SpecOp op[MAXB];
  int nsrc[MAXB];
  const double2* tw;
  long long sstride;
  int keep = -1;
  int remap = 0;   // 1: XCD-contiguous tile order (xcd_tile)
  // ymap (y passes): field f of the launch reads scratch spectrum yin[f],
  // writes yout[f] and, with yik[f], multiplies its input by i K_y (the y
  // part of a factored derivative); else in place on spectrum f
  int ymap = 0;
  int yin[MAXB];
  int yout[MAXB];
  int yik[MAXB];
  // forward z pass of a pruned launch: 1 stores the zeros of the masked
  // spectrum outside the kept cube, 0 stores the kept cube only (for outputs
  // whose every reader stays inside it)
  int kzero = 1;
  // z passes of a pruned launch that touch kept columns only (set by the
  // launcher): the grid covers the kept tiles alone, 2 keep + 1 kept ky rows
  // times ceil((keep + 1) / C) column groups
  int kcomp = 0;
};

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  Remap the linear workgroup id so that XCD k runs a contiguous range of
// logical tiles (in order): tiles that share 128-byte lines (adjacent column
// groups) then go through one L2 instead of being fetched once per XCD.
constexpr unsigned NUM_XCD = 8;
__device__ __forceinline__ void xcd_tile(bool on, unsigned& bx, unsigned& by) {
  const unsigned nx = gridDim.x, total = nx * gridDim.y;
  const unsigned pid = blockIdx.y * nx + blockIdx.x;
  const unsigned lid =
      (!on || (total % NUM_XCD)) ? pid : (pid % NUM_XCD) * (total / NUM_XCD) + pid / NUM_XCD;
  bx = lid % nx;
  by = lid / nx;
}

// |m| of FFT index k on an n-point axis (Nyquist counted +n/2)
__device__ __forceinline__ int mabs(int k, int n) { return k <= (n >> 1) ? k : n - k; }
// FFT index n in [0, N) lies outside the band |m| <= keep, for 0 <= keep < N/2:
// keep < n < N - keep, as one unsigned compare (a per-element mabs test costs
// the strided passes registers)
__device__ __forceinline__ bool out_band(int n, int keep, int N) {
  return (unsigned)(n - keep - 1) < (unsigned)(N - 2 * keep - 1);
}

// Spectral factor of one half-spectrum mode (see SpecOp in mlq_common.hpp).
__device__ __forceinline__ double2 spec_factor(const SpecOp& o, int kx, int ky, int kz, int n) {
  const int h = n >> 1;
  const int mx = kx, my = ky <= h ? ky : ky - n, mz = kz <= h ? kz : kz - n;
  double2 f = make_double2(o.scale, 0.0);
  const int fl = o.flags;
  if (fl == 0) return f;
  if (fl & (SP_MASK | SP_HSCALE)) {
    const int mm = max(abs(mx), max(abs(my), abs(mz)));
    if ((fl & SP_MASK) && mm > o.mcut) return make_double2(0.0, 0.0);
    if ((fl & SP_HSCALE) && mm > o.mcut) f.x *= o.hs;
  }
  if ((fl & SP_ZEROK0) && mx == 0 && my == 0 && mz == 0) return make_double2(0.0, 0.0);
  const double K1 = TWO_PI * mx, K2 = TWO_PI * my, K3 = TWO_PI * mz;
  const double ksq = K1 * K1 + K2 * K2 + K3 * K3;
  if (fl & SP_NEGKSQ) f.x *= -ksq;
  if (fl & SP_INVKSQ) {
    if (ksq == 0.0) return make_double2(0.0, 0.0);
    f.x *= -1.0 / ksq;
  }
  if (fl & SP_IK) {
    const int md = o.dim == 0 ? mx : (o.dim == 1 ? my : mz);
    if (md == h) return make_double2(0.0, 0.0);  // own-dimension Nyquist
    const double Kd = TWO_PI * md;
    f = make_double2(0.0, f.x * Kd);
  }
  return f;
}

// The spectral factor along one z column (fixed kx, ky): the same values as
// spec_factor for every op whose i K part (if any) is along z, with the
// column's terms precomputed (cz: the column's part of the mask and |K|^2).
struct ColFactor {
  int fl, mcut, mxy;   // mxy = max(|m_x|, |m_y|)
  double s, kxy, hs;   // scale, K_x^2 + K_y^2, SP_HSCALE factor
};
__device__ __forceinline__ bool col_factor_ok(const SpecOp& o) {
  return !(o.flags & SP_IK) || o.dim == 2;
}
__device__ __forceinline__ ColFactor col_factor_init(const SpecOp& o, double scale, int kx, int ky,
                                                     int n) {
  const int h = n >> 1, my = ky <= h ? ky : ky - n;
  const double K1 = TWO_PI * kx, K2 = TWO_PI * my;
  return ColFactor{o.flags, o.mcut, max(kx, abs(my)), scale, K1 * K1 + K2 * K2, o.hs};
}
__device__ __forceinline__ double2 col_factor(const ColFactor& c, int kz, int n) {
  const int h = n >> 1, mz = kz <= h ? kz : kz - n;
  const int fl = c.fl;
  double f = c.s;
  if ((fl & SP_MASK) && max(c.mxy, abs(mz)) > c.mcut) return make_double2(0.0, 0.0);
  if ((fl & SP_HSCALE) && max(c.mxy, abs(mz)) > c.mcut) f *= c.hs;
  const double Kz = TWO_PI * mz, ksq = c.kxy + Kz * Kz;
  if ((fl & SP_ZEROK0) && ksq == 0.0) return make_double2(0.0, 0.0);
  if (fl & SP_NEGKSQ) f *= -ksq;
  if (fl & SP_INVKSQ) {
    if (ksq == 0.0) return make_double2(0.0, 0.0);
    f *= -1.0 / ksq;
  }
  if (fl & SP_IK) return make_double2(0.0, mz == h ? 0.0 : f * Kz);   // i K_z
  return make_double2(f, 0.0);
}

template <typename T> __device__ __forceinline__ c2<T> apply_factor(c2<T> v, double2 f) {
  const T fx = (T)f.x, fy = (T)f.y;
  return cmk<T>(v.x * fx - v.y * fy, v.x * fy + v.y * fx);
}

// a, or zero when !in (component selects: a select of whole array elements
// can make the compiler keep the array in scratch memory)
template <typename T> __device__ __forceinline__ c2<T> keep_or_zero(bool in, c2<T> a) {
  return cmk<T>(in ? a.x : (T)0, in ? a.y : (T)0);
}

// complex value from lane src of a width-wide lane group
__device__ __forceinline__ float2 shfl_c2(float2 v, int src, int width) {
  return cmk<float>(__shfl(v.x, src, width), __shfl(v.y, src, width));
}
__device__ __forceinline__ double2 shfl_c2(double2 v, int src, int width) {
  return cmk<double>(__shfl(v.x, src, width), __shfl(v.y, src, width));
}

// The field a transform of a row-view launch belongs to.  When every wave
// holds whole groups of consecutive row pairs of one field (the transforms per
// wave divide both the transforms per workgroup and the row pairs per field),
// the index is wave-uniform and is moved to a scalar register, so the
// per-field kernel arguments indexed by it load through the scalar unit
// instead of as a vector load the wave has to wait for.
template <int N, int NF> __device__ __forceinline__ int wave_field(int b) {
  constexpr int per_wave = 64 / FftShape<N>::Tf;
  if constexpr (NF % per_wave == 0 && ((long long)N * N / 2) % per_wave == 0)
    return __builtin_amdgcn_readfirstlane(b);
  else
    return b;
}

// Register type of the x-pass transforms: the field precision, so that an
// FP32 transform's first radix pass (no twiddles) runs in FP32 as in the
// column passes; fft_pass keeps every later pass FP64 in registers.
#ifndef MLQ_XP0_F32
#define MLQ_XP0_F32 1
#endif
template <typename T>
using XReg = typename std::conditional<MLQ_XP0_F32 != 0, T, double>::type;

// ---------------------------------------------------------------------------
// x pass, forward: two real rows packed as one complex row (a + i b), one
// N-point FFT, split into the two half spectra.
// ---------------------------------------------------------------------------
// Transform of the packed pair in v (row view L of the thread's row pair) and
// the split into the half spectra at d (row0) and d + PX (row0 + 1); only
// kx <= kmax is stored.
template <typename T, int N, bool ZB = false>
__device__ __forceinline__ void xfwd_pair(c2<XReg<T>> (&v)[FftShape<N>::E], int t,
                                          const double2* __restrict__ tw, const LdsRow<T>& L,
                                          c2<T>* d, int kmax) {
  constexpr int E = FftShape<N>::E, Tf = FftShape<N>::Tf, PX = spec_pitch<T>(N);
  // ZB (kmax < 3N/8): entries 3N/8 .. 5N/8 - 1 of the packed row are never
  // read by the split below
  fft_core<XReg<T>, N, E, -1, LdsRow<T>, ZB>(v, t, TwPow{tw}, L);
  static_assert(Tf <= 64 && (64 % Tf) == 0, "one transform lies within one wave");
  // The split needs Z[k] and Z[N - k] (rounded to the field precision, as an
  // LDS exchange would).  For k = t + m Tf, m < E/2, element N - k sits in
  // slot E - 1 - m of lane Tf - t of the same transform (t >= 1) or in the
  // thread's own slot (E - m) mod E (t = 0): a lane exchange inside the wave.
  const int p = (Tf - t) & (Tf - 1);
  c2<T> z[E];
#pragma unroll
  for (int m = 0; m < E; ++m) z[m] = cmk<T>((T)v[m].x, (T)v[m].y);
#pragma unroll
  for (int m = 0; m < E / 2; ++m) {
    const c2<T> sh = shfl_c2(z[E - 1 - m], p, Tf);
    const int k = t + m * Tf;
    if (k > kmax) continue;
    // in the field precision: the sum rounds once and the halving is exact,
    // as the rounding of the sum formed in FP64
    const c2<T> zk = z[m], zn = t == 0 ? z[(E - m) % E] : sh;
    gst<nt_st<PC_XF>()>(d + k, cmk<T>((T)0.5 * (zk.x + zn.x), (T)0.5 * (zk.y - zn.y)));
    gst<nt_st<PC_XF>()>(d + PX + k, cmk<T>((T)0.5 * (zk.y + zn.y), (T)-0.5 * (zk.x - zn.x)));
  }
  if (t == 0 && kmax == N / 2) {
    gst<nt_st<PC_XF>()>(d + N / 2, cmk<T>(z[E / 2].x, (T)0));
    gst<nt_st<PC_XF>()>(d + PX + N / 2, cmk<T>(z[E / 2].y, (T)0));
  }
}

template <typename T, int N, int NF, bool PR, bool ZB = false>
__global__ __launch_bounds__(NF* FftShape<N>::Tf) void k_xfwd(XArgs<T> a) {
  constexpr int E = FftShape<N>::E, Tf = FftShape<N>::Tf, PX = spec_pitch<T>(N);
  constexpr int LP = LdsRow<T>::pitch(N);
  __shared__ c2<T> sm[NF * LP];
  const int f = threadIdx.x / Tf, t = threadIdx.x % Tf;
  const long long pairs_per = (long long)N * N / 2;
  const long long gp = (long long)blockIdx.x * NF + f;
  const int b = wave_field<N, NF>((int)(gp / pairs_per));
  const long long lp = gp - (long long)b * pairs_per;
  const long long row0 = 2 * lp;
  const T* s = a.rsrc[b] + row0 * N;
  const T* mu = a.rmul[b];
  c2<XReg<T>> v[E];
  if (mu) {   // uniform within the field's waves
    mu += row0 * N;
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const T p0 = gld<nt_ld<PC_XF>()>(s + t + m * Tf) * gld<nt_ld<PC_XF>()>(mu + t + m * Tf);
      const T p1 =
          gld<nt_ld<PC_XF>()>(s + N + t + m * Tf) * gld<nt_ld<PC_XF>()>(mu + N + t + m * Tf);
      v[m] = cmk<XReg<T>>((XReg<T>)p0, (XReg<T>)p1);
    }
  } else {
#pragma unroll
    for (int m = 0; m < E; ++m)
      v[m] = cmk<XReg<T>>((XReg<T>)gld<nt_ld<PC_XF>()>(s + t + m * Tf),
                          (XReg<T>)gld<nt_ld<PC_XF>()>(s + N + t + m * Tf));
  }
  LdsRow<T> L{sm + f * LP};
  const int kmax = PR ? a.keep : N / 2;   // PR: pruned launch (a.keep >= 0)
  xfwd_pair<T, N, ZB>(v, t, a.tw, L, a.spec + b * a.sstride + row0 * PX, kmax);
}

// Forward x pass of NB fields produced in the kernel (one row pair per
// workgroup): P computes the NB values of each of the pair's 2N points from
// its inputs (the values of one point share their loads), stages them in the
// LDS rows, and the NB transforms run on the staged rows.  Replaces a
// pointwise kernel that writes NB fields plus the x pass that reads them back.
template <class P, typename T, int N>
constexpr bool xfwd_p_fits() {
  return (long long)P::NB * LdsRow<T>::pitch(N) * (long long)sizeof(c2<T>) <= 160 * 1024 &&
         P::NB * FftShape<N>::Tf <= 1024;
}
template <typename T, int N, bool PR, class P, bool ZB = false>
__global__ __launch_bounds__(P::NB* FftShape<N>::Tf) void k_xfwd_p(P prod, XArgs<T> a) {
  constexpr int NB = P::NB, E = FftShape<N>::E, Tf = FftShape<N>::Tf, PX = spec_pitch<T>(N);
  constexpr int LP = LdsRow<T>::pitch(N);
  static_assert(2 * N <= 2 * LP, "a row pair's 2N reals fit its LDS row");
  __shared__ c2<T> sm[NB * LP];
  T* stg = reinterpret_cast<T*>(sm);   // field f: 2N reals at stg + 2 LP f
  const long long row0 = 2 * (long long)blockIdx.x;
  for (int p = threadIdx.x; p < 2 * N; p += NB * Tf) {
    T o[NB];
    prod(row0 * N + p, o);
#pragma unroll
    for (int f = 0; f < NB; ++f) stg[2 * LP * f + p] = o[f];
  }
  __syncthreads();
  const int f = threadIdx.x / Tf, t = threadIdx.x % Tf;
  const T* s = stg + 2 * LP * f;
  c2<XReg<T>> v[E];
#pragma unroll
  for (int m = 0; m < E; ++m)
    v[m] = cmk<XReg<T>>((XReg<T>)s[t + m * Tf], (XReg<T>)s[N + t + m * Tf]);
  __syncthreads();   // the transform rewrites the staged rows
  LdsRow<T> L{sm + f * LP};
  const int kmax = PR ? a.keep : N / 2;
  xfwd_pair<T, N, ZB>(v, t, a.tw, L, a.spec + f * a.sstride + row0 * PX, kmax);
}

// x pass, inverse (C2R) of one row pair, rows row0 and row0 + 1 of the
// spectrum whose row row0 starts at sp: rebuild the packed complex row from
// the two half spectra (Hermitian extension, real parts taken at kx = 0 and
// kx = N/2) and inverse FFT; v[m] = (row0, row0 + 1) at x = t + m Tf,
// unscaled.  xik: multiply by i K_x first; kmax: modes kx > kmax are zero.
template <typename T, int N, bool PR, bool ZB>
__device__ __forceinline__ void xinv_pair(const c2<T>* __restrict__ sp, int kmax, bool xik, int t,
                                          const double2* tw, const LdsRow<T>& L,
                                          c2<XReg<T>> (&v)[FftShape<N>::E]) {
  using R = XReg<T>;
  constexpr int E = FftShape<N>::E, Tf = FftShape<N>::Tf, PX = spec_pitch<T>(N);
  static_assert(Tf <= 64 && (64 % Tf) == 0, "one transform lies within one wave");
  // Each mode is loaded once, by the thread that owns kx = t + m Tf (m < E/2)
  // in the packed row; slot E/2 of thread 0 holds kx = N/2.  Modes beyond the
  // kept cube are zero: their loads are redirected to kx = 0 and discarded.
  // Every load of the thread is issued before any of them is used (one
  // memory round trip per transform, no wait inside the load sequence).
  c2<T> A[E / 2 + 1], B[E / 2 + 1];
#pragma unroll
  for (int m = 0; m < E / 2; ++m) {
    const int k = t + m * Tf;
    const int ks = k <= kmax ? k : 0;
    if (!PR || m * Tf <= kmax) {   // uniform: a slot run wholly past kmax loads nothing
      A[m] = gld<nt_ld<PC_XI>()>(sp + ks);
      B[m] = gld<nt_ld<PC_XI>()>(sp + PX + ks);
    } else {
      A[m] = cmk<T>((T)0, (T)0);
      B[m] = A[m];
    }
  }
  // Nyquist kx = N/2 (zero under i K_x, the REAL(ifft) rule), used by thread
  // 0 only; the other lanes load kx = 0 again, the line of their slot 0
  const bool nyq = t == 0 && N / 2 <= kmax && !xik;
  A[E / 2] = gld<nt_ld<PC_XI>()>(sp + (nyq ? N / 2 : 0));
  B[E / 2] = gld<nt_ld<PC_XI>()>(sp + PX + (nyq ? N / 2 : 0));
  A[E / 2] = keep_or_zero<T>(nyq, A[E / 2]);
  B[E / 2] = keep_or_zero<T>(nyq, B[E / 2]);
#pragma unroll
  for (int m = 0; m < E / 2; ++m) {
    const int k = t + m * Tf;
    if (PR) {
      A[m] = keep_or_zero<T>(k <= kmax, A[m]);
      B[m] = keep_or_zero<T>(k <= kmax, B[m]);
    }
    if (xik) {   // i K_x, K_x = 2 pi k (k < N/2 here)
      const double K = TWO_PI * k;
      A[m] = cmk<T>((T)(-K * (double)A[m].y), (T)(K * (double)A[m].x));
      B[m] = cmk<T>((T)(-K * (double)B[m].y), (T)(K * (double)B[m].x));
    }
  }
  // The Hermitian extension k = t + m Tf > N/2 (m >= E/2) needs mode N - k.
  // For t >= 1 that is kx = (Tf - t) + (E - 1 - m) Tf, owned by lane Tf - t
  // of the same transform (a lane exchange inside the wave, no LDS round
  // trip); for t = 0 it is kx = (E - m) Tf, the thread's own slot E - m.
  const int p = (Tf - t) & (Tf - 1);
#pragma unroll
  for (int m = 0; m < E; ++m) {
    c2<T> As, Bs;
    if (m < E / 2) {
      As = A[m];
      Bs = B[m];
    } else {
      const c2<T> ash = shfl_c2(A[E - 1 - m], p, Tf), bsh = shfl_c2(B[E - 1 - m], p, Tf);
      As = t == 0 ? A[E - m] : ash;
      Bs = t == 0 ? B[E - m] : bsh;
    }
    const int k = t + m * Tf;
    R ax = As.x, ay = As.y, bx = Bs.x, by = Bs.y;
    if (k == 0 || k == N / 2) { ay = (R)0; by = (R)0; }
    if (k > N / 2) { ay = -ay; by = -by; }
    v[m] = cmk<R>(ax - by, ay + bx);
  }
  // ZB (pruned launches with keep < 3N/8): elements 3N/8 .. 5N/8 - 1 of the
  // packed row are exact zeros
  fft_core<R, N, E, +1, LdsRow<T>, ZB>(v, t, TwPow{tw}, L);
}

// x pass, inverse: one row pair per transform (xinv_pair), scaled and stored.
template <typename T, int N, int NF, bool PR, bool ZB = false>
__global__ __launch_bounds__(NF* FftShape<N>::Tf) void k_xinv(XArgs<T> a) {
  constexpr int E = FftShape<N>::E, Tf = FftShape<N>::Tf, PX = spec_pitch<T>(N);
  constexpr int LP = LdsRow<T>::pitch(N);
  __shared__ c2<T> sm[NF * LP];
  const int f = threadIdx.x / Tf, t = threadIdx.x % Tf;
  const long long pairs_per = (long long)N * N / 2;
  const long long gp = (long long)blockIdx.x * NF + f;
  const int b = wave_field<N, NF>((int)(gp / pairs_per));
  const long long lp = gp - (long long)b * pairs_per;
  const long long row0 = 2 * lp;
  const int slot = a.xmap ? a.xin[b] : b;
  const bool xik = a.xmap && a.xik[b];
  const int kmax = PR ? a.keep : N / 2;   // PR: pruned launch (a.keep >= 0)
  c2<XReg<T>> v[E];
  xinv_pair<T, N, PR, ZB>(a.spec + slot * a.sstride + row0 * PX, kmax, xik, t, a.tw,
                          LdsRow<T>{sm + f * LP}, v);
  T* o = a.rdst[b] + row0 * N;
  const XReg<T> sc = (XReg<T>)a.scale;   // 1/N^3, a power of two: exact in either precision
#pragma unroll
  for (int m = 0; m < E; ++m) {
    gst<nt_st<PC_XI>()>(o + t + m * Tf, (T)(v[m].x * sc));
    gst<nt_st<PC_XI>()>(o + N + t + m * Tf, (T)(v[m].y * sc));
  }
}

// Inverse x pass, pointwise producer and forward x pass of one row pair in
// one workgroup (the x direction is local to a row).  Wave f (f < P::NI)
// inverse-transforms input f (ai, as k_xinv: the spectra the inverse y pass
// left in the scratch slots), stores it to ai.rdst[f] when that is set and
// stages it in its own LDS row; P computes the NB output values of each of
// the pair's 2N points from the staged inputs and its own loads, staged in
// place; wave f then transforms output f forward into scratch slot f (as
// k_xfwd_p).  Every read of the pair's input rows precedes the barrier after
// the inverse phase, every write of its output rows follows it, and no other
// workgroup touches these rows, so inputs and outputs may share scratch
// slots.  Replaces an inverse x pass that writes the NI fields and a producer
// pass that reads them back.
template <class P, typename T, int N>
constexpr bool xinv_fwd_p_fits() {
  return P::NI <= P::NB && xfwd_p_fits<P, T, N>();
}
template <typename T, int N, class P>
__global__ __launch_bounds__(P::NB* FftShape<N>::Tf) void k_xinv_fwd_p(P prod, XArgs<T> ai,
                                                                      XArgs<T> ao) {
  constexpr int NB = P::NB, NI = P::NI, E = FftShape<N>::E, Tf = FftShape<N>::Tf;
  constexpr int PX = spec_pitch<T>(N), LP = LdsRow<T>::pitch(N);
  static_assert(NI <= NB && 2 * N <= 2 * LP, "staged rows fit the LDS rows");
  __shared__ c2<T> sm[NB * LP];
  T* stg = reinterpret_cast<T*>(sm);   // field f: 2N reals at stg + 2 LP f
  const long long row0 = 2 * (long long)blockIdx.x;
  const int f = threadIdx.x / Tf, t = threadIdx.x % Tf;
  const LdsRow<T> L{sm + f * LP};
  if (f < NI) {   // uniform within each transform's lanes
    c2<XReg<T>> v[E];
    xinv_pair<T, N, true, true>(ai.spec + ai.xin[f] * ai.sstride + row0 * PX, ai.keep,
                                ai.xik[f] != 0, t, ai.tw, L, v);
    // the transform's last LDS reads precede its final wave barrier: its own
    // row is free for the staged values
    T* o = ai.rdst[f];
    T* s = stg + 2 * LP * f;
    const XReg<T> sc = (XReg<T>)ai.scale;   // 1/N^3, a power of two
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const T x0 = (T)(v[m].x * sc), x1 = (T)(v[m].y * sc);
      s[t + m * Tf] = x0;
      s[N + t + m * Tf] = x1;
      if (o) {
        gst<nt_st<PC_XI>()>(o + row0 * N + t + m * Tf, x0);
        gst<nt_st<PC_XI>()>(o + row0 * N + N + t + m * Tf, x1);
      }
    }
  }
  __syncthreads();
  for (int p = threadIdx.x; p < 2 * N; p += NB * Tf) {
    T o[NB];
    prod(row0 * N + p, stg + p, 2 * LP, o);   // staged input g at stg[p + 2 LP g]
#pragma unroll
    for (int g = 0; g < NB; ++g) stg[2 * LP * g + p] = o[g];
  }
  __syncthreads();
  const T* s = stg + 2 * LP * f;
  c2<XReg<T>> v[E];
#pragma unroll
  for (int m = 0; m < E; ++m)
    v[m] = cmk<XReg<T>>((XReg<T>)s[t + m * Tf], (XReg<T>)s[N + t + m * Tf]);
  __syncthreads();   // the transform rewrites the staged rows
  xfwd_pair<T, N, true>(v, t, ao.tw, L, ao.spec + f * ao.sstride + row0 * PX, ao.keep);
}

// ---------------------------------------------------------------------------
// Strided pass along y or z: C adjacent columns per workgroup, each column a
// length-N transform.  MODE 0: plain in-place pass on the scratch spectra
// (y pass).  MODE 1: inverse z pass reading sum_i coef_i * src_i times the
// spectral factor (prologue), writing scratch.  MODE 2: forward z pass
// reading scratch, writing factor * result to dst (epilogue).
// ---------------------------------------------------------------------------
// FP32 inverse z pass at E = 8: compiled for 8 waves per SIMD (at most 64
// VGPRs; the full-transform variant spills 12 bytes, the band-limited one
// fits) instead of 72 VGPRs, which at 16 columns (1024 threads) let only one
// workgroup per CU run; measured 0.4 ms faster per mixed iteration
#ifndef MLQ_ZI_WAVES
#define MLQ_ZI_WAVES 8
#endif
// ZB (band-limited launches with keep < 3N/8): an inverse pass's rows
// 3N/8 .. 5N/8 - 1 are zero on entry and the first radix-8 pass skips them; a
// forward pass never stores those rows (the last pass leaves them zero).
template <typename T, int N, int E, int C, int DIR, int MODE, bool ZB = false>
__global__ __launch_bounds__(C*(N / E))
__attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 && MODE == 1 && E == 8 ? MLQ_ZI_WAVES : 1)))
void k_strided(SArgs<T> a, int ncols) {
  constexpr int Tf = N / E, PP = spec_pitch<T>(N);
  constexpr bool NTL = nt_ld<strided_class<MODE, DIR>()>(), NTS = nt_st<strided_class<MODE, DIR>()>();
  // element and plane strides are fixed by the pass (y: rows of one z plane;
  // z: planes of one spectrum); element offsets within a column fit 32 bits
  constexpr int estride = MODE == 0 ? PP : N * PP;
  constexpr long long pstride = MODE == 0 ? (long long)N * PP : (long long)N * N * PP;
  __shared__ c2<T> sm[N * C];
  __shared__ double2 smtw[N];   // twiddle table (read from the second radix pass on)
  const int c = threadIdx.x % C, t = threadIdx.x / C;
  TwStage<N, C * Tf> tws;
  tws.load(a.tw);
  unsigned bx, by;
  xcd_tile(a.remap != 0, bx, by);
  const int keep = a.keep;
  int col = (int)bx * C + c;
  if (MODE != 0 && a.kcomp) {
    const unsigned nt = (unsigned)(keep + C) / C, jy = bx / nt, jt = bx - jy * nt;
    const int kyt = (int)jy <= keep ? (int)jy : (int)jy + N - 2 * keep - 1, kxt = (int)jt * C + c;
    col = kxt < PP ? kyt * PP + kxt : ncols;   // a lane past the row pitch is invalid
  }
  bool valid = col < ncols;
  const int plane = (int)by;
  c2<T> v[E];
  int kx = 0, ky = 0;
  // kept: the column carries modes inside the kept cube (z passes); rowin(n):
  // element n of the column is read (inverse passes skip masked rows)
  bool kept = true;
  if (MODE != 0) {
    ky = col / PP;
    kx = col - ky * PP;
    valid = valid && kx <= N / 2;   // padding columns carry no data
    kept = keep < 0 || (kx <= keep && mabs(ky, N) <= keep);
  }
  const bool ld_ok = valid && kept;
  // z passes of a pruned launch: a tile with no kept column skips its
  // transform (block-uniform test).  The inverse pass leaves such columns
  // unwritten (the pruned y pass reads kept columns only); the forward pass
  // stores the zeros of the masked spectrum.
  if (MODE != 0 && keep >= 0 && !__syncthreads_or(ld_ok)) {
    if (MODE == 2 && valid && a.kzero) {
      c2<T>* d = a.dst[plane] + col;
#pragma unroll
      for (int m = 0; m < E; ++m) gst<NTS>(d + (t + m * Tf) * estride, cmk<T>((T)0, (T)0));
    }
    return;
  }
  const bool band = keep >= 0 && 2 * keep < N;   // else no row is outside
  auto rowin = [&](int n) { return DIR < 0 || !band || !out_band(n, keep, N); };
  // rows outside the kept cube: the load is redirected to row 0 of the same
  // column (always kept, in bounds) and its value replaced by zero, so the E
  // loads stay unconditional (no per-element branch)
  auto ldrow = [&](const c2<T>* base, int n) {
    const bool r = rowin(n);
    const c2<T> x = gld<NTL>(base + (r ? n : 0) * estride);
    return r ? x : cmk<T>((T)0, (T)0);
  };
  if (MODE == 1) {
    const int bb = plane;
    const SpecOp op = a.op[bb];
    const int ns = a.nsrc[bb];
    const c2<T>* s0 = a.src[bb][0] + col;
    if (ld_ok) {
#pragma unroll
      for (int m = 0; m < E; ++m) v[m] = ldrow(s0, t + m * Tf);
    } else {
#pragma unroll
      for (int m = 0; m < E; ++m) v[m] = cmk<T>((T)0, (T)0);
    }
    // c0 stays pending until the first combination (one rounding per add) or
    // is folded into the spectral factor
    double c0 = a.coef[bb][0];
    for (int i = 1; i < ns; ++i) {
      const c2<T>* si = a.src[bb][i] + col;
      const double ci = a.coef[bb][i];
      // half the column at a time: loading all E elements of the source
      // while v is live cost the kernel 16 registers (5 instead of 7 waves
      // per SIMD at E = 8)
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        c2<T> ld[E / 2];
#pragma unroll
        for (int mm = 0; mm < E / 2; ++mm)
          ld[mm] = ld_ok ? ldrow(si, t + (h * E / 2 + mm) * Tf) : cmk<T>((T)0, (T)0);
#pragma unroll
        for (int mm = 0; mm < E / 2; ++mm) {
          const int m = h * E / 2 + mm;
          v[m] = cmk<T>((T)(c0 * (double)v[m].x + ci * (double)ld[mm].x),
                        (T)(c0 * (double)v[m].y + ci * (double)ld[mm].y));
        }
        __builtin_amdgcn_sched_barrier(0);
      }
      c0 = 1.0;
    }
    // Spectral factor (uniform branch: op and c0 are per plane).  Factors
    // whose i K part (if any) is along z are applied in registers from the
    // column's precomputed terms.  Others (i K_x or i K_y, unfactored): the
    // tile goes through the thread's own LDS slots so that the factor code
    // runs in a rolled loop, one element at a time, and does not multiply the
    // register footprint by E; the barrier orders these slot accesses before
    // the first FFT pass writes other threads' slots.
    const double fs = c0 * op.scale;
    if ((op.flags & ~SP_MASK) == 0) {
      // the mask and a scale: a real factor, fs or 0, per element
      if (op.flags != 0 || fs != 1.0) {
        const int mxy = max(kx, mabs(ky, N));
#pragma unroll
        for (int m = 0; m < E; ++m) {
          const bool in = !(op.flags & SP_MASK) || max(mxy, mabs(t + m * Tf, N)) <= op.mcut;
          v[m] = in ? (fs == 1.0 ? v[m]
                                 : cmk<T>((T)((double)v[m].x * fs), (T)((double)v[m].y * fs)))
                    : cmk<T>((T)0, (T)0);
        }
      }
    } else if (col_factor_ok(op)) {
      {
        const ColFactor cf = col_factor_init(op, fs, kx, ky, N);
#pragma unroll
        for (int m = 0; m < E; ++m) {
          const double2 f = col_factor(cf, t + m * Tf, N);
          const c2<double> r = cmul(ccvt<double>(v[m]), f);
          v[m] = cmk<T>((T)r.x, (T)r.y);
        }
      }
    } else {
#pragma unroll
      for (int m = 0; m < E; ++m) sm[(t + m * Tf) * C + c] = v[m];
#pragma unroll 1
      for (int m = 0; m < E; ++m) {
        const int kz = t + m * Tf;
        const double2 f = spec_factor(op, kx, ky, kz, N);
        const c2<double> r = cmul(ccvt<double>(sm[kz * C + c]), make_double2(c0 * f.x, c0 * f.y));
        sm[kz * C + c] = cmk<T>((T)r.x, (T)r.y);
      }
#pragma unroll
      for (int m = 0; m < E; ++m) v[m] = sm[(t + m * Tf) * C + c];
      __syncthreads();
    }
  } else {
    const c2<T>* s = a.spec + plane * pstride + col;
    bool yik = false;
    if (MODE == 0 && a.ymap) {
      const int f = plane / N, z = plane - f * N;
      s = a.spec + a.yin[f] * a.sstride + z * pstride + col;
      yik = a.yik[f] != 0;
    }
    if (ld_ok) {
#pragma unroll
      for (int m = 0; m < E; ++m) v[m] = ldrow(s, t + m * Tf);
    } else {
#pragma unroll
      for (int m = 0; m < E; ++m) v[m] = cmk<T>((T)0, (T)0);
    }
    if (yik) {   // i K_y, zero at the y Nyquist (REAL(ifft) rule)
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int n = t + m * Tf;
        const double K = n == N / 2 ? 0.0 : TWO_PI * (n < N / 2 ? n : n - N);
        v[m] = cmk<T>((T)(-K * (double)v[m].y), (T)(K * (double)v[m].x));
      }
    }
  }
  LdsCol<T, C> L{sm, c};
  tws.store(smtw);   // ordered before the first twiddle read by the first pass's barrier
  fft_core<T, N, E, DIR, LdsCol<T, C>, ZB>(v, t, TwTab{smtw}, L);
  if (!valid) return;
  if (MODE == 1 && !kept) return;   // the pruned inverse y pass never reads it
  if (MODE == 2) {
    const int bb = plane;
    const SpecOp op = a.op[bb];
    c2<T>* d = a.dst[bb] + col;
    // kzero = 0 (a pruned launch whose readers stay inside the kept cube):
    // nothing outside the cube is stored
    const bool zout = a.kzero || keep < 0;
    if (!kept) {   // outside the kept cube: the masked spectrum is zero
      if (zout) {
#pragma unroll
        for (int m = 0; m < E; ++m) gst<NTS>(d + (t + m * Tf) * estride, cmk<T>((T)0, (T)0));
      }
    } else if ((op.flags & ~SP_MASK) == 0) {
      // the mask and a scale: a real factor, op.scale or 0, per element (a
      // plain transform stores its values as they are)
      const int mxy = max(kx, mabs(ky, N));
      const double fs = op.scale;
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int n = t + m * Tf;
        if (!zout && band && out_band(n, keep, N)) continue;
        const bool in = !(op.flags & SP_MASK) || max(mxy, mabs(n, N)) <= op.mcut;
        const c2<T> r = in ? (fs == 1.0 ? v[m]
                                        : cmk<T>((T)((double)v[m].x * fs), (T)((double)v[m].y * fs)))
                           : cmk<T>((T)0, (T)0);
        gst<NTS>(d + n * estride, r);
      }
    } else if (col_factor_ok(op)) {
      // closed form along the column, in registers
      const ColFactor cf = col_factor_init(op, op.scale, kx, ky, N);
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int n = t + m * Tf;
        if (!zout && band && out_band(n, keep, N)) continue;
        const c2<double> p = cmul(ccvt<double>(v[m]), col_factor(cf, n, N));
        gst<NTS>(d + n * estride, cmk<T>((T)p.x, (T)p.y));
      }
    } else {
      // rolled epilogue through the thread's own LDS slots (every read of the
      // last FFT pass precedes its barrier, so the slots are free)
#pragma unroll
      for (int m = 0; m < E; ++m) sm[(t + m * Tf) * C + c] = v[m];
#pragma unroll 1
      for (int m = 0; m < E; ++m) {
        const int kz = t + m * Tf;
        const c2<double> p = cmul(ccvt<double>(sm[kz * C + c]), spec_factor(op, kx, ky, kz, N));
        gst<NTS>(d + kz * estride, cmk<T>((T)p.x, (T)p.y));
      }
    }
  } else {
    // y passes and the inverse z pass: the forward y pass stores only the
    // kept ky rows (the pruned z pass reads nothing else)
    c2<T>* d = a.spec + plane * pstride + col;
    if (MODE == 0 && a.ymap) {
      const int f = plane / N, z = plane - f * N;
      d = a.spec + a.yout[f] * a.sstride + z * pstride + col;
    }
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int n = t + m * Tf;
      if (MODE == 0 && DIR < 0 && band && out_band(n, keep, N)) continue;
      gst<NTS>(d + n * estride, v[m]);
    }
  }
}

// ---------------------------------------------------------------------------
// Fused z round trip: REAL(IFFT(op * FFT(f))) along z for the columns that the
// forward x and y passes left in the scratch spectra (kx, ky transformed, z
// physical).  Each workgroup loads C adjacent columns, runs the forward z
// transform, multiplies by the column's spectral factor, runs the inverse z
// transform and stores the result for the inverse y and x passes.  The
// spectrum of the round trip never leaves the workgroup: every butterfly,
// twiddle product and LDS exchange of both transforms is FP64 and the column
// is rounded to the field precision once, on its store.  In FP32 this removes
// the rounding of fully transformed coefficients, whose errors are coherent
// across the field: at 128^3, q = 6, it accounts for 1e-7 .. 2e-7 of the
// relative error of R(u), against 1e-8 .. 2e-8 for the rounding after the x
// and y passes (FP32-storage emulation of the oracle).
// Plane p of the launch reads scratch slot zin[p] and writes zout[p] (equal
// only when no other plane reads that slot).  With keep >= 0 every op masks
// to the cube max_j |m_j| <= keep: columns outside it are never read by the
// pruned inverse y pass and are skipped.
// ---------------------------------------------------------------------------
template <typename T> struct ZArgs {
  c2<T>* spec;        // scratch spectra, stride sstride
  const double2* tw;
  long long sstride;
  int keep = -1;
  int remap = 0;
  int kcomp = 0;      // grid over the kept tiles only (band-limited launches)
  int zin[MAXB];
  int zout[MAXB];
  SpecOp op[MAXB];    // factors along z only (col_factor_ok)
};
// LDS element of the round trip's column tile: FP64 (one rounding, at the
// store), except for FP32 fields at N >= 1024, where the FP64 tile would
// leave one 512-thread workgroup per CU (2 waves per SIMD): FP32 rows between
// the radix passes there, as in the separate z passes
template <typename T, int N>
using ZrtLds = typename std::conditional<(sizeof(T) == 4 && N >= 1024), float, double>::type;
// columns per workgroup: 16 (128-byte FP32 segments) while the column tile
// N x C and the twiddle table fit the 160 KB of LDS, else 8
template <typename T, int N> constexpr int zrt_C() {
  return (long long)N * (16 * (long long)sizeof(c2<ZrtLds<T, N>>) + 16) <= 160 * 1024 ? 16 : 8;
}

template <typename T, int N, int E, int C>
__global__ __launch_bounds__(C*(N / E)) void k_zrt(ZArgs<T> a, int ncols) {
  constexpr int Tf = N / E, PP = spec_pitch<T>(N);
  constexpr int estride = N * PP;
  using LS = ZrtLds<T, N>;
  __shared__ c2<LS> sm[N * C];
  // twiddles from an LDS table while it takes one entry per thread (the
  // registers that carry it to LDS spilled the 1024-point kernel)
  constexpr bool TAB = C * Tf >= N;
  __shared__ double2 smtw[TAB ? N : 1];   // read from the second radix pass on
  const int c = threadIdx.x % C, t = threadIdx.x / C;
  TwStage<TAB ? N : 1, C * Tf> tws;
  if constexpr (TAB) tws.load(a.tw);
  unsigned bx, by;
  xcd_tile(a.remap != 0, bx, by);
  const int keep = a.keep;
  int col = (int)bx * C + c;
  if (a.kcomp) {   // tile bx of the kept rows: (keep + C) / C tiles per kept ky
    const unsigned nt = (unsigned)(keep + C) / C, jy = bx / nt, jt = bx - jy * nt;
    const int kyt = (int)jy <= keep ? (int)jy : (int)jy + N - 2 * keep - 1, kxt = (int)jt * C + c;
    col = kxt < PP ? kyt * PP + kxt : ncols;   // a lane past the row pitch is invalid
  }
  const int plane = (int)by;
  const int ky = col / PP, kx = col - ky * PP;
  const bool valid = col < ncols && kx <= N / 2;
  const bool ok = valid && (keep < 0 || (kx <= keep && mabs(ky, N) <= keep));
  if (!__syncthreads_or(ok)) return;   // block-uniform: no column of the tile is read later
  constexpr bool NTL = nt_ld<PC_ZI>(), NTS = nt_st<PC_ZI>();
  const c2<T>* s = a.spec + a.zin[plane] * a.sstride + col;
  c2<double> v[E];
#pragma unroll
  for (int m = 0; m < E; ++m)
    v[m] = ok ? ccvt<double>(gld<NTL>(s + (t + m * Tf) * estride)) : cmk<double>(0.0, 0.0);
  LdsCol<LS, C> L{sm, c};
  const SpecOp op = a.op[plane];
  const ColFactor cf = col_factor_init(op, op.scale, kx, ky, N);
  auto round_trip = [&](const auto& tw) {
    fft_core<double, N, E, -1, LdsCol<LS, C>>(v, t, tw, L);
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = cmul(v[m], col_factor(cf, t + m * Tf, N));
    // the last forward pass read its LDS rows before its barrier: the first
    // inverse pass may overwrite them
    fft_core<double, N, E, +1, LdsCol<LS, C>>(v, t, tw, L);
  };
  if constexpr (TAB) {
    tws.store(smtw);   // ordered before the first twiddle read by the first pass's barrier
    round_trip(TwTab{smtw});
  } else {
    round_trip(TwPow{a.tw});
  }
  if (!ok) return;
  c2<T>* d = a.spec + a.zout[plane] * a.sstride + col;
#pragma unroll
  for (int m = 0; m < E; ++m) gst<NTS>(d + (t + m * Tf) * estride, cmk<T>((T)v[m].x, (T)v[m].y));
}

// Column copy with the memory access pattern of a strided pass and none of
// its arithmetic: C adjacent columns per workgroup, E elements per thread at
// the pass's element stride (AXIS 1: along y inside one z plane; AXIS 2:
// along z across the planes), the pass's load and store hints, padding
// columns skipped.  Its rate is the ceiling that pattern sets for the pass.
template <typename T, int N, int E, int C, int AXIS>
__global__ __launch_bounds__(C*(N / E)) void k_colcopy(const c2<T>* __restrict__ src,
                                                       c2<T>* __restrict__ dst, int ncols,
                                                       int remap) {
  constexpr int Tf = N / E, PP = spec_pitch<T>(N);
  constexpr int estride = AXIS == 1 ? PP : N * PP;
  constexpr long long pstride = AXIS == 1 ? (long long)N * PP : (long long)N * N * PP;
  constexpr int CLS = AXIS == 1 ? PC_YF : PC_ZF;
  const int c = threadIdx.x % C, t = threadIdx.x / C;
  unsigned bx, by;
  xcd_tile(remap != 0, bx, by);
  const int col = (int)bx * C + c;
  if (col >= ncols || col % PP > N / 2) return;   // no barrier follows
  const long long off = (long long)by * pstride + col;
  c2<T> v[E];
#pragma unroll
  for (int m = 0; m < E; ++m) v[m] = gld<nt_ld<CLS>()>(src + off + (long long)(t + m * Tf) * estride);
#pragma unroll
  for (int m = 0; m < E; ++m) gst<nt_st<CLS>()>(dst + off + (long long)(t + m * Tf) * estride, v[m]);
}

// ---------------------------------------------------------------------------
// FP64 reductions: per-block partial sums (wavefront shuffles + LDS), then a
// single-block final sum; fixed block count, so results are deterministic.
// ---------------------------------------------------------------------------
constexpr int RED_BLOCKS = 4096;   // 16 workgroups per CU: enough loads in flight
constexpr int RED_THREADS = 256;
// k_reduce: each workgroup sums contiguous chunks of RED_CHUNK * RED_THREADS
// calls (0: a grid-stride loop over RED_BLOCKS workgroups); the workgroups
// resident at one time then stream one dense address range
#ifndef MLQ_RED_CHUNK
#define MLQ_RED_CHUNK 8
#endif
constexpr int RED_CHUNK = MLQ_RED_CHUNK;
constexpr int RED_MAXBLK = 16384;   // chunked launches: workgroups at most
static_assert(RED_CHUNK % 2 == 0, "chunks run two calls per trip");
inline long long red_blocks(long long n) {   // workgroups of a k_reduce launch over n calls
  if (RED_CHUNK == 0) return RED_BLOCKS;
  const long long ch = (long long)RED_CHUNK * RED_THREADS, b = (n + ch - 1) / ch;
  return b < 1 ? 1 : (b > RED_MAXBLK ? RED_MAXBLK : b);
}
constexpr int RED_MAXK = 48;   // widest reduction: 2 sums per step size x 12 (+ spare)

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

template <int K> __device__ __forceinline__ void block_reduce_store(double (&acc)[K], double* out) {
  __shared__ double sh[RED_THREADS / 64][K];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double s = wave_sum(acc[k]);
    if (lane == 0) sh[w][k] = s;
  }
  __syncthreads();
  if (threadIdx.x < K) {
    double s = 0.0;
    for (int i = 0; i < RED_THREADS / 64; ++i) s += sh[i][threadIdx.x];
    out[threadIdx.x] = s;
  }
}

// 16-byte vector loads: w consecutive values of one field
template <typename T> struct V16;
template <> struct V16<float> { using type = float4; static constexpr int w = 4; };
template <> struct V16<double> { using type = double2; static constexpr int w = 2; };
// (streaming: the nontemporal hint unless MLQ_NT_VEC=0)
#ifndef MLQ_NT_VEC
#define MLQ_NT_VEC 1
#endif
template <typename T> struct V16N;
template <> struct V16N<float> { typedef float type __attribute__((ext_vector_type(4))); };
template <> struct V16N<double> { typedef double type __attribute__((ext_vector_type(2))); };
template <typename T>
__device__ __forceinline__ void ld16(const T* p, long long iv, T (&o)[V16<T>::w]) {
  using NV = typename V16N<T>::type;
  const NV* q = reinterpret_cast<const NV*>(p) + iv;
  const NV v = MLQ_NT_VEC ? __builtin_nontemporal_load(q) : *q;
  __builtin_memcpy(o, &v, sizeof(v));
}
template <typename T>
__device__ __forceinline__ void st16(T* p, long long iv, const T (&o)[V16<T>::w]) {
  using NV = typename V16N<T>::type;
  NV v;
  __builtin_memcpy(&v, o, sizeof(v));
  NV* q = reinterpret_cast<NV*>(p) + iv;
  if (MLQ_NT_VEC) __builtin_nontemporal_store(v, q);
  else *q = v;
}
// points per functor call of a map or reduction (`static constexpr int vec`,
// default 1): the functor's index then counts vectors of that many
// consecutive points
template <class F, class = void> struct VecW { static constexpr int v = 1; };
template <class F> struct VecW<F, std::void_t<decltype(F::vec)>> {
  static constexpr int v = F::vec;
};

// algorithmic bytes per element of a map or reduction: one read of every
// input element and one write of every output element the functor touches
// (host method `double abytes() const`); 0 when the functor does not state it.
// For a producer of the fused forward x pass, `inbytes()` is the bytes it
// reads (and writes besides its outputs) per point.
template <class F, class = void> struct AlgB {
  static double of(const F&) { return 0.0; }
};
template <class F> struct AlgB<F, std::void_t<decltype(std::declval<const F&>().abytes())>> {
  static double of(const F& f) { return f.abytes(); }
};
template <class F, class = void> struct ProdB {
  static double of(const F&) { return 0.0; }
};
template <class F> struct ProdB<F, std::void_t<decltype(std::declval<const F&>().inbytes())>> {
  static double of(const F& f) { return f.inbytes(); }
};

// elements per grid-stride trip: 2 unless the functor says otherwise
// (`static constexpr int unroll`, for functors heavy in registers)
template <class F, class = void> struct RedUnroll { static constexpr int v = 2; };
template <class F> struct RedUnroll<F, std::void_t<decltype(F::unroll)>> {
  static constexpr int v = F::unroll;
};

template <int K, class F>
__global__ __launch_bounds__(RED_THREADS) void k_reduce(F f, long long n, double* part) {
  double acc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) acc[k] = 0.0;
  if constexpr (RED_CHUNK > 0) {
    constexpr long long CH = (long long)RED_CHUNK * RED_THREADS;
    for (long long c0 = (long long)blockIdx.x * CH; c0 < n; c0 += (long long)gridDim.x * CH) {
      const long long i0 = c0 + threadIdx.x;
      if (c0 + CH <= n) {
        if constexpr (RedUnroll<F>::v >= 2) {   // two calls per trip: their loads overlap
          for (int j = 0; j < RED_CHUNK; j += 2) {
            f(i0 + (long long)j * RED_THREADS, acc);
            f(i0 + (long long)(j + 1) * RED_THREADS, acc);
          }
        } else {
#pragma unroll 1
          for (int j = 0; j < RED_CHUNK; ++j) f(i0 + (long long)j * RED_THREADS, acc);
        }
      } else {
#pragma unroll 1
        for (int j = 0; j < RED_CHUNK; ++j) {
          const long long i = i0 + (long long)j * RED_THREADS;
          if (i < n) f(i, acc);
        }
      }
    }
  } else {
    // two grid-stride elements per trip: their loads are independent and
    // overlap (twice the bytes in flight per thread)
    const long long st = (long long)gridDim.x * RED_THREADS;
    long long i = (long long)blockIdx.x * RED_THREADS + threadIdx.x;
    if constexpr (RedUnroll<F>::v >= 2) {
      for (; i + st < n; i += 2 * st) {
        f(i, acc);
        f(i + st, acc);
      }
    }
    for (; i < n; i += st) f(i, acc);
  }
  block_reduce_store<K>(acc, part + (long long)blockIdx.x * K);
}

__global__ void k_reduce_final(const double* part, int nblk, int K, double* out);

// Gather of up to FETCH_MAX complex values from arbitrary device addresses
// into one buffer (one host transfer for all of them instead of one each)
constexpr int FETCH_MAX = 16;
template <typename T> struct FetchArgs { const c2<T>* p[FETCH_MAX]; int n; };
template <typename T> __global__ void k_fetch(FetchArgs<T> a, double* out) {
  const int i = threadIdx.x;
  if (i < a.n) {
    const c2<T> v = *a.p[i];
    out[2 * i] = (double)v.x;
    out[2 * i + 1] = (double)v.y;
  }
}

}  // namespace mlq
