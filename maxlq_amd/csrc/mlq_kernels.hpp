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
#include "mlq_fft.hpp"

namespace mlq {

constexpr int MAXB = 12;
constexpr double TWO_PI = 6.283185307179586476925286766559;

// Row pitch of a half spectrum in elements: n/2+1 rounded up so that every
// row starts on a 128-byte boundary (16 complex floats, 8 complex doubles).
// The padding entries kx in [n/2+1, pitch) are never read or written.
template <typename T> constexpr int spec_pitch(int n) {
  constexpr int a = 128 / (int)sizeof(c2<T>);
  return ((n / 2 + 1) + a - 1) / a * a;
}

// Decode a flat half-spectrum index (pitch pp) into mode indices; false on
// padding.  m_y, m_z in (-n/2, n/2] (Nyquist counted +n/2).
struct KIdx { int kx, ky, kz, mx, my, mz; };
__device__ __forceinline__ bool kdecode(long long i, int n, int pp, KIdx& k) {
  const long long row = i / pp;
  k.kx = (int)(i - row * pp);
  if (k.kx > (n >> 1)) return false;
  k.ky = (int)(row % n);
  k.kz = (int)(row / n);
  const int h = n >> 1;
  k.mx = k.kx;
  k.my = k.ky <= h ? k.ky : k.ky - n;
  k.mz = k.kz <= h ? k.kz : k.kz - n;
  return true;
}

template <typename T> struct XArgs {
  const T* rsrc[MAXB];
  T* rdst[MAXB];
  c2<T>* spec;        // batch-contiguous spectra, stride sstride
  const double2* tw;  // exp(-2 pi i k/N), FP64 for every field precision
  long long sstride;
  int nb;
  double scale;
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
};

// Spectral factor of one half-spectrum mode (see SpecOp in mlq_common.hpp).
__device__ __forceinline__ double2 spec_factor(const SpecOp& o, int kx, int ky, int kz, int n) {
  const int h = n >> 1;
  const int mx = kx, my = ky <= h ? ky : ky - n, mz = kz <= h ? kz : kz - n;
  double2 f = make_double2(o.scale, 0.0);
  const int fl = o.flags;
  if (fl == 0) return f;
  if (fl & SP_MASK) {
    int mm = max(abs(mx), max(abs(my), abs(mz)));
    if (mm > o.mcut) return make_double2(0.0, 0.0);
  }
  if ((fl & SP_ZEROK0) && mx == 0 && my == 0 && mz == 0) return make_double2(0.0, 0.0);
  const double K1 = TWO_PI * mx, K2 = TWO_PI * my, K3 = TWO_PI * mz;
  const double ksq = K1 * K1 + K2 * K2 + K3 * K3;
  if (fl & SP_NEGKSQ) f.x *= -ksq;
  if (fl & SP_INVKSQ) {
    if (ksq == 0.0) return make_double2(0.0, 0.0);
    f.x *= -1.0 / ksq;
  }
  if (fl & SP_HILBERT) f.x /= (1.0 + pow(o.hil_a * sqrt(ksq), o.hil_e));
  if (fl & SP_IK) {
    const int md = o.dim == 0 ? mx : (o.dim == 1 ? my : mz);
    if (md == h) return make_double2(0.0, 0.0);  // own-dimension Nyquist
    const double Kd = TWO_PI * md;
    f = make_double2(0.0, f.x * Kd);
  }
  return f;
}

template <typename T> __device__ __forceinline__ c2<T> apply_factor(c2<T> v, double2 f) {
  const T fx = (T)f.x, fy = (T)f.y;
  return cmk<T>(v.x * fx - v.y * fy, v.x * fy + v.y * fx);
}

// ---------------------------------------------------------------------------
// x pass, forward: two real rows packed as one complex row (a + i b), one
// N-point FFT, split into the two half spectra.
// ---------------------------------------------------------------------------
template <typename T, int N, int NF>
__global__ __launch_bounds__(NF* FftShape<N>::Tf) void k_xfwd(XArgs<T> a) {
  constexpr int E = FftShape<N>::E, Tf = FftShape<N>::Tf, PX = spec_pitch<T>(N);
  constexpr int LP = LdsRow<T>::pitch(N);
  __shared__ c2<T> sm[NF * LP];
  const int f = threadIdx.x / Tf, t = threadIdx.x % Tf;
  const long long pairs_per = (long long)N * N / 2;
  const long long gp = (long long)blockIdx.x * NF + f;
  const int b = (int)(gp / pairs_per);
  const long long lp = gp - (long long)b * pairs_per;
  const long long row0 = 2 * lp;
  const T* s = a.rsrc[b] + row0 * N;
  c2<double> v[E];
#pragma unroll
  for (int m = 0; m < E; ++m) v[m] = cmk<double>((double)s[t + m * Tf], (double)s[N + t + m * Tf]);
  LdsRow<T> L{sm + f * LP};
  fft_core<double, N, E, -1>(v, t, a.tw, L);
#pragma unroll
  for (int m = 0; m < E; ++m) L.st(t + m * Tf, v[m]);
  __syncthreads();
  c2<T>* d = a.spec + b * a.sstride + row0 * PX;
#pragma unroll
  for (int m = 0; m < E / 2; ++m) {
    const int k = t + m * Tf;
    const c2<double> zk = L.ld(k), zn = L.ld((N - k) & (N - 1));
    d[k] = cmk<T>((T)(0.5 * (zk.x + zn.x)), (T)(0.5 * (zk.y - zn.y)));
    d[PX + k] = cmk<T>((T)(0.5 * (zk.y + zn.y)), (T)(-0.5 * (zk.x - zn.x)));
  }
  if (t == 0) {
    const c2<double> z = L.ld(N / 2);
    d[N / 2] = cmk<T>((T)z.x, (T)0);
    d[PX + N / 2] = cmk<T>((T)z.y, (T)0);
  }
}

// x pass, inverse (C2R): rebuild the packed complex row from the two half
// spectra (Hermitian extension, real parts taken at kx = 0 and kx = N/2),
// inverse FFT, unpack, scale.
template <typename T, int N, int NF>
__global__ __launch_bounds__(NF* FftShape<N>::Tf) void k_xinv(XArgs<T> a) {
  constexpr int E = FftShape<N>::E, Tf = FftShape<N>::Tf, PX = spec_pitch<T>(N);
  constexpr int LP = LdsRow<T>::pitch(N);
  __shared__ c2<T> sm[NF * LP];
  const int f = threadIdx.x / Tf, t = threadIdx.x % Tf;
  const long long pairs_per = (long long)N * N / 2;
  const long long gp = (long long)blockIdx.x * NF + f;
  const int b = (int)(gp / pairs_per);
  const long long lp = gp - (long long)b * pairs_per;
  const long long row0 = 2 * lp;
  const c2<T>* sp = a.spec + b * a.sstride + row0 * PX;
  c2<double> v[E];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int k = t + m * Tf;
    const int kk = k <= N / 2 ? k : N - k;
    const c2<T> As = sp[kk], Bs = sp[PX + kk];
    double ax = As.x, ay = As.y, bx = Bs.x, by = Bs.y;
    if (k == 0 || k == N / 2) { ay = 0.0; by = 0.0; }
    if (k > N / 2) { ay = -ay; by = -by; }
    v[m] = cmk<double>(ax - by, ay + bx);
  }
  LdsRow<T> L{sm + f * LP};
  fft_core<double, N, E, +1>(v, t, a.tw, L);
  T* o = a.rdst[b] + row0 * N;
  const double sc = a.scale;
#pragma unroll
  for (int m = 0; m < E; ++m) {
    o[t + m * Tf] = (T)(v[m].x * sc);
    o[N + t + m * Tf] = (T)(v[m].y * sc);
  }
}

// ---------------------------------------------------------------------------
// Strided pass along y or z: C adjacent columns per workgroup, each column a
// length-N transform.  MODE 0: plain in-place pass on the scratch spectra
// (y pass).  MODE 1: inverse z pass reading sum_i coef_i * src_i times the
// spectral factor (prologue), writing scratch.  MODE 2: forward z pass
// reading scratch, writing factor * result to dst (epilogue).
// ---------------------------------------------------------------------------
template <typename T, int N, int E, int C, int DIR, int MODE>
__global__ __launch_bounds__(C*(N / E)) void k_strided(SArgs<T> a, int ncols, long long estride,
                                                       long long pstride) {
  constexpr int Tf = N / E, PP = spec_pitch<T>(N);
  __shared__ c2<T> sm[N * C];
  const int c = threadIdx.x % C, t = threadIdx.x / C;
  const int col = blockIdx.x * C + c;
  bool valid = col < ncols;
  const int plane = blockIdx.y;
  c2<double> v[E];
  int kx = 0, ky = 0;
  if (MODE != 0) {
    ky = col / PP;
    kx = col - ky * PP;
    valid = valid && kx <= N / 2;   // padding columns carry no data
  }
  if (MODE == 1) {
    const int bb = plane;
    const SpecOp op = a.op[bb];
    const int ns = a.nsrc[bb];
    c2<T> ld[E];
    // all E loads of a source are issued before any arithmetic on them
    {
      const c2<T>* s0 = a.src[bb][0] + col;
      const double c0 = a.coef[bb][0];
#pragma unroll
      for (int m = 0; m < E; ++m)
        ld[m] = valid ? s0[(long long)(t + m * Tf) * estride] : cmk<T>((T)0, (T)0);
#pragma unroll
      for (int m = 0; m < E; ++m) v[m] = cmk<double>(c0 * (double)ld[m].x, c0 * (double)ld[m].y);
    }
    for (int i = 1; i < ns; ++i) {
      const c2<T>* si = a.src[bb][i] + col;
      const double ci = a.coef[bb][i];
#pragma unroll
      for (int m = 0; m < E; ++m)
        ld[m] = valid ? si[(long long)(t + m * Tf) * estride] : cmk<T>((T)0, (T)0);
#pragma unroll
      for (int m = 0; m < E; ++m) {
        v[m].x += ci * (double)ld[m].x;
        v[m].y += ci * (double)ld[m].y;
      }
    }
    if (op.flags != 0 || op.scale != 1.0) {
#pragma unroll
      for (int m = 0; m < E; ++m) v[m] = cmul(v[m], spec_factor(op, kx, ky, t + m * Tf, N));
    }
  } else {
    const c2<T>* s = a.spec + plane * pstride + col;
    c2<T> ld[E];
#pragma unroll
    for (int m = 0; m < E; ++m)
      ld[m] = valid ? s[(long long)(t + m * Tf) * estride] : cmk<T>((T)0, (T)0);
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = cmk<double>((double)ld[m].x, (double)ld[m].y);
  }
  LdsCol<T, C> L{sm, c};
  fft_core<double, N, E, DIR>(v, t, a.tw, L);
  if (!valid) return;
  if (MODE == 2) {
    const int bb = plane;
    const SpecOp op = a.op[bb];
    c2<T>* d = a.dst[bb] + col;
    const bool mul = op.flags != 0 || op.scale != 1.0;
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int kz = t + m * Tf;
      const c2<double> r = mul ? cmul(v[m], spec_factor(op, kx, ky, kz, N)) : v[m];
      d[(long long)kz * estride] = cmk<T>((T)r.x, (T)r.y);
    }
  } else {
    c2<T>* d = a.spec + plane * pstride + col;
#pragma unroll
    for (int m = 0; m < E; ++m) d[(long long)(t + m * Tf) * estride] = cmk<T>((T)v[m].x, (T)v[m].y);
  }
}

// ---------------------------------------------------------------------------
// FP64 reductions: per-block partial sums (wavefront shuffles + LDS), then a
// single-block final sum; fixed block count, so results are deterministic.
// ---------------------------------------------------------------------------
constexpr int RED_BLOCKS = 1024;
constexpr int RED_THREADS = 256;
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

template <int K, class F>
__global__ __launch_bounds__(RED_THREADS) void k_reduce(F f, long long n, double* part) {
  double acc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) acc[k] = 0.0;
  for (long long i = (long long)blockIdx.x * RED_THREADS + threadIdx.x; i < n;
       i += (long long)gridDim.x * RED_THREADS)
    f(i, acc);
  block_reduce_store<K>(acc, part + (long long)blockIdx.x * K);
}

__global__ void k_reduce_final(const double* part, int nblk, int K, double* out);

}  // namespace mlq
