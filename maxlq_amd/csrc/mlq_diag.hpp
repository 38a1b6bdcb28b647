// Diagnostics of a velocity field (diagnosticScalars, function_ops.f90:
// 2503-2602, with Energy :2394-2412, Helicity :2479-2497, calculate_ring_data
// :2837-2900 and calculate_geometric_data :2681-2832) and the perturbations of
// the kappa test (kappa_test_pert, function_ops.f90:613-977).
#pragma once
#include <algorithm>
#include <limits>

#include "mlq_engine_core.hpp"

namespace mlq {

// ---- one pass over u and grad u ----------------------------------------------------
// Per point (G[i*3+j] = d_j u_i, w = curl u as vel2vort forms it from the same
// derivatives, Q the Q criterion of calculate_geometric_data):
//   sums  0-2 u_c^2/2, 3-5 w_c^2/2, 6 u.w, 7 (div u)^2, 8 [Q > 0]
//   maxima 9-11 u_c, 12-14 w_c, 15 |u|, 16 |w|, 17 u.w, 18 Q;  minimum 19 u.w
constexpr int DIAG_K = 20, DIAG_NSUM = 9, DIAG_NMAX = 10;
constexpr int DIAG_BLOCKS = 2048, DIAG_THREADS = 256;

__device__ __forceinline__ void diag_point(const double u[3], const double g[9], double w[3],
                                           double& Q) {
  w[0] = g[7] - g[5];   // d_y u_z - d_z u_y
  w[1] = g[2] - g[6];   // d_z u_x - d_x u_z
  w[2] = g[3] - g[1];   // d_x u_y - d_y u_x
  const double ww = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
  const double a = g[1] + g[3], b = g[2] + g[6], c = g[5] + g[7];
  Q = fmax(0.25 * ww - 0.25 * a * a - 0.25 * b * b - 0.25 * c * c -
               0.5 * (g[0] * g[0] + g[4] * g[4] + g[8] * g[8]),
           0.0);
}

template <typename T>
__global__ __launch_bounds__(DIAG_THREADS) void k_diag(const T* __restrict__ u,
                                                       const T* __restrict__ G, long long nr,
                                                       double* __restrict__ part) {
  double v[DIAG_K];
  for (int k = 0; k < DIAG_NSUM; ++k) v[k] = 0.0;
  for (int k = DIAG_NSUM; k < DIAG_NSUM + DIAG_NMAX; ++k) v[k] = -INFINITY;
  v[DIAG_K - 1] = INFINITY;
  for (long long i = (long long)blockIdx.x * DIAG_THREADS + threadIdx.x; i < nr;
       i += (long long)gridDim.x * DIAG_THREADS) {
    double uu[3], g[9], w[3], Q;
    for (int c = 0; c < 3; ++c) uu[c] = (double)u[c * nr + i];
    for (int k = 0; k < 9; ++k) g[k] = (double)G[k * nr + i];
    diag_point(uu, g, w, Q);
    const double h = uu[0] * w[0] + uu[1] * w[1] + uu[2] * w[2];
    const double dv = g[0] + g[4] + g[8];
    for (int c = 0; c < 3; ++c) {
      v[c] += 0.5 * uu[c] * uu[c];
      v[3 + c] += 0.5 * w[c] * w[c];
      v[9 + c] = fmax(v[9 + c], uu[c]);
      v[12 + c] = fmax(v[12 + c], w[c]);
    }
    v[6] += h;
    v[7] += dv * dv;
    if (Q > 0.0) v[8] += 1.0;
    v[15] = fmax(v[15], sqrt(uu[0] * uu[0] + uu[1] * uu[1] + uu[2] * uu[2]));
    v[16] = fmax(v[16], sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]));
    v[17] = fmax(v[17], h);
    v[18] = fmax(v[18], Q);
    v[19] = fmin(v[19], h);
  }
  __shared__ double sh[DIAG_THREADS / 64][DIAG_K];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int k = 0; k < DIAG_K; ++k) {
    double x = v[k];
    for (int o = 32; o > 0; o >>= 1) {
      const double y = __shfl_down(x, o, 64);
      x = k < DIAG_NSUM ? x + y : (k < DIAG_NSUM + DIAG_NMAX ? fmax(x, y) : fmin(x, y));
    }
    if (lane == 0) sh[wv][k] = x;
  }
  __syncthreads();
  if (threadIdx.x < DIAG_K) {
    const int k = threadIdx.x;
    double x = sh[0][k];
    for (int i = 1; i < DIAG_THREADS / 64; ++i)
      x = k < DIAG_NSUM ? x + sh[i][k] : (k < DIAG_NSUM + DIAG_NMAX ? fmax(x, sh[i][k]) : fmin(x, sh[i][k]));
    part[(long long)blockIdx.x * DIAG_K + k] = x;
  }
}

// the four fields diagnosticScalars saves: |u|, |w|, u.w and Q / max Q
template <typename T> struct F_diag_fields {
  const T* u; const T* G; long long nr; double* o; double Qmax;
  __device__ void operator()(long long i) const {
    double uu[3], g[9], w[3], Q;
    for (int c = 0; c < 3; ++c) uu[c] = (double)u[c * nr + i];
    for (int k = 0; k < 9; ++k) g[k] = (double)G[k * nr + i];
    diag_point(uu, g, w, Q);
    o[i] = sqrt(uu[0] * uu[0] + uu[1] * uu[1] + uu[2] * uu[2]);
    o[nr + i] = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    o[2 * nr + i] = uu[0] * w[0] + uu[1] * w[1] + uu[2] * w[2];
    o[3 * nr + i] = Q / Qmax;
  }
};

// calculate_ring_data: the points where |w| / max |w| > 0.95, listed in the
// reference's loop order (linear index, x fastest).  Each block covers
// RING_PTS consecutive points, 4 per thread; k_ring_count counts a block's
// points, the host turns the counts into offsets, k_ring_write places each
// point at its block offset plus its rank inside the block.
constexpr int RING_THREADS = 256, RING_PER = 4, RING_PTS = RING_THREADS * RING_PER;

template <typename T>
__device__ __forceinline__ int ring_flags(const T* __restrict__ G, long long nr, double magWmax,
                                          long long i0) {
  int f = 0;
  for (int k = 0; k < RING_PER; ++k) {
    const long long i = i0 + k;
    if (i >= nr) break;
    const double w0 = (double)G[7 * nr + i] - (double)G[5 * nr + i];
    const double w1 = (double)G[2 * nr + i] - (double)G[6 * nr + i];
    const double w2 = (double)G[3 * nr + i] - (double)G[1 * nr + i];
    if (sqrt(w0 * w0 + w1 * w1 + w2 * w2) / magWmax > 0.95) f |= 1 << k;
  }
  return f;
}

// exclusive prefix of v over the block's threads; *total gets the block sum
__device__ __forceinline__ int block_excl_scan(int v, int* total) {
  __shared__ int wsum[RING_THREADS / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int x = v;
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[wv] = x;
  __syncthreads();
  int before = 0, all = 0;
  for (int k = 0; k < RING_THREADS / 64; ++k) {
    if (k < wv) before += wsum[k];
    all += wsum[k];
  }
  *total = all;
  return before + x - v;
}

template <typename T>
__global__ __launch_bounds__(RING_THREADS) void k_ring_count(const T* __restrict__ G, long long nr,
                                                             double magWmax,
                                                             unsigned* __restrict__ cnt) {
  const long long i0 = (long long)blockIdx.x * RING_PTS + (long long)threadIdx.x * RING_PER;
  const int f = ring_flags(G, nr, magWmax, i0);
  int total = 0;
  (void)block_excl_scan(__popc(f), &total);
  if (threadIdx.x == 0) cnt[blockIdx.x] = (unsigned)total;
}

template <typename T>
__global__ __launch_bounds__(RING_THREADS) void k_ring_write(const T* __restrict__ G, long long nr,
                                                             double magWmax,
                                                             const long long* __restrict__ off,
                                                             long long* __restrict__ idx,
                                                             long long cap) {
  const long long i0 = (long long)blockIdx.x * RING_PTS + (long long)threadIdx.x * RING_PER;
  const int f = ring_flags(G, nr, magWmax, i0);
  int total = 0;
  long long pos = off[blockIdx.x] + block_excl_scan(__popc(f), &total);
  for (int k = 0; k < RING_PER; ++k)
    if (f & (1 << k)) {
      if (pos < cap) idx[pos] = i0 + k;
      ++pos;
    }
}

template <typename T>
void Engine<T>::diag_scalars(const T* u, double* out, double* fields, double* ring, long long cap,
                             long long* nring) {
  RT Gb = rtmp(9);
  grad9_rt(u, Gb);
  const size_t pbytes = sizeof(double) * DIAG_BLOCKS * DIAG_K;
  double* part = (double*)dmalloc(pbytes);
  prof_begin(PC_REDUCE);
  hipLaunchKernelGGL((k_diag<T>), dim3(DIAG_BLOCKS), dim3(DIAG_THREADS), 0, st_, u, Gb.get(), nr_,
                     part);
  MLQ_HIP(hipGetLastError());
  prof_end(PC_REDUCE, 12.0 * real_bytes());
  std::vector<double> h(DIAG_BLOCKS * DIAG_K);
  MLQ_HIP(hipMemcpyAsync(h.data(), part, pbytes, hipMemcpyDeviceToHost, st_));
  sync();
  dfree(part, pbytes);
  double v[DIAG_K];
  for (int k = 0; k < DIAG_K; ++k) {
    v[k] = h[k];
    for (int b = 1; b < DIAG_BLOCKS; ++b) {
      const double x = h[(size_t)b * DIAG_K + k];
      v[k] = k < DIAG_NSUM ? v[k] + x : (k < DIAG_NSUM + DIAG_NMAX ? std::max(v[k], x) : std::min(v[k], x));
    }
  }
  const double dV = 1.0 / (double)nr_;
  // out: K[3], E[3], Umax[3], Wmax[3], vorCore[3], magUmax, magWmax, H, maxHel,
  // minHel, ||div u||^2 (the mlq_diag layout)
  for (int c = 0; c < 3; ++c) {
    out[c] = v[c] * dV;
    out[3 + c] = v[3 + c] * dV;
    out[6 + c] = v[9 + c];
    out[9 + c] = v[12 + c];
  }
  out[12] = v[8] * dV;     // volume where Q > 0
  out[13] = v[18];         // max Q
  out[14] = 0.0;           // the reference leaves this entry unset
  out[15] = v[15];
  out[16] = v[16];
  out[17] = v[6] * dV;
  out[18] = v[17];
  out[19] = v[19];
  out[20] = v[7] * dV;
  if (fields) {
    const size_t fb = sizeof(double) * 4 * (size_t)nr_;
    double* f = (double*)dmalloc(fb);
    map(F_diag_fields<T>{u, Gb.get(), nr_, f, v[18]}, nr_);
    MLQ_HIP(hipMemcpyAsync(fields, f, fb, hipMemcpyDeviceToHost, st_));
    sync();
    dfree(f, fb);
  }
  if (nring) {
    const long long nb = (nr_ + RING_PTS - 1) / RING_PTS;
    const long long ucap = cap > 0 ? cap : 0;
    const size_t cb = sizeof(unsigned) * (size_t)nb, ob = sizeof(long long) * (size_t)nb;
    const size_t ib = sizeof(long long) * (size_t)std::max<long long>(1, ucap);
    unsigned* cnt = (unsigned*)dmalloc(cb);
    hipLaunchKernelGGL((k_ring_count<T>), dim3((unsigned)nb), dim3(RING_THREADS), 0, st_, Gb.get(),
                       nr_, v[16], cnt);
    MLQ_HIP(hipGetLastError());
    std::vector<unsigned> hc((size_t)nb);
    MLQ_HIP(hipMemcpyAsync(hc.data(), cnt, cb, hipMemcpyDeviceToHost, st_));
    sync();
    dfree(cnt, cb);
    std::vector<long long> ho((size_t)nb);
    long long c = 0;
    for (long long b = 0; b < nb; ++b) {
      ho[(size_t)b] = c;
      c += hc[(size_t)b];
    }
    *nring = c;
    const long long m = std::min(c, ucap);
    if (ring && m > 0) {
      long long* off = (long long*)dmalloc(ob);
      long long* idx = (long long*)dmalloc(ib);
      MLQ_HIP(hipMemcpyAsync(off, ho.data(), ob, hipMemcpyHostToDevice, st_));
      hipLaunchKernelGGL((k_ring_write<T>), dim3((unsigned)nb), dim3(RING_THREADS), 0, st_,
                         Gb.get(), nr_, v[16], off, idx, ucap);
      MLQ_HIP(hipGetLastError());
      std::vector<long long> li((size_t)m);
      MLQ_HIP(hipMemcpyAsync(li.data(), idx, sizeof(long long) * (size_t)m, hipMemcpyDeviceToHost,
                             st_));
      sync();
      dfree(off, ob);
      dfree(idx, ib);
      const double dx = 1.0 / (double)n_;
      for (long long p = 0; p < m; ++p) {
        const long long i = li[(size_t)p];
        ring[3 * p] = (double)(i % n_) * dx;
        ring[3 * p + 1] = (double)((i / n_) % n_) * dx;
        ring[3 * p + 2] = (double)(i / ((long long)n_ * n_)) * dx;
      }
    }
  }
}

// ---- kappa_test_pert ----------------------------------------------------------------
// seeded uniform values in [-1, 1) (the role of random_number, 2 x - 1)
template <typename T> struct F_uniform {
  T* o; unsigned long long base; long long npts; int comp;
  __device__ void operator()(long long i) const {
    const unsigned long long h =
        splitmix64(base + (unsigned long long)comp * (unsigned long long)npts + (unsigned long long)i);
    o[i] = (T)(2.0 * ((double)(h >> 11) * 0x1.0p-53) - 1.0);
  }
};
// sin/cos perturbations ("sine", "divfree-sine")
template <typename T> struct F_sine_pert {
  T* o; long long nr; int n; double m1, m2, m3;
  __device__ void operator()(long long i) const {
    const double dx = 1.0 / (double)n;
    const double X = (double)(i % n) * dx, Y = (double)((i / n) % n) * dx,
                 Z = (double)(i / ((long long)n * n)) * dx;
    const double sx = sin(TWO_PI * m1 * X), cx = cos(TWO_PI * m1 * X);
    const double sy = sin(TWO_PI * m2 * Y), cy = cos(TWO_PI * m2 * Y);
    const double sz = sin(TWO_PI * m3 * Z), cz = cos(TWO_PI * m3 * Z);
    o[i] = (T)(sx * cy * cz);
    o[nr + i] = (T)(-cx * sy * cz);
    o[2 * nr + i] = (T)(cx * cy * sz);
  }
};
// spectral reshaping of the loaded/random perturbations, |K| = 2 pi |m| with
// the Nyquist counted +pi N: 0 poly 10^m2 (|K|/(N/4))^m1, 1 exp 10^(-|K|/5),
// 2 the constant c; every kind zeroes |K| < MACH_EPSILON (:779-791, 883-895)
template <typename T> struct S_pert {
  c2<T>* a; long long ns; int n, pp, kind; double m1, m2, c;
  __device__ void operator()(long long i) const {
    KIdx k;
    if (!kdecode(i, n, pp, k)) return;
    const double K1 = TWO_PI * k.mx, K2 = TWO_PI * k.my, K3 = TWO_PI * k.mz;
    const double nk = sqrt(K1 * K1 + K2 * K2 + K3 * K3);
    double f;
    if (nk < 2.0e-16) f = 0.0;
    else if (kind == 0) f = pow(10.0, m2) * pow(nk / ((double)n / 4.0), m1);
    else if (kind == 1) f = pow(10.0, -nk / 5.0);
    else f = c;
    for (int cc = 0; cc < 3; ++cc) {
      const c2<T> v = a[cc * ns + i];
      a[cc * ns + i] = cmk<T>((T)(f * (double)v.x), (T)(f * (double)v.y));
    }
  }
};

template <typename T>
void Engine<T>::kappa_pert(int out, int kind, double m1, double m2, double m3,
                           unsigned long long seed, const double* loaded) {
  if (fld(out).ncomp != 3) throw Error(MLQ_EINVAL, "kappa_test_pert: a vector field");
  T* o = fphys(out);
  // fftfwdv, the mode factor, fftbwdv, REAL (the factor is even in K)
  auto reshape = [&](int sk, double c) {
    ST s = stmp(3);
    C* d[3] = {s.get(), s.get() + ns_, s.get() + 2 * ns_};
    spec3(o, d);
    smap(S_pert<T>{s.get(), ns_, n_, pp_, sk, m1, m2, c});
    InvSrc is[3];
    SpecOp none[3] = {sp_none(), sp_none(), sp_none()};
    T* dst[3] = {o, o + nr_, o + 2 * nr_};
    for (int j = 0; j < 3; ++j) is[j] = InvSrc{{d[j], nullptr, nullptr}, {1.0, 0, 0}, 1};
    inv(3, is, none, dst);
  };
  const bool loads = kind >= MLQ_PERT_LOAD_RAW && kind <= MLQ_PERT_LOAD_EXP;
  if (loads) {
    if (!loaded) throw Error(MLQ_EINVAL, "kappa_test_pert: this kind needs the loaded field");
    upload(out, loaded);
  }
  // 2 x - 1 with x uniform on [0, 1), zero mode removed, projected,
  // divided by 100 <phi, phi>_L2 (the start of "create-random-poly-b" and
  // "save-random")
  auto random_normalised = [&]() {
    const unsigned long long base = seed * 0x9E3779B97F4A7C15ULL;
    for (int c = 0; c < 3; ++c) map(F_uniform<T>{o + c * nr_, base, nr_, c}, nr_);
    reshape(2, 1.0);
    divAvg_free(out);
    double ss = 0.0;   // global_summed_field_inner_product(phi, phi, "L2")
    reduce<1>(R_dot<T>{o, o}, 3 * nr_, &ss);
    const double c = 100.0 * ss / (double)nr_;
    map(F_divide<T>{o, c}, 3 * nr_);
  };
  switch (kind) {
    case MLQ_PERT_SINE:                                             // :660-673
      map(F_sine_pert<T>{o, nr_, n_, m1, m2, m3}, nr_);
      divAvg_free(out);
      break;
    case MLQ_PERT_DIVFREE_SINE:                                     // :675-689
      map(F_sine_pert<T>{o, nr_, n_, m1, m1 + m3, m3}, nr_);
      break;
    case MLQ_PERT_LOAD_RAW:                                         // :655-658
      break;
    case MLQ_PERT_LOAD:                                             // :692-702
      divAvg_free(out);
      break;
    case MLQ_PERT_LOAD_SMOOTH:                                      // :704-713
    case MLQ_PERT_LOAD_SMOOTH_RAW: {                                // :927-935
      T* f[3] = {o, o + nr_, o + 2 * nr_};
      dealias_dev(f, 3, 20.0);
      if (kind == MLQ_PERT_LOAD_SMOOTH) divAvg_free(out);
      break;
    }
    case MLQ_PERT_LOAD_K:                                           // :715-741
      reshape(2, m1);
      divAvg_free(out);
      break;
    case MLQ_PERT_LOAD_POLY:                                        // :769-795
      reshape(0, 0.0);
      divAvg_free(out);
      break;
    case MLQ_PERT_LOAD_EXP:                                         // :873-899
      reshape(1, 0.0);
      divAvg_free(out);
      break;
    case MLQ_PERT_RANDOM_POLY:                                      // :797-844
      random_normalised();
      reshape(0, 0.0);
      divAvg_free(out);
      break;
    case MLQ_PERT_RANDOM:                                           // :937-966
      random_normalised();
      break;
    default:
      throw Error(MLQ_EINVAL, "kappa_test_pert: unknown kind");
  }
}

}  // namespace mlq
