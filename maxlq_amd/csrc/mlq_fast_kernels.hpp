// Kernels of the resident-state RCG iteration: the products that make the
// line-search quantities quadratic polynomials in the step size, their
// global polynomial coefficients, and spectral reductions over the modes the
// 2/3 mask keeps.
#pragma once
#include "mlq_pointwise.hpp"

namespace mlq {

constexpr int RED_MAXK_BIG = 48;

// Products of u, d and their gradients (G[i*3+j] = d_j u_i, Gd likewise)
// whose dealiased transforms are the tau-coefficients of the line search:
//   s(tau)   = D(phi.phi)                     -> S1' = 2 u.d, S2' = d.d
//   w_j(tau) = D(sum_i phi_i d_j phi_i)       -> b'_j, e'_j
//   X(tau)   = sum_ij d_i phi_j d_j phi_i     -> X1, X2 (pressure source)
//   Q(tau)   = |grad phi|^2                   -> Q1, Q2
// The direction is d = dB + gam u (dB band-limited, gam a scalar) and its
// gradient Gd = grad dB + gam G: d and Gd here hold the dB parts.
// uu = u, dd = d, g = grad u, h = grad d at one point
template <typename T>
__device__ __forceinline__ void ls_products_core(const T (&uu)[3], const T (&dd)[3], const T (&g)[9],
                                                 const T (&h)[9], T (&o)[12]) {
  o[0] = (T)2 * (uu[0] * dd[0] + uu[1] * dd[1] + uu[2] * dd[2]);
  o[1] = dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    T bj = 0, ej = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      bj += uu[a] * h[a * 3 + j] + dd[a] * g[a * 3 + j];
      ej += dd[a] * h[a * 3 + j];
    }
    o[2 + j] = bj;
    o[5 + j] = ej;
  }
  T x1 = 0, x2 = 0, q1 = 0, q2 = 0;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      x1 += g[c * 3 + a] * h[a * 3 + c] + h[c * 3 + a] * g[a * 3 + c];
      x2 += h[c * 3 + a] * h[a * 3 + c];
      q1 += g[a * 3 + c] * h[a * 3 + c];
      q2 += h[a * 3 + c] * h[a * 3 + c];
    }
  o[8] = x1;
  o[9] = x2;
  o[10] = (T)2 * q1;
  o[11] = q2;
}
template <typename T>
__device__ __forceinline__ void ls_products_at(const T* u, const T* d, const T* G, const T* Gd,
                                               long long nr, T gam, long long i, T (&o)[12]) {
  T uu[3], dd[3], g[9], h[9];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    uu[c] = gld<NT_PROD>(u + c * nr + i);
    dd[c] = gld<NT_PROD>(d + c * nr + i) + gam * uu[c];
  }
#pragma unroll
  for (int a = 0; a < 9; ++a) {
    g[a] = gld<NT_PROD>(G + a * nr + i);
    h[a] = gld<NT_PROD>(Gd + a * nr + i) + gam * g[a];
  }
  ls_products_core(uu, dd, g, h, o);
}
// the 12 coefficient fields of the line search, in the order
// S1, S2, b (3), e (3), X1, X2, Q1, Q2 (pointwise version: one field each)
template <typename T> struct F_ls_products {
  const T* u; const T* d; const T* G; const T* Gd; long long nr;
  T* S1; T* S2; T* b; T* e; T* X1; T* X2; T* Q1; T* Q2;   // b, e: 3 comps (stride nr)
  T gam;
  __device__ void operator()(long long i) const {
    T o[12];
    ls_products_at(u, d, G, Gd, nr, gam, i, o);
    S1[i] = o[0];
    S2[i] = o[1];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      b[j * nr + i] = o[2 + j];
      e[j * nr + i] = o[5 + j];
    }
    X1[i] = o[8];
    X2[i] = o[9];
    Q1[i] = o[10];
    Q2[i] = o[11];
  }
};
// the same 12 values as a producer of the fused forward x pass (k_xfwd_p)
template <typename T> struct P_ls_products {
  static constexpr int NB = 12;
  const T* u; const T* d; const T* G; const T* Gd; long long nr; T gam;
  double inbytes() const { return 24.0 * sizeof(T); }   // u, d, grad u, grad d
  __device__ void operator()(long long i, T (&o)[12]) const {
    ls_products_at(u, d, G, Gd, nr, gam, i, o);
  }
};

// the same 12 values with the dB parts of d and grad d staged by the inverse
// x pass of the same workgroup (k_xinv_fwd_p): staged input g at s[g * ss],
// d_0..2 then grad d in the layout of Gd
template <typename T> struct PI_ls_products {
  static constexpr int NB = 12, NI = 12;
  const T* u; const T* G; long long nr; T gam;
  double inbytes() const { return 12.0 * sizeof(T); }   // u, grad u
  __device__ void operator()(long long i, const T* s, int ss, T (&o)[12]) const {
    T uu[3], dd[3], g[9], h[9];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      uu[c] = gld<NT_PROD>(u + c * nr + i);
      dd[c] = s[c * ss] + gam * uu[c];
    }
#pragma unroll
    for (int a = 0; a < 9; ++a) {
      g[a] = gld<NT_PROD>(G + a * nr + i);
      h[a] = s[(3 + a) * ss] + gam * g[a];
    }
    ls_products_core(uu, dd, g, h, o);
  }
};

// ---------------------------------------------------------------------------
// Line-search fields with a small footprint (the velocity gradients one
// column at a time): the same fields as ls_products_at and the state's
// s, w, |grad u|^2 and pressure source, for phi = u + tau d with d a full
// physical field.
// S1 = 2 u.d, S2 = d.d (the tau and tau^2 coefficients of phi.phi)
template <typename T> struct F_uddots {
  const T* u; const T* d; long long nr; T* S1; T* S2;
  __device__ void operator()(long long i) const {
    T ud = 0, dd = 0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const T uc = u[c * nr + i], dc = d[c * nr + i];
      ud += uc * dc;
      dd += dc * dc;
    }
    S1[i] = (T)2 * ud;
    S2[i] = dd;
  }
};
// Column j of grad u and grad d, Gj = {d_j u_0..2, d_j d_0..2}: the j
// components of w(tau) = a + tau b + tau^2 e (w_j = sum_i phi_i d_j phi_i),
// and the column's part of |grad phi|^2 = Q0 + tau Q1 + tau^2 Q2 and of the
// pressure source X(tau) = sum_ac d_a phi_c d_c phi_a (its a = c = j term),
// written to Q and X (3 coefficient fields each) when first, else added
template <typename T> struct F_ls_column {
  const T* u; const T* d; const T* Gj; long long nr; int j; int first;
  T* a; T* b; T* e; T* Q; T* X;
  __device__ void operator()(long long i) const {
    T g[3], h[3], uu[3], dd[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      uu[c] = u[c * nr + i];
      dd[c] = d[c * nr + i];
      g[c] = Gj[c * nr + i];
      h[c] = Gj[(3 + c) * nr + i];
    }
    T aj = 0, bj = 0, ej = 0, q0 = 0, q1 = 0, q2 = 0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      aj += uu[c] * g[c];
      bj += uu[c] * h[c] + dd[c] * g[c];
      ej += dd[c] * h[c];
      q0 += g[c] * g[c];
      q1 += g[c] * h[c];
      q2 += h[c] * h[c];
    }
    a[i] = aj;
    b[i] = bj;
    e[i] = ej;
    const T gj = j == 0 ? g[0] : (j == 1 ? g[1] : g[2]);
    const T hj = j == 0 ? h[0] : (j == 1 ? h[1] : h[2]);
    const T qv[3] = {q0, (T)2 * q1, q2}, xv[3] = {gj * gj, (T)2 * gj * hj, hj * hj};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      Q[k * nr + i] = first ? qv[k] : Q[k * nr + i] + qv[k];
      X[k * nr + i] = first ? xv[k] : X[k * nr + i] + xv[k];
    }
  }
};
// the a != c terms of the pressure source for one pair a < c, from
// P4 = {d_a u_c, d_c u_a, d_a d_c, d_c d_a} (the (a, c) and (c, a) terms are
// equal), added to X
template <typename T> struct F_ls_pair {
  const T* P4; long long nr; T* X;
  __device__ void operator()(long long i) const {
    const T g1 = P4[i], g2 = P4[nr + i], h1 = P4[2 * nr + i], h2 = P4[3 * nr + i];
    X[i] += (T)2 * (g1 * g2);
    X[nr + i] += (T)2 * (g1 * h2 + h1 * g2);
    X[2 * nr + i] += (T)2 * (h1 * h2);
  }
};

// q = 4 polynomial coefficients (11 global sums):
//   sum_j w_j(tau)^2 with w = a + tau b + tau^2 e          -> c0..c4
//   p(tau) * sum_j phi_j w_j, p = P0 + tau P1 + tau^2 P2   -> r0..r5
template <typename T> struct R_ls_poly {
  static constexpr int unroll = 1;   // 18 input streams: enough loads in flight
  static constexpr int vec = V16<T>::w;   // consecutive points per call (16-byte loads)
  const T* a; const T* b; const T* e; const T* u; const T* d;
  const T* P0; const T* P1; const T* P2; long long nr; double gam;   // d + gam u
  double abytes() const { return 18.0 * sizeof(T); }   // per point
  __device__ void operator()(long long iv, double (&acc)[11]) const {
    constexpr int W = vec;
    double c0 = 0, c1 = 0, c2 = 0, c3 = 0, c4 = 0;
    double v0[W], v1[W], v2[W], v3[W];
#pragma unroll
    for (int l = 0; l < W; ++l) v0[l] = v1[l] = v2[l] = v3[l] = 0.0;
    const long long nv = nr / W;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      T av[W], bv[W], ev[W], uv[W], dv[W];
      ld16(a, j * nv + iv, av);
      ld16(b, j * nv + iv, bv);
      ld16(e, j * nv + iv, ev);
      ld16(u, j * nv + iv, uv);
      ld16(d, j * nv + iv, dv);
#pragma unroll
      for (int l = 0; l < W; ++l) {
        const double aj = av[l], bj = bv[l], ej = ev[l];
        const double uj = uv[l], dj = (double)dv[l] + gam * uj;
        c0 += aj * aj;
        c1 += 2.0 * aj * bj;
        c2 += bj * bj + 2.0 * aj * ej;
        c3 += 2.0 * bj * ej;
        c4 += ej * ej;
        v0[l] += uj * aj;
        v1[l] += uj * bj + dj * aj;
        v2[l] += uj * ej + dj * bj;
        v3[l] += dj * ej;
      }
    }
    T p0v[W], p1v[W], p2v[W];
    ld16(P0, iv, p0v);
    ld16(P1, iv, p1v);
    ld16(P2, iv, p2v);
    acc[0] += c0; acc[1] += c1; acc[2] += c2; acc[3] += c3; acc[4] += c4;
#pragma unroll
    for (int l = 0; l < W; ++l) {
      const double p0 = p0v[l], p1 = p1v[l], p2 = p2v[l];
      acc[5] += p0 * v0[l];
      acc[6] += p0 * v1[l] + p1 * v0[l];
      acc[7] += p0 * v2[l] + p1 * v1[l] + p2 * v0[l];
      acc[8] += p0 * v3[l] + p1 * v2[l] + p2 * v1[l];
      acc[9] += p1 * v3[l] + p2 * v2[l];
      acc[10] += p2 * v3[l];
    }
  }
};

// ---------------------------------------------------------------------------
// q = 6 line search by polynomials.  Where max(s(tau), 0) does not clip
// (s = S0 + tau S1 + tau^2 S2), the chain of J gives |phi|^4 = D(s^2),
// |phi|^2 = s and the pressure term D(p(tau) s(tau)): R1, R2, R3 are
// polynomials in tau whose coefficients are sums over the grid of products of
// coefficient fields.  C_k = sum_{i+j=k} S_i S_j and E_k = sum_{i+j=k} P_i S_j
// (k = 0..4), dealiased by the caller.
template <typename T> struct F_ls6_prod {
  const T* S0; const T* S1; const T* S2; const T* P0; const T* P1; const T* P2; long long nr;
  T* C; T* E;   // 5 fields each (stride nr)
  __device__ void operator()(long long i) const {
    const T s0 = S0[i], s1 = S1[i], s2 = S2[i], p0 = P0[i], p1 = P1[i], p2 = P2[i];
    C[i] = s0 * s0;
    C[nr + i] = (T)2 * s0 * s1;
    C[2 * nr + i] = s1 * s1 + (T)2 * s0 * s2;
    C[3 * nr + i] = (T)2 * s1 * s2;
    C[4 * nr + i] = s2 * s2;
    E[i] = p0 * s0;
    E[nr + i] = p0 * s1 + p1 * s0;
    E[2 * nr + i] = p0 * s2 + p1 * s1 + p2 * s0;
    E[3 * nr + i] = p1 * s2 + p2 * s1;
    E[4 * nr + i] = p2 * s2;
  }
};
// acc[3k + j] = sum D(C_k) Q_j (the R1 sums)
template <typename T> struct R_ls6q {
  static constexpr int unroll = 1;
  const T* DC; const T* Q0; const T* Q1; const T* Q2; long long nr;
  double abytes() const { return 8.0 * sizeof(T); }
  __device__ void operator()(long long i, double (&acc)[15]) const {
    const double q[3] = {(double)Q0[i], (double)Q1[i], (double)Q2[i]};
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      const double c = DC[k * nr + i];
#pragma unroll
      for (int j = 0; j < 3; ++j) acc[3 * k + j] += c * q[j];
    }
  }
};
// acc[5i + n] = sum S_i c_n with sum_j w_j(tau)^2 = sum_n c_n tau^n (the R2 sums)
template <typename T> struct R_ls6w {
  static constexpr int unroll = 1;
  const T* S0; const T* S1; const T* S2; const T* a; const T* b; const T* e; long long nr;
  double abytes() const { return 12.0 * sizeof(T); }
  __device__ void operator()(long long i, double (&acc)[15]) const {
    const double sv[3] = {(double)S0[i], (double)S1[i], (double)S2[i]};
    double c0 = 0, c1 = 0, c2 = 0, c3 = 0, c4 = 0;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const double aj = a[j * nr + i], bj = b[j * nr + i], ej = e[j * nr + i];
      c0 += aj * aj;
      c1 += 2.0 * aj * bj;
      c2 += bj * bj + 2.0 * aj * ej;
      c3 += 2.0 * bj * ej;
      c4 += ej * ej;
    }
    const double cn[5] = {c0, c1, c2, c3, c4};
#pragma unroll
    for (int m = 0; m < 3; ++m)
#pragma unroll
      for (int n = 0; n < 5; ++n) acc[5 * m + n] += sv[m] * cn[n];
  }
};
// acc[4k + l] = sum D(E_k) v_l with phi.w(tau) = sum_l v_l tau^l, phi = u +
// tau (d + gam u) (the R3 sums)
template <typename T> struct R_ls6b {
  static constexpr int unroll = 1;
  const T* DE; const T* u; const T* d; const T* a; const T* b; const T* e; long long nr;
  double gam;
  double abytes() const { return 20.0 * sizeof(T); }
  __device__ void operator()(long long i, double (&acc)[20]) const {
    double v0 = 0, v1 = 0, v2 = 0, v3 = 0;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const double aj = a[j * nr + i], bj = b[j * nr + i], ej = e[j * nr + i];
      const double uj = u[j * nr + i], dj = (double)d[j * nr + i] + gam * uj;
      v0 += uj * aj;
      v1 += uj * bj + dj * aj;
      v2 += uj * ej + dj * bj;
      v3 += dj * ej;
    }
    const double vl[4] = {v0, v1, v2, v3};
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      const double c = DE[k * nr + i];
#pragma unroll
      for (int l = 0; l < 4; ++l) acc[4 * k + l] += c * vl[l];
    }
  }
};
template <typename T> struct R_sum3 {
  const T* a; const T* b; const T* c;
  __device__ void operator()(long long i, double (&acc)[3]) const {
    acc[0] += (double)a[i];
    acc[1] += (double)b[i];
    acc[2] += (double)c[i];
  }
};

// g_b = sqrt(max(S0 + tau_b S1 + tau_b^2 S2, 0)) and gg_b = g_b * g_b for a
// batch of step sizes (calc_uk, function_ops.f90:2332-2339, at phi = u + tau d)
template <typename T> struct F_ls_g {
  const T* S0; const T* S1; const T* S2;
  T* g[MAXB]; T* gg[MAXB]; double tau[MAXB]; int nb;
  __device__ void operator()(long long i) const {
    const T s0 = S0[i], s1 = S1[i], s2 = S2[i];
    for (int b = 0; b < nb; ++b) {
      const T t = (T)tau[b];
      const T s = s0 + t * (s1 + t * s2);
      const T gv = s <= (T)0 ? (T)0 : sqrt(s);
      if (g[b]) g[b][i] = gv;
      gg[b][i] = gv * gv;
    }
  }
};

// phi_c = a u_c + b d_c: Au_c = A phi_c (the normal transform's input) and
// un_c = sc phi_c (the retracted field), three components
template <typename T> struct F_step {
  const T* u; const T* d; const T* A; long long nr; T a, b, sc; T* Au; T* un;
  __device__ void operator()(long long i) const {
    const T w = A[i];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const T phi = a * u[c * nr + i] + b * d[c * nr + i];
      Au[c * nr + i] = w * phi;
      un[c * nr + i] = sc * phi;
    }
  }
};
// producer form: A phi_c feeds the fused forward x pass, sc phi_c is stored;
// with As, asc * A is stored too (As may be A: the same point, read first)
template <typename T> struct P_step {
  static constexpr int NB = 3;
  const T* u; const T* d; const T* A; long long nr; T a, b, sc; T* un;
  T* As = nullptr; T asc = 1;
  double inbytes() const { return (As ? 11.0 : 10.0) * sizeof(T); }   // A, u, d read, u_new (and As) written
  __device__ void operator()(long long i, T (&o)[3]) const {
    const T w = gld<NT_PROD>(A + i);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const T phi = a * gld<NT_PROD>(u + c * nr + i) + b * gld<NT_PROD>(d + c * nr + i);
      o[c] = w * phi;
      gst<NT_PROD>(un + c * nr + i, (T)(sc * phi));
    }
    if (As) gst<NT_PROD>(As + i, asc * w);
  }
};

// o = x0 + t x1 + t^2 x2 stored, and the sum of o^2 (the q = 4 step: |phi|^2
// and ||phi||_4^4 in one pass)
template <typename T> struct R_quad_sq {
  static constexpr int vec = V16<T>::w;   // 16-byte accesses
  const T* x0; const T* x1; const T* x2; T* o; T t;
  double abytes() const { return 4.0 * sizeof(T); }
  __device__ void operator()(long long iv, double (&acc)[1]) const {
    T a[vec], b[vec], c[vec], r[vec];
    ld16(x0, iv, a);
    ld16(x1, iv, b);
    ld16(x2, iv, c);
#pragma unroll
    for (int l = 0; l < vec; ++l) r[l] = a[l] + t * (b[l] + t * c[l]);
    st16(o, iv, r);
#pragma unroll
    for (int l = 0; l < vec; ++l) acc[0] += (double)r[l] * (double)r[l];
  }
};

// o = sc * (x0 + t x1 + t^2 x2)
template <typename T> struct F_quad {
  static constexpr int vec = V16<T>::w;   // 16-byte accesses; o may alias x0
  const T* x0; const T* x1; const T* x2; T* o; T t, sc;
  double abytes() const { return 4.0 * sizeof(T); }
  __device__ void operator()(long long iv) const {
    T a[vec], b[vec], c[vec], r[vec];
    ld16(x0, iv, a);
    ld16(x1, iv, b);
    ld16(x2, iv, c);
#pragma unroll
    for (int l = 0; l < vec; ++l) r[l] = sc * (a[l] + t * (b[l] + t * c[l]));
    st16(o, iv, r);
  }
};
// g = sqrt(max(s, 0)) of s = sc (x0 + t x1 + t^2 x2) as F_quad then
// F_sqrtclip compute it, without storing s
template <typename T> struct F_quad_sqrtclip {
  static constexpr int vec = V16<T>::w;
  const T* x0; const T* x1; const T* x2; T* o; T t, sc;
  double abytes() const { return 4.0 * sizeof(T); }
  __device__ void operator()(long long iv) const {
    T a[vec], b[vec], c[vec], r[vec];
    ld16(x0, iv, a);
    ld16(x1, iv, b);
    ld16(x2, iv, c);
#pragma unroll
    for (int l = 0; l < vec; ++l) {
      const T v = sc * (a[l] + t * (b[l] + t * c[l]));
      r[l] = v <= (T)0 ? (T)0 : sqrt(v);
    }
    st16(o, iv, r);
  }
};
template <typename T> struct S_quad {
  const c2<T>* x0; const c2<T>* x1; const c2<T>* x2; c2<T>* o; double t, sc;
  __device__ void operator()(long long i) const {
    const c2<T> a = x0[i], b = x1[i], c = x2[i];
    o[i] = cmk<T>((T)(sc * (a.x + t * (b.x + t * c.x))), (T)(sc * (a.y + t * (b.y + t * c.y))));
  }
};

// Y_i = q2 * aux2 * u_i - pc * A * grad_i p + 2 pc * sum_j d_i u_j d_j phi + vnu * A * Lap u_i
// (the terms of GradL2ForLq that are dealiased before the final sum,
// function_ops.f90:1140-1143, 1165-1169, 1178-1190); with gp null, lap holds
// the force vnu Lap u - pc grad p
template <typename T> struct F_grad_terms {
  const T* aux2; const T* u; const T* A; const T* gp; const T* G; const T* gphi; const T* lap;
  long long nr; T q2, pc, vnu; T* Y;
  // pending state update of G (optional): G := ga G + gb Gd0, written to Gout
  // by the thread that uses it (the same point, so in place is safe)
  const T* Gd0 = nullptr; T ga = 1, gb = 0; T* Gout = nullptr;
  __device__ void at(long long i, T (&y)[3]) const {
    T g[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) g[k] = gld<NT_PROD>(G + k * nr + i);
    if (Gd0) {
#pragma unroll
      for (int k = 0; k < 9; ++k) {
        g[k] = ga * g[k] + gb * gld<NT_PROD>(Gd0 + k * nr + i);
        gst<NT_PROD>(Gout + k * nr + i, g[k]);
      }
    }
    const T a2 = gld<NT_PROD>(aux2 + i), Ai = A ? gld<NT_PROD>(A + i) : (T)1;
    const T f0 = gld<NT_PROD>(gphi + i), f1 = gld<NT_PROD>(gphi + nr + i),
            f2 = gld<NT_PROD>(gphi + 2 * nr + i);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const T t5 = g[0 * 3 + c] * f0 + g[1 * 3 + c] * f1 + g[2 * 3 + c] * f2;
      if (gp) {
        y[c] = q2 * a2 * gld<NT_PROD>(u + c * nr + i) - pc * Ai * gld<NT_PROD>(gp + c * nr + i) +
               (T)2 * pc * t5 + vnu * Ai * gld<NT_PROD>(lap + c * nr + i);
      } else {
        y[c] = q2 * a2 * gld<NT_PROD>(u + c * nr + i) + (T)2 * pc * t5 +
               Ai * gld<NT_PROD>(lap + c * nr + i);
      }
    }
  }
  __device__ void operator()(long long i) const {
    T y[3];
    at(i, y);
#pragma unroll
    for (int c = 0; c < 3; ++c) Y[c * nr + i] = y[c];
  }
};
// the same three components as a producer of the fused forward x pass
template <typename T> struct P_grad_terms {
  static constexpr int NB = 3;
  F_grad_terms<T> f;   // Y unused
  double inbytes() const {
    // grad u, aux2, A, grad phi, u, grad p, Lap u (or the force); the pending
    // update of grad u reads Gd0 and writes grad u back
    return ((f.A ? 23.0 : 22.0) - (f.gp ? 0.0 : 3.0) + (f.Gd0 ? 18.0 : 0.0)) * sizeof(T);
  }
  __device__ void operator()(long long i, T (&o)[3]) const { f.at(i, o); }
};

// Hilbert weight then Leray projection on three half spectra, in place
// (HilbertGradient :3611 followed by divAvg_free :2879-2909), optional scale.
template <typename T> struct S_hilbert_leray {
  c2<T>* u0; c2<T>* u1; c2<T>* u2; int n, px; double lam, e2s, scale;
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
    const double w = scale / (1.0 + pow(lam * sqrt(ksq), e2s));
    const double e1 = mx == h ? 0.0 : K1, e2 = my == h ? 0.0 : K2, e3 = mz == h ? 0.0 : K3;
    const c2<T> a = u0[i], b = u1[i], c = u2[i];
    const double ar = w * a.x, ai = w * a.y, br = w * b.x, bi = w * b.y, cr = w * c.x, ci = w * c.y;
    const double dr = (e1 * ar + e2 * br + e3 * cr) / ksq;
    const double di = (e1 * ai + e2 * bi + e3 * ci) / ksq;
    u0[i] = cmk<T>((T)(ar - e1 * dr), (T)(ai - e1 * di));
    u1[i] = cmk<T>((T)(br - e2 * dr), (T)(bi - e2 * di));
    u2[i] = cmk<T>((T)(cr - e3 * dr), (T)(ci - e3 * di));
  }
};

// Spectral part of the gradient step, fused: with P the Leray projection
// (Nyquist-zeroed K', k = 0 removed, divAvg_free :2879-2909) and H the
// Sobolev weight 1/(1 + (lam |K|)^e2s) (HilbertGradient :3611),
//   V  <- P H (V - nu |K|^2 W)      (the Laplacian term of GradL2ForLq :1186-1190)
//   nH <- P H (sW W)                (the normal, optimization.f90:162-169)
// V, W, nH: three spectra each (stride ns).
template <typename T> struct S_dir_prep {
  c2<T>* V; const c2<T>* W; c2<T>* nH; long long ns; int n, px; double lam, e2s, nu, sW;
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
#pragma unroll
      for (int c = 0; c < 3; ++c) { V[c * ns + i] = cmk<T>(0, 0); nH[c * ns + i] = cmk<T>(0, 0); }
      return;
    }
    const double hw = 1.0 / (1.0 + pow(lam * sqrt(ksq), e2s));
    const double e[3] = {mx == h ? 0.0 : K1, my == h ? 0.0 : K2, mz == h ? 0.0 : K3};
    const double f = -nu * ksq;
    double vr[3], vi[3], wr[3], wi[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const c2<T> v = V[c * ns + i], w = W[c * ns + i];
      vr[c] = hw * ((double)v.x + f * (double)w.x);
      vi[c] = hw * ((double)v.y + f * (double)w.y);
      wr[c] = hw * (sW * (double)w.x);
      wi[c] = hw * (sW * (double)w.y);
    }
    const double dvr = (e[0] * vr[0] + e[1] * vr[1] + e[2] * vr[2]) / ksq;
    const double dvi = (e[0] * vi[0] + e[1] * vi[1] + e[2] * vi[2]) / ksq;
    const double dwr = (e[0] * wr[0] + e[1] * wr[1] + e[2] * wr[2]) / ksq;
    const double dwi = (e[0] * wi[0] + e[1] * wi[1] + e[2] * wi[2]) / ksq;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      V[c * ns + i] = cmk<T>((T)(vr[c] - e[c] * dvr), (T)(vi[c] - e[c] * dvi));
      nH[c * ns + i] = cmk<T>((T)(wr[c] - e[c] * dwr), (T)(wi[c] - e[c] * dwi));
    }
  }
};

// Every inner product of the CG step in one pass over three-component
// spectra (H_l weight m (1 + (lam |K|)^s)^2, half-spectrum multiplicity m):
//   acc[0..5] = <V,V>, <V,N>, <N,N>, <U,U>, <V,G0>, <N,G0>  (H_l; G0 may be null)
//   acc[6..8] = sum m |K'.V|^2, m Re(K'.V conj K'.N), m |K'.N|^2  (divergence)
//   acc[9..10] = <V,U>, <N,U>  (H_l)
template <typename T> struct R_dir {
  static constexpr int unroll = 1;   // one pow per mode: heavy in registers
  const c2<T>* V; const c2<T>* N; const c2<T>* U; const c2<T>* G0; long long ns; int n, px;
  double lam, s;
  __device__ void operator()(long long i, double (&acc)[11]) const {
    const RowIdx ri = row_split(i, px);   // px: row pitch
    const int kx = ri.kx;
    if (kx > (n >> 1)) return;             // row padding
    const int ky = row_ky(ri.row, n), kz = row_kz(ri.row, n);
    const int h = n >> 1;
    const int mx = kx, my = ky <= h ? ky : ky - n, mz = kz <= h ? kz : kz - n;
    const double K = TWO_PI * sqrt((double)mx * mx + (double)my * my + (double)mz * mz);
    const double w = 1.0 + pow(lam * K, s);
    const double m = (kx == 0 || kx == h) ? 1.0 : 2.0;
    const double mw = m * w * w;
    const double e[3] = {mx == h ? 0.0 : TWO_PI * mx, my == h ? 0.0 : TWO_PI * my,
                         mz == h ? 0.0 : TWO_PI * mz};
    double vv = 0, vn = 0, nn = 0, uu = 0, vg = 0, ng = 0, vu = 0, nu = 0;
    double dvr = 0, dvi = 0, dnr = 0, dni = 0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const c2<T> v = V[c * ns + i], a = N[c * ns + i], u = U[c * ns + i];
      const double vx = v.x, vy = v.y, ax = a.x, ay = a.y, ux = u.x, uy = u.y;
      vv += vx * vx + vy * vy;
      vn += vx * ax + vy * ay;
      nn += ax * ax + ay * ay;
      uu += ux * ux + uy * uy;
      vu += vx * ux + vy * uy;
      nu += ax * ux + ay * uy;
      if (G0) {
        const c2<T> g = G0[c * ns + i];
        vg += vx * (double)g.x + vy * (double)g.y;
        ng += ax * (double)g.x + ay * (double)g.y;
      }
      dvr += e[c] * vx; dvi += e[c] * vy;
      dnr += e[c] * ax; dni += e[c] * ay;
    }
    acc[0] += mw * vv; acc[1] += mw * vn; acc[2] += mw * nn; acc[3] += mw * uu;
    acc[4] += mw * vg; acc[5] += mw * ng;
    acc[6] += m * (dvr * dvr + dvi * dvi);
    acc[7] += m * (dvr * dnr + dvi * dni);
    acc[8] += m * (dnr * dnr + dni * dni);
    acc[9] += mw * vu;
    acc[10] += mw * nu;
  }
};

// ---- kept-cube spectral kernels ----------------------------------------------------
// Index space: the modes the 2/3 mask keeps, kx in [0, M], my, mz in [-M, M]
// (kx fastest; M < n/2, so no Nyquist mode is inside).  Spectra produced by
// masked transforms are zero outside this cube, and these kernels leave the
// outside of their outputs untouched: their consumers read only the cube
// (kept-cube kernels and the band-limited transform passes).  Weights of
// |K| = 2 pi |m| come from tables indexed by |m|^2.
// Each call handles a run of kw<T>() consecutive kx of one kept row through
// one 16-byte access per spectrum (2 modes in FP32, 1 in FP64; at 512^3 the
// FP32 kernels with one 8-byte access per mode ran 1.3-1.6x faster than
// their FP64 twins instead of 2x).  A row of M + 1 modes takes
// kept_calls_row calls; the last run of an FP32 row may hold kx = M + 1,
// outside the cube (M + 1 <= n/2, in bounds), which is read as zero and
// never written.
template <typename T> __host__ __device__ constexpr int kw() { return 16 / (int)sizeof(c2<T>); }
__host__ __device__ constexpr long long kept_calls_row(int M, int w) { return (M + w) / w; }
template <typename T> inline long long kept_calls(int M) {
  return kept_calls_row(M, kw<T>()) * (2LL * M + 1) * (2LL * M + 1);
}
// algorithmic bytes per call of a kernel that moves per_mode bytes per kept mode
template <typename T> inline double kept_call_bytes(int M, double per_mode) {
  return per_mode * (M + 1) / (double)kept_calls_row(M, kw<T>());
}
struct KeptRun { int kx0, my, mz, r2; long long off; };   // off: flat offset of kx0
template <int Wd> __device__ __forceinline__ KeptRun kept_run(long long i, int n, int px, int M) {
  const unsigned P = (unsigned)((M + Wd) / Wd), W = 2 * M + 1;
  KeptRun k;
  const unsigned ii = (unsigned)i, r = ii / P;
  k.kx0 = (int)(ii - r * P) * Wd;
  const unsigned rz = r / W;
  k.my = (int)(r - rz * W) - M;
  k.mz = (int)rz - M;
  const int ky = k.my < 0 ? k.my + n : k.my, kz = k.mz < 0 ? k.mz + n : k.mz;
  k.off = ((long long)kz * n + ky) * px + k.kx0;
  k.r2 = k.my * k.my + k.mz * k.mz;
  return k;
}
// the modes of a run of one spectrum; those past the cube read as zero
template <typename T>
__device__ __forceinline__ void kld(const c2<T>* p, int kx0, int M, c2<T> (&v)[kw<T>()]) {
  using NV = typename V16N<T>::type;
  const NV* q = reinterpret_cast<const NV*>(p);
  const NV x = NT_PROD ? __builtin_nontemporal_load(q) : *q;
  __builtin_memcpy(v, &x, sizeof(x));
#pragma unroll
  for (int l = 1; l < kw<T>(); ++l) v[l] = keep_or_zero<T>(kx0 + l <= M, v[l]);
}
// store a run; a mode past the cube is left untouched
template <typename T>
__device__ __forceinline__ void kst(c2<T>* p, int kx0, int M, const c2<T> (&v)[kw<T>()]) {
  using NV = typename V16N<T>::type;
  if (kw<T>() == 1 || kx0 + kw<T>() - 1 <= M) {
    NV x;
    __builtin_memcpy(&x, v, sizeof(x));
    NV* q = reinterpret_cast<NV*>(p);
    if (NT_PROD) __builtin_nontemporal_store(x, q);
    else *q = x;
  } else {
    gst<NT_PROD>(p, v[0]);
  }
}

// Direction set-up on the kept cube (V, W, G0 band-limited; U's kept part
// suffices for <V,U>, <N,U>), with hwl[m2] = (hw, hl), hw = 1 / (1 + (lam |K|)^e2s)
// the Hilbert weight and hl = (1 + (lam |K|)^s)^2 the squared H_l weight:
//   V := Leray(hw (V - nu |K|^2 W)),  nH := Leray(hw sW W)
// stored in place, and from the stored (rounded) values in the same pass
// acc = {<V,V>, <V,N>, <N,N>, <V,G0>, <N,G0>, <V,U>, <N,U>, |K.V|^2,
// Re(K.V conj K.N), |K.N|^2} (H_l products for the first seven)
template <typename T> struct R_dir_prep_k {
  c2<T>* V; const c2<T>* W; c2<T>* nH; const c2<T>* U; const c2<T>* G0; long long ns; int n, px, M;
  const double2* hwl;
  double nu, sW;
  static constexpr int unroll = 1;   // one call per trip: 15-18 loads of 16 bytes in flight
  double abytes() const { return kept_call_bytes<T>(M, (G0 ? 18.0 : 15.0) * sizeof(c2<T>)); }
  __device__ void operator()(long long i, double (&acc)[10]) const {
    constexpr int w = kw<T>();
    const KeptRun k = kept_run<w>(i, n, px, M);
    c2<T> v[3][w], a[3][w], u[3][w], g[3][w];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      kld<T>(V + c * ns + k.off, k.kx0, M, v[c]);
      kld<T>(W + c * ns + k.off, k.kx0, M, a[c]);
      kld<T>(U + c * ns + k.off, k.kx0, M, u[c]);
      if (G0) kld<T>(G0 + c * ns + k.off, k.kx0, M, g[c]);
    }
    c2<T> ov[3][w], on[3][w];
#pragma unroll
    for (int l = 0; l < w; ++l) {
      const int kx = k.kx0 + l, m2 = kx * kx + k.r2;
      const double e[3] = {TWO_PI * kx, TWO_PI * k.my, TWO_PI * k.mz};
      const double ksq = e[0] * e[0] + e[1] * e[1] + e[2] * e[2];
      const double2 hw = hwl[m2];
      // the k = 0 mode: both outputs zero (h = 0, and no division by ksq = 0)
      const double h = m2 == 0 ? 0.0 : hw.x, f = -nu * ksq, kd = m2 == 0 ? 1.0 : ksq;
      double vr[3], vi[3], wr[3], wi[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        vr[c] = h * ((double)v[c][l].x + f * (double)a[c][l].x);
        vi[c] = h * ((double)v[c][l].y + f * (double)a[c][l].y);
        wr[c] = h * (sW * (double)a[c][l].x);
        wi[c] = h * (sW * (double)a[c][l].y);
      }
      const double dvr = (e[0] * vr[0] + e[1] * vr[1] + e[2] * vr[2]) / kd;
      const double dvi = (e[0] * vi[0] + e[1] * vi[1] + e[2] * vi[2]) / kd;
      const double dwr = (e[0] * wr[0] + e[1] * wr[1] + e[2] * wr[2]) / kd;
      const double dwi = (e[0] * wi[0] + e[1] * wi[1] + e[2] * wi[2]) / kd;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        ov[c][l] = cmk<T>((T)(vr[c] - e[c] * dvr), (T)(vi[c] - e[c] * dvi));
        on[c][l] = cmk<T>((T)(wr[c] - e[c] * dwr), (T)(wi[c] - e[c] * dwi));
      }
      // the sums of the stored values (a mode past the cube is zero)
      const bool in = kx <= M;
      const double m = kx == 0 ? 1.0 : 2.0;
      const double mw = in ? m * hw.y : 0.0;
      double vv = 0, vn = 0, nn = 0, vg = 0, ng = 0, vu = 0, nu2 = 0;
      double dr = 0, di = 0, dnr = 0, dni = 0;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double vx = ov[c][l].x, vy = ov[c][l].y, ax = on[c][l].x, ay = on[c][l].y;
        const double ux = u[c][l].x, uy = u[c][l].y;
        vv += vx * vx + vy * vy;
        vn += vx * ax + vy * ay;
        nn += ax * ax + ay * ay;
        vu += vx * ux + vy * uy;
        nu2 += ax * ux + ay * uy;
        if (G0) {
          vg += vx * (double)g[c][l].x + vy * (double)g[c][l].y;
          ng += ax * (double)g[c][l].x + ay * (double)g[c][l].y;
        }
        dr += e[c] * vx; di += e[c] * vy;
        dnr += e[c] * ax; dni += e[c] * ay;
      }
      const double mk = in ? m : 0.0;
      acc[0] += mw * vv; acc[1] += mw * vn; acc[2] += mw * nn;
      acc[3] += mw * vg; acc[4] += mw * ng; acc[5] += mw * vu; acc[6] += mw * nu2;
      acc[7] += mk * (dr * dr + di * di);
      acc[8] += mk * (dr * dnr + di * dni);
      acc[9] += mk * (dnr * dnr + dni * dni);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      kst<T>(V + c * ns + k.off, k.kx0, M, ov[c]);
      kst<T>(nH + c * ns + k.off, k.kx0, M, on[c]);
    }
  }
};

// gu1 = cg (V - alpha N); d = gu1 + beta D0 on the kept cube
template <typename T> struct S_dir_update_k {
  const c2<T>* V; const c2<T>* N; const c2<T>* D0; c2<T>* gu1; c2<T>* d; long long ns; int n, px, M;
  double cg, alpha, beta;
  double abytes() const { return kept_call_bytes<T>(M, (D0 ? 15.0 : 12.0) * sizeof(c2<T>)); }
  __device__ void operator()(long long i) const {
    constexpr int w = kw<T>();
    const KeptRun k = kept_run<w>(i, n, px, M);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const long long o = c * ns + k.off;
      c2<T> v[w], a[w], q[w], g[w], r[w];
      kld<T>(V + o, k.kx0, M, v);
      kld<T>(N + o, k.kx0, M, a);
      if (D0) kld<T>(D0 + o, k.kx0, M, q);
#pragma unroll
      for (int l = 0; l < w; ++l) {
        const double gr = cg * ((double)v[l].x - alpha * (double)a[l].x);
        const double gi = cg * ((double)v[l].y - alpha * (double)a[l].y);
        g[l] = cmk<T>((T)gr, (T)gi);
        r[l] = D0 ? cmk<T>((T)(gr + beta * (double)q[l].x), (T)(gi + beta * (double)q[l].y)) : g[l];
      }
      kst<T>(gu1 + o, k.kx0, M, g);
      kst<T>(d + o, k.kx0, M, r);
    }
  }
};

// o = sc (x0 + t x1 + t^2 x2) on the kept cube (one spectrum)
template <typename T> struct S_quad_k {
  const c2<T>* x0; const c2<T>* x1; const c2<T>* x2; c2<T>* o; int n, px, M; double t, sc;
  double abytes() const { return kept_call_bytes<T>(M, 4.0 * sizeof(c2<T>)); }
  __device__ void operator()(long long i) const {
    constexpr int w = kw<T>();
    const KeptRun k = kept_run<w>(i, n, px, M);
    c2<T> a[w], b[w], c[w], r[w];
    kld<T>(x0 + k.off, k.kx0, M, a);
    kld<T>(x1 + k.off, k.kx0, M, b);
    kld<T>(x2 + k.off, k.kx0, M, c);
#pragma unroll
    for (int l = 0; l < w; ++l)
      r[l] = cmk<T>((T)(sc * (a[l].x + t * (b[l].x + t * c[l].x))),
                    (T)(sc * (a[l].y + t * (b[l].y + t * c[l].y))));
    kst<T>(o + k.off, k.kx0, M, r);
  }
};

// i K . W (divergence of a band-limited vector spectrum) on the kept cube
template <typename T> struct S_div_k {
  const c2<T>* W; c2<T>* o; long long ns; int n, px, M;
  double abytes() const { return kept_call_bytes<T>(M, 4.0 * sizeof(c2<T>)); }
  __device__ void operator()(long long i) const {
    constexpr int w = kw<T>();
    const KeptRun k = kept_run<w>(i, n, px, M);
    c2<T> a[w], b[w], c[w], r[w];
    kld<T>(W + k.off, k.kx0, M, a);
    kld<T>(W + ns + k.off, k.kx0, M, b);
    kld<T>(W + 2 * ns + k.off, k.kx0, M, c);
    const double e2 = TWO_PI * k.my, e3 = TWO_PI * k.mz;
#pragma unroll
    for (int l = 0; l < w; ++l) {
      const double e1 = TWO_PI * (k.kx0 + l);
      const double re = e1 * a[l].x + e2 * b[l].x + e3 * c[l].x;
      const double im = e1 * a[l].y + e2 * b[l].y + e3 * c[l].y;
      r[l] = cmk<T>((T)(-im), (T)re);
    }
    kst<T>(o + k.off, k.kx0, M, r);
  }
};

// Spectrum of the force nu Lap u - pc grad p over the whole half spectrum:
// F_c = -vnu |K|^2 h U_c - pc [kept] i K_c P, h the pending factor of U
// outside the kept cube (1 inside); P (the masked pressure spectrum) is read
// inside the cube only.  One inverse transform of F replaces those of Lap u
// and grad p, and the consumers read 3 fields instead of 6.
template <typename T> struct S_force {
  const c2<T>* U; const c2<T>* P; c2<T>* F; long long ns; int n, px, M; double vnu, pc, hs;
  // per element of the padded index space: U read and F written at every
  // mode (P, read inside the cube only, is not counted)
  double abytes() const { return 6.0 * sizeof(c2<T>) * (n / 2 + 1) / px; }
  __device__ void operator()(long long i) const {
    KIdx k;
    if (!kdecode(i, n, px, k)) return;   // row padding
    const bool cube = max(k.mx, max(abs(k.my), abs(k.mz))) <= M;
    const double K[3] = {TWO_PI * k.mx, TWO_PI * k.my, TWO_PI * k.mz};
    const double ksq = K[0] * K[0] + K[1] * K[1] + K[2] * K[2];
    const double fu = -vnu * ksq * (cube ? 1.0 : hs);
    const c2<T> p = cube ? P[i] : cmk<T>((T)0, (T)0);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const c2<T> u = U[c * ns + i];
      // - pc i K_c (p.x + i p.y) = pc K_c p.y - i pc K_c p.x (no Nyquist mode in the cube)
      const double pk = pc * K[c];
      F[c * ns + i] = cmk<T>((T)(fu * (double)u.x + pk * (double)p.y),
                             (T)(fu * (double)u.y - pk * (double)p.x));
    }
  }
};

// u_hat <- a u_hat + b dB over the whole half spectrum (dB read only on the
// kept cube when band-limited) and <u_hat, u_hat>_H_l of the result
// u_hat := a u_hat + b D over the whole half spectrum, with the modes
// outside the cube max|m_j| <= M multiplied by ah instead of a (their stored
// values carry a pending factor); D is added inside the cube only when dlim
// (a band-limited D), everywhere otherwise.  acc[0], acc[1]: <u_hat, u_hat>_H_l
// inside and outside the cube (unnormalised).
template <typename T> struct R_uh_update {
  c2<T>* U; const c2<T>* D; long long ns; int n, px, M; bool dlim; double a, ah, b;
  const double* hl;
  __device__ void operator()(long long i, double (&acc)[2]) const {
    const RowIdx ri = row_split(i, px);   // px: row pitch
    const int kx = ri.kx;
    if (kx > (n >> 1)) return;             // row padding
    const int ky = row_ky(ri.row, n), kz = row_kz(ri.row, n);
    const int h = n >> 1;
    const int my = ky <= h ? ky : ky - n, mz = kz <= h ? kz : kz - n;
    const bool cube = kx <= M && my <= M && my >= -M && mz <= M && mz >= -M;
    const bool addD = D && (!dlim || cube);
    const double am = cube ? a : ah;
    const double m = (kx == 0 || kx == h) ? 1.0 : 2.0;
    const double w = m * hl[kx * kx + my * my + mz * mz];
    double s = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const c2<T> u = U[c * ns + i];
      double ur = am * (double)u.x, ui = am * (double)u.y;
      if (addD) {
        const c2<T> d = D[c * ns + i];
        ur += b * (double)d.x;
        ui += b * (double)d.y;
      }
      const c2<T> r = cmk<T>((T)ur, (T)ui);
      U[c * ns + i] = r;
      s += (double)r.x * (double)r.x + (double)r.y * (double)r.y;
    }
    acc[cube ? 0 : 1] += w * s;
  }
};
// gu1 = cg (V - alpha N); d = gu1 + beta D0 (D0 null: d = gu1)  (optimization.f90:177, 211-214, 240)
template <typename T> struct S_dir_update {
  const c2<T>* V; const c2<T>* N; const c2<T>* D0; c2<T>* gu1; c2<T>* d; long long ns;
  double cg, alpha, beta;
  __device__ void operator()(long long i) const {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const c2<T> v = V[c * ns + i], a = N[c * ns + i];
      const double gr = cg * ((double)v.x - alpha * (double)a.x);
      const double gi = cg * ((double)v.y - alpha * (double)a.y);
      gu1[c * ns + i] = cmk<T>((T)gr, (T)gi);
      if (D0) {
        const c2<T> q = D0[c * ns + i];
        d[c * ns + i] = cmk<T>((T)(gr + beta * (double)q.x), (T)(gi + beta * (double)q.y));
      } else {
        d[c * ns + i] = cmk<T>((T)gr, (T)gi);
      }
    }
  }
};

// o = a*x + b*y on three spectra at once (x, y, o: 3 comps, stride ns)
template <typename T> struct S_axpby3 {
  const c2<T>* x; const c2<T>* y; c2<T>* o; long long ns; double a, b;
  __device__ void operator()(long long i) const {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const c2<T> p = x[c * ns + i], q = y ? y[c * ns + i] : cmk<T>(0, 0);
      o[c * ns + i] = cmk<T>((T)(a * p.x + b * q.x), (T)(a * p.y + b * q.y));
    }
  }
};
// V_hat = Vm + c * (-|K|^2) W   (aux7 of GradL2ForLq added in spectral space)
template <typename T> struct S_add_lap {
  c2<T>* V; const c2<T>* W; long long ns; int n, px; double c;
  __device__ void operator()(long long i) const {
    const RowIdx ri = row_split(i, px);   // px: row pitch
    const int kx = ri.kx;
    if (kx > (n >> 1)) return;             // row padding
    const int ky = row_ky(ri.row, n), kz = row_kz(ri.row, n);
    const int h = n >> 1;
    const int my = ky <= h ? ky : ky - n, mz = kz <= h ? kz : kz - n;
    const double ksq = TWO_PI * TWO_PI * ((double)kx * kx + (double)my * my + (double)mz * mz);
    const double f = -c * ksq;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const c2<T> v = V[k * ns + i], w = W[k * ns + i];
      V[k * ns + i] = cmk<T>((T)(v.x + f * w.x), (T)(v.y + f * w.y));
    }
  }
};

// Weighted Parseval sums over three-component spectra:
//   sum_c sum_k m(kx) w(k)^2 Re(a_c conj b_c), kind as R_specdot.
template <typename T> struct R_specdot3 {
  static constexpr int unroll = 1;
  const c2<T>* a; const c2<T>* b; long long ns; int n, px, kind; double lam, s;
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
    double t = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const c2<T> x = a[c * ns + i], y = b[c * ns + i];
      t += (double)x.x * (double)y.x + (double)x.y * (double)y.y;
    }
    acc[0] += m * w * w * t;
  }
};
// sum_k |sum_j i K'_j g_j|^2 with half-spectrum weights (||div g||^2 by Parseval)
template <typename T> struct R_divnorm {
  const c2<T>* g; long long ns; int n, px;
  __device__ void operator()(long long i, double (&acc)[1]) const {
    const RowIdx ri = row_split(i, px);   // px: row pitch
    const int kx = ri.kx;
    if (kx > (n >> 1)) return;             // row padding
    const int ky = row_ky(ri.row, n), kz = row_kz(ri.row, n);
    const int h = n >> 1;
    const int mx = kx, my = ky <= h ? ky : ky - n, mz = kz <= h ? kz : kz - n;
    const double e1 = mx == h ? 0.0 : TWO_PI * mx, e2 = my == h ? 0.0 : TWO_PI * my,
                 e3 = mz == h ? 0.0 : TWO_PI * mz;
    const c2<T> a = g[i], b = g[ns + i], c = g[2 * ns + i];
    const double re = e1 * a.x + e2 * b.x + e3 * c.x, im = e1 * a.y + e2 * b.y + e3 * c.y;
    const double m = (kx == 0 || kx == h) ? 1.0 : 2.0;
    acc[0] += m * (re * re + im * im);
  }
};

// Line-search spectral sums over the modes kept by the 2/3 mask, for a batch
// of step sizes: for each b,
//   acc[2b]   = sum_k m |A_b|^2                              (||phi||_q^q)
//   acc[2b+1] = sum_k m Re(conj(A_b) (Q0 + t Q1 + t^2 Q2))   (R1 numerator)
// The index space is the kept cube kx <= M, |ky| <= M, |kz| <= M.
template <typename T> struct R_ls_spec {
  const c2<T>* A[MAXB]; const c2<T>* Q0; const c2<T>* Q1; const c2<T>* Q2;
  double tau[MAXB]; int nb, n, px, M;
  __device__ void operator()(long long i, double (&acc)[2 * MAXB]) const {
    const int W = 2 * M + 1;
    const unsigned ii = (unsigned)i, r = ii / (unsigned)(M + 1);
    const int kx = (int)(ii - r * (unsigned)(M + 1));
    const unsigned rz = r / (unsigned)W;
    const int jy = (int)(r - rz * (unsigned)W), jz = (int)rz;
    const int my = jy - M, mz = jz - M;
    const int ky = my < 0 ? my + n : my, kz = mz < 0 ? mz + n : mz;
    const long long off = ((long long)kz * n + ky) * px + kx;
    const double m = (kx == 0 || kx == n / 2) ? 1.0 : 2.0;
    const c2<T> q0 = Q0[off], q1 = Q1[off], q2 = Q2[off];
    for (int b = 0; b < nb; ++b) {
      const c2<T> a = A[b][off];
      const double t = tau[b];
      const double qr = q0.x + t * (q1.x + t * q2.x), qi = q0.y + t * (q1.y + t * q2.y);
      acc[2 * b] += m * ((double)a.x * a.x + (double)a.y * a.y);
      acc[2 * b + 1] += m * ((double)a.x * qr + (double)a.y * qi);
    }
  }
};

// Generic per-step sums for exponents other than 4 (physical space):
//   acc[0] = sum A (Q0 + t Q1 + t^2 Q2)
//   acc[1] = sum Bw * sum_j w_j^2,  w = a + t b + t^2 e   (Bw = 1 when null)
//   acc[2] = sum Pd * sum_j phi_j w_j,  phi = u + t d
//   acc[3] = sum L^2                                      (||phi||_q^q)
template <typename T> struct R_ls_generic {
  const T* A; const T* Bw; const T* Pd; const T* Lh;
  const T* Q0; const T* Q1; const T* Q2; const T* a; const T* b; const T* e;
  const T* u; const T* d; long long nr; double t; double gam;   // d + gam u
  __device__ void operator()(long long i, double (&acc)[4]) const {
    const double q = (double)Q0[i] + t * ((double)Q1[i] + t * (double)Q2[i]);
    acc[0] += (double)A[i] * q;
    double w2 = 0.0, pw = 0.0;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const double wj = (double)a[j * nr + i] + t * ((double)b[j * nr + i] + t * (double)e[j * nr + i]);
      w2 += wj * wj;
      const double uj = u[j * nr + i];
      pw += (uj + t * ((double)d[j * nr + i] + gam * uj)) * wj;
    }
    acc[1] += (Bw ? (double)Bw[i] : 1.0) * w2;
    acc[2] += (double)Pd[i] * pw;
    acc[3] += (double)Lh[i] * (double)Lh[i];
  }
};

// p(t) * Bq, for the dealiased pressure term at exponents other than 4
template <typename T> struct F_pB {
  const T* P0; const T* P1; const T* P2; const T* Bq; T* o; T t;
  __device__ void operator()(long long i) const {
    o[i] = (P0[i] + t * (P1[i] + t * P2[i])) * (Bq ? Bq[i] : (T)1);
  }
};

// sum_c sum_k m Re(conj(W_c) xi_c) over three-component spectra restricted to
// the kept cube (W masked), for three right-hand sides at once.
template <typename T> struct R_kept_dot3 {
  const c2<T>* W; const c2<T>* x; const c2<T>* y; const c2<T>* z; long long ns; int n, px, M;
  double abytes() const { return kept_call_bytes<T>(M, 12.0 * sizeof(c2<T>)); }
  __device__ void operator()(long long i, double (&acc)[3]) const {
    constexpr int w = kw<T>();
    const KeptRun k = kept_run<w>(i, n, px, M);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const long long o = c * ns + k.off;
      c2<T> wv[w], a[w], b[w], zz[w];
      kld<T>(W + o, k.kx0, M, wv);
      kld<T>(x + o, k.kx0, M, a);
      kld<T>(y + o, k.kx0, M, b);
      kld<T>(z + o, k.kx0, M, zz);
#pragma unroll
      for (int l = 0; l < w; ++l) {
        const int kx = k.kx0 + l;
        const double m = (kx == 0 || kx == n / 2) ? 1.0 : 2.0;
        acc[0] += m * ((double)wv[l].x * a[l].x + (double)wv[l].y * a[l].y);
        acc[1] += m * ((double)wv[l].x * b[l].x + (double)wv[l].y * b[l].y);
        acc[2] += m * ((double)wv[l].x * zz[l].x + (double)wv[l].y * zz[l].y);
      }
    }
  }
};

// R_kept_dot3 for (X, D, old U) and the update of U inside the cube in one
// pass (W masked, D band-limited; the modes outside the cube are left with
// their pending factor): acc[0..2] = sum m Re(conj(W_c) xi_c) for xi = X, D
// and the old U; U := a U + b D with acc[3] = <U_new, U_new>_H_l inside the
// cube, hl[m2] = (1 + (lam |K|)^s)^2
template <typename T> struct R_step_k {
  const c2<T>* W; const c2<T>* X; c2<T>* U; const c2<T>* D; long long ns; int n, px, M;
  double a, b; const double* hl;
  static constexpr int unroll = 1;   // one call per trip: 12 loads of 16 bytes in flight
  double abytes() const { return kept_call_bytes<T>(M, 15.0 * sizeof(c2<T>)); }
  __device__ void operator()(long long i, double (&acc)[4]) const {
    constexpr int w = kw<T>();
    const KeptRun k = kept_run<w>(i, n, px, M);
    double s[w];
#pragma unroll
    for (int l = 0; l < w; ++l) s[l] = 0.0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const long long o = c * ns + k.off;
      c2<T> wv[w], xv[w], uv[w], dv[w], r[w];
      kld<T>(W + o, k.kx0, M, wv);
      kld<T>(X + o, k.kx0, M, xv);
      kld<T>(U + o, k.kx0, M, uv);
      kld<T>(D + o, k.kx0, M, dv);
#pragma unroll
      for (int l = 0; l < w; ++l) {
        const int kx = k.kx0 + l;
        const double m = (kx == 0 || kx == n / 2) ? 1.0 : 2.0;
        acc[0] += m * ((double)wv[l].x * xv[l].x + (double)wv[l].y * xv[l].y);
        acc[1] += m * ((double)wv[l].x * dv[l].x + (double)wv[l].y * dv[l].y);
        acc[2] += m * ((double)wv[l].x * uv[l].x + (double)wv[l].y * uv[l].y);
        const double ur = a * (double)uv[l].x + b * (double)dv[l].x;
        const double ui = a * (double)uv[l].y + b * (double)dv[l].y;
        r[l] = cmk<T>((T)ur, (T)ui);
        s[l] += (double)r[l].x * (double)r[l].x + (double)r[l].y * (double)r[l].y;
      }
      kst<T>(U + o, k.kx0, M, r);
    }
#pragma unroll
    for (int l = 0; l < w; ++l) {
      const int kx = k.kx0 + l;
      const double m = (kx == 0 || kx == (n >> 1)) ? 1.0 : 2.0;
      acc[3] += m * hl[kx * kx + k.r2] * s[l];
    }
  }
};

// ---------------------------------------------------------------------------
// Line-search polynomials (q = 4).  Where g^2 = max(s(tau), 0) does not clip,
// s(tau) = S0 + tau S1 + tau^2 S2 is band-limited and A(tau) = S0h + tau S1h
// + tau^2 S2h (masked spectra), so both line-search sums are degree-4
// polynomials in tau:
//   acc[j]     : sum_kept m |A|^2             = sum_j acc[j] tau^j
//   acc[5 + j] : sum_kept m Re(conj(A) Q)     = sum_j acc[5+j] tau^j
// with Q = Q0 + tau Q1 + tau^2 Q2.  Kept-cube index space as R_ls_spec.
template <typename T> struct R_ls_poly4 {
  const c2<T>* S0; const c2<T>* S1; const c2<T>* S2;
  const c2<T>* Q0; const c2<T>* Q1; const c2<T>* Q2; int n, px, M;
  double abytes() const { return kept_call_bytes<T>(M, 6.0 * sizeof(c2<T>)); }
  __device__ void operator()(long long i, double (&acc)[10]) const {
    constexpr int w = kw<T>();
    const KeptRun k = kept_run<w>(i, n, px, M);
    c2<T> a0[w], a1[w], a2[w], q0[w], q1[w], q2[w];
    kld<T>(S0 + k.off, k.kx0, M, a0);
    kld<T>(S1 + k.off, k.kx0, M, a1);
    kld<T>(S2 + k.off, k.kx0, M, a2);
    kld<T>(Q0 + k.off, k.kx0, M, q0);
    kld<T>(Q1 + k.off, k.kx0, M, q1);
    kld<T>(Q2 + k.off, k.kx0, M, q2);
    auto re = [](c2<T> x, c2<T> y) { return (double)x.x * (double)y.x + (double)x.y * (double)y.y; };
#pragma unroll
    for (int l = 0; l < w; ++l) {
      const int kx = k.kx0 + l;
      const double m = (kx == 0 || kx == n / 2) ? 1.0 : 2.0;
      acc[0] += m * re(a0[l], a0[l]);
      acc[1] += m * 2.0 * re(a0[l], a1[l]);
      acc[2] += m * (re(a1[l], a1[l]) + 2.0 * re(a0[l], a2[l]));
      acc[3] += m * 2.0 * re(a1[l], a2[l]);
      acc[4] += m * re(a2[l], a2[l]);
      acc[5] += m * re(a0[l], q0[l]);
      acc[6] += m * (re(a0[l], q1[l]) + re(a1[l], q0[l]));
      acc[7] += m * (re(a0[l], q2[l]) + re(a1[l], q1[l]) + re(a2[l], q0[l]));
      acc[8] += m * (re(a1[l], q2[l]) + re(a2[l], q1[l]));
      acc[9] += m * re(a2[l], q2[l]);
    }
  }
};

// Points where g^2 = max(s(tau), 0) can clip for some tau.  Where S0 > 0,
// S2 > 0 and S1^2 <= 4 S0 S2 the quadratic s(tau) is positive for every tau
// (for the undealiased products this is Cauchy-Schwarz); every other point
// is appended to `list` as (S0, S1, S2) (at most cap entries; *count counts
// all of them).
template <typename T>
__global__ __launch_bounds__(RED_THREADS) void k_ls_risky(const T* S0, const T* S1, const T* S2,
                                                           long long n, unsigned* count,
                                                           double* list, unsigned cap) {
  for (long long i = (long long)blockIdx.x * RED_THREADS + threadIdx.x; i < n;
       i += (long long)gridDim.x * RED_THREADS) {
    const double s0 = S0[i], s1 = S1[i], s2 = S2[i];
    if (!(s0 > 0.0 && s2 > 0.0 && s1 * s1 <= 4.0 * s0 * s2)) {
      const unsigned k = atomicAdd(count, 1u);
      if (k < cap) {
        list[3 * (size_t)k] = s0;
        list[3 * (size_t)k + 1] = s1;
        list[3 * (size_t)k + 2] = s2;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Fused q = 4 line-search sums.  For a batch of step sizes tau_b:
//   A_b = M FFT(g_b^2),  g_b = sqrt(max(S0 + tau_b S1 + tau_b^2 S2, 0))
//   out[2b]   = sum_kept m |A_b|^2                            (||phi||_q^q)
//   out[2b+1] = sum_kept m Re(conj(A_b) (Q0 + tau Q1 + tau^2 Q2))
// (the same quantities as F_ls_g -> masked forward transform -> R_ls_spec).
// k_xfwd_ls forms g_b^2 from the coefficient rows and runs the x pass for
// every step of the batch, storing only kx <= keep; the pruned forward y pass
// follows; k_zls runs the z pass over the kept (kx, ky) columns only and
// reduces instead of storing the spectra.
// ---------------------------------------------------------------------------
template <typename T> struct LsArgs {
  const T* S0; const T* S1; const T* S2;
  const c2<T>* Q0; const c2<T>* Q1; const c2<T>* Q2;
  c2<T>* spec;          // scratch spectra, one per step, stride sstride
  long long sstride;
  const double2* tw;
  double tau[MAXB];
  int nb, keep, remap;
  double* part;         // [z-pass block][2 nb] partial sums
};

template <typename T> __device__ __forceinline__ T ls_gg(T s0, T s1, T s2, T t) {
  const T s = s0 + t * (s1 + t * s2);
  const T gv = s <= (T)0 ? (T)0 : sqrt(s);
  return gv * gv;
}

// ZB (keep < 3N/8): the transform's entries 3N/8 .. 5N/8 - 1 are never read
template <typename T, int N, int NF, bool ZB = false>
__global__ __launch_bounds__(NF* FftShape<N>::Tf) void k_xfwd_ls(LsArgs<T> a) {
  constexpr int E = FftShape<N>::E, Tf = FftShape<N>::Tf, PX = spec_pitch<T>(N);
  constexpr int LP = LdsRow<T>::pitch(N);
  __shared__ c2<T> sm[NF * LP];
  const int f = threadIdx.x / Tf, t = threadIdx.x % Tf;
  const long long row0 = 2 * ((long long)blockIdx.x * NF + f);
  const T* r0 = a.S0 + row0 * N;
  const T* r1 = a.S1 + row0 * N;
  const T* r2 = a.S2 + row0 * N;
  LdsRow<T> L{sm + f * LP};
  for (int b = 0; b < a.nb; ++b) {
    const T tb = (T)a.tau[b];
    c2<XReg<T>> v[E];
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int i = t + m * Tf;
      v[m] = cmk<XReg<T>>((XReg<T>)ls_gg(r0[i], r1[i], r2[i], tb),
                          (XReg<T>)ls_gg(r0[N + i], r1[N + i], r2[N + i], tb));
    }
    fft_core<XReg<T>, N, E, -1, LdsRow<T>, ZB>(v, t, TwPow{a.tw}, L);
#pragma unroll
    for (int m = 0; m < E; ++m) L.st(t + m * Tf, ccvt<double>(v[m]));
    __syncthreads();
    c2<T>* d = a.spec + b * a.sstride + row0 * PX;
#pragma unroll
    for (int m = 0; m < E / 2; ++m) {
      const int k = t + m * Tf;
      if (k > a.keep) continue;
      const c2<T> zk = L.ld(k), zn = L.ld((N - k) & (N - 1));
      d[k] = cmk<T>((T)0.5 * (zk.x + zn.x), (T)0.5 * (zk.y - zn.y));
      d[PX + k] = cmk<T>((T)0.5 * (zk.y + zn.y), (T)-0.5 * (zk.x - zn.x));
    }
    __syncthreads();   // the LDS rows are reused by the next step's transform
  }
}

// ZB (keep < 3N/8): rows 3N/8 .. 5N/8 - 1 enter the sums with weight zero
template <typename T, int N, int E, int C, bool ZB = false>
__global__ __launch_bounds__(C*(N / E)) void k_zls(LsArgs<T> a, int ncols, long long estride) {
  constexpr int Tf = N / E, PP = spec_pitch<T>(N);
  constexpr int NW = C * Tf / 64;
  __shared__ c2<T> sm[N * C];
  __shared__ double2 smtw[N];   // twiddle table (read from the second radix pass on)
  __shared__ double red[NW > 0 ? NW : 1][2];
  const int c = threadIdx.x % C, t = threadIdx.x / C;
  TwStage<N, C * Tf> tws;
  tws.load(a.tw);
  unsigned bx, by;
  xcd_tile(a.remap != 0, bx, by);
  const int col = (int)bx * C + c;
  const bool valid = col < ncols;
  const int b = (int)by;
  const int K1 = a.keep + 1;
  const int kx = valid ? col % K1 : 0;
  const int my = (valid ? col / K1 : a.keep) - a.keep;
  const int ky = my < 0 ? my + N : my;
  const long long off = (long long)ky * PP + kx;   // offset of the column in a plane
  const c2<T>* s = a.spec + b * a.sstride + off;
  c2<T> v[E];
  if (valid) {
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = s[(long long)(t + m * Tf) * estride];
  } else {
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = cmk<T>((T)0, (T)0);
  }
  LdsCol<T, C> L{sm, c};
  tws.store(smtw);   // ordered before the first twiddle read by the first pass's barrier
  fft_core<T, N, E, -1, LdsCol<T, C>, ZB>(v, t, TwTab{smtw}, L);
  double s0 = 0.0, s1 = 0.0;
  if (valid) {
    const double w = (kx == 0 || kx == N / 2) ? 1.0 : 2.0;
    const double tb = a.tau[b];
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int kz = t + m * Tf;
      const bool in = mabs(kz, N) <= a.keep;
      const long long o = (long long)(in ? kz : 0) * estride + off;   // unconditional load
      const c2<T> q0 = a.Q0[o], q1 = a.Q1[o], q2 = a.Q2[o];
      const double qr = q0.x + tb * (q1.x + tb * q2.x), qi = q0.y + tb * (q1.y + tb * q2.y);
      const double ax = v[m].x, ay = v[m].y;
      const double wi = in ? w : 0.0;
      s0 += wi * (ax * ax + ay * ay);
      s1 += wi * (ax * qr + ay * qi);
    }
  }
  s0 = wave_sum(s0);
  s1 = wave_sum(s1);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) { red[wv][0] = s0; red[wv][1] = s1; }
  __syncthreads();
  if (threadIdx.x < 2) {
    double r = 0.0;
    for (int i = 0; i < NW; ++i) r += red[i][threadIdx.x];
    a.part[(long long)bx * (2 * a.nb) + 2 * b + threadIdx.x] = r;
  }
}

}  // namespace mlq
