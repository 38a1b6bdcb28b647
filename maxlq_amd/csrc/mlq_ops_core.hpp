// Engine<T>: the reference field operators (function_ops.f90) on device data.
// Each operator keeps the reference's dealias chain (every dealias of a
// nonlinear product is kept; masks that are the identity are skipped since
// their round trip only adds round-off) but computes shared quantities once:
// the velocity gradient, |u| and the powers of the calc_gk_order2 recursion.
#pragma once
#include "mlq_engine_core.hpp"

namespace mlq {

inline int dealias_mcut(int n, double p) {
  // kept iff max_j |K_j| <= k_cut, K = 2 pi m (function_ops.f90:3374, 3391)
  const double pi = 4.0 * std::atan2(1.0, 1.0);
  const double kcut = 2.0 * pi * (double)n / (p + 1.0);
  int m = 0;
  while (m <= n && 2.0 * pi * (double)(m + 1) <= kcut) ++m;
  return m;
}

// ---- chain: g = sqrt(max(D(u.u), 0)) and g^k (function_ops.f90:2281-2343) ----
template <typename T> void Engine<T>::chain_init(Chain& ch, const T* u, const T* s_pre) {
  ch.g = rtmp();
  if (s_pre) {
    map(F_sqrtclip<T>{s_pre, ch.g}, nr_);
  } else {
    map(F_dot3self<T>{u, nr_, ch.g}, nr_);
    if (p_.toDealias) { T* f = ch.g; dealias_dev(&f, 1, 2.0); }
    map(F_sqrtclip<T>{ch.g, ch.g}, nr_);
  }
  ch.haveG2 = false;
  ch.memo.clear();
}

template <typename T> const T* Engine<T>::chain_pow(Chain& ch, double k) {
  const long long key = std::llround(k * 1.0e6);
  auto it = ch.memo.find(key);
  if (it != ch.memo.end()) return it->second.get();
  RT out = rtmp();
  const double eps = 2.0e-16;
  if (k + eps < 0.0) {                                         // :2291-2301
    map(F_pow<T>{ch.g, out, k, 1}, nr_);
  } else if (k + eps < 2.0) {                                  // :2302-2304
    if (k == 1.0 && p_.toDealias && !mask_is_identity(mcut_of(k))) {
      // g^1 = g: the round trip reads g itself
      const T* src = ch.g;
      const RtOut o{0, sp_mask(mcut_of(k)), out};
      rtrip(1, &src, 1, &o);
    } else {
      map(F_pow<T>{ch.g, out, k, 0}, nr_);
      if (p_.toDealias) { T* f = out; dealias_dev(&f, 1, k); }
    }
  } else {                                                     // :2305-2317
    if (!ch.haveG2) {
      ch.G2 = rtmp();
      mul_dealias(ch.g, ch.g, ch.G2, p_.toDealias);
      ch.haveG2 = true;
    }
    const T* prev = chain_pow(ch, k - 2.0);
    mul_dealias(ch.G2, prev, out, p_.toDealias);
  }
  T* r = out.get();
  ch.memo.emplace(key, std::move(out));
  return r;
}

template <typename T> typename Engine<T>::RT Engine<T>::chain_take(Chain& ch, double k) {
  chain_pow(ch, k);
  auto it = ch.memo.find(std::llround(k * 1.0e6));
  RT r = std::move(it->second);
  ch.memo.erase(it);
  return r;
}

template <typename T> void Engine<T>::spec3(const T* u, C* const* uh) {
  const T* src[3] = {u, u + nr_, u + 2 * nr_};
  SpecOp ops[3] = {sp_none(), sp_none(), sp_none()};
  fwd(3, src, uh, ops);
}

template <typename T> void Engine<T>::grad9(C* const* uh, T* const* G) {
  InvSrc is[9];
  SpecOp ops[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      is[i * 3 + j] = InvSrc{{uh[i], nullptr, nullptr}, {1.0, 0, 0}, 1};
      ops[i * 3 + j] = sp_ik(j);
    }
  inv(9, is, ops, G);
}

// G[i*3+j] = d_j u_i of a physical vector field by round trips (no stored
// spectrum: d_j u_i = REAL(IFFT(i K_j FFT(u_i))) with one rounding per pass)
template <typename T> void Engine<T>::grad9_rt(const T* u, T* G) {
  const T* src[3] = {u, u + nr_, u + 2 * nr_};
  RtOut o[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) o[i * 3 + j] = RtOut{i, sp_ik(j), G + (i * 3 + j) * nr_};
  rtrip(3, src, 9, o);
}

template <typename T> double Engine<T>::lq_norm_dev(const T* u) {
  Chain ch;
  chain_init(ch, u);
  const T* a = chain_pow(ch, p_.q / 2.0);                     // :1686
  return std::pow(dot(a, a) / (double)nr_, 1.0 / p_.q);        // :1687-1688
}

// R(u) = (R, R1+R2, R3): calc_local_dLqdt_inclParts, function_ops.f90:1304-1358
template <typename T> void Engine<T>::dLqdt_dev(const T* u, double q, double out[3]) {
  const double nu = p_.visc, vc = p_.viscCoefficient, pc = p_.pressureCoefficient;
  const double dV = 1.0 / (double)nr_;
  const bool D = p_.toDealias != 0;
  RT Gb = rtmp(9);                                               // G[i*3+j] = d_j u_i
  T* Gp = Gb.get();
  grad9_rt(u, Gp);
  RT s = rtmp();
  map(F_dot3self<T>{u, nr_, s}, nr_);
  if (D) { T* f = s; dealias_dev(&f, 1, 2.0); }
  Chain ch;
  chain_init(ch, u, s);
  const T* A = chain_pow(ch, q - 2.0);                          // :1314
  RT X = rtmp();
  map(F_gradsq<T>{Gp, nr_, X}, nr_);                             // :1316
  if (D) { T* f = X; dealias_dev(&f, 1, 2.0); }
  const double R1 = -vc * nu * dot(A, X) * dV;                   // :1318
  const bool ide = D && ((!use_e_u_) || p_.dealiase_if_mult_by_e_u);
  RT w0 = rtmp(), w1 = rtmp(), w2 = rtmp();
  RT eb;                                                         // e_u (3 comps) when used
  const T* v = u;
  if (use_e_u_) {
    eb = rtmp(3);
    map(F_unit<T>{u, s, nr_, eb}, nr_);                          // :1323 (|u| from D(u.u))
    v = eb;
  }
  map(F_vgradu<T>{v, Gp, nr_, w0, w1, w2}, nr_);
  if (ide) { T* f[3] = {w0, w1, w2}; dealias_dev(f, 3, 2.0); }
  RT Y = rtmp();
  map(F_sumsq3<T>{w0, w1, w2, Y}, nr_);
  if (D) { T* f = Y; dealias_dev(&f, 1, 2.0); }
  const T* Bq = use_e_u_ ? A : chain_pow(ch, q - 4.0);           // :1331
  const double R2 = -vc * (q - 2.0) * nu * dot(Bq, Y) * dV;      // :1328/:1333
  // pressure: p = solve_poisson(-D(grad u : grad u^T))            :1339-1340
  RT Z = rtmp();
  map(F_gradgradT<T>{Gp, nr_, Z}, nr_);
  RT pr = rtmp();
  {
    SpecOp op;
    op.flags = (D ? SP_MASK : 0) | SP_INVKSQ;
    op.mcut = mcut_of(2.0);
    op.scale = -1.0;
    const T* z = Z;
    const RtOut o{0, op, pr};
    rtrip(1, &z, 1, &o);
  }
  RT P2 = rtmp();
  mul_dealias(pr, Bq, P2, D);                                     // :1343/:1349
  // v.(v.grad)u with the same inner products w_j                   :1345/:1351
  RT V = rtmp();
  map(F_dot3<T>{v, nr_, w0, w1, w2, V}, nr_);
  if (ide) { T* f = V; dealias_dev(&f, 1, 2.0); }
  const double R3 = pc * (q - 2.0) * dot(P2, V) * dV;             // :1346/:1352
  out[0] = R1 + R2 + R3;
  out[1] = R1 + R2;
  out[2] = R3;
}

// L2 gradient of R: GradL2ForLq, function_ops.f90:1066-1211.  Written for a
// small footprint: the velocity gradient (9 fields) lives only while the
// pressure source is formed; the transposed-gradient term recomputes it three
// columns at a time, and the dealiased products overwrite their inputs.
template <typename T> void Engine<T>::gradL2_dev(const T* u, T* const* Vout) {
  const double q = p_.q, nu = p_.visc, vc = p_.viscCoefficient, pc = p_.pressureCoefficient;
  const bool D = p_.toDealias != 0;
  ST uh[3] = {stmp(), stmp(), stmp()};
  C* uhp[3] = {uh[0], uh[1], uh[2]};
  spec3(u, uhp);
  RT s = rtmp();
  map(F_dot3self<T>{u, nr_, s}, nr_);
  if (D) { T* f = s; dealias_dev(&f, 1, 2.0); }
  Chain ch;
  chain_init(ch, u, s);
  const T* A = chain_pow(ch, q - 2.0);                          // :1102
  // pressure and its gradient                                       :1105-1109
  ST ph = stmp();
  {
    RT Gb = rtmp(9);
    T* G[9];
    for (int a = 0; a < 9; ++a) G[a] = Gb.get() + a * nr_;
    grad9(uhp, G);
    RT Z = rtmp();
    map(F_gradgradT<T>{Gb.get(), nr_, Z}, nr_);
    SpecOp op;
    op.flags = (D ? SP_MASK : 0) | SP_INVKSQ;
    op.mcut = mcut_of(2.0);
    op.scale = -1.0;
    fwd1(Z, ph, op);
  }
  RT gradP = rtmp(3), DeltaU = rtmp(3);
  {
    InvSrc is[6];
    SpecOp ops[6];
    T* dst[6];
    for (int j = 0; j < 3; ++j) {
      is[j] = InvSrc{{ph, nullptr, nullptr}, {1.0, 0, 0}, 1};
      ops[j] = sp_ik(j);
      dst[j] = gradP + j * nr_;
      is[3 + j] = InvSrc{{uhp[j], nullptr, nullptr}, {1.0, 0, 0}, 1};
      SpecOp o; o.flags = SP_NEGKSQ; ops[3 + j] = o;            // :1111-1112
      dst[3 + j] = DeltaU + j * nr_;
    }
    inv(6, is, ops, dst);
  }
  ph = ST();
  RT aux3v = rtmp(3);
  {
    RT aux2 = rtmp();
    if (use_e_u_) {
      RT eb = rtmp(3);
      T* ev = eb;
      map(F_unit<T>{u, s, nr_, ev}, nr_);                          // :1117
      map(F_udotforce<T>{ev, DeltaU, gradP, nr_, (T)(vc * nu), (T)pc, aux2}, nr_);
      if (D && p_.dealiase_if_mult_by_e_u) { T* f = aux2; dealias_dev(&f, 1, 2.0); }
      mul_dealias(A, aux2, aux2, D);
      {
        const T* a3[3] = {aux2, aux2, aux2};
        const T* b3[3] = {ev, ev + nr_, ev + 2 * nr_};
        T* o3[3] = {aux3v, aux3v + nr_, aux3v + 2 * nr_};
        mul_dealias(3, a3, b3, o3, D && p_.dealiase_if_mult_by_e_u);
      }
    } else {
      map(F_udotforce<T>{u, DeltaU, gradP, nr_, (T)(vc * nu), (T)pc, aux2}, nr_);  // :1131-1133
      if (D) { T* f = aux2; dealias_dev(&f, 1, 2.0); }
      const T* Bq = chain_pow(ch, q - 4.0);                        // :1136
      mul_dealias(Bq, aux2, aux2, D);
      {
        const T* a3[3] = {aux2, aux2, aux2};
        const T* b3[3] = {u, u + nr_, u + 2 * nr_};
        T* o3[3] = {aux3v, aux3v + nr_, aux3v + 2 * nr_};
        mul_dealias(3, a3, b3, o3, D);
      }
    }
  }
  // grad of Delta^-1 div(|u|^(q-2)u) and Delta of D(|u|^(q-2)u)       :1150-1158, 1172-1173
  RT gphi = rtmp(3), aux7 = rtmp(3);
  {
    ST W[3] = {stmp(), stmp(), stmp()};
    C* Wp[3] = {W[0], W[1], W[2]};
    {
      RT t0 = rtmp(3);
      T* tp[3] = {t0, t0 + nr_, t0 + 2 * nr_};
      for (int i = 0; i < 3; ++i) map(F_mul<T>{A, u + i * nr_, tp[i]}, nr_);
      SpecOp m = D ? sp_mask(mcut_of(2.0)) : sp_none();
      SpecOp ops[3] = {m, m, m};
      const T* src[3] = {tp[0], tp[1], tp[2]};
      fwd(3, src, Wp, ops);
    }
    ST dv = stmp();
    smap(S_div<T>{Wp[0], Wp[1], Wp[2], dv, n_, pp_});
    InvSrc is[6];
    SpecOp ops[6];
    T* dst[6];
    for (int j = 0; j < 3; ++j) {
      is[j] = InvSrc{{dv, nullptr, nullptr}, {1.0, 0, 0}, 1};
      SpecOp o; o.flags = SP_IK | SP_INVKSQ; o.dim = j; o.scale = 1.0;
      ops[j] = o;
      dst[j] = gphi + j * nr_;
      is[3 + j] = InvSrc{{Wp[j], nullptr, nullptr}, {1.0, 0, 0}, 1};
      SpecOp l; l.flags = SP_NEGKSQ; ops[3 + j] = l;
      dst[3 + j] = aux7 + j * nr_;
    }
    inv(6, is, ops, dst);
  }
  // aux5_c = sum_a d_c u_a gphi_a (:1165-1169): column c of the gradient
  // recomputed from the spectrum of u
  RT aux5 = rtmp(3);
  {
    RT col = rtmp(3);
    for (int c = 0; c < 3; ++c) {
      InvSrc is[3];
      SpecOp ops[3];
      T* dst[3];
      for (int a = 0; a < 3; ++a) {
        is[a] = InvSrc{{uhp[a], nullptr, nullptr}, {1.0, 0, 0}, 1};
        ops[a] = sp_ik(c);
        dst[a] = col + a * nr_;
      }
      inv(3, is, ops, dst);
      map(F_dot3<T>{col, nr_, gphi, gphi + nr_, gphi + 2 * nr_, aux5 + c * nr_}, nr_);
    }
  }
  if (D) { T* f[3] = {aux5, aux5 + nr_, aux5 + 2 * nr_}; dealias_dev(f, 3, 2.0); }
  // aux4 = D(|u|^(q-2) grad p), aux6 = D(|u|^(q-2) Delta u)          :1178-1185
  // (formed in place of grad p and Delta u)
  for (int i = 0; i < 3; ++i) {
    map(F_mul<T>{A, gradP + i * nr_, gradP + i * nr_}, nr_);
    map(F_mul<T>{A, DeltaU + i * nr_, DeltaU + i * nr_}, nr_);
  }
  if (D) {
    T* f[6] = {gradP, gradP + nr_, gradP + 2 * nr_, DeltaU, DeltaU + nr_, DeltaU + 2 * nr_};
    dealias_dev(f, 6, 2.0);
  }
  for (int i = 0; i < 3; ++i)
    map(F_gradcombine<T>{aux3v + i * nr_, gradP + i * nr_, aux5 + i * nr_, DeltaU + i * nr_,
                         aux7 + i * nr_, Vout[i], (T)(q - 2.0), (T)pc, (T)(vc * nu)}, nr_);
}

// ---------------------------------------------------------------------------
// IEngine operators on field handles
// ---------------------------------------------------------------------------
template <typename T> void Engine<T>::gradient(int u_sc, int out_vec) {
  ST h = stmp();
  fwd1(fphys(u_sc), h);
  InvSrc is[3];
  SpecOp ops[3];
  T* dst[3];
  for (int j = 0; j < 3; ++j) {
    is[j] = InvSrc{{h, nullptr, nullptr}, {1.0, 0, 0}, 1};
    ops[j] = sp_ik(j);
    dst[j] = fphys(out_vec, j);
  }
  inv(3, is, ops, dst);
}

template <typename T> void Engine<T>::divergence(int u_vec, int out_sc) {
  ST uh[3] = {stmp(), stmp(), stmp()};
  C* uhp[3] = {uh[0], uh[1], uh[2]};
  spec3(fphys(u_vec), uhp);
  ST dv = stmp();
  smap(S_div<T>{uhp[0], uhp[1], uhp[2], dv, n_, pp_});
  inv1(dv, fphys(out_sc));
}

template <typename T> void Engine<T>::laplacian(int f_vec) {
  ST uh[3] = {stmp(), stmp(), stmp()};
  C* uhp[3] = {uh[0], uh[1], uh[2]};
  spec3(fphys(f_vec), uhp);
  InvSrc is[3];
  SpecOp ops[3];
  T* dst[3];
  for (int j = 0; j < 3; ++j) {
    is[j] = InvSrc{{uhp[j], nullptr, nullptr}, {1.0, 0, 0}, 1};
    SpecOp o; o.flags = SP_NEGKSQ; ops[j] = o;
    dst[j] = fphys(f_vec, j);
  }
  inv(3, is, ops, dst);
}

template <typename T> void Engine<T>::solve_poisson(int f_sc, double l, int out_sc) {
  ST h = stmp();
  SpecOp op;
  op.flags = SP_INVKSQ;
  op.scale = 1.0 / (l * l);    // phi_hat = -f_hat / (l^2 ksq)  (:1980)
  fwd1(fphys(f_sc), h, op);
  inv1(h, fphys(out_sc));
}

template <typename T> void Engine<T>::dealias_scalar(int g_sc, double order) {
  T* f = fphys(g_sc);
  dealias_dev(&f, 1, order);
}

template <typename T> void Engine<T>::divAvg_free(int f_vec) {
  ST uh[3] = {stmp(), stmp(), stmp()};
  C* uhp[3] = {uh[0], uh[1], uh[2]};
  spec3(fphys(f_vec), uhp);
  smap(S_leray<T>{uhp[0], uhp[1], uhp[2], n_, pp_});
  InvSrc is[3];
  SpecOp ops[3] = {sp_none(), sp_none(), sp_none()};
  T* dst[3];
  for (int j = 0; j < 3; ++j) {
    is[j] = InvSrc{{uhp[j], nullptr, nullptr}, {1.0, 0, 0}, 1};
    dst[j] = fphys(f_vec, j);
  }
  inv(3, is, ops, dst);
}

template <typename T> void Engine<T>::HilbertGradient(int f_vec, double order) {
  ST uh[3] = {stmp(), stmp(), stmp()};
  C* uhp[3] = {uh[0], uh[1], uh[2]};
  spec3(fphys(f_vec), uhp);
  smap(S_hilbert3<T>{uhp[0], uhp[1], uhp[2], n_, pp_, p_.lambda1, 2.0 * order});
  InvSrc is[3];
  SpecOp none[3] = {sp_none(), sp_none(), sp_none()};
  T* dst[3];
  for (int j = 0; j < 3; ++j) {
    is[j] = InvSrc{{uhp[j], nullptr, nullptr}, {1.0, 0, 0}, 1};
    dst[j] = fphys(f_vec, j);
  }
  inv(3, is, none, dst);
}

template <typename T> void Engine<T>::calc0thFourierModes(int v, double out[3]) {
  for (int c = 0; c < 3; ++c) out[c] = sum(fphys(v, c));   // unnormalised k=0 coefficient
}

// calculate_spectral_data (function_ops.f90:2614-2677): shell sums of
// |fftfwd(u_c)|^2 over the three components; out is the reference's
// spectral_data(1:n, 1:2) in Fortran order (out[i] = 2 pi (i+1), out[n+i] =
// shell i), scaled by ||u||_2^2 / (2 pi sum) when normalize
// (normalizeSpectrumByL2Norm, :2667-2671).
template <typename T> void Engine<T>::spectral_data(int u_vec, int normalize, double* out) {
  if (fld(u_vec).ncomp != 3) throw Error(MLQ_EINVAL, "spectral data of a vector field");
  spectral_data_dev(fphys(u_vec), normalize != 0, out);
}
template <typename T>
void Engine<T>::spectral_data_dev(const T* u, bool normalize, double* out) {
  const int n = n_;
  ST s = stmp(3);
  C* d[3] = {s.get(), s.get() + ns_, s.get() + 2 * ns_};
  const T* src[3] = {u, u + nr_, u + 2 * nr_};
  const SpecOp ops[3] = {sp_none(), sp_none(), sp_none()};
  fwd(3, src, d, ops);
  constexpr long long MAXBLK = 2048;
  const long long chunk0 = 256LL * 32;
  long long nblk = std::min(MAXBLK, (ns_ + chunk0 - 1) / chunk0);
  long long chunk = (ns_ + nblk - 1) / nblk;
  chunk = (chunk + 255) / 256 * 256;
  nblk = (ns_ + chunk - 1) / chunk;
  const size_t pbytes = sizeof(double) * (size_t)nblk * (size_t)n + sizeof(double) * n;
  double* part = (double*)dmalloc(pbytes);
  double* shells = part + nblk * n;
  prof_begin(PC_REDUCE);
  hipLaunchKernelGGL((k_shell_sum<T>), dim3((unsigned)nblk), dim3(256), sizeof(double) * n, st_,
                     d[0], d[1], d[2], n, pp_, ns_, chunk, part);
  MLQ_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_shell_final, dim3(n), dim3(256), 0, st_, part, (int)nblk, n, shells);
  MLQ_HIP(hipGetLastError());
  prof_end(PC_REDUCE, 3.0 * spec_bytes());
  std::vector<double> h(n);
  MLQ_HIP(hipMemcpyAsync(h.data(), shells, sizeof(double) * n, hipMemcpyDeviceToHost, st_));
  MLQ_HIP(hipStreamSynchronize(st_));
  dfree(part, pbytes);
  double tot = 0.0;
  for (int i = 0; i < n; ++i) {
    out[i] = 2.0 * (4.0 * std::atan2(1.0, 1.0)) * (double)(i + 1);   // kk_max, :2643
    out[n + i] = h[i];
    tot += h[i];
  }
  if (normalize) {   // :2667-2671
    double l2 = 0.0;   // sum_c ||u_c||^2 / nr: one reduction over the 3 contiguous components
    reduce<1>(R_dot<T>{u, u}, 3 * nr_, &l2);
    l2 /= (double)nr_;
    const double spectral_ener = 2.0 * (4.0 * std::atan2(1.0, 1.0)) * tot;
    for (int i = 0; i < n; ++i) out[n + i] = l2 / spectral_ener * out[n + i];
  }
}

// calc_global_summed_enstrophy (function_ops.f90:2455-2467): Enstrophy
// (:2422-2442) of the vorticity from vel2vort (:2982-3042), 1/2 sum_c
// ||w_c||^2 dV with w = REAL(ifft(i K x u_hat)) (i K_j zero at the own-dimension
// Nyquist, which REAL() discards).
template <typename T> double Engine<T>::enstrophy_dev(const T* u) {
  ST uh[3] = {stmp(), stmp(), stmp()};
  C* uhp[3] = {uh[0], uh[1], uh[2]};
  spec3(u, uhp);
  // w_x = d_y u_z - d_z u_y, w_y = d_z u_x - d_x u_z, w_z = d_x u_y - d_y u_x
  static const int comp[6] = {2, 1, 0, 2, 1, 0}, dim[6] = {1, 2, 2, 0, 0, 1};
  RT dw = rtmp(6);
  InvSrc is[6];
  SpecOp ops[6];
  T* dst[6];
  for (int k = 0; k < 6; ++k) {
    is[k] = InvSrc{{uhp[comp[k]], nullptr, nullptr}, {1.0, 0, 0}, 1};
    ops[k] = sp_ik(dim[k]);
    dst[k] = dw + (long long)k * nr_;
  }
  inv(6, is, ops, dst);
  double e = 0.0;
  for (int c = 0; c < 3; ++c) {
    T* w = dw + (long long)(2 * c) * nr_;
    map(F_axpby<T>{w, w + nr_, w, (T)1, (T)-1}, nr_);
    e += 0.5 * dot(w, w) / (double)nr_;
  }
  return e;
}
// calc_global_dEnstrophydt (function_ops.f90:1417-1481):
//   R_1 = -viscCoefficient visc ||Lap u||_2^2,  R_2 = <D((u.grad) u), Lap u>_L2
// with gradU(i, j) = d_j u_i from gradient() and D the order-2 dealias
// (toDealias); one component of grad u at a time.
template <typename T> double Engine<T>::dEnstrophydt_dev(const T* u) {
  ST uh[3] = {stmp(), stmp(), stmp()};
  C* uhp[3] = {uh[0], uh[1], uh[2]};
  spec3(u, uhp);
  RT lap = rtmp(3), adv = rtmp(3);
  {
    InvSrc is[3];
    SpecOp ops[3];
    T* dst[3];
    for (int j = 0; j < 3; ++j) {
      is[j] = InvSrc{{uhp[j], nullptr, nullptr}, {1.0, 0, 0}, 1};
      SpecOp o;
      o.flags = SP_NEGKSQ;
      ops[j] = o;
      dst[j] = lap + (long long)j * nr_;
    }
    inv(3, is, ops, dst);
  }
  double r1 = 0.0;
  for (int c = 0; c < 3; ++c) {
    const T* l = lap + (long long)c * nr_;
    r1 += dot(l, l) / (double)nr_;
  }
  r1 *= -p_.viscCoefficient * p_.visc;
  {
    RT g = rtmp(3);
    for (int i = 0; i < 3; ++i) {
      InvSrc is[3];
      SpecOp ops[3];
      T* dst[3];
      for (int j = 0; j < 3; ++j) {
        is[j] = InvSrc{{uhp[i], nullptr, nullptr}, {1.0, 0, 0}, 1};
        ops[j] = sp_ik(j);
        dst[j] = g + (long long)j * nr_;
      }
      inv(3, is, ops, dst);
      map(F_udotgrad<T>{u, g, nr_, adv + (long long)i * nr_}, nr_);
    }
  }
  if (p_.toDealias) {
    T* a[3] = {adv, adv + nr_, adv + 2 * nr_};
    dealias_dev(a, 3, 2.0);
  }
  double r2 = 0.0;
  for (int c = 0; c < 3; ++c)
    r2 += dot(adv + (long long)c * nr_, lap + (long long)c * nr_) / (double)nr_;
  return r1 + r2;
}
template <typename T> double Engine<T>::enstrophy(int u_vec) {
  if (fld(u_vec).ncomp != 3) throw Error(MLQ_EINVAL, "enstrophy of a vector field");
  return enstrophy_dev(fphys(u_vec));
}
template <typename T> double Engine<T>::dEnstrophydt(int u_vec) {
  if (fld(u_vec).ncomp != 3) throw Error(MLQ_EINVAL, "dEnstrophydt of a vector field");
  return dEnstrophydt_dev(fphys(u_vec));
}

template <typename T> void Engine<T>::calc_uk(int u_vec, double k, int out_sc) {
  Chain ch;
  chain_init(ch, fphys(u_vec));
  const T* r = chain_pow(ch, k);
  map(F_copy<T>{r, fphys(out_sc)}, nr_);
}

template <typename T> void Engine<T>::calc_gk_order2(int g_sc, double k, int out_sc) {
  Chain ch;
  ch.g = rtmp();
  map(F_copy<T>{fphys(g_sc), ch.g}, nr_);
  const T* r = chain_pow(ch, k);
  map(F_copy<T>{r, fphys(out_sc)}, nr_);
}

template <typename T> double Engine<T>::Lq_norm(int u_vec) { return lq_norm_dev(fphys(u_vec)); }

template <typename T> void Engine<T>::rescale(int u_vec, double B) {
  const double L = lq_norm_dev(fphys(u_vec));
  map(F_scale<T>{fphys(u_vec), (T)(B / L)}, nr_ * 3);
}

template <typename T> void Engine<T>::constraint_derivative(int u_vec, int out_vec) {
  const double q = p_.q;
  const T* u = fphys(u_vec);
  Chain ch;
  chain_init(ch, u);
  const double nrm = std::pow(dot(chain_pow(ch, q / 2.0), chain_pow(ch, q / 2.0)) / (double)nr_,
                              1.0 / q);                                     // :993
  const T* A = chain_pow(ch, q - 2.0);                                      // :995
  T* o[3] = {fphys(out_vec, 0), fphys(out_vec, 1), fphys(out_vec, 2)};
  {
    const T* a3[3] = {A, A, A};
    const T* b3[3] = {u, u + nr_, u + 2 * nr_};
    mul_dealias(3, a3, b3, o, p_.toDealias && (q - 2.0 > 2.0e-16));
  }
  map(F_scale<T>{o[0], (T)std::pow(nrm, 1.0 - q)}, nr_ * 3);               // :1003
}

template <typename T>
void Engine<T>::vector_transport(int u, double q, double B, int eta, int xi, int out) {
  RT up[3] = {rtmp(), rtmp(), rtmp()};
  RT upb = rtmp(3);
  T* upe = upb.get();
  map(F_axpby<T>{fphys(u), fphys(eta), upe, (T)1, (T)1}, nr_ * 3);        // :1733
  Chain ch;
  chain_init(ch, upe);
  const T* A = chain_pow(ch, q - 2.0);                                     // :1735
  T* aux[3] = {up[0], up[1], up[2]};
  {
    const T* a3[3] = {A, A, A};
    const T* b3[3] = {upe, upe + nr_, upe + 2 * nr_};
    mul_dealias(3, a3, b3, aux, p_.toDealias && (q - 2.0 > 2.0e-16));   // :1739
  }
  double ir = 0.0;
  for (int i = 0; i < 3; ++i) ir += dot(aux[i], fphys(xi, i)) / (double)nr_;   // :1742
  const T* a2 = chain_pow(ch, q / 2.0);
  const double L = std::pow(dot(a2, a2) / (double)nr_, 1.0 / q);          // :1744
  map(F_transport<T>{fphys(xi), upe, fphys(out), (T)(std::pow(L, -q) * ir), (T)(B / L)},
      nr_ * 3);                                                            // :1746-1747
}

template <typename T> double Engine<T>::inner(int f, int g, int kind) {
  if (kind == MLQ_IP_OLD_H || kind == MLQ_IP_OLD_HL)
    throw Error(MLQ_EBADKIND, "old inner product exponent");             // :3485-3493
  if (kind < 0 || kind > 2) throw Error(MLQ_EBADKIND, "inner product type not defined");
  const int nc = fld(f).ncomp;
  if (fld(g).ncomp != nc) throw Error(MLQ_EINVAL, "component mismatch");
  double s = 0.0;
  if (kind == MLQ_IP_L2) {
    for (int c = 0; c < nc; ++c) s += dot(fphys(f, c), fphys(g, c)) / (double)nr_;
    return s;
  }
  // Sobolev products by Parseval: (1/N^6) sum_k w(k)^2 Re(f_k conj g_k)
  ST a = stmp(), b = stmp();
  const int sk = kind == MLQ_IP_HL ? 1 : 2;
  for (int c = 0; c < nc; ++c) {
    fwd1(fphys(f, c), a);
    fwd1(fphys(g, c), b);
    s += spec_dot(a, b, sk) / ((double)nr_ * (double)nr_);
  }
  return s;
}

template <typename T> void Engine<T>::dLqdt_parts(int u_vec, double q, double out[3]) {
  dLqdt_dev(fphys(u_vec), q, out);
  if (std::isnan(out[0]) || std::isnan(out[1]) || std::isnan(out[2]))
    throw Error(MLQ_ENAN, "nan in R(u)");
}

template <typename T> void Engine<T>::grad_L2(int u_vec, int out_vec) {
  T* o[3] = {fphys(out_vec, 0), fphys(out_vec, 1), fphys(out_vec, 2)};
  gradL2_dev(fphys(u_vec), o);
}

template <typename T> double Engine<T>::eval_J(int f_vec, int system) {
  double parts[3];
  switch (system) {
    case MLQ_J_MAXDLQDT:
      dLqdt_dev(fphys(f_vec), p_.q, parts);
      if (std::isnan(parts[0])) throw Error(MLQ_ENAN, "nan in R(u)");
      return parts[0];
    case MLQ_J_LINEMIN: {
      RT ab = rtmp(3);
      T* a = ab.get();
      map(F_copy<T>{fphys(f_vec), a}, nr_ * 3);                       // :531-532
      const double L = lq_norm_dev(a);
      map(F_scale<T>{a, (T)(p_.B / L)}, nr_ * 3);                    // :538
      dLqdt_dev(a, p_.q, parts);
      if (std::isnan(parts[0])) throw Error(MLQ_ENAN, "nan in R(u)");
      return -1.0 * parts[0];                                        // :539
    }
    case MLQ_J_LQNORM:
      return lq_norm_dev(fphys(f_vec));
    default:
      return 0.0;
  }
}

template <typename T>
void Engine<T>::eval_J_batch_ref(int u, int d, const double* taus, int ntau, double* J) {
  RT pb = rtmp(3);
  T* phi = pb.get();
  for (int i = 0; i < ntau; ++i) {
    map(F_axpby<T>{fphys(u), fphys(d), phi, (T)1, (T)taus[i]}, nr_ * 3);
    const double L = lq_norm_dev(phi);
    map(F_scale<T>{phi, (T)(p_.B / L)}, nr_ * 3);
    double parts[3];
    dLqdt_dev(phi, p_.q, parts);
    J[i] = -parts[0];
  }
}

template <typename T> void Engine<T>::seeded(int f_vec, unsigned long long seed, int kind, double a) {
  const unsigned long long base = seed * 0xD1B54A32D192ED03ULL;
  for (int c = 0; c < 3; ++c) map(F_noise<T>{fphys(f_vec, c), base, nr_, c}, nr_);
  ST uh[3] = {stmp(), stmp(), stmp()};
  C* uhp[3] = {uh[0], uh[1], uh[2]};
  spec3(fphys(f_vec), uhp);
  smap(S_envelope_leray<T>{uhp[0], uhp[1], uhp[2], n_, pp_, kind, a});
  InvSrc is[3];
  SpecOp ops[3] = {sp_none(), sp_none(), sp_none()};
  T* dst[3];
  for (int j = 0; j < 3; ++j) {
    is[j] = InvSrc{{uhp[j], nullptr, nullptr}, {1.0, 0, 0}, 1};
    dst[j] = fphys(f_vec, j);
  }
  inv(3, is, ops, dst);
  rescale(f_vec, p_.B);
}

// complex 3-D transforms of host arrays via two real transforms (real and
// imaginary parts), expanded to the full spectrum with Hermitian symmetry
template <typename T> void Engine<T>::fft_c2c(const double* in, double* out, bool inverse, int ncomp) {
  const long long NN = nr_;
  std::vector<double> re(NN), im(NN), hre(2 * ns_), him(2 * ns_);
  for (int c = 0; c < ncomp; ++c) {
    const double* src = in + 2 * NN * c;
    double* dst = out + 2 * NN * c;
    if (!inverse) {
      for (long long i = 0; i < NN; ++i) { re[i] = src[2 * i]; im[i] = src[2 * i + 1]; }
      fft_r2c_host(re.data(), hre.data());
      fft_r2c_host(im.data(), him.data());
      const int n = n_;
      for (int kz = 0; kz < n; ++kz)
        for (int ky = 0; ky < n; ++ky)
          for (int kx = 0; kx < n; ++kx) {
            double ar, ai, br, bi;
            if (kx <= n / 2) {
              const long long h = ((long long)kz * n + ky) * px_ + kx;
              ar = hre[2 * h]; ai = hre[2 * h + 1]; br = him[2 * h]; bi = him[2 * h + 1];
            } else {
              const int cz = (n - kz) % n, cy = (n - ky) % n, cx = n - kx;
              const long long h = ((long long)cz * n + cy) * px_ + cx;
              ar = hre[2 * h]; ai = -hre[2 * h + 1]; br = him[2 * h]; bi = -him[2 * h + 1];
            }
            const long long o = ((long long)kz * n + ky) * n + kx;
            dst[2 * o] = ar - bi;
            dst[2 * o + 1] = ai + br;
          }
    } else {
      // inverse of a general complex spectrum: split F = A + iB with A, B the
      // Hermitian parts, A(k) = (F(k) + conj F(-k))/2, B(k) = (F(k) - conj F(-k))/(2i)
      const int n = n_;
      for (int kz = 0; kz < n; ++kz)
        for (int ky = 0; ky < n; ++ky)
          for (int kx = 0; kx <= n / 2; ++kx) {
            const long long o = ((long long)kz * n + ky) * n + kx;
            const long long m = ((long long)((n - kz) % n) * n + (n - ky) % n) * n + (n - kx) % n;
            const double fr = src[2 * o], fi = src[2 * o + 1], gr = src[2 * m], gi = -src[2 * m + 1];
            const long long h = ((long long)kz * n + ky) * px_ + kx;
            hre[2 * h] = 0.5 * (fr + gr); hre[2 * h + 1] = 0.5 * (fi + gi);
            him[2 * h] = 0.5 * (fi - gi); him[2 * h + 1] = -0.5 * (fr - gr);
          }
      fft_c2r_host(hre.data(), re.data());
      fft_c2r_host(him.data(), im.data());
      for (long long i = 0; i < NN; ++i) { dst[2 * i] = re[i]; dst[2 * i + 1] = im[i]; }
    }
  }
}

template <typename T> void Engine<T>::fft_r2c_host(const double* in, double* out) {
  RT r = rtmp();
  ST s = stmp();
  std::vector<T> tr(nr_);
  for (long long i = 0; i < nr_; ++i) tr[i] = (T)in[i];
  MLQ_HIP(hipMemcpyAsync(r.get(), tr.data(), sizeof(T) * nr_, hipMemcpyHostToDevice, st_));
  fwd1(r, s);
  const long long nh = (long long)n_ * n_ * px_;   // host layout: rows of n/2+1
  std::vector<C> ts(nh);
  MLQ_HIP(hipMemcpy2DAsync(ts.data(), sizeof(C) * px_, s.get(), sizeof(C) * pp_, sizeof(C) * px_,
                           (size_t)n_ * n_, hipMemcpyDeviceToHost, st_));
  sync();
  for (long long i = 0; i < nh; ++i) { out[2 * i] = ts[i].x; out[2 * i + 1] = ts[i].y; }
}

template <typename T> void Engine<T>::fft_c2r_host(const double* in, double* out) {
  RT r = rtmp();
  ST s = stmp();
  const long long nh = (long long)n_ * n_ * px_;   // host layout: rows of n/2+1
  std::vector<C> ts(nh);
  for (long long i = 0; i < nh; ++i) { ts[i].x = (T)in[2 * i]; ts[i].y = (T)in[2 * i + 1]; }
  MLQ_HIP(hipMemcpy2DAsync(s.get(), sizeof(C) * pp_, ts.data(), sizeof(C) * px_, sizeof(C) * px_,
                           (size_t)n_ * n_, hipMemcpyHostToDevice, st_));
  inv1(s, r);
  std::vector<T> tr(nr_);
  MLQ_HIP(hipMemcpyAsync(tr.data(), r.get(), sizeof(T) * nr_, hipMemcpyDeviceToHost, st_));
  sync();
  for (long long i = 0; i < nr_; ++i) out[i] = tr[i];
}

}  // namespace mlq
