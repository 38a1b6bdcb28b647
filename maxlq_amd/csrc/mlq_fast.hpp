// Resident-state RCG iteration (the fast path of maxdLqdt, q >= 4 without the
// e_u variant).  Same mathematics as optimization.f90:7-514 with the
// reference operators; what changes is where the work happens:
//
//  * u, grad u, D(u.u), |u|^(q-2), the pressure spectrum, the spectra of
//    D(|u|^(q-2) u) and of |grad u|^2 are kept in HBM across iterations; after
//    the step they follow from the line-search fields by linear combination
//    and by the homogeneity of every operator under u -> s u.
//  * Gradients, Hilbert weights, Leray projections, H_l inner products, the
//    CG direction and the vector transport stay in spectral space.
//  * For phi = u + tau d every dealiased quadratic quantity of the line
//    search (D(phi.phi), w_j = D(sum_i phi_i d_j phi_i), the pressure,
//    |grad phi|^2) is a quadratic polynomial in tau whose coefficient fields
//    are computed once per iteration; J(tau) then needs only the nonlinear
//    |phi|^k chain (for q = 4 a single forward transform of g^2, with R2 and
//    R3 polynomials in tau and R1, ||phi||_q^q spectral sums) and the
//    retraction B phi/||phi||_q enters through R(s phi) = s^q (R1+R2)(phi) +
//    s^(q+1) R3(phi).
#pragma once
#include <chrono>
#include <map>
#include <vector>

#include "mlq_fast_kernels.hpp"
#include "mlq_linesearch.hpp"
#include "mlq_ops_core.hpp"

namespace mlq {

template <typename T> class FastRcg {
  using E_ = Engine<T>;
  using C = c2<T>;
  using RT = typename E_::RT;
  using ST = typename E_::ST;
  using InvSrc = typename E_::InvSrc;

 public:
  FastRcg(E_& e) : e_(e), p_(e.p_) {
    q_ = p_.q;
    B_ = p_.B;
    nuvc_ = p_.visc * p_.viscCoefficient;
    pc_ = p_.pressureCoefficient;
    n_ = e.n_;
    pp_ = e.pp_;
    nr_ = e.nr_;
    ns_ = e.ns_;
    M_ = e.mcut_of(2.0);
    q4_ = std::fabs(q_ - 4.0) < 1e-14;
    poly_ok_ = q4_ && M_ < n_ / 2;
    q6_ = std::fabs(q_ - 6.0) < 1e-14 && M_ < n_ / 2;
    N3_ = (double)nr_;
    N6_ = N3_ * N3_;
    nkc_ = kept_calls<T>(M_);
    // weights of |K| = 2 pi |m| by |m|^2: Hilbert 1/(1 + (lam |K|)^(2s)) and the
    // squared H_l weight (1 + (lam |K|)^s)^2, s = 3/2 - 1/q (function_ops.f90:3611, 3449)
    const int m2max = 3 * (n_ / 2) * (n_ / 2);
    std::vector<double> hw(m2max + 1), hl(m2max + 1);
    const double sx = 1.5 - 1.0 / q_, lam = p_.lambda1;
    for (int m2 = 0; m2 <= m2max; ++m2) {
      const double K = TWO_PI * std::sqrt((double)m2);
      hw[m2] = 1.0 / (1.0 + std::pow(lam * K, 2.0 * sx));
      const double w = 1.0 + std::pow(lam * K, sx);
      hl[m2] = w * w;
    }
    tab_bytes_ = sizeof(double) * (size_t)(m2max + 1);
    hlt_ = (double*)e_.dmalloc(tab_bytes_);
    MLQ_HIP(hipMemcpy(hlt_, hl.data(), tab_bytes_, hipMemcpyHostToDevice));
    // both weights of a mode in one 16-byte entry (one gather per mode)
    std::vector<double2> hwl(m2max + 1);
    for (int m2 = 0; m2 <= m2max; ++m2) hwl[m2] = make_double2(hw[m2], hl[m2]);
    hwlt_ = (double2*)e_.dmalloc(2 * tab_bytes_);
    MLQ_HIP(hipMemcpy(hwlt_, hwl.data(), 2 * tab_bytes_, hipMemcpyHostToDevice));
  }
  ~FastRcg() {
    try {
      e_.dfree(hlt_, tab_bytes_);
      e_.dfree(hwlt_, 2 * tab_bytes_);
    } catch (...) {
    }
  }
  FastRcg(const FastRcg&) = delete;
  FastRcg& operator=(const FastRcg&) = delete;

  // ---- state from a physical field ----------------------------------------------
  void init_state(const T* u_in) {
    g_pending_ = false;
    // Held like the buffers that later iterations keep alive across the
    // direction update (the deferred update's operand; the previous
    // gradient and direction spectra of the CG recursion) and released by
    // the first direction: the first iteration then reaches the steady-state
    // footprint and later ones allocate nothing (a multi-GB hipMalloc
    // inside an iteration costs ~0.1 s).
    Gd_prev_ = e_.rtmp(9);
    warm_[0] = e_.stmp(3);
    warm_[1] = e_.stmp(3);
    u_ = e_.rtmp(3);
    e_.map(F_copy<T>{u_in, u_}, nr_ * 3);
    uh_ = e_.stmp(3);
    spec3(u_, uh_);
    {
      // <u_hat, u_hat>_H_l inside and outside the kept cube (the update with
      // a = 1, b = 0 rewrites u_hat unchanged)
      double r[2];
      e_.template reduce<2>(R_uh_update<T>{uh_, nullptr, ns_, n_, pp_, M_, true, 1.0, 1.0, 0.0,
                                           hlt_}, ns_, r);
      uu_ = (r[0] + r[1]) / N6_;
      hsum_ = r[1];
      hk_ = 1.0;
    }
    G_ = e_.rtmp(9);
    e_.grad9_rt(u_, G_);   // from u itself: no rounded spectrum in between
    s_ = e_.rtmp();
    e_.map(F_dot3self<T>{u_, nr_, s_}, nr_);
    if (poly_ok_) {
      // D(u.u) with its masked spectrum kept for the line-search polynomials
      S0h_ = e_.stmp();
      e_.fwd1(s_, S0h_, sp_mask(M_));
      InvSrc is{{S0h_.get(), nullptr, nullptr}, {1.0, 0, 0}, 1};
      SpecOp none = sp_none();
      T* dst = s_;
      e_.inv(1, &is, &none, &dst, M_);
    } else {
      dealias1(s_);
    }
    chain_state_from_s();
    // pressure spectrum: M FFT(grad u : grad u^T) / |K|^2  (solve_poisson(-aux))
    ph_ = e_.stmp();
    {
      RT X = e_.rtmp();
      e_.map(F_gradgradT<T>{G_, nr_, X}, nr_);
      e_.fwd1(X, ph_, poisson_op());
    }
    // W_i = M FFT(|u|^(q-2) u_i)
    W_ = e_.stmp(3);
    fwd_Au(A_, u_, W_);
    // w_j = D(sum_i u_i d_j u_i)
    a_ = e_.rtmp(3);
    e_.map(F_vgradu<T>{u_, G_, nr_, a_, a_ + nr_, a_ + 2 * nr_}, nr_);
    dealias3(a_);
    // |grad u|^2: masked spectrum (q = 4) or physical (other q)
    Q0_ = e_.rtmp();
    e_.map(F_gradsq<T>{G_, nr_, Q0_}, nr_);
    Q0h_ = e_.stmp();
    e_.fwd1(Q0_, Q0h_, sp_mask(M_));
  }

  void download_u(T* dst) { e_.map(F_copy<T>{u_, dst}, nr_ * 3); }
  const T* u_field() const { return u_; }
  // the direction d1 = d_ + gam_ u_ (the state keeps its u component as the
  // scalar gam_), between precompute_ls and the step
  RT direction_field() {
    RT d1 = e_.rtmp(3);
    e_.map(F_axpby<T>{d_, u_, d1, (T)1, (T)gam_}, 3 * nr_);
    return d1;
  }
  // |u|^(q-2) of the state (at q = 4 without clipping it is D(u.u) = s_)
  const T* Acur() const { return A_.get() ? A_.get() : s_.get(); }

  // direction d given as a physical field (for batched J evaluations through
  // the C ABI)
  void set_direction_phys(const T* d) {
    dB_ = e_.stmp(3);
    spec3(d, dB_);
    gam_ = 0.0;
    dB_limited_ = false;   // an arbitrary field: full inverse transforms
  }

  // ---- one RCG iteration up to the CG direction (optimization.f90:146-261) --------
  // Returns the message codes raised by the checks.
  struct DirOut { double beta; std::vector<int> msgs; };
  DirOut direction(int iter, int reset_every) {
    DirOut out;
    const double s = 1.5 - 1.0 / q_;
    // the force nu Lap u - pc grad p: one spectral pass combines the Laplacian
    // of u_hat (full spectrum, its pending factor outside the cube) and the
    // gradient of the masked pressure spectrum, one full inverse transform
    RT force = e_.rtmp(3);
    {
      ST fh = e_.stmp(3);
      e_.map(S_force<T>{uh_, ph_, fh, ns_, n_, pp_, M_, nuvc_, pc_, hk_}, ns_, PC_SPECTRAL);
      InvSrc is[3];
      SpecOp ops[3];
      T* dst[3];
      for (int j = 0; j < 3; ++j) {
        is[j] = InvSrc{{fh + j * ns_, nullptr, nullptr}, {1.0, 0, 0}, 1};
        ops[j] = sp_none();
        dst[j] = force + j * nr_;
      }
      e_.inv(3, is, ops, dst);
    }
    // aux2 = D(|u|^(q-4) D(u.(nu Lap u - grad p)))           (:1131-1138)
    RT aux2 = e_.rtmp();
    {
      // D(u.force): the product goes straight into its masked spectrum
      const F_udotforce<T> uf{u_, force, nullptr, nr_, (T)nuvc_, (T)pc_, aux2};
      ST h = e_.stmp();
      C* hd = h;
      const SpecOp op = sp_mask(M_);
      if (e_.fwd_prod(P_udotforce<T>{uf}, &hd, &op)) {
        InvSrc is{{hd, nullptr, nullptr}, {1.0, 0, 0}, 1};
        const SpecOp none = sp_none();
        T* o = aux2;
        e_.inv(1, &is, &none, &o, M_);
      } else {
        e_.map(uf, nr_);
        dealias1(aux2);
      }
    }
    if (!q4_) {
      e_.map(F_mul<T>{Bq_, aux2, aux2}, nr_);
      dealias1(aux2);
    }
    // grad phi, phi = Lap^-1 div D(|u|^(q-2) u)               (:1155-1158)
    RT gphi = e_.rtmp(3);
    {
      ST dv = e_.stmp();
      e_.map(S_div_k<T>{W_, dv, ns_, n_, pp_, M_}, nkc_, PC_SPECTRAL);
      InvSrc is[3];
      SpecOp ops[3];
      T* dst[3];
      for (int j = 0; j < 3; ++j) {
        is[j] = InvSrc{{dv.get(), nullptr, nullptr}, {1.0, 0, 0}, 1};
        SpecOp o; o.flags = SP_IK | SP_INVKSQ; o.dim = j; ops[j] = o;
        dst[j] = gphi + j * nr_;
      }
      e_.inv(3, is, ops, dst, M_);   // div of the masked W: band-limited
    }
    // V_hat = M FFT(Y); then, in one spectral pass, V_hat - nu |K|^2 W (the
    // final sum of GradL2ForLq, :1186-1190), Hilbert + Leray of it (:150-153)
    // and of the normal ||u||_q^(1-q) W (:162-169)
    ST Vh = e_.stmp(3);
    {
      F_grad_terms<T> gt{aux2, u_, Acur(), nullptr, G_, gphi, force, nr_, (T)(q_ - 2.0), (T)pc_,
                         (T)nuvc_, nullptr};
      if (g_pending_) {   // the last step's update of G rides along
        gt.Gd0 = Gd_prev_;
        gt.ga = (T)gA_;
        gt.gb = (T)gB_;
        gt.Gout = G_;
      }
      C* vd[3] = {Vh, Vh + ns_, Vh + 2 * ns_};
      const SpecOp ops[3] = {sp_mask(M_), sp_mask(M_), sp_mask(M_)};
      if (!e_.fwd_prod(P_grad_terms<T>{gt}, vd, ops)) {
        flush_G();
        RT Y = e_.rtmp(3);
        F_grad_terms<T> g2{aux2, u_, Acur(), nullptr, G_, gphi, force, nr_, (T)(q_ - 2.0), (T)pc_,
                           (T)nuvc_, Y};
        e_.map(g2, nr_);
        fwd3(Y, Vh, sp_mask(M_));
      }
      g_pending_ = false;
      Gd_prev_ = RT();
    }
    ST nH = e_.stmp(3);
    // the Hilbert weight and Leray projection of V and of the normal, and in
    // the same pass every H_l product of the step; the
    // tangential projection gproj = Vh - alpha nH (:177), the normalisation
    // (:211-214) and the Polak-Ribiere numerator (:226) follow by bilinearity
    // (V, nH and gradJused0 are band-limited: the kept cube carries them; <u,u>
    // comes from the last update of u_hat)
    double r[10];
    e_.template reduce<10>(R_dir_prep_k<T>{Vh, W_, nH, uh_, have_cg_ ? gu0_.get() : nullptr, ns_, n_,
                                           pp_, M_, hwlt_, nuvc_, std::pow(L_, 1.0 - q_)},
                           nkc_, r);
    for (double& x : r) x /= N6_;
    // transported gradient vT_g0 = vg_s gradJused0 + vg_u u (see transport_and_update)
    const double vv = r[0], vn = r[1], nn = r[2], uu = uu_;
    const double vg = vg_s_ * r[3] + vg_u_ * r[5], ng = vg_s_ * r[4] + vg_u_ * r[6];
    const double alpha = vn / nn;                                 // :170-171
    const double a1 = 1.0 - alpha;
    const double t1 = std::fabs((vn - alpha * nn) / (vv + 2.0 * a1 * vn + a1 * a1 * nn));   // :180
    if (t1 > 9.999999960041972e-13) out.msgs.push_back(42);
    const double gg = vv - 2.0 * alpha * vn + alpha * alpha * nn;   // <gproj, gproj>
    const double dn = r[7] - 2.0 * alpha * r[8] + alpha * alpha * r[9];
    if (dn / gg > 1.0000000036274937e-15) out.msgs.push_back(43);   // :188-189
    const double c = std::sqrt(uu / gg);                          // :211-214
    const double gu1n = c * c * gg;                               // <gu1, gu1>
    double beta = 0.0;
    if (have_cg_) {                                               // :226-232
      beta = gu1n - c * (vg - alpha * ng);                        // <gu1, gu1 - vT_g0>
      const double g0n = gu0n_;
      if (g0n < MACH_EPSILON || beta < MACH_EPSILON) beta = 0.0;
      else beta = beta / g0n;
    }
    if (reset_every > 0 && iter % reset_every == 0) beta = 0.0;   // :235
    // d1 = gradJused1 + beta vT_d0 with vT_d0 = vd_B dB + vd_u u (:240): the
    // band-limited part and the multiple of u are kept apart
    gu1_ = e_.stmp(3);
    ST dBn = e_.stmp(3);
    e_.map(S_dir_update_k<T>{Vh, nH, have_cg_ ? dB_.get() : nullptr, gu1_, dBn, ns_, n_, pp_, M_,
                             c, alpha, beta * vd_B_}, nkc_, PC_SPECTRAL);
    gam_ = have_cg_ ? beta * vd_u_ : 0.0;
    dB_ = std::move(dBn);
    warm_[0] = ST();
    warm_[1] = ST();
    dB_limited_ = true;
    gu1n_ = gu1n;
    // zero-mode checks (:251-261): the k = 0 coefficient is the unnormalised mean
    {
      double m0 = k0_abs_sum(dB_, gam_);
      if (m0 > 1.0000000474974513e-03 * B_) out.msgs.push_back(44);
      m0 = k0_abs_sum(uh_);
      if (m0 > 1.0000000474974513e-03 * B_) out.msgs.push_back(44);
    }
    out.beta = beta;
    return out;
  }

  // ---- line-search coefficient fields for phi = u + tau d --------------------------
  void precompute_ls() {
    flush_G();
    d_ = e_.rtmp(3);
    Gd_ = e_.rtmp(9);
    // the band-limited part of d and its gradient
    InvSrc dis[12];
    SpecOp dops[12];
    T* ddst[12];
    for (int i = 0; i < 3; ++i) {
      dis[i] = InvSrc{{dB_ + i * ns_, nullptr, nullptr}, {1.0, 0, 0}, 1};
      dops[i] = sp_none();
      ddst[i] = d_ + i * nr_;
      for (int j = 0; j < 3; ++j) {
        dis[3 + i * 3 + j] = dis[i];
        dops[3 + i * 3 + j] = sp_ik(j);
        ddst[3 + i * 3 + j] = Gd_ + (i * 3 + j) * nr_;
      }
    }
    const int dkeep = dB_limited_ ? M_ : -1;
    S1_ = e_.rtmp(); S2_ = e_.rtmp(); b_ = e_.rtmp(3); ee_ = e_.rtmp(3);
    if (!poly_ok_) e_.inv(12, dis, dops, ddst, dkeep);
    if (poly_ok_) {
      // the 12 coefficient fields go straight into their masked spectra (one
      // fused forward pass, no intermediate fields): S1, S2 (kept for the
      // line-search polynomials), b, e (dealiased below), X1, X2 (the
      // Poisson solves P1, P2) and Q1, Q2 (|grad phi|^2 coefficients)
      S12h_ = e_.stmp(2);
      P12h_ = e_.stmp(2);
      Q12h_ = e_.stmp(2);
      ST beh = e_.stmp(6);
      C* dst[12] = {S12h_, S12h_ + ns_, beh, beh + ns_, beh + 2 * ns_, beh + 3 * ns_,
                    beh + 4 * ns_, beh + 5 * ns_, P12h_, P12h_ + ns_, Q12h_, Q12h_ + ns_};
      SpecOp ops[12];
      for (int k = 0; k < 12; ++k) ops[k] = (k == 8 || k == 9) ? poisson_op() : sp_mask(M_);
      // d and grad d come out of the same x pass that forms the products from
      // them (one read of u and grad u, none of d and grad d) when it can;
      // else their inverse transforms, then the products
      const PI_ls_products<T> pi{u_, G_, nr_, (T)gam_};
      if (!e_.inv_fwd_prod(dis, dops, ddst, dkeep, pi, dst, ops)) {
        e_.inv(12, dis, dops, ddst, dkeep);
        if (!e_.fwd_prod(P_ls_products<T>{u_, d_, G_, Gd_, nr_, (T)gam_}, dst, ops)) {
          RT X = e_.rtmp(12);
          e_.map(F_ls_products<T>{u_, d_, G_, Gd_, nr_, X, X + nr_, X + 2 * nr_, X + 5 * nr_,
                                  X + 8 * nr_, X + 9 * nr_, X + 10 * nr_, X + 11 * nr_, (T)gam_},
                 nr_);
          const T* src[12];
          for (int k = 0; k < 12; ++k) src[k] = X + k * nr_;
          e_.fwd(12, src, dst, ops);
        }
      }
      // dealiased S1, S2, b, e and the pressure coefficients P0, P1, P2 in
      // one batch of band-limited inverses
      P_ = e_.rtmp(3);
      const C* in[11] = {S12h_, S12h_ + ns_, beh, beh + ns_, beh + 2 * ns_, beh + 3 * ns_,
                         beh + 4 * ns_, beh + 5 * ns_, ph_, P12h_, P12h_ + ns_};
      T* out[11] = {S1_, S2_, b_, b_ + nr_, b_ + 2 * nr_, ee_, ee_ + nr_, ee_ + 2 * nr_,
                    P_, P_ + nr_, P_ + 2 * nr_};
      InvSrc is[11];
      SpecOp none[11];
      for (int k = 0; k < 11; ++k) {
        is[k] = InvSrc{{in[k], nullptr, nullptr}, {1.0, 0, 0}, 1};
        none[k] = sp_none();
      }
      e_.inv(11, is, none, out, M_);
    } else {
      RT X12 = e_.rtmp(2);
      Q1_ = e_.rtmp(); Q2_ = e_.rtmp();
      e_.map(F_ls_products<T>{u_, d_, G_, Gd_, nr_, S1_, S2_, b_, ee_, X12, X12 + nr_, Q1_, Q2_,
                              (T)gam_}, nr_);
      T* f[8] = {S1_, S2_, b_, b_ + nr_, b_ + 2 * nr_, ee_, ee_ + nr_, ee_ + 2 * nr_};
      e_.dealias_dev(f, 8, 2.0);
      // pressure coefficients P1, P2 (spectra kept for the state update) and P0
      P12h_ = e_.stmp(2);
      {
        const T* src[2] = {X12, X12 + nr_};
        C* dst[2] = {P12h_, P12h_ + ns_};
        SpecOp ops[2] = {poisson_op(), poisson_op()};
        e_.fwd(2, src, dst, ops);
      }
      P_ = e_.rtmp(3);
      {
        InvSrc is[3] = {InvSrc{{ph_.get(), nullptr, nullptr}, {1.0, 0, 0}, 1},
                        InvSrc{{P12h_.get(), nullptr, nullptr}, {1.0, 0, 0}, 1},
                        InvSrc{{P12h_ + ns_, nullptr, nullptr}, {1.0, 0, 0}, 1}};
        SpecOp ops[3] = {sp_none(), sp_none(), sp_none()};
        T* dst[3] = {P_, P_ + nr_, P_ + 2 * nr_};
        e_.inv(3, is, ops, dst, M_);   // pressure spectra are masked
      }
      // |grad phi|^2 coefficient spectra
      Q12h_ = e_.stmp(2);
      {
        const T* src[2] = {Q1_, Q2_};
        C* dst[2] = {Q12h_, Q12h_ + ns_};
        SpecOp ops[2] = {sp_mask(M_), sp_mask(M_)};
        e_.fwd(2, src, dst, ops);
      }
    }
    if (q4_) {
      double acc[11];
      e_.template reduce<11>(R_ls_poly<T>{a_, b_, ee_, u_, d_, P_, P_ + nr_, P_ + 2 * nr_, nr_,
                                          gam_}, nr_, acc);
      for (int k = 0; k < 5; ++k) cpoly_[k] = acc[k];
      for (int k = 0; k < 6; ++k) rpoly_[k] = acc[5 + k];
      if (poly_ok_) {
        double pa[10];
        e_.template reduce<10>(R_ls_poly4<T>{S0h_, S12h_, S12h_ + ns_, Q0h_, Q12h_, Q12h_ + ns_,
                                              n_, pp_, M_},
                               nkc_, pa);
        for (int k = 0; k < 5; ++k) { lpoly_[k] = pa[k]; qpoly_[k] = pa[5 + k]; }
        e_.ls_risky(s_, S1_, S2_, risky_, nrisky_);
        // k = 0 coefficients: sum over the grid of S0, S1, S2
        const C* p[3] = {S0h_.get(), S12h_.get(), S12h_ + ns_};
        std::complex<double> z[3];
        e_.fetch(3, p, z);
        for (int k = 0; k < 3; ++k) ssum_[k] = z[k].real();
      }
    }
    cache_.clear();
    g64_fresh_ = false;
    if (q6_) {
      // q = 6: the R sums as polynomials in tau (FP64, on the twin for FP32
      // fields) and the points that can clip
      if constexpr (std::is_same<T, double>::value) {
        e_.ls6_coefs(LsGen<T>{s_, S1_, S2_, P_, Q0_, Q1_, Q2_, a_, b_, ee_, u_, d_, gam_}, r6_[0],
                     r6_[1], r6_[2]);
      } else {
        Engine<double>* tw = e_.fp64_twin();
        load_twin(tw);
        tw->ls6_coefs(gen64_, r6_[0], r6_[1], r6_[2]);
      }
      e_.ls_risky(s_, S1_, S2_, risky_, nrisky_);
      e_.template reduce<3>(R_sum3<T>{s_, S1_, S2_}, nr_, ssum_);
    }
  }

  // ---- J_LineMin(tau) = -R(B phi/||phi||_q) for a batch of step sizes ---------------
  void eval_batch(const double* taus, int nb, double* J, double* Lout = nullptr) {
    for (int b0 = 0; b0 < nb; b0 += e_.sb_) {
      const int k = std::min(e_.sb_, nb - b0);
      if (q4_) eval_q4(taus + b0, k, J + b0, Lout ? Lout + b0 : nullptr);
      else
        for (int i = 0; i < k; ++i) eval_generic(taus[b0 + i], J + b0 + i, Lout ? Lout + b0 + i : nullptr);
    }
  }

  // J from the cache (evaluated if absent) without counting it as a line-search
  // request: J1 of the iteration is -J_LineMin(tau1) (optimization.f90:349)
  double J_value(double tau) {
    auto it = cache_.find(tau);
    if (it != cache_.end()) return it->second;
    const std::vector<double> none;
    const double v = J(tau, none);
    --nJ_;
    return v;
  }

  // cached J with optional speculation list
  double J(double tau, const std::vector<double>& spec) {
    auto it = cache_.find(tau);
    if (it != cache_.end()) { ++nJ_; return it->second; }
    std::vector<double> batch{tau};
    for (double t : spec) {
      if ((int)batch.size() >= batch_) break;
      if (cache_.count(t) || std::find(batch.begin(), batch.end(), t) != batch.end()) continue;
      batch.push_back(t);
    }
    std::vector<double> Jv(batch.size());
    eval_batch(batch.data(), (int)batch.size(), Jv.data());
    nJc_ += (long long)batch.size();
    ++nJ_;
    for (size_t i = 0; i < batch.size(); ++i) {
      if (std::isnan(Jv[i])) throw Error(MLQ_ENAN, "nan in J");
      cache_[batch[i]] = Jv[i];
    }
    return Jv[0];
  }

  // ---- vector transport at the old u and the retraction (:326-346) -----------------
  // Returns ||u + tau1 d||_q.
  double transport_and_update(double tau1) {
    const double q = q_;
    // phi = u + tau1 d1 = (1 + tau1 gam) u + tau1 dB and the |phi|^k chain
    const double opg = 1.0 + tau1 * gam_;
    RT s1 = e_.rtmp();
    typename E_::Chain ch;
    // q = 4 without clipping at tau1: |phi|^2 = D(D(g g) 1), g = sqrt(max(s1, 0)),
    // is s1 itself (s1 is band-limited and positive; each dealias round trip
    // returns its input up to rounding), and ||phi||_4^4 the sum of s1^2 formed
    // with it
    const bool direct = poly_ok_ && poly_exact(tau1);
    const T* A1;
    const T* L1h;
    double Lq_sum = 0.0;
    if (direct) {
      e_.template reduce<1>(R_quad_sq<T>{s_, S1_, S2_, s1, (T)tau1}, nr_, &Lq_sum);
      A1 = s1;
      L1h = s1;
    } else {
      e_.map(F_quad<T>{s_, S1_, S2_, s1, (T)tau1, (T)1}, nr_);
      e_.chain_init(ch, nullptr, s1);
      A1 = e_.chain_pow(ch, q - 2.0);
      L1h = e_.chain_pow(ch, q / 2.0);
      Lq_sum = e_.dot(L1h, L1h);
    }
    const double L1 = std::pow(Lq_sum / N3_, 1.0 / q);
    const double sc = B_ / L1;
    const double Lmq = std::pow(L1, -q);
    // spectrum of D(|phi|^(q-2) phi), already scaled to the retracted state
    // (the normal spectrum W of u_new = sc phi is sc^(q-1) times it)
    const double wsc = std::pow(sc, q - 1.0);
    // one pass over u, d: the transform input |phi|^(q-2) phi and the
    // retracted field sc phi
    ST Wphi = e_.stmp(3);
    RT unew = e_.rtmp(3);
    // s1 of the new state, sc^2 s1, stored by the step pass when s1 is A1
    bool s1_scaled = false;
    {
      C* wd[3] = {Wphi, Wphi + ns_, Wphi + 2 * ns_};
      const SpecOp ops[3] = {sp_mask(M_, wsc), sp_mask(M_, wsc), sp_mask(M_, wsc)};
      P_step<T> ps{u_, d_, A1, nr_, (T)opg, (T)tau1, (T)sc, unew};
      if (direct) { ps.As = s1; ps.asc = (T)(sc * sc); }
      if (e_.fwd_prod(ps, wd, ops)) {
        s1_scaled = direct;
      } else {
        RT Au = e_.rtmp(3);
        e_.map(F_step<T>{u_, d_, A1, nr_, (T)opg, (T)tau1, (T)sc, Au, unew}, nr_);
        fwd3(Au, Wphi, sp_mask(M_, wsc));
      }
    }
    // <D(|phi|^(q-2) phi), xi>_L2 for xi = g_used1 and d1 = dB + gam u; with dB
    // band-limited the same pass applies the update of u_hat inside the cube
    // (see below)
    if (std::fabs(opg) < 1.0e-8)
      throw Error(MLQ_EUNSUP, "step along a direction dominated by -u/tau (1 + tau gam = 0)");
    double ir[3], uh_in = 0.0;
    if (dB_limited_) {
      double r4[4];
      e_.template reduce<4>(R_step_k<T>{Wphi, gu1_, uh_, dB_, ns_, n_, pp_, M_, sc * opg, sc * tau1,
                                        hlt_}, nkc_, r4);
      for (int k = 0; k < 3; ++k) ir[k] = r4[k];
      uh_in = r4[3];
    } else {
      e_.template reduce<3>(R_kept_dot3<T>{Wphi, gu1_, dB_, uh_, ns_, n_, pp_, M_}, nkc_, ir);
    }
    const double ir0 = ir[0] / (N6_ * wsc), ir1 = (ir[1] + gam_ * ir[2]) / (N6_ * wsc);
    // vector transport Gamma(u, tau1 d1, xi) = sc xi - c_xi phi_hat (:1746-1747)
    // with phi_hat = u_hat_new / sc, and the retraction u_hat_new = sc phi_hat:
    //   vT_g0 = sc gradJused1 - (c0/sc) u_new
    //   vT_d0 = sc dB / (1 + tau1 gam) + (gam / (1 + tau1 gam) - c1/sc) u_new
    // (u_old = (u_new/sc - tau1 dB) / (1 + tau1 gam)); only scalars are kept
    const double c0 = sc * Lmq * ir0, c1 = sc * Lmq * ir1;
    vg_s_ = sc;
    vg_u_ = -c0 / sc;
    vd_B_ = sc / opg;
    vd_u_ = gam_ / opg - c1 / sc;
    // u_hat_new = sc opg u_hat + sc tau1 dB.  With dB band-limited only the
    // kept cube changes: the modes outside are only rescaled, which is kept as
    // the pending factor hk_ (their stored values times hk_), and their
    // H_l norm is hk_^2 hsum_.  Otherwise the whole spectrum is updated and the
    // pending factor applied.
    if (dB_limited_) {   // u_hat updated inside the cube by the pass above
      hk_ *= sc * opg;
      uu_ = (uh_in + hk_ * hk_ * hsum_) / N6_;
    } else {
      double r[2];
      e_.template reduce<2>(R_uh_update<T>{uh_, dB_, ns_, n_, pp_, M_, false, sc * opg,
                                           sc * opg * hk_, sc * tau1, hlt_}, ns_, r);
      uu_ = (r[0] + r[1]) / N6_;
      hsum_ = r[1];
      hk_ = 1.0;
    }
    gu0n_ = gu1n_;
    gu0_ = std::move(gu1_);
    have_cg_ = true;
    // retraction and state update by homogeneity
    u_ = std::move(unew);
    // G = grad u_new = sc opg G + sc tau1 grad dB: deferred into the next
    // gradient-terms pass (which reads G anyway)
    Gd_prev_ = std::move(Gd_);
    gA_ = sc * opg;
    gB_ = sc * tau1;
    g_pending_ = true;
    if (!s1_scaled) e_.map(F_scale<T>{s1, (T)(sc * sc)}, nr_);
    std::swap(s_, s1);
    e_.map(F_quad<T>{a_, b_, ee_, a_, (T)tau1, (T)(sc * sc)}, 3 * nr_);   // 3 components
    // masked spectra: only the kept cube changes
    e_.map(S_quad_k<T>{ph_, P12h_, P12h_ + ns_, ph_, n_, pp_, M_, tau1, sc * sc}, nkc_, PC_SPECTRAL);
    e_.map(S_quad_k<T>{Q0h_, Q12h_, Q12h_ + ns_, Q0h_, n_, pp_, M_, tau1, sc * sc}, nkc_,
           PC_SPECTRAL);
    if (poly_ok_)
      e_.map(S_quad_k<T>{S0h_, S12h_, S12h_ + ns_, S0h_, n_, pp_, M_, tau1, sc * sc}, nkc_,
             PC_SPECTRAL);
    if (!q4_) e_.map(F_quad<T>{Q0_, Q1_, Q2_, Q0_, (T)tau1, (T)(sc * sc)}, nr_);
    W_ = std::move(Wphi);
    // |u|^(q-2), |u|^(q-4) and ||u||_q of the new state from the chain at tau1
    if (direct) {
      A_ = RT();   // |u_new|^2 = sc^2 s1 is s_ itself
    } else {
      A_ = e_.chain_take(ch, q - 2.0);
      e_.map(F_scale<T>{A_, (T)std::pow(sc, q - 2.0)}, nr_);
    }
    if (!q4_) {
      const T* Bq1 = e_.chain_pow(ch, q - 4.0);
      Bq_ = e_.rtmp();
      e_.map(F_copy<T>{Bq1, Bq_}, nr_);
      e_.map(F_scale<T>{Bq_, (T)std::pow(sc, q - 4.0)}, nr_);
    }
    L_ = sc * L1;
    release_ls();
    return L1;
  }

  // apply a deferred update of G now
  void flush_G() {
    if (!g_pending_) return;
    e_.map(F_axpby<T>{G_, Gd_prev_, G_, (T)gA_, (T)gB_}, nr_ * 9);
    g_pending_ = false;
    Gd_prev_ = RT();
  }

  void release_ls() {
    d_ = RT(); Gd_ = RT(); S1_ = RT(); S2_ = RT(); b_ = RT(); ee_ = RT(); Q1_ = RT(); Q2_ = RT();
    P_ = RT(); P12h_ = ST(); Q12h_ = ST(); S12h_ = ST();
  }

  double L() const { return L_; }
  long long nJ_ = 0, nJc_ = 0, nJpoly_ = 0;
  int batch_ = 1;
  bool poly_enabled_ = true;   // false: every J through the batched transform

 private:
  // ---- helpers ---------------------------------------------------------------------
  SpecOp poisson_op() const {
    SpecOp op;
    op.flags = SP_MASK | SP_INVKSQ;
    op.mcut = M_;
    op.scale = -1.0;
    return op;
  }
  void spec3(const T* u, C* uh) {
    const T* src[3] = {u, u + nr_, u + 2 * nr_};
    C* dst[3] = {uh, uh + ns_, uh + 2 * ns_};
    SpecOp ops[3] = {sp_none(), sp_none(), sp_none()};
    e_.fwd(3, src, dst, ops);
  }
  void fwd3(const T* f, C* h, SpecOp op) {
    const T* src[3] = {f, f + nr_, f + 2 * nr_};
    C* dst[3] = {h, h + ns_, h + 2 * ns_};
    SpecOp ops[3] = {op, op, op};
    e_.fwd(3, src, dst, ops);
  }
  void grad9(const C* uh, T* G) {
    InvSrc is[9];
    SpecOp ops[9];
    T* dst[9];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        is[i * 3 + j] = InvSrc{{uh + i * ns_, nullptr, nullptr}, {1.0, 0, 0}, 1};
        ops[i * 3 + j] = sp_ik(j);
        dst[i * 3 + j] = G + (i * 3 + j) * nr_;
      }
    e_.inv(9, is, ops, dst);
  }
  void dealias1(T* f) { e_.dealias_dev(&f, 1, 2.0); }
  void dealias3(T* f) {
    T* g[3] = {f, f + nr_, f + 2 * nr_};
    e_.dealias_dev(g, 3, 2.0);
  }
  // W_i = M FFT(A u_i)
  // W_i = scale * M FFT(A u_i)
  void fwd_Au(const T* A, const T* u, C* W, double scale = 1.0) {
    RT t = e_.rtmp(3);
    for (int i = 0; i < 3; ++i) e_.map(F_mul<T>{A, u + i * nr_, t + i * nr_}, nr_);
    fwd3(t, W, sp_mask(M_, scale));
  }
  double hl(const C* a, const C* b) {
    double r;
    e_.template reduce<1>(R_specdot3<T>{a, b, ns_, n_, pp_, 1, p_.lambda1, 1.5 - 1.0 / q_}, ns_, &r);
    return r / N6_;
  }
  // sum_c |Re (h + g uh)_c(k = 0)|
  double k0_abs_sum(const C* h, double g = 0.0) {
    const C* p[6];
    for (int c = 0; c < 3; ++c) { p[c] = h + c * ns_; p[3 + c] = uh_ + c * ns_; }
    std::complex<double> v[6];
    e_.fetch(6, p, v);
    double r = 0.0;
    for (int c = 0; c < 3; ++c) r += std::fabs(v[c].real() + g * v[3 + c].real());
    return r;
  }
  void chain_state_from_s() {
    typename E_::Chain ch;
    e_.chain_init(ch, nullptr, s_);
    const T* A = e_.chain_pow(ch, q_ - 2.0);
    A_ = e_.rtmp();
    e_.map(F_copy<T>{A, A_}, nr_);
    if (!q4_) {
      Bq_ = e_.rtmp();
      e_.map(F_copy<T>{e_.chain_pow(ch, q_ - 4.0), Bq_}, nr_);
    }
    const T* Lh = e_.chain_pow(ch, q_ / 2.0);
    L_ = std::pow(e_.dot(Lh, Lh) / N3_, 1.0 / q_);
  }

  void finish_J(double tau, double R12, double R3, double Lq_q, double* J, double* Lout) {
    const double L = std::pow(Lq_q, 1.0 / q_);
    const double sc = B_ / L;
    *J = -(std::pow(sc, q_) * R12 + std::pow(sc, q_ + 1.0) * R3);
    if (Lout) *Lout = L;
    (void)tau;
  }

  void eval_q4(const double* taus, int nb, double* J, double* Lout) {
    double acc[2 * MAXB];
    // step sizes where no grid point clips: the exact polynomials; the others
    // in one fused batched transform
    int ex[MAXB], nex = 0;
    double tex[MAXB];
    for (int b = 0; b < nb; ++b) {
      const double t = taus[b];
      if (poly_ok_ && poly_enabled_ && poly_exact(t)) {
        double la = 0.0, lb = 0.0;
        for (int k = 4; k >= 0; --k) { la = la * t + lpoly_[k]; lb = lb * t + qpoly_[k]; }
        acc[2 * b] = la;
        acc[2 * b + 1] = lb;
        ++nJpoly_;
      } else {
        ex[nex] = b;
        tex[nex++] = t;
      }
    }
    if (nex == 0) {
    } else if (M_ < n_ / 2) {
      double ae[2 * MAXB];
      e_.ls_q4_sums(s_, S1_, S2_, Q0h_, Q12h_, Q12h_ + ns_, tex, nex, M_, ae);
      for (int i = 0; i < nex; ++i) { acc[2 * ex[i]] = ae[2 * i]; acc[2 * ex[i] + 1] = ae[2 * i + 1]; }
    } else {
      // the mask keeps every mode: unfused path (g^2 fields, transform, sums)
      std::vector<RT> gg;
      std::vector<ST> Ah;
      F_ls_g<T> fg{};
      fg.S0 = s_; fg.S1 = S1_; fg.S2 = S2_; fg.nb = nb;
      std::vector<const T*> src(nb);
      std::vector<C*> dst(nb);
      std::vector<SpecOp> ops(nb, sp_mask(M_));
      for (int b = 0; b < nb; ++b) {
        gg.push_back(e_.rtmp());
        Ah.push_back(e_.stmp());
        fg.g[b] = nullptr;
        fg.gg[b] = gg[b];
        fg.tau[b] = taus[b];
        src[b] = gg[b];
        dst[b] = Ah[b];
      }
      e_.map(fg, nr_, PC_LS_POINT);
      e_.fwd(nb, src.data(), dst.data(), ops.data());
      R_ls_spec<T> rf{};
      for (int b = 0; b < nb; ++b) { rf.A[b] = Ah[b]; rf.tau[b] = taus[b]; }
      rf.Q0 = Q0h_; rf.Q1 = Q12h_; rf.Q2 = Q12h_ + ns_;
      rf.nb = nb; rf.n = n_; rf.px = pp_; rf.M = M_;
      e_.template reduce<2 * MAXB>(rf, (long long)(M_ + 1) * (2 * M_ + 1) * (2 * M_ + 1), acc);
    }
    for (int b = 0; b < nb; ++b) {
      const double t = taus[b];
      const double Lq_q = acc[2 * b] / N6_;
      const double R1 = -nuvc_ * acc[2 * b + 1] / N6_;
      double cp = 0.0, rp = 0.0;
      for (int k = 4; k >= 0; --k) cp = cp * t + cpoly_[k];
      for (int k = 5; k >= 0; --k) rp = rp * t + rpoly_[k];
      const double R2 = -nuvc_ * (q_ - 2.0) * cp / N3_;
      const double R3 = pc_ * (q_ - 2.0) * rp / N3_;
      finish_J(t, R1 + R2, R3, Lq_q, J + b, Lout ? Lout + b : nullptr);
    }
  }

  // generic line search (q != 4): the sums of R(phi) at tau from the
  // coefficient fields.  FP32 fields: the |phi|^k chain runs on the FP64 twin
  // engine from exact FP64 copies of the fields (made once per precompute),
  // so J(tau) carries the FP32 rounding of the fields but no rounding that
  // depends on tau; brent then sees a smooth J and stops after as many
  // requests as in FP64.
  void eval_generic(double t, double* J, double* Lout) {
    // q = 6 where no point clips (as at q = 4: the clipped mass below 1e-14 of
    // sum s(tau) for FP64 fields, 1e-9 for FP32 fields, whose own rounding is
    // larger): the R sums from their polynomials, the norm from two dealias
    // round trips (FP64, on the twin for FP32 fields)
    if (q6_ && poly_enabled_ && poly_exact(t)) {
      double lq;
      if constexpr (std::is_same<T, double>::value) {
        lq = e_.ls6_norm(LsGen<T>{s_, S1_, S2_, P_, Q0_, Q1_, Q2_, a_, b_, ee_, u_, d_, gam_}, t);
      } else {
        Engine<double>* tw = e_.fp64_twin();
        if (!g64_fresh_) load_twin(tw);
        lq = tw->ls6_norm(gen64_, t);
      }
      auto pv = [t](const double* c, int n) {
        double v = 0.0;
        for (int k = n - 1; k >= 0; --k) v = v * t + c[k];
        return v;
      };
      const double R1 = -nuvc_ * pv(r6_[0], 7) / N3_;
      const double R2 = -nuvc_ * (q_ - 2.0) * pv(r6_[1], 7) / N3_;
      const double R3 = pc_ * (q_ - 2.0) * pv(r6_[2], 8) / N3_;
      ++nJpoly_;
      finish_J(t, R1 + R2, R3, lq / N3_, J, Lout);
      return;
    }
    double acc[4];
    if constexpr (std::is_same<T, double>::value) {
      e_.ls_generic(LsGen<T>{s_, S1_, S2_, P_, Q0_, Q1_, Q2_, a_, b_, ee_, u_, d_, gam_}, t, acc);
    } else {
      Engine<double>* tw = e_.fp64_twin();
      if (!g64_fresh_) load_twin(tw);
      tw->ls_generic(gen64_, t, acc);
    }
    const double R1 = -nuvc_ * acc[0] / N3_;
    const double R2 = -nuvc_ * (q_ - 2.0) * acc[1] / N3_;
    const double R3 = pc_ * (q_ - 2.0) * acc[2] / N3_;
    finish_J(t, R1 + R2, R3, acc[3] / N3_, J, Lout);
  }
  // exact FP64 copies of the generic line-search fields on the twin
  template <typename W = T> void load_twin(Engine<double>* tw) {
    static_assert(std::is_same<W, float>::value, "FP32 fields only");
    if (!g64_.get()) g64_ = tw->rtmp(24);
    const W* src[12] = {s_, S1_, S2_, P_, Q0_, Q1_, Q2_, a_, b_, ee_, u_, d_};
    const int nc[12] = {1, 1, 1, 3, 1, 1, 1, 3, 3, 3, 3, 3};
    const double* p[12];
    double* dst = g64_;
    for (int k = 0; k < 12; ++k) {
      e_.map(F_widen{src[k], dst}, nc[k] * nr_);
      p[k] = dst;
      dst += nc[k] * nr_;
    }
    e_.sync();   // the twin runs on its own stream
    gen64_ = LsGen<double>{p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[8], p[9], p[10], p[11],
                           gam_};
    g64_fresh_ = true;
  }

  E_& e_;
  const mlq_params p_;
  double q_, B_, nuvc_, pc_, N3_, N6_;
  int n_, pp_, M_;
  long long nr_, ns_;
  bool q4_, poly_ok_;
  bool q6_ = false;                   // q = 6: the line-search sums by polynomials
  double r6_[3][8] = {};              // their coefficients: R1, R2 (degree 6), R3 (degree 7)
  long long nkc_;                     // calls of a kept-cube kernel (kw<T>() modes each)
  double* hlt_ = nullptr;             // squared H_l weight by |m|^2
  double2* hwlt_ = nullptr;           // (Hilbert, squared H_l) weights by |m|^2
  size_t tab_bytes_ = 0;
  double uu_ = 0.0;                   // <u_hat, u_hat>_H_l of the state
  // u_hat outside the kept cube is stored up to the factor hk_; hsum_ is the
  // stored part's <., .>_H_l there (unnormalised)
  double hk_ = 1.0, hsum_ = 0.0;
  // state
  RT u_, G_, s_, A_, Bq_, a_, Q0_;
  ST uh_, ph_, W_, Q0h_;
  double L_ = 1.0;
  // CG state
  // CG state: gradJused of this and the last step; the direction's band-limited
  // part dB and its multiple of u (d1 = dB + gam u); the transported vectors
  // as combinations vT_g0 = vg_s gu0 + vg_u u, vT_d0 = vd_B dB + vd_u u
  ST gu0_, gu1_, dB_;
  double gam_ = 0.0;
  bool dB_limited_ = true;
  double vg_s_ = 0.0, vg_u_ = 0.0, vd_B_ = 0.0, vd_u_ = 0.0;
  double gu0n_ = 0.0, gu1n_ = 0.0;   // <gradJused, gradJused>_H_l of the last two steps
  bool have_cg_ = false;
  // line-search coefficients
  RT d_, Gd_, S1_, S2_, b_, ee_, Q1_, Q2_, P_;
  // deferred update G := gA_ G + gB_ Gd_prev_ (applied by the next direction)
  RT Gd_prev_;
  ST warm_[2];   // see init_state
  double gA_ = 1.0, gB_ = 0.0;
  bool g_pending_ = false;
  ST P12h_, Q12h_;
  double cpoly_[5] = {0}, rpoly_[6] = {0};
  // line-search polynomials of ||phi||_q^q and the R1 numerator (valid where
  // poly_exact); masked spectra of D(u.u) (state) and of S1, S2
  ST S0h_, S12h_;
  double lpoly_[5] = {0}, qpoly_[5] = {0};
  // points that can clip (S0, S1, S2 triples; nrisky_ of them, the list is
  // empty when more than its capacity) and the grid sums of S0, S1, S2
  std::vector<double> risky_;
  long long nrisky_ = 0;
  double ssum_[3] = {0, 0, 0};
  // the polynomials are used at tau when the mass clipped by max(s(tau), 0)
  // is negligible against sum s(tau): 1e-14 (FP64) or 1e-9 (FP32) relative
  bool poly_exact(double t) const {
    return poly_exact(t, sizeof(T) == sizeof(double) ? 1.0e-14 : 1.0e-9);
  }
  bool poly_exact(double t, double tol) const {
    if (nrisky_ > 0 && risky_.empty()) return false;   // too many points to account for
    double m = 0.0;
    for (size_t i = 0; i < risky_.size(); i += 3) {
      const double v = risky_[i] + t * (risky_[i + 1] + t * risky_[i + 2]);
      if (v < 0.0) m -= v;
    }
    const double tot = ssum_[0] + t * (ssum_[1] + t * ssum_[2]);
    return tot > 0.0 && m <= tol * tot;
  }
  std::map<double, double> cache_;
  // FP64 copies of the generic line-search fields (FP32 engine, q != 4)
  Tmp<double, double> g64_;
  LsGen<double> gen64_{};
  bool g64_fresh_ = false;
};

// ---------------------------------------------------------------------------
// The solve loop on the resident state (optimization.f90:7-514)
// ---------------------------------------------------------------------------
template <typename T>
void rcg_solve_fast(Engine<T>& e, int u, const mlq_rcg_cfg& cfg, mlq_rcg_result& res,
                    const LogSink& log) {
  auto t0 = std::chrono::steady_clock::now();
  const double B = e.p_.B;
  auto msg = [&](int iter, int code) { log.emit(MLQ_LOG_MSG, iter, code, 0, nullptr); };
  e.rescale(u, B);                                                 // :61
  e.divAvg_free(u);                                                // :62
  KappaInSolve<T> kap(e, cfg, log);
  kap.initial(cfg.start_iter, u);                                  // :69-72
  double J0 = e.eval_J(u, MLQ_J_MAXDLQDT);                         // :78
  double J1 = J0;
  FastRcg<T> F(e);
  F.batch_ = cfg.batch_linesearch > 0 ? std::min(e.sb_, std::max(1, cfg.batch_linesearch)) : 1;
  F.init_state(e.fphys(u));
  int iter = cfg.start_iter;
  const int start = cfg.start_iter;
  double tau0 = 100.0, tau1 = tau0;
  const int nroll = std::max(1, cfg.rolling);
  std::vector<double> rel(nroll, 0.0);
  double avg = 1.0;
  int while_flag = 0, stop = 0;
  {
    double d[6] = {0.0, 0.0, J0, 0.0, 0.0, 0.0};
    log.emit(MLQ_LOG_ITER, iter, iter, 0, d);
  }
  if (e.spectra_due(cfg, iter)) e.emit_spectrum(log, cfg, iter, 0, e.fphys(u));   // :119-121
  std::vector<double> dspec;
  while (std::fabs(avg) > cfg.optim_tol && iter < cfg.max_iter && while_flag < 1) {
    iter += 1;
    if (kap.every > 0 && iter % kap.every == 0) {                  // :147-149
      F.download_u(e.fphys(u));
      kap.periodic(iter, u);
    }
    auto dir = F.direction(iter, cfg.reset_every);
    for (int m : dir.msgs) msg(iter, m);
    const double beta = dir.beta;
    F.precompute_ls();
    if (kap.ends) {
      auto d1 = F.direction_field();
      kap.keep_d1(d1);
    }
    // the spectrum of d1 (:381) is taken while its physical field exists and
    // delivered after the iteration's record
    if (e.spectra_due(cfg, iter)) {
      auto d1 = F.direction_field();
      e.spectrum_of(cfg, d1, dspec);
    }
    // speculative step sizes: the first mnbrak requests (0, tau0, GOLD*tau0)
    const double tB0 = std::max(tau0, MACH_EPSILON);
    std::vector<double> spec = {0.0, tB0, MNBRAK_GOLD * tB0, MNBRAK_CGOLD * tB0,
                                MNBRAK_GOLD * (MNBRAK_GOLD * tB0), MNBRAK_CGOLD * (MNBRAK_CGOLD * tB0)};
    JFun J = [&](double tau) { return F.J(tau, spec); };
    LineminLog llog = [&](int mode, int bi, double tA, double tB, double tC, double FA, double FB,
                          double FC) {
      double d[6] = {tA, tB, tC, FA, FB, FC};
      log.emit(MLQ_LOG_LINEMIN, iter, bi, mode, d);
    };
    BrentLog blog = [&](int j, double X, double FX) {
      double d[2] = {X, FX};
      log.emit(MLQ_LOG_BRENT, iter, j, 0, d);
    };
    msg(iter, 20);
    const long long nJ_it = F.nJ_;
    BracketResult br = mnbrak(J, 0.0, tau0, llog);                 // :269
    if (br.flag != 0) {
      if (br.flag == 1) msg(iter, 22);
      log.emit(MLQ_LOG_ERR, iter, br.flag, 0, nullptr);
      double d[6] = {tau1, beta, J0, 0.0, 0.0, (double)(F.nJ_ - nJ_it)};
      log.emit(MLQ_LOG_ITER, iter, iter, 0, d);
      while_flag = 1;
      stop = br.flag == 1 ? 2 : 3;
      break;
    }
    msg(iter, 21);
    msg(iter, 30);
    spec.clear();
    BrentResult bres = brent(J, br.tA, br.tC, blog);               // :297
    tau1 = bres.X;
    if (tau1 > br.tC) {
      if (tau1 > 2.0 * br.tC) msg(iter, 32);
    } else {
      msg(iter, 31);
    }
    if (tau1 > TAU_MAX) {                                          // :310-318
      msg(iter, 32);
      while_flag = 1;
      tau1 = 0.0;
      stop = 4;
    } else if (tau1 > std::sqrt(TAU_MAX)) {
      msg(iter, 41);
    }
    // J requests of mnbrak and brent (J_value below is not one)
    const long long nJ_ls = F.nJ_ - nJ_it;
    // J1 = R(rescale(u + tau1 d1)) = -J_LineMin(tau1)                 (:349)
    J1 = -F.J_value(tau1);
    kap.keep_u0(F.u_field());
    F.transport_and_update(tau1);
    const double deltaJ = (J1 - J0) / std::fabs(J0);
    if (deltaJ < -MACH_EPSILON) {                                  // :351-365
      double d[6] = {tau1, beta, J0, br.tA, br.tC, (double)(nJ_ls)};
      log.emit(MLQ_LOG_ITER, iter, iter + 1, 0, d);
      msg(iter, 0);
      while_flag = 1;
      stop = 5;
      break;
    }
    {
      double d[6] = {tau1, beta, J1, br.tA, br.tC, (double)(nJ_ls)};
      log.emit(MLQ_LOG_ITER, iter, iter, 0, d);
    }
    if (e.spectra_due(cfg, iter)) {                                // :377-381
      e.emit_spectrum(log, cfg, iter, 0, F.u_field());
      log.emit(MLQ_LOG_SPECTRUM, iter, 1, 0, dspec.data());
    }
    if (cfg.save_every > 0 && iter % cfg.save_every == 0 && log.cb) {   // :388-390
      F.download_u(e.fphys(u));
      e.emit_field(log, iter, u);
    }
    J0 = J1;
    tau0 = tau1;
    for (int i = nroll - 1; i > 0; --i) rel[i] = rel[i - 1];
    rel[0] = deltaJ;
    double sr = 0.0;
    for (double x : rel) sr += x;
    avg = sr / (double)(iter - start);
  }
  F.download_u(e.fphys(u));
  const bool ok = (iter < cfg.max_iter) && while_flag < 1;
  if (stop == 0 && !ok) stop = 1;
  msg(iter, 1);
  if (kap.ends) kap.final(iter, ok, (stop == 2 || stop == 3) ? u : kap.u0);   // :457-469
  res.J1 = J1;
  res.iters = iter;
  res.successful = ok ? 1 : 0;
  res.while_flag = while_flag;
  res.stop_reason = stop;
  res.n_J = (int)F.nJ_;
  res.n_J_computed = (int)F.nJc_;
  res.n_J_poly = (int)F.nJpoly_;
  res.enstrophy = e.enstrophy_dev(e.fphys(u));                     // :491-492
  res.dEnstrophydt = e.dEnstrophydt_dev(e.fphys(u));
  res.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace mlq

namespace mlq {

template <typename T>
void Engine<T>::ls_generic(const LsGen<T>& f, double t, double out[4], double s2) {
  const double q = p_.q;
  Chain ch;   // g = sqrt(max(s(t), 0)) in one pass (s itself is not needed)
  ch.g = rtmp();
  map(F_quad_sqrtclip<T>{f.S0, f.S1, f.S2, ch.g, (T)t, (T)s2}, nr_);
  const T* A = chain_pow(ch, q - 2.0);
  const T* Bq = chain_pow(ch, q - 4.0);
  const T* Lh = chain_pow(ch, q / 2.0);
  // D(Bq) unless |phi|^(q-4) is already band-limited by the 2/3 mask
  RT Bw;
  const T* Bwp = Bq;
  if (q - 4.0 + 2.0e-16 < 2.0) {
    Bw = rtmp();
    map(F_copy<T>{Bq, Bw}, nr_);
    T* b = Bw;
    dealias_dev(&b, 1, 2.0);
    Bwp = Bw;
  }
  RT Pd = rtmp();
  map(F_pB<T>{f.P, f.P + nr_, f.P + 2 * nr_, Bq, Pd, (T)t}, nr_);
  {
    T* pd = Pd;
    dealias_dev(&pd, 1, 2.0);
  }
  reduce<4>(R_ls_generic<T>{A, Bwp, Pd, Lh, f.Q0, f.Q1, f.Q2, f.a, f.b, f.e, f.u, f.d, nr_, t,
                            f.gam}, nr_, out);
}

// q = 6: the coefficients of the R1, R2 and R3 sums of J as polynomials in
// tau (unnormalised, r1 and r2 of degree 6, r3 of degree 7), valid where
// max(s(tau), 0) does not clip (see F_ls6_prod): |phi|^4 = D(s^2) from the
// dealiased C_k, the pressure term D(p s) from the dealiased E_k, |phi|^2 = s
template <typename T>
void Engine<T>::ls6_coefs(const LsGen<T>& f, double r1[7], double r2[7], double r3[8]) {
  RT C = rtmp(5), E = rtmp(5);
  map(F_ls6_prod<T>{f.S0, f.S1, f.S2, f.P, f.P + nr_, f.P + 2 * nr_, nr_, C, E}, nr_);
  {
    T* fl[10];
    for (int k = 0; k < 5; ++k) {
      fl[k] = C + k * nr_;
      fl[5 + k] = E + k * nr_;
    }
    dealias_dev(fl, 10, 2.0);
  }
  double aq[15], aw[15], b[20];
  reduce<15>(R_ls6q<T>{C, f.Q0, f.Q1, f.Q2, nr_}, nr_, aq);
  reduce<15>(R_ls6w<T>{f.S0, f.S1, f.S2, f.a, f.b, f.e, nr_}, nr_, aw);
  reduce<20>(R_ls6b<T>{E, f.u, f.d, f.a, f.b, f.e, nr_, f.gam}, nr_, b);
  for (int m = 0; m < 7; ++m) r1[m] = r2[m] = 0.0;
  for (int m = 0; m < 8; ++m) r3[m] = 0.0;
  for (int k = 0; k < 5; ++k)
    for (int j = 0; j < 3; ++j) r1[k + j] += aq[3 * k + j];
  for (int i = 0; i < 3; ++i)
    for (int n = 0; n < 5; ++n) r2[i + n] += aw[5 * i + n];
  for (int k = 0; k < 5; ++k)
    for (int l = 0; l < 4; ++l) r3[k + l] += b[4 * k + l];
}

// q = 6: sum over the grid of |phi|^6 = (|phi|^3)^2 as the chain forms it
// where s(tau) does not clip: |phi|^3 = D(D(g g) D_1(g)), g = sqrt(s), with
// D(g g) = s (s is band-limited) and D_1 the order-1 dealias
template <typename T> double Engine<T>::ls6_norm(const LsGen<T>& f, double t) {
  RT s1 = rtmp(), g = rtmp();
  map(F_quad<T>{f.S0, f.S1, f.S2, s1, (T)t, (T)1}, nr_);
  map(F_sqrtclip<T>{s1, g}, nr_);
  {
    T* p = g;
    dealias_dev(&p, 1, 1.0);
  }
  mul_dealias(s1, g, g, true);
  return dot(g, g);
}

template <typename T> Engine<double>* Engine<T>::fp64_twin() {
  if constexpr (std::is_same<T, double>::value) {
    return this;
  } else {
    if (!twin_) twin_.reset(new Engine<double>(n_, p_));
    Engine<double>* e = static_cast<Engine<double>*>(twin_.get());
    e->set_params(p_);
    return e;
  }
}

template <typename T>
void Engine<T>::rcg_solve(int u, const mlq_rcg_cfg& cfg, mlq_rcg_result& res, const LogSink& log) {
  // the kappa tests' fields (perturbation and gradient; uvec0 and d1 for the
  // end tests) count against the resident-state footprint
  struct Reset { int& k; ~Reset() { k = 0; } } reset{kappa_fields_};
  kappa_fields_ = (cfg.flags & MLQ_RCG_KAPPA) ? 4 : (cfg.kappa_every > 0 ? 2 : 0);
  if ((cfg.flags & MLQ_RCG_REFERENCE) || !fast_path_ok()) rcg_solve_ref(u, cfg, res, log);
  else rcg_solve_fast(*this, u, cfg, res, log);
}

// J_LineMin at a batch of step sizes for phi = u + tau d: one coefficient
// precompute, then the batched evaluation.
template <typename T>
void Engine<T>::eval_J_batch(int u, int d, const double* taus, int ntau, double* J) {
  if (!fast_path_ok()) { eval_J_batch_ref(u, d, taus, ntau, J); return; }
  FastRcg<T> F(*this);
  F.init_state(fphys(u));
  F.set_direction_phys(fphys(d));
  F.precompute_ls();
  F.eval_batch(taus, ntau, J);
}

template <typename T>
void Engine<T>::time_linesearch(int u, int d, const double* taus, int ntau, int batch, double* J,
                                double* t_ms) {
  using clk = std::chrono::steady_clock;
  auto ms = [](clk::time_point a, clk::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
  };
  sync();
  if (batch == 0 || !fast_path_ok()) {
    const auto t0 = clk::now();
    eval_J_batch_ref(u, d, taus, ntau, J);
    sync();
    t_ms[0] = 0.0;
    t_ms[1] = ms(t0, clk::now());
    return;
  }
  FastRcg<T> F(*this);
  F.init_state(fphys(u));
  F.set_direction_phys(fphys(d));
  F.poly_enabled_ = batch < 0;
  if (batch < 0) batch = ntau;
  sync();
  const auto t0 = clk::now();
  F.precompute_ls();
  sync();
  const auto t1 = clk::now();
  for (int i = 0; i < ntau; i += batch) F.eval_batch(taus + i, std::min(batch, ntau - i), J + i);
  sync();
  t_ms[0] = ms(t0, t1);
  t_ms[1] = ms(t1, clk::now());
}

// ---------------------------------------------------------------------------
// Fused q = 4 line-search sums (kernels in mlq_fast_kernels.hpp)
// ---------------------------------------------------------------------------
template <typename T, int N, int V>
void launch_zls_v(hipStream_t st, const LsArgs<T>& la, int ncols, long long es, int nb) {
  using S = SVar<T, N, V>;
  if (MLQ_FWD_ZB && la.keep >= 0 && 8 * la.keep < 3 * N)   // zero-band last radix pass
    hipLaunchKernelGGL((k_zls<T, N, S::E, S::C, true>), dim3((ncols + S::C - 1) / S::C, nb),
                       dim3(S::C * (N / S::E)), 0, st, la, ncols, es);
  else
    hipLaunchKernelGGL((k_zls<T, N, S::E, S::C>), dim3((ncols + S::C - 1) / S::C, nb),
                       dim3(S::C * (N / S::E)), 0, st, la, ncols, es);
}

template <typename T> template <int N>
void Engine<T>::ls_q4_sums_n(const T* S0, const T* S1, const T* S2, const C* Q0, const C* Q1,
                             const C* Q2, const double* taus, int nb, int keep, double* out) {
  constexpr int NF = xpass_NF<N>();
  constexpr int Tf = FftShape<N>::Tf;
  constexpr int PP = spec_pitch<T>(N);
  const int var = svar_[5];   // tile of the band-limited forward z pass
  int C0 = 0;
  switch (var & 3) {
    case 1: C0 = SVar<T, N, 1>::C; break;
    case 2: C0 = SVar<T, N, 2>::C; break;
    case 3: C0 = SVar<T, N, 3>::C; break;
    default: C0 = SVar<T, N, 0>::C; break;
  }
  const int ncols = (keep + 1) * (2 * keep + 1);
  const int nbx = (ncols + C0 - 1) / C0;
  const size_t need = sizeof(double) * (size_t)nbx * 2 * MAXB;
  if (need > lspart_bytes_) {
    if (lspart_) dfree(lspart_, lspart_bytes_);
    lspart_ = (double*)dmalloc(need);
    lspart_bytes_ = need;
  }
  LsArgs<T> la{};
  la.S0 = S0; la.S1 = S1; la.S2 = S2;
  la.Q0 = Q0; la.Q1 = Q1; la.Q2 = Q2;
  la.spec = scratch_;
  la.sstride = ns_;
  la.tw = tw_;
  for (int b = 0; b < nb; ++b) la.tau[b] = taus[b];
  la.nb = nb;
  la.keep = keep;
  la.remap = (var >> 2) & 1;
  la.part = lspart_;
  prof_begin(PC_XFWD);
  if (MLQ_FWD_ZB && keep >= 0 && 8 * keep < 3 * N)   // zero-band last radix pass
    hipLaunchKernelGGL((k_xfwd_ls<T, N, NF, true>), dim3((unsigned)((long long)N * N / 2 / NF)),
                       dim3(NF * Tf), 0, st_, la);
  else
    hipLaunchKernelGGL((k_xfwd_ls<T, N, NF>), dim3((unsigned)((long long)N * N / 2 / NF)),
                       dim3(NF * Tf), 0, st_, la);
  MLQ_HIP(hipGetLastError());
  prof_end(PC_XFWD, 3.0 * real_bytes() + nb * ab_xcols(keep));
  SArgs<T> sa{};
  sa.spec = scratch_;
  sa.tw = tw_;
  sa.sstride = ns_;
  sa.keep = keep;
  prof_begin(PC_YFWD);
  launch_strided<T, N, -1, 0>(svar_[4], st_, sa, keep + 1, N * nb);
  MLQ_HIP(hipGetLastError());
  prof_end(PC_YFWD, nb * (ab_xcols(keep) + ab_kcols(keep)));
  prof_begin(PC_ZFWD);
  switch (var & 3) {
    case 1: launch_zls_v<T, N, 1>(st_, la, ncols, (long long)N * PP, nb); break;
    case 2: launch_zls_v<T, N, 2>(st_, la, ncols, (long long)N * PP, nb); break;
    case 3: launch_zls_v<T, N, 3>(st_, la, ncols, (long long)N * PP, nb); break;
    default: launch_zls_v<T, N, 0>(st_, la, ncols, (long long)N * PP, nb); break;
  }
  MLQ_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_reduce_final, dim3(2 * nb), dim3(RED_THREADS), 0, st_, lspart_, nbx, 2 * nb,
                     red_dev_);
  MLQ_HIP(hipGetLastError());
  // the z pass reads the y outputs and the kept cubes of Q0..Q2, and stores nothing
  prof_end(PC_ZFWD, nb * ab_kcols(keep) + 3.0 * ab_cube(keep));
  MLQ_HIP(hipMemcpyAsync(red_host_, red_dev_, sizeof(double) * 2 * nb, hipMemcpyDeviceToHost, st_));
  wait_spin();
  for (int k = 0; k < 2 * nb; ++k) out[k] = red_host_[k];
}

template <typename T>
void Engine<T>::ls_risky(const T* S0, const T* S1, const T* S2, std::vector<double>& list,
                         long long& count) {
  constexpr unsigned CAP = 1u << 16;
  const size_t need = sizeof(double) * 3 * CAP + 256;
  if (need > lsrisky_bytes_) {
    if (lsrisky_) dfree(lsrisky_, lsrisky_bytes_);
    lsrisky_ = (double*)dmalloc(need);
    lsrisky_bytes_ = need;
  }
  unsigned* cnt = (unsigned*)(lsrisky_ + 3 * CAP);
  MLQ_HIP(hipMemsetAsync(cnt, 0, sizeof(unsigned), st_));
  prof_begin(PC_REDUCE);
  hipLaunchKernelGGL((k_ls_risky<T>), dim3(RED_BLOCKS), dim3(RED_THREADS), 0, st_, S0, S1, S2, nr_,
                     cnt, lsrisky_, CAP);
  MLQ_HIP(hipGetLastError());
  prof_end(PC_REDUCE, 3.0 * real_bytes());
  // the count through the pinned reduction buffer (a pageable destination
  // makes the copy synchronous with a round trip of its own)
  MLQ_HIP(hipMemcpyAsync(red_host_, cnt, sizeof(unsigned), hipMemcpyDeviceToHost, st_));
  wait_spin();
  unsigned c = 0;
  std::memcpy(&c, red_host_, sizeof(unsigned));
  count = c;
  list.clear();
  if (c > 0 && c <= CAP) {
    list.resize(3 * (size_t)c);
    MLQ_HIP(hipMemcpy(list.data(), lsrisky_, sizeof(double) * 3 * c, hipMemcpyDeviceToHost));
  }
}

template <typename T>
void Engine<T>::ls_q4_sums(const T* S0, const T* S1, const T* S2, const C* Q0, const C* Q1,
                           const C* Q2, const double* taus, int nb, int keep, double* out) {
  if (nb < 1 || nb > sb_ || keep < 0 || keep >= n_ / 2)
    throw Error(MLQ_EINVAL, "ls_q4_sums: batch or mask out of range");
  MLQ_DISPATCH_N(ls_q4_sums_n, S0, S1, S2, Q0, Q1, Q2, taus, nb, keep, out);
}

}  // namespace mlq
