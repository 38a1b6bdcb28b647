// The RCG solve for one constraint value B: maxdLqdt, optimization.f90:7-514.
// Algorithmic body only (the per-iteration prints, spectra, field saves and
// kappa tests of the reference are diagnostics); every record the reference
// writes from inside the loop is delivered to the log sink.
#pragma once
#include <limits>
#include <chrono>
#include <vector>

#include "mlq_lean_ls.hpp"
#include "mlq_linesearch.hpp"
#include "mlq_ops_core.hpp"

namespace mlq {

// Kappa tests inside the solve: kappaTest at the start (gradient in L2 at the
// rescaled, projected start field, optimization.f90:69-72) and the end
// (:457-469), kappaTestEveryXiteration inside the loop (:147-149).  The
// perturbation is made once per solve: "load-random-poly-b" (m1 = -5) of the
// caller's field, else "create-random-poly-b" (m1 = -5) with the caller's
// seed (:1019-1031).  At the end the reference tests at uvec0, the iterate
// before the last update, with gradJ1, the H_l gradient (HilbertGradient and
// divAvg_free, :146-153) of the last iteration's field, and d1, its last
// direction: after a failed bracket (:269-290) that field is the current
// iterate, which has not been updated, while uvec0 is the previous
// iteration's; before any update uvec0 is the start field here (the
// reference's array is not yet assigned).
template <typename T> struct KappaInSolve {
  enum { INITIAL = 0, PERIODIC = 1, END_GRAD = 2, END_D1 = 3 };
  Engine<T>& e;
  const LogSink& log;
  bool ends = false;
  int every = 0;
  int pert = -1, grad = -1, u0 = -1, d1 = -1;
  KappaInSolve(Engine<T>& eng, const mlq_rcg_cfg& cfg, const LogSink& lg) : e(eng), log(lg) {
    ends = (cfg.flags & MLQ_RCG_KAPPA) != 0;
    every = cfg.kappa_every;
    if (!ends && every < 1) return;
    pert = e.field_alloc(3);
    grad = e.field_alloc(3);
    if (ends) {
      u0 = e.field_alloc(3);
      d1 = e.field_alloc(3);
    }
    if (cfg.kappa_field) e.kappa_pert(pert, MLQ_PERT_LOAD_POLY, -5.0, 0.0, 0.0, 0, cfg.kappa_field);
    else e.kappa_pert(pert, MLQ_PERT_RANDOM_POLY, -5.0, 0.0, 0.0, cfg.kappa_seed, nullptr);
  }
  ~KappaInSolve() {
    for (int h : {pert, grad, u0, d1})
      if (h >= 0) try { e.field_free(h); } catch (...) {}
  }
  KappaInSolve(const KappaInSolve&) = delete;
  KappaInSolve& operator=(const KappaInSolve&) = delete;
  void run(int iter, int id, bool adjusted, int phi, int g, int ip) {
    double d[83];
    double frac = 0.0, best = 0.0;
    int row = 0;
    e.kappa_test(phi, g, pert, MLQ_J_MAXDLQDT, ip, d, &frac, &best, &row);
    d[80] = frac;
    d[81] = best;
    d[82] = (double)row;
    log.emit(MLQ_LOG_KAPPA, iter, id, adjusted ? 1 : 0, d);
  }
  // the start field u (rescaled and projected)
  void initial(int iter, int u) {
    if (!ends) return;
    e.grad_L2(u, grad);
    run(iter, INITIAL, false, u, grad, MLQ_IP_L2);
    copy(e.fphys(u), u0);
  }
  void periodic(int iter, int u) {
    if (every < 1 || iter % every != 0) return;
    e.grad_L2(u, grad);
    run(iter, PERIODIC, false, u, grad, MLQ_IP_L2);
  }
  void keep_u0(const T* u) { if (ends) copy(u, u0); }
  void keep_d1(const T* d) { if (ends) copy(d, d1); }
  // gfield: the field the last gradient was taken at; converged: the solve
  // ended successfully
  void final(int iter, bool converged, int gfield) {
    if (!ends) return;
    e.grad_L2(gfield, grad);
    e.HilbertGradient(grad, 1.5 - 1.0 / e.p_.q);
    e.divAvg_free(grad);
    if (!converged) run(iter, END_D1, true, u0, d1, MLQ_IP_HL);
    run(iter, END_GRAD, true, u0, grad, MLQ_IP_HL);
  }
  void copy(const T* src, int h) { e.map(F_copy<T>{src, e.fphys(h)}, 3 * e.nr_); }
};

template <typename T>
void Engine<T>::rcg_solve_ref(int u, const mlq_rcg_cfg& cfg, mlq_rcg_result& res,
                              const LogSink& log) {
  auto t0 = std::chrono::steady_clock::now();
  const double q = p_.q, B = p_.B;
  const long long n3 = nr_ * 3;
  // state fields (optimization.f90:15, 28-50); gradJ0, diff_gradJ and uvec0 are
  // never read by the algorithm and are not kept.  One buffer carries the
  // gradient, then its tangential part, then the normalised gradient
  // (gradJ1 -> gradJ1projected -> gradJused1, each the only later reader of
  // the one before); gradJused0 enters only through its H_l norm, kept as a
  // number; the step eta shares the scratch vector.
  const int g1 = field_alloc(3), nH = field_alloc(3), d1 = field_alloc(3), vg0 = field_alloc(3),
            vd0 = field_alloc(3), tmp = field_alloc(3), fs = field_alloc(1);
  const int gproj = g1, gu1 = g1, eta = tmp;
  double gu0_norm = 0.0;                                           // <gradJused0, gradJused0>_H_l
  struct Cleanup {
    Engine* e; std::vector<int> h;
    ~Cleanup() { for (int x : h) try { e->field_free(x); } catch (...) {} }
  } cleanup{this, {g1, nH, d1, vg0, vd0, tmp, fs}};
  auto axpby = [&](int x, double a, int y, double b, int o) {
    map(F_axpby<T>{fphys(x), fphys(y), fphys(o), (T)a, (T)b}, n3);
  };
  auto msg = [&](int iter, int code) { log.emit(MLQ_LOG_MSG, iter, code, 0, nullptr); };
  const int HL = MLQ_IP_HL;

  rescale(u, B);                                                    // :61
  divAvg_free(u);                                                   // :62
  KappaInSolve<T> kap(*this, cfg, log);
  kap.initial(cfg.start_iter, u);                                   // :69-72
  double J0 = eval_J(u, MLQ_J_MAXDLQDT);                            // :78
  double J1 = J0;
  int iter = cfg.start_iter;
  const int start = cfg.start_iter;
  double tau0 = 100.0, tau1 = tau0;                                 // :92-94
  const int nroll = std::max(1, cfg.rolling);
  std::vector<double> rel(nroll, 0.0);
  double avg = 1.0;
  int while_flag = 0, stop = 0;
  long long nJ = 0, nJc = 0;
  {
    double d[6] = {0.0, 0.0, J0, 0.0, 0.0, 0.0};
    log.emit(MLQ_LOG_ITER, iter, iter, 0, d);                       // :117
  }
  if (spectra_due(cfg, iter)) emit_spectrum(log, cfg, iter, 0, fphys(u));   // :119-121
  const double s = 1.5 - 1.0 / q;
  while (std::fabs(avg) > cfg.optim_tol && iter < cfg.max_iter && while_flag < 1) {
    iter += 1;
    grad_L2(u, g1);                                                 // :146
    kap.periodic(iter, u);                                          // :147-149
    HilbertGradient(g1, s);                                         // :151
    divAvg_free(g1);                                                // :153
    constraint_derivative(u, nH);                                   // :162
    HilbertGradient(nH, s);                                         // :164
    divAvg_free(nH);                                                // :169
    const double alpha = inner(nH, g1, HL) / inner(nH, nH, HL);     // :170-171
    axpby(g1, 1.0, nH, -alpha, gproj);                              // :177
    axpby(gproj, 1.0, nH, 1.0, tmp);
    const double t1 = std::fabs(inner(gproj, nH, HL) / inner(tmp, tmp, HL));   // :180
    if (t1 > 9.9999999600419720e-13) msg(iter, 42);
    divergence(gproj, fs);                                          // :188
    const double t2 = inner(fs, fs, MLQ_IP_L2) / inner(gproj, gproj, HL);
    if (t2 > 1.0000000036274937e-15) msg(iter, 43);
    // normalised tangential gradient (:201-214)
    const double nrm = std::sqrt(inner(u, u, HL) / inner(gproj, gproj, HL));
    axpby(gproj, nrm, gproj, 0.0, gu1);
    // Polak-Ribiere with transported vectors (:226-240)
    axpby(gu1, 1.0, vg0, -1.0, tmp);
    double beta = inner(gu1, tmp, HL);
    const double g0n = gu0_norm;                                    // :227
    if (g0n < MACH_EPSILON || beta < MACH_EPSILON) beta = 0.0;
    else beta = beta / g0n;
    if (cfg.reset_every > 0 && iter % cfg.reset_every == 0) beta = 0.0;
    axpby(gu1, 1.0, vd0, beta, d1);
    kap.keep_d1(fphys(d1));
    double m0[3];
    calc0thFourierModes(d1, m0);                                    // :251-261
    if (std::fabs(m0[0]) + std::fabs(m0[1]) + std::fabs(m0[2]) > 1.0000000474974513e-03 * B) msg(iter, 44);
    calc0thFourierModes(u, m0);
    if (std::fabs(m0[0]) + std::fabs(m0[1]) + std::fabs(m0[2]) > 1.0000000474974513e-03 * B) msg(iter, 44);
    // line search on J(tau) = eval_J(u + tau d1, "LineMin"): every value
    // computed from the field as eval_J does (optimization.f90:531-539), or,
    // where the resident state applies but does not fit (and with
    // MLQ_RCG_LEAN), from the coefficient fields of u and d1 formed once here
    std::unique_ptr<LeanLs<T>> lean;
    if (!(cfg.flags & MLQ_RCG_REFERENCE) &&
        ((cfg.flags & MLQ_RCG_LEAN) ? fast_path_applies() : lean_ls_applies()) && lean_ls_fits(0)) {
      lean.reset(new LeanLs<T>(*this));
      lean->build(fphys(u), fphys(d1));
    }
    JFun J = [&](double tau) {
      double v;
      if (lean) v = lean->J(tau);
      else eval_J_batch_ref(u, d1, &tau, 1, &v);
      ++nJ; ++nJc;
      if (std::isnan(v)) throw Error(MLQ_ENAN, "nan in J");
      return v;
    };
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
    const long long nJ_it = nJ;
    BracketResult br = mnbrak(J, 0.0, tau0, llog);                  // :269
    if (br.flag != 0) {
      if (br.flag == 1) msg(iter, 22);
      log.emit(MLQ_LOG_ERR, iter, br.flag, 0, nullptr);
      double d[6] = {tau1, beta, J0, 0.0, 0.0, (double)(nJ - nJ_it)};
      log.emit(MLQ_LOG_ITER, iter, iter, 0, d);                     // :285
      while_flag = 1;
      stop = br.flag == 1 ? 2 : 3;
      break;
    }
    msg(iter, 21);
    msg(iter, 30);
    BrentResult bres = brent(J, br.tA, br.tC, blog);                // :297
    lean.reset();
    tau1 = bres.X;
    if (tau1 > br.tC) {
      if (tau1 > 2.0 * br.tC) msg(iter, 32);
    } else {
      msg(iter, 31);
    }
    if (tau1 > TAU_MAX) {                                           // :310-318
      msg(iter, 32);
      while_flag = 1;
      tau1 = 0.0;
      stop = 4;
    } else if (tau1 > std::sqrt(TAU_MAX)) {
      msg(iter, 41);
    }
    // vector transport at the old u with eta = tau1 d1 (:326-327)
    axpby(d1, tau1, d1, 0.0, eta);
    vector_transport(u, q, B, eta, gu1, vg0);
    vector_transport(u, q, B, eta, d1, vd0);
    // update (:344-350)
    kap.keep_u0(fphys(u));
    axpby(u, 1.0, d1, tau1, u);
    rescale(u, B);
    J1 = eval_J(u, MLQ_J_MAXDLQDT);
    const double deltaJ = (J1 - J0) / std::fabs(J0);
    if (deltaJ < -MACH_EPSILON) {                                   // :351-365
      double d[6] = {tau1, beta, J0, br.tA, br.tC, (double)(nJ - nJ_it)};
      log.emit(MLQ_LOG_ITER, iter, iter + 1, 0, d);
      msg(iter, 0);
      while_flag = 1;
      stop = 5;
      break;
    }
    {
      double d[6] = {tau1, beta, J1, br.tA, br.tC, (double)(nJ - nJ_it)};
      log.emit(MLQ_LOG_ITER, iter, iter, 0, d);                     // :375
    }
    if (spectra_due(cfg, iter)) {                                   // :377-381
      emit_spectrum(log, cfg, iter, 0, fphys(u));
      emit_spectrum(log, cfg, iter, 1, fphys(d1));
    }
    if (cfg.save_every > 0 && iter % cfg.save_every == 0 && log.cb)   // :388-390
      emit_field(log, iter, u);
    J0 = J1;                                                        // :432-436
    gu0_norm = inner(gu1, gu1, HL);
    tau0 = tau1;
    for (int i = nroll - 1; i > 0; --i) rel[i] = rel[i - 1];        // :439-445
    rel[0] = deltaJ;
    double sr = 0.0;
    for (double x : rel) sr += x;
    avg = sr / (double)(iter - start);
  }
  const bool ok = (iter < cfg.max_iter) && while_flag < 1;          // :451
  if (stop == 0 && !ok) stop = 1;
  msg(iter, 1);
  if (kap.ends) kap.final(iter, ok, (stop == 2 || stop == 3) ? u : kap.u0);   // :457-469
  res.J1 = J1;
  res.iters = iter;
  res.successful = ok ? 1 : 0;
  res.while_flag = while_flag;
  res.stop_reason = stop;
  res.n_J = (int)nJ;
  res.n_J_computed = (int)nJc;
  res.n_J_poly = 0;
  res.enstrophy = enstrophy_dev(fphys(u));                          // :491-492
  res.dEnstrophydt = dEnstrophydt_dev(fphys(u));
  res.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// kappa test (optimization.f90:957-1129) with the perturbation given by the
// caller (the reference loads or draws its own, :1001-1011): 16 step sizes
// eps = 10^(ii - 15), kappa = (J(phi + eps pert) - J(phi)) / (eps <grad, pert>),
// the best row by the reference's criterion (:1035, including its use of the
// current kappa on both sides) and the 4-point adjusted normalisation
// (:1075-1097).  table[ii*5 + c] = {eps, <grad, pert>, J1 - J0, kappa,
// kappa adjusted}; best_row is 1-based.
template <typename T>
void Engine<T>::kappa_test(int phi, int grad, int pert, int system, int kind, double* table,
                           double* frac_out, double* best_kappa, int* best_row) {
  constexpr int NK = 16, ADJ = 4;
  const double innerProd = inner(grad, pert, kind);
  const double J0 = eval_J(phi, system);
  const int tmp = field_alloc(3);
  double bestKappa = std::numeric_limits<double>::max();
  int bestii = 0;
  try {
    for (int ii = 1; ii <= NK; ++ii) {
      const double myexp = -7.0 - (double)NK / 2.0 + (double)ii;
      const double eps = std::pow(10.0, myexp);
      map(F_axpby<T>{fphys(phi), fphys(pert), fphys(tmp), (T)1, (T)eps}, nr_ * 3);
      const double J1 = eval_J(tmp, system);
      const double kappa = (J1 - J0) / (eps * innerProd);
      if (std::max(std::fabs(kappa - 1.0), std::fabs(1.0 / kappa - 1.0)) <
          std::max(std::fabs(bestKappa - 1.0), std::fabs(1.0 / kappa - 1.0))) {
        bestKappa = kappa;
        bestii = ii;
      }
      double* row = table + (ii - 1) * 5;
      row[0] = eps; row[1] = innerProd; row[2] = J1 - J0; row[3] = kappa; row[4] = kappa;
    }
  } catch (...) {
    field_free(tmp);
    throw;
  }
  field_free(tmp);
  int iAdjMin = 1;
  double fracAdjMin = 9999.0;
  for (int ii = 1; ii <= NK - (ADJ - 1) - 1; ++ii) {
    double frac = 0.0;
    for (int jj = 0; jj < ADJ; ++jj)
      frac += std::fabs(1.0 - table[(ii + jj - 1) * 5 + 4] / table[(ii + jj) * 5 + 4]);
    if (frac < fracAdjMin) { fracAdjMin = frac; iAdjMin = ii; }
  }
  double frac = 1.0;
  for (int jj = 0; jj < ADJ; ++jj) frac *= std::fabs(table[(iAdjMin + jj - 1) * 5 + 4]);
  frac = std::copysign(std::pow(frac, 1.0 / ADJ), table[(iAdjMin - 1) * 5 + 4]);
  for (int ii = 0; ii < NK; ++ii) table[ii * 5 + 4] /= frac;
  *frac_out = frac;
  *best_kappa = bestKappa;
  *best_row = bestii;
}

}  // namespace mlq
