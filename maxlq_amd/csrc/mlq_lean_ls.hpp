// Line search of the reference-structured iteration from coefficient fields
// (optimization.f90:531-539 evaluated by homogeneity): for phi = u + tau d,
// D(phi.phi), w_j = D(sum_i phi_i d_j phi_i), the pressure and |grad phi|^2
// are quadratic in tau.  Their 18 coefficient fields are formed once per
// iteration with a small footprint (the velocity gradients one column at a
// time, never the 9-component gradients of u and d together); each J(tau)
// is then one evaluation of the |phi|^k chain and the R sums from them
// (Engine::ls_generic) instead of a full R evaluation of the field, with the
// retraction B phi / ||phi||_q through R(s phi) = s^q (R1 + R2)(phi) +
// s^(q+1) R3(phi).  This is what the iteration runs when the resident state
// does not fit (1024^3 at q = 9).
#pragma once
#include <cmath>
#include <set>

#include "mlq_fast_kernels.hpp"
#include "mlq_ops_core.hpp"

namespace mlq {

template <typename T> struct LeanLs {
  using RT = typename Engine<T>::RT;
  using RtOut = typename Engine<T>::RtOut;
  Engine<T>& e;
  long long nr;
  const T* u = nullptr;
  const T* d = nullptr;
  RT S, Wa, Wb, We, P, Q;   // S: D(u.u), S1, S2; W: a, b, e (3 components each); P, Q: 3 coefficients

  explicit LeanLs(Engine<T>& eng) : e(eng), nr(eng.nr_) {}

  // |phi|^k fields the chain of one J(tau) holds (ls_generic: k = q-2, q-4,
  // q/2, each with its k-2, k-4, ... down to k < 2), plus g and D(g^2)
  static int chain_fields(double q) {
    std::set<long long> keys;
    for (double k0 : {q - 2.0, q - 4.0, q / 2.0}) {
      for (double k = k0;; k -= 2.0) {
        keys.insert(std::llround(k * 1.0e6));
        if (k + 2.0e-16 < 2.0) break;
      }
    }
    return (int)keys.size() + 2;
  }
  // device bytes at the peak: the 18 coefficient fields and the larger of
  // the build's temporaries (6 fields) and one J's (chain, s(tau), D(|phi|^(q-4)),
  // the pressure product)
  static double bytes(const Engine<T>& e) {
    return (18.0 + std::max(6, chain_fields(e.p_.q) + 3)) * e.real_bytes();
  }

  double dn_un = 0.0;   // ||d|| / ||u|| (L2)

  void build(const T* u_in, const T* d_in) {
    u = u_in;
    d = d_in;
    {
      double dd, uu;
      e.template reduce<1>(R_dot<T>{d, d}, 3 * nr, &dd);
      e.template reduce<1>(R_dot<T>{u, u}, 3 * nr, &uu);
      dn_un = std::sqrt(dd / uu);
    }
    const SpecOp pois = [&] {
      SpecOp o;
      o.flags = SP_MASK | SP_INVKSQ;
      o.mcut = e.mcut_of(2.0);
      o.scale = -1.0;
      return o;
    }();
    // D(u.u), D(2 u.d), D(d.d)
    S = e.rtmp(3);
    e.map(F_dot3self<T>{u, nr, S}, nr);
    e.map(F_uddots<T>{u, d, nr, S + nr, S + 2 * nr}, nr);
    {
      T* f[3] = {S, S + nr, S + 2 * nr};
      e.dealias_dev(f, 3, 2.0);
    }
    Wa = e.rtmp(3);
    Wb = e.rtmp(3);
    We = e.rtmp(3);
    Q = e.rtmp(3);
    P = e.rtmp(3);   // the pressure source X until the Poisson solve
    const T* src[6] = {u, u + nr, u + 2 * nr, d, d + nr, d + 2 * nr};
    {
      // column j of grad u and grad d by round trips (d_j f = REAL(IFFT(i K_j FFT(f))))
      RT Gj = e.rtmp(6);
      for (int j = 0; j < 3; ++j) {
        RtOut o[6];
        for (int k = 0; k < 6; ++k) o[k] = RtOut{k, sp_ik(j), Gj + k * nr};
        e.rtrip(6, src, 6, o);
        e.map(F_ls_column<T>{u, d, Gj, nr, j, j == 0, Wa + j * nr, Wb + j * nr, We + j * nr, Q, P},
              nr);
      }
    }
    {
      RT P4 = e.rtmp(4);
      const int pr[3][2] = {{0, 1}, {0, 2}, {1, 2}};
      for (const auto& ac : pr) {
        const int a = ac[0], c = ac[1];
        const T* s4[4] = {u + c * nr, u + a * nr, d + c * nr, d + a * nr};
        const RtOut o[4] = {RtOut{0, sp_ik(a), P4}, RtOut{1, sp_ik(c), P4 + nr},
                            RtOut{2, sp_ik(a), P4 + 2 * nr}, RtOut{3, sp_ik(c), P4 + 3 * nr}};
        e.rtrip(4, s4, 4, o);
        e.map(F_ls_pair<T>{P4, nr, P}, nr);
      }
    }
    {
      T* f[9];
      for (int k = 0; k < 3; ++k) {
        f[k] = Wa + k * nr;
        f[3 + k] = Wb + k * nr;
        f[6 + k] = We + k * nr;
      }
      e.dealias_dev(f, 9, 2.0);
    }
    {
      // P_k = REAL(IFFT(M FFT(X_k) / |K|^2)) in place (solve_poisson(-aux))
      const T* xs[3] = {P, P + nr, P + 2 * nr};
      RtOut o[3];
      for (int k = 0; k < 3; ++k) o[k] = RtOut{k, pois, P + k * nr};
      e.rtrip(3, xs, 3, o);
    }
  }

  // J_LineMin(tau) = -R(B phi / ||phi||_q).  The chain runs on phi' = s0 phi,
  // s0 = 1 / max(1, |tau| ||d|| / ||u||) (L2 norms): at large step sizes
  // |phi|^(q-2) of phi itself would exceed the FP32 range inside the
  // transforms of its dealiasing (the reference evaluates R of the retracted
  // field, of norm B).  R(phi') and ||phi'||_q follow from the sums by
  // homogeneity (|grad phi|^2 and the pressure scale by s0^2, w by s0^2,
  // phi.w by s0^3), then R(B phi / ||phi||_q) = R(sc phi'), sc = B / ||phi'||_q.
  double J(double tau) {
    const double s0 = 1.0 / std::max(1.0, std::fabs(tau) * dn_un);
    double acc[4];
    e.ls_generic(LsGen<T>{S, S + nr, S + 2 * nr, P, Q, Q + nr, Q + 2 * nr, Wa, Wb, We, u, d, 0.0},
                 tau, acc, s0 * s0);
    const double s2 = s0 * s0;
    acc[0] *= s2;
    acc[1] *= s2 * s2;
    acc[2] *= s2 * s2 * s0;
    const mlq_params& p = e.p_;
    const double q = p.q, N3 = (double)nr, nuvc = p.visc * p.viscCoefficient;
    const double R1 = -nuvc * acc[0] / N3;
    const double R2 = -nuvc * (q - 2.0) * acc[1] / N3;
    const double R3 = p.pressureCoefficient * (q - 2.0) * acc[2] / N3;
    const double sc = p.B / std::pow(acc[3] / N3, 1.0 / q);
    const double v = -(std::pow(sc, q) * (R1 + R2) + std::pow(sc, q + 1.0) * R3);
    if (std::isnan(v))
      throw Error(MLQ_ENAN, "nan in J at tau = " + std::to_string(tau) + " (phi scaled by " +
                                std::to_string(s0) + ")");
    return v;
  }
};

template <typename T> bool Engine<T>::lean_ls_fits(int held) const {
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) { (void)hipGetLastError(); return false; }
  return LeanLs<T>::bytes(*this) + held * real_bytes() <= 0.95 * ((double)fr + (double)pooled_);
}

}  // namespace mlq
