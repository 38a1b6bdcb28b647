// Scalar line search of the RCG solve: mnbrak (optimization.f90:613-778) and
// brent (optimization.f90:784-951), restated with the reference's constants
// and branch order so that, given the same J values, the same step sizes are
// requested.  J(tau) is a callback; requests may be served from a cache of
// speculatively evaluated step sizes (see RcgDriver).
#pragma once
#include <cmath>
#include <functional>

namespace mlq {

constexpr double MACH_EPSILON = 2.0e-16;   // global_variables.f90:28
constexpr double TAU_MAX = 1.0e2;          // :29
// optimization.f90:804-806: single-precision literals promoted to double
constexpr double BRENT_TOL = 9.999999974752427e-07;
constexpr double BRENT_ZEPS = 9.99999993922529e-09;
constexpr double BRENT_CGOLD = 0.3819659948348999;
constexpr int BRENT_ITMAX = 300;
// optimization.f90:630-634
const double MNBRAK_GOLD = (1.0 + std::sqrt(5.0)) / 2.0;
const double MNBRAK_CGOLD = 1.0 / MNBRAK_GOLD;
constexpr double MNBRAK_GLIMIT = 10.0;
constexpr int MNBRAK_ITMAX = 100;

inline double fsign(double a, double b) { return std::copysign(std::fabs(a), b); }

using JFun = std::function<double(double)>;
// log of save_linemin_data: (mode 0 replace / 1 append, brakIter, tA, tB, tC, FA, FB, FC)
using LineminLog = std::function<void(int, int, double, double, double, double, double, double)>;
using BrentLog = std::function<void(int, double, double)>;

struct BracketResult { double tA = 0, tC = 0; int flag = 0; };

inline BracketResult mnbrak(const JFun& J, double tA0, double tB0, const LineminLog& log) {
  const double GOLD = MNBRAK_GOLD, CGOLD = MNBRAK_CGOLD, GLIMIT = MNBRAK_GLIMIT;
  int brakIter = 0;
  double tA = tA0, tB = std::max(tB0, MACH_EPSILON), tC = 0.0;
  double FC = 0.0;
  double FA = J(tA);
  double FB = J(tB);
  if (log) log(0, brakIter, tA, tB, tC, FA, FB, FC);
  while (FB > FA && tB > MACH_EPSILON) {                          // :677-683
    tB = CGOLD * tB;
    FB = J(tB);
    if (log) log(1, brakIter, tA, tB, tC, FA, FB, FC);
  }
  BracketResult r;
  if (tB <= MACH_EPSILON) { r.flag = 1; return r; }               // :685-688
  tC = GOLD * tB;
  FC = J(tC);
  if (log) log(1, brakIter, tA, tB, tC, FA, FB, FC);
  while (FB >= FC && brakIter < MNBRAK_ITMAX) {                   // :697
    brakIter += 1;
    tC = GOLD * tC;
    FC = J(tC);
    const double R = (tB - tA) * (FB - FC);
    const double Q = (tB - tC) * (FB - FA);
    double tP = tB - 0.5 * ((tB - tC) * Q - (tB - tA) * R) /
                         fsign(std::max(std::fabs(Q - R), MACH_EPSILON), Q - R);
    const double Pmax = tB + GLIMIT * (tC - tB);
    double FP;
    if ((tB - tP) * (tP - tC) > 0) {                                // :710
      FP = J(tP);
      if (FP < FC) {
        tA = tB; FA = FB; tB = tP; FB = FP;
        break;
      } else if (FP > FB) {
        tC = tP; FC = FP;
        break;
      }
      tP = tC + GOLD * (tC - tB);
      FP = J(tP);
    } else if ((tC - tP) * (tP - Pmax) > 0) {                       // :730
      FP = J(tP);
      if (FP < FC) {
        tB = tC; tC = tP; FB = FC; FC = FP;
        tP = tC + GOLD * (tC - tB);
        FP = J(tP);
      }
    } else if ((tP - Pmax) * (Pmax - tC) >= 0) {                    // :744
      tP = Pmax;
      FP = J(tP);
    } else {
      tP = tC + GOLD * (tC - tB);
      FP = J(tP);
    }
    tA = tB; tB = tC; tC = tP;
    FA = FB; FB = FC; FC = FP;
    if (log) log(1, brakIter, tA, tB, tC, FA, FB, FC);
  }
  r.tA = tA;
  r.tC = tC;
  r.flag = brakIter >= MNBRAK_ITMAX ? 2 : 0;                      // :769-773
  return r;
}

struct BrentResult { double X = 0, FX = 0; };

inline BrentResult brent(const JFun& J, double b0, double b1, const BrentLog& log) {
  const double TOL = BRENT_TOL, ZEPS = BRENT_ZEPS, CGOLD = BRENT_CGOLD;
  double D = 0.0;
  double A = std::min(b0, b1), B = std::max(b0, b1);
  double V = b1 * CGOLD, W = V, X = V, E = 0.0;                    // :818-824
  double FX = J(D);                                                // :833-834
  if (log) log(0, D, FX);
  FX = J(X);                                                       // :844-845
  double FV = FX, FW = FX;
  double P = 0, Q = 0, R = 0, U = 0;
  for (int j = 1; j <= BRENT_ITMAX; ++j) {
    if (log) log(j, X, FX);
    const double XM = 0.5 * (A + B);
    const double TOL1 = TOL * std::fabs(X) + ZEPS;
    const double TOL2 = 2.0 * TOL1;
    if (std::fabs(X - XM) <= (TOL2 - 0.5 * (B - A))) break;         // :865
    int FLAG = 1;
    if (std::fabs(E) > TOL1) {
      R = (X - W) * (FX - FV);
      Q = (X - V) * (FX - FW);
      P = (X - V) * Q - (X - W) * R;
      Q = 2.0 * (Q - R);
      if (Q > 0.0) P = -P;
      Q = std::fabs(Q);
      const double ETEMP = E;
      E = D;
      if ((std::fabs(P) >= std::fabs(0.5 * Q * ETEMP)) || (P <= Q * (A - X)) || (P >= Q * (B - X)))
        FLAG = 1;
      else
        FLAG = 2;
    }
    if (FLAG == 1) {
      E = (X >= XM) ? (A - X) : (B - X);
      D = CGOLD * E;
    } else {
      D = P / Q;
      U = X + D;
      if ((U - A < TOL2) || (B - U < TOL2)) D = fsign(TOL1, XM - X);
    }
    U = (std::fabs(D) >= TOL1) ? X + D : X + fsign(TOL1, D);
    const double FU = J(U);
    if (FU <= FX) {
      if (U >= X) A = X; else B = X;
      V = W; FV = FW; W = X; FW = FX; X = U; FX = FU;
    } else {
      if (U < X) A = U; else B = U;
      if ((FU <= FW) || (W == X)) {
        V = W; FV = FW; W = U; FW = FU;
      } else if ((FU <= FV) || (V == X) || (V == W)) {
        V = U; FV = FU;
      }
    }
  }
  if (log) log(999, X, FX);
  BrentResult r;
  r.X = X;
  r.FX = FX;
  return r;
}

}  // namespace mlq
