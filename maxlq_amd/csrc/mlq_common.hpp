// Common types for the maxLq MI355X library: precision traits, the spectral
// multiplier descriptor, status codes and the HIP error check.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

#include "../../include/maxlq.h"

namespace mlq {

template <typename T> struct V2;
template <> struct V2<float> { using type = float2; };
template <> struct V2<double> { using type = double2; };
template <typename T> using c2 = typename V2<T>::type;

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define MLQ_HIP(call)                                                        \
  do {                                                                       \
    hipError_t mlq_hip_rc_ = (call);                                         \
    if (mlq_hip_rc_ != hipSuccess)                                           \
      throw ::mlq::Error(MLQ_EHIP, std::string(#call " failed: ") +          \
                                       hipGetErrorString(mlq_hip_rc_));      \
  } while (0)

// Spectral multiplier applied per half-spectrum mode (kx, ky, kz).  All wave
// numbers are K_j = 2 pi m_j with m in (-N/2, N/2] (Nyquist counted +N/2),
// initialize.f90:170-192.  The factor is the product of the enabled parts:
//   MASK    keep iff max_j |m_j| <= mcut            (function_ops.f90:3391)
//   IK      i K_dim, zero at the own-dimension Nyquist (REAL(ifft) rule)
//   NEGKSQ  -|K|^2                                   (laplacian, :3272)
//   INVKSQ  -1/|K|^2, 0 at k = 0                     (solve_poisson, :1977-1981)
//   ZEROK0  zero the k = 0 mode                      (divAvg_free, :2891-2893)
//   HSCALE  times ``hs`` outside the cube max_j |m_j| <= mcut (a spectrum
//           whose high modes are stored up to a pending scalar factor)
// times the real ``scale``.  (The Sobolev weight of HilbertGradient needs a
// pow per mode and runs as its own spectral map, so that the pass kernels
// carrying these factors stay within 64 VGPRs.)
enum SpecFlags : int {
  SP_MASK = 1,
  SP_IK = 2,
  SP_NEGKSQ = 4,
  SP_INVKSQ = 8,
  SP_ZEROK0 = 32,
  SP_HSCALE = 64,
};

struct SpecOp {
  int flags = 0;
  int dim = 0;
  int mcut = 1 << 30;
  int pad = 0;
  double scale = 1.0;
  double hs = 1.0;   // SP_HSCALE factor
};

inline SpecOp sp_none(double s = 1.0) { SpecOp o; o.scale = s; return o; }
inline SpecOp sp_mask(int mcut, double s = 1.0) {
  SpecOp o; o.flags = SP_MASK; o.mcut = mcut; o.scale = s; return o;
}
inline SpecOp sp_ik(int dim, double s = 1.0) {
  SpecOp o; o.flags = SP_IK; o.dim = dim; o.scale = s; return o;
}

// Largest integer m with 2 pi m <= 2 pi N/(p+1) evaluated as the reference
// does (k_cut = 2.0*PI*real(n)/(p+1); mode kept iff max|K| <= k_cut).
int dealias_mcut(int n, double p);

}  // namespace mlq
