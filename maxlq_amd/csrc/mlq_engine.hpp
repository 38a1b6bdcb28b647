// Type-erased interface of the device engine (FP64 or mixed FP32/FP64) used
// by the C ABI and the host RCG driver.  Host-facing values are FP64.
#pragma once
#include <memory>
#include <string>
#include <vector>

#include "mlq_common.hpp"

namespace mlq {

// A device field: ncomp physical components (real) and, once computed, their
// half spectra.  Owned by the engine's field table.
struct DevField {
  int ncomp = 0;
  void* phys = nullptr;   // ncomp * nreal elements of T
  void* spec = nullptr;   // ncomp * nspec elements of c2<T>, or null
  bool spec_valid = false;
};

enum ProfClass : int {
  PC_XFWD = 0, PC_YFWD, PC_ZFWD, PC_ZINV, PC_YINV, PC_XINV,
  PC_POINTWISE, PC_SPECTRAL, PC_REDUCE, PC_LS_POINT, PC_ZRT, PC_COUNT
};

struct LogSink {
  mlq_log_cb cb = nullptr;
  void* user = nullptr;
  void emit(int kind, int iter, int i0, int i1, const double* d) const {
    if (cb) cb(user, kind, iter, i0, i1, d);
  }
};

class IEngine {
 public:
  virtual ~IEngine() = default;
  virtual int n() const = 0;
  virtual const mlq_params& params() const = 0;
  virtual void set_params(const mlq_params& p) = 0;
  virtual void sync() = 0;
  virtual double real_bytes() const = 0;
  virtual double spec_bytes() const = 0;

  // fields
  virtual int field_alloc(int ncomp) = 0;
  virtual void field_free(int h) = 0;
  virtual void upload(int h, const double* host) = 0;
  virtual void download(int h, double* host) = 0;
  virtual int field_ncomp(int h) const = 0;

  // fftwfunction
  virtual void fft_c2c(const double* in, double* out, bool inverse, int ncomp) = 0;
  virtual void fft_r2c_host(const double* in, double* out) = 0;
  virtual void fft_c2r_host(const double* in, double* out) = 0;

  // function_ops on device handles (scalar fields have ncomp = 1)
  virtual void gradient(int u_sc, int out_vec) = 0;
  virtual void divergence(int u_vec, int out_sc) = 0;
  virtual void laplacian(int f_vec) = 0;
  virtual void solve_poisson(int f_sc, double l, int out_sc) = 0;
  virtual void dealias_scalar(int g_sc, double order) = 0;
  virtual void divAvg_free(int f_vec) = 0;
  virtual void HilbertGradient(int f_vec, double order) = 0;
  virtual void calc0thFourierModes(int v, double out[3]) = 0;
  virtual void calc_uk(int u_vec, double k, int out_sc) = 0;
  virtual void calc_gk_order2(int g_sc, double k, int out_sc) = 0;
  virtual double Lq_norm(int u_vec) = 0;
  virtual void rescale(int u_vec, double B) = 0;
  virtual void constraint_derivative(int u_vec, int out_vec) = 0;
  virtual void vector_transport(int u, double q, double B, int eta, int xi, int out) = 0;
  virtual double inner(int f, int g, int kind) = 0;
  virtual void dLqdt_parts(int u_vec, double q, double out[3]) = 0;
  virtual void spectral_data(int u_vec, int normalize, double* out) = 0;
  virtual double enstrophy(int u_vec) = 0;
  virtual double dEnstrophydt(int u_vec) = 0;
  virtual void grad_L2(int u_vec, int out_vec) = 0;
  virtual double eval_J(int f_vec, int system) = 0;
  virtual void eval_J_batch(int u, int d, const double* taus, int ntau, double* J) = 0;
  // one coefficient precompute, then the step sizes in launches of `batch`
  // (0: reference-structured evaluations); t_ms = {precompute, evaluation}
  virtual void time_linesearch(int u, int d, const double* taus, int ntau, int batch, double* J,
                               double* t_ms) = 0;
  // kappa test (optimization.f90:957-1129) with a given perturbation: table
  // rows {eps, <grad, pert>, J1 - J0, kappa, kappa adjusted}, 16 rows
  virtual void kappa_test(int phi, int grad, int pert, int system, int kind, double* table,
                          double* frac, double* best_kappa, int* best_row) = 0;
  virtual void seeded(int f_vec, unsigned long long seed, int kind, double a) = 0;
  virtual void diagnostic_scalars(int u_vec, double* out, double* fields, double* ring,
                                  long long cap, long long* nring) = 0;
  virtual void kappa_test_pert(int out, int kind, double m1, double m2, double m3,
                               unsigned long long seed, const double* loaded) = 0;

  // whether rcg_solve / eval_J_batch take the resident-state iteration
  virtual bool uses_fast_path() const = 0;
  // the iteration rcg_solve takes: 1 resident state, 2 reference-structured
  // with the line search from coefficient fields, 0 reference-structured
  virtual int rcg_path() const = 0;
  // the RCG solve (optimization.f90:7-514)
  virtual void rcg_solve(int u, const mlq_rcg_cfg& cfg, mlq_rcg_result& res,
                         const LogSink& log) = 0;

  // measurement
  virtual void time_fft(int batch, int reps, double* t_ms) = 0;
  virtual void time_copy(int nfields, int reps, double* t_ms) = 0;
  virtual void time_column_copy(int axis, int var, int nfields, int reps, double* t_ms,
                                double* bytes) = 0;
  // time the tile candidates of the strided passes and keep the fastest;
  // times[pass*8+cand] in ms per launch, choice[pass]
  virtual void autotune(int batch, double* times, int* choice) = 0;
  virtual void set_tiles(const int* choice) = 0;
  virtual void prof_enable(bool on) = 0;
  virtual void prof_get(int cls, long long* count, double* ms, double* bytes) = 0;
  virtual void mem_info(double* used, double* peak) = 0;
  // device bytes in use (allocated minus cached in the pool) and their peak
  virtual void mem_live(double* live, double* peak) = 0;
  virtual void transform_counts(long long* fwd, long long* inv) const = 0;
  virtual hipStream_t stream() const = 0;          // the context's HIP stream
  virtual const double2* twiddles() const = 0;     // exp(-2 pi i k/n), k < n, on the device
};

std::unique_ptr<IEngine> make_engine_fp64(int n, const mlq_params& p);
// mlq_noalloc.hip: device address of the no-allocation scratch on the
// current device (once per device), and fftfwdr through it
void noalloc_bind();
void fftfwdr_noalloc(int n, hipStream_t st, const double2* tw, const double* u, double* fu);
std::unique_ptr<IEngine> make_engine_fp32(int n, const mlq_params& p);

}  // namespace mlq
