// extern "C" entry points of libmaxlq.so (declared in include/maxlq.h).
// Host arrays use the reference layout (FP64, x fastest, component blocks).
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "mlq_engine.hpp"

using namespace mlq;

struct mlq_ctx {
  std::unique_ptr<IEngine> eng;
  int n = 0;
  std::string err;
  int scratch_vec[4] = {-1, -1, -1, -1};
  int scratch_sc[2] = {-1, -1};
};

namespace {

template <class F> int guarded(mlq_ctx* c, F&& f) {
  if (!c) return MLQ_EINVAL;
  try {
    f();
    c->err.clear();
    return MLQ_OK;
  } catch (const Error& e) {
    c->err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    c->err = "host allocation failed";
    return MLQ_EOOM;
  } catch (const std::exception& e) {
    c->err = e.what();
    return MLQ_EINVAL;
  }
}

int vec(mlq_ctx* c, int i) {
  if (c->scratch_vec[i] < 0) c->scratch_vec[i] = c->eng->field_alloc(3);
  return c->scratch_vec[i];
}
int sca(mlq_ctx* c, int i) {
  if (c->scratch_sc[i] < 0) c->scratch_sc[i] = c->eng->field_alloc(1);
  return c->scratch_sc[i];
}
long long nreal(mlq_ctx* c) { return (long long)c->n * c->n * c->n; }

void check_nan(const double* a, long long n) {
  for (long long i = 0; i < n; ++i)
    if (std::isnan(a[i])) throw Error(MLQ_ENAN, "nan in result");
}

}  // namespace

extern "C" {

void mlq_default_params(mlq_params* p, double q) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->q = q;
  p->B = 1.0;
  p->visc = 1.0;
  p->lambda1 = 0.1;
  p->viscCoefficient = 1.0;
  p->pressureCoefficient = 1.0;
  p->toDealias = 1;
  p->use_e_u = -1;
  p->dealiase_if_mult_by_e_u = 1;
  p->precision = MLQ_FP64;
}

int mlq_init(mlq_ctx** ctx, int n1, int n2, int n3, const mlq_params* p) {
  if (!ctx || !p) return MLQ_EINVAL;
  *ctx = nullptr;
  if (n1 != n2 || n2 != n3) return MLQ_EUNSUP;   // cubic grids (initialize.f90:142-147)
  const int n = n1;
  if (n < 8 || n > 1024 || (n & (n - 1))) return MLQ_EUNSUP;
  auto* c = new mlq_ctx();
  c->n = n;
  int rc = guarded(c, [&] {
    c->eng = p->precision == MLQ_MIXED ? make_engine_fp32(n, *p) : make_engine_fp64(n, *p);
    noalloc_bind();
    // MLQ_AUTOTUNE: 0 skips the tile autotuner, a positive value sets the
    // number of spectra it times each candidate on (default 3)
    const char* at = std::getenv("MLQ_AUTOTUNE");
    const int ab = at ? std::atoi(at) : 3;
    if (n >= 128 && !(at && at[0] == '0')) c->eng->autotune(ab > 0 ? ab : 3, nullptr, nullptr);
  });
  if (rc != MLQ_OK) { delete c; return rc; }
  *ctx = c;
  return MLQ_OK;
}

int mlq_finalize(mlq_ctx* ctx) {
  if (!ctx) return MLQ_EINVAL;
  delete ctx;
  return MLQ_OK;
}

int mlq_set_params(mlq_ctx* ctx, const mlq_params* p) {
  return guarded(ctx, [&] {
    if (!p) throw Error(MLQ_EINVAL, "null params");
    if (p->precision != ctx->eng->params().precision)
      throw Error(MLQ_EINVAL, "precision is fixed at mlq_init");
    ctx->eng->set_params(*p);
  });
}

const char* mlq_last_error(const mlq_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int mlq_local_size_3d(const mlq_ctx* ctx, int* local_n, int* local_k_offset) {
  if (!ctx || !local_n || !local_k_offset) return MLQ_EINVAL;
  *local_n = ctx->n;
  *local_k_offset = 0;
  return MLQ_OK;
}

intptr_t mlq_fftw_mpi_local_size_3d(intptr_t n0, intptr_t n1, intptr_t n2, int32_t comm,
                                    intptr_t* local_n0, intptr_t* local_0_start) {
  (void)comm;
  if (local_n0) *local_n0 = n0;
  if (local_0_start) *local_0_start = 0;
  return n0 * n1 * n2;
}

int mlq_diagnosticScalars(mlq_ctx* ctx, const double* u, mlq_diag* out, double* fields,
                          double* ring_xyz, long long ring_cap, long long* ring_points) {
  return guarded(ctx, [&] {
    if (!u || !out) throw Error(MLQ_EINVAL, "null argument");
    IEngine& e = *ctx->eng;
    e.upload(vec(ctx, 0), u);
    long long nr = 0;
    e.diagnostic_scalars(vec(ctx, 0), reinterpret_cast<double*>(out), fields, ring_xyz, ring_cap,
                         ring_points ? &nr : nullptr);
    if (ring_points) *ring_points = nr;
  });
}
int mlq_kappa_test_pert(mlq_ctx* ctx, mlq_field out, int kind, double m1, double m2, double m3,
                        unsigned long long seed, const double* loaded) {
  return guarded(ctx, [&] { ctx->eng->kappa_test_pert(out, kind, m1, m2, m3, seed, loaded); });
}

int mlq_sync(mlq_ctx* ctx) { return guarded(ctx, [&] { ctx->eng->sync(); }); }

// ---- fftwfunction ----
int mlq_fftfwd(mlq_ctx* ctx, const double* u, double* fu) {
  return guarded(ctx, [&] {
    ctx->eng->fft_c2c(u, fu, false, 1);
    check_nan(fu, 2 * nreal(ctx));
  });
}
int mlq_fftbwd(mlq_ctx* ctx, const double* fu, double* u) {
  return guarded(ctx, [&] { ctx->eng->fft_c2c(fu, u, true, 1); });
}
int mlq_fftfwdV(mlq_ctx* ctx, const double* v, double* fv) {
  return guarded(ctx, [&] {
    ctx->eng->fft_c2c(v, fv, false, 3);
    check_nan(fv, 6 * nreal(ctx));
  });
}
int mlq_fftbwdV(mlq_ctx* ctx, const double* fv, double* v) {
  return guarded(ctx, [&] { ctx->eng->fft_c2c(fv, v, true, 3); });
}
int mlq_fftfwdr_noalloc(mlq_ctx* ctx, const double* u_dev, double* fu_dev) {
  return guarded(ctx, [&] {
    fftfwdr_noalloc(ctx->n, ctx->eng->stream(), ctx->eng->twiddles(), u_dev, fu_dev);
  });
}
int mlq_fftfwdr(mlq_ctx* ctx, const double* u, double* fu) {
  return guarded(ctx, [&] { ctx->eng->fft_r2c_host(u, fu); });
}
int mlq_fftbwdr(mlq_ctx* ctx, const double* fu, double* u) {
  return guarded(ctx, [&] { ctx->eng->fft_c2r_host(fu, u); });
}

// ---- function_ops on host arrays ----
int mlq_gradient(mlq_ctx* ctx, const double* u, double* g) {
  return guarded(ctx, [&] {
    IEngine& e = *ctx->eng;
    e.upload(sca(ctx, 0), u);
    e.gradient(sca(ctx, 0), vec(ctx, 0));
    e.download(vec(ctx, 0), g);
  });
}
int mlq_divergence(mlq_ctx* ctx, const double* u, double* d) {
  return guarded(ctx, [&] {
    IEngine& e = *ctx->eng;
    e.upload(vec(ctx, 0), u);
    e.divergence(vec(ctx, 0), sca(ctx, 0));
    e.download(sca(ctx, 0), d);
  });
}
int mlq_laplacian(mlq_ctx* ctx, double* f) {
  return guarded(ctx, [&] {
    IEngine& e = *ctx->eng;
    e.upload(vec(ctx, 0), f);
    e.laplacian(vec(ctx, 0));
    e.download(vec(ctx, 0), f);
  });
}
int mlq_solve_poisson(mlq_ctx* ctx, const double* f, double l, double* phi) {
  return guarded(ctx, [&] {
    IEngine& e = *ctx->eng;
    e.upload(sca(ctx, 0), f);
    e.solve_poisson(sca(ctx, 0), l, sca(ctx, 1));
    e.download(sca(ctx, 1), phi);
  });
}
int mlq_dealias_scalar(mlq_ctx* ctx, double* g, double order) {
  return guarded(ctx, [&] {
    IEngine& e = *ctx->eng;
    e.upload(sca(ctx, 0), g);
    e.dealias_scalar(sca(ctx, 0), order);
    e.download(sca(ctx, 0), g);
  });
}
int mlq_divAvg_free(mlq_ctx* ctx, double* f) {
  return guarded(ctx, [&] {
    IEngine& e = *ctx->eng;
    e.upload(vec(ctx, 0), f);
    e.divAvg_free(vec(ctx, 0));
    e.download(vec(ctx, 0), f);
  });
}
int mlq_HilbertGradient(mlq_ctx* ctx, double* g, double order) {
  return guarded(ctx, [&] {
    IEngine& e = *ctx->eng;
    e.upload(vec(ctx, 0), g);
    e.HilbertGradient(vec(ctx, 0), order);
    e.download(vec(ctx, 0), g);
  });
}
int mlq_calc0thFourierModes(mlq_ctx* ctx, const double* v, double out[3]) {
  return guarded(ctx, [&] {
    IEngine& e = *ctx->eng;
    e.upload(vec(ctx, 0), v);
    e.calc0thFourierModes(vec(ctx, 0), out);
  });
}
int mlq_calc_uk(mlq_ctx* ctx, const double* u, double k, double* uk) {
  return guarded(ctx, [&] {
    IEngine& e = *ctx->eng;
    e.upload(vec(ctx, 0), u);
    e.calc_uk(vec(ctx, 0), k, sca(ctx, 0));
    e.download(sca(ctx, 0), uk);
  });
}
int mlq_calc_gk_order2(mlq_ctx* ctx, const double* g, double k, double* gk) {
  return guarded(ctx, [&] {
    IEngine& e = *ctx->eng;
    e.upload(sca(ctx, 0), g);
    e.calc_gk_order2(sca(ctx, 0), k, sca(ctx, 1));
    e.download(sca(ctx, 1), gk);
  });
}
int mlq_calculate_spectral_data(mlq_ctx* ctx, const double* u, int normalize,
                                double* spectral_data) {
  return guarded(ctx, [&] {
    IEngine& e = *ctx->eng;
    e.upload(vec(ctx, 0), u);
    e.spectral_data(vec(ctx, 0), normalize, spectral_data);
  });
}
int mlq_calc_global_summed_enstrophy(mlq_ctx* ctx, const double* u, double* enstrophy) {
  return guarded(ctx, [&] {
    IEngine& e = *ctx->eng;
    e.upload(vec(ctx, 0), u);
    *enstrophy = e.enstrophy(vec(ctx, 0));
  });
}
int mlq_calc_global_dEnstrophydt(mlq_ctx* ctx, const double* u, double* dEdt) {
  return guarded(ctx, [&] {
    IEngine& e = *ctx->eng;
    e.upload(vec(ctx, 0), u);
    *dEdt = e.dEnstrophydt(vec(ctx, 0));
  });
}
int mlq_spectral_data_field(mlq_ctx* ctx, mlq_field u, int normalize, double* spectral_data) {
  return guarded(ctx, [&] { ctx->eng->spectral_data(u, normalize, spectral_data); });
}
int mlq_calc_global_Lq_norm(mlq_ctx* ctx, const double* u, double* norm) {
  return guarded(ctx, [&] {
    IEngine& e = *ctx->eng;
    e.upload(vec(ctx, 0), u);
    *norm = e.Lq_norm(vec(ctx, 0));
  });
}
int mlq_rescaleLqNorm(mlq_ctx* ctx, double* u, double B) {
  return guarded(ctx, [&] {
    IEngine& e = *ctx->eng;
    e.upload(vec(ctx, 0), u);
    e.rescale(vec(ctx, 0), B);
    e.download(vec(ctx, 0), u);
  });
}
int mlq_calcConstraintDerivativeL2(mlq_ctx* ctx, const double* u, double* out) {
  return guarded(ctx, [&] {
    IEngine& e = *ctx->eng;
    e.upload(vec(ctx, 0), u);
    e.constraint_derivative(vec(ctx, 0), vec(ctx, 1));
    e.download(vec(ctx, 1), out);
  });
}
int mlq_vectorTransport(mlq_ctx* ctx, const double* u, double q, double B, const double* eta,
                        const double* xi, double* out) {
  return guarded(ctx, [&] {
    IEngine& e = *ctx->eng;
    e.upload(vec(ctx, 0), u);
    e.upload(vec(ctx, 1), eta);
    e.upload(vec(ctx, 2), xi);
    e.vector_transport(vec(ctx, 0), q, B, vec(ctx, 1), vec(ctx, 2), vec(ctx, 3));
    e.download(vec(ctx, 3), out);
  });
}
int mlq_inner_product(mlq_ctx* ctx, const double* f, const double* g, int kind, double* out) {
  return guarded(ctx, [&] {
    IEngine& e = *ctx->eng;
    e.upload(sca(ctx, 0), f);
    e.upload(sca(ctx, 1), g);
    *out = e.inner(sca(ctx, 0), sca(ctx, 1), kind);
  });
}
int mlq_global_inner_product(mlq_ctx* ctx, const double* f, const double* g, int kind,
                             double* out) {
  return mlq_inner_product(ctx, f, g, kind, out);
}
int mlq_global_summed_field_inner_product(mlq_ctx* ctx, const double* f, const double* g,
                                          int kind, double* out) {
  return guarded(ctx, [&] {
    IEngine& e = *ctx->eng;
    e.upload(vec(ctx, 0), f);
    e.upload(vec(ctx, 1), g);
    *out = e.inner(vec(ctx, 0), vec(ctx, 1), kind);
  });
}
int mlq_calc_local_dLqdt_inclParts(mlq_ctx* ctx, const double* u, double q, double out[3]) {
  return guarded(ctx, [&] {
    IEngine& e = *ctx->eng;
    e.upload(vec(ctx, 0), u);
    e.dLqdt_parts(vec(ctx, 0), q, out);
  });
}
int mlq_calc_global_dLqdt_inclParts(mlq_ctx* ctx, const double* u, double q, double out[3]) {
  return mlq_calc_local_dLqdt_inclParts(ctx, u, q, out);
}
int mlq_GradL2ForLq(mlq_ctx* ctx, const double* U, double* V) {
  return guarded(ctx, [&] {
    IEngine& e = *ctx->eng;
    e.upload(vec(ctx, 0), U);
    e.grad_L2(vec(ctx, 0), vec(ctx, 1));
    e.download(vec(ctx, 1), V);
    check_nan(V, 3 * nreal(ctx));
  });
}
int mlq_eval_J(mlq_ctx* ctx, const double* field, int sys, double* J) {
  return guarded(ctx, [&] {
    IEngine& e = *ctx->eng;
    e.upload(vec(ctx, 0), field);
    *J = e.eval_J(vec(ctx, 0), sys);
  });
}

// ---- device-resident fields ----
int mlq_field_alloc(mlq_ctx* ctx, int ncomp, mlq_field* f) {
  return guarded(ctx, [&] {
    if (!f) throw Error(MLQ_EINVAL, "null handle");
    *f = ctx->eng->field_alloc(ncomp);
  });
}
int mlq_field_free(mlq_ctx* ctx, mlq_field f) {
  return guarded(ctx, [&] { ctx->eng->field_free(f); });
}
int mlq_upload(mlq_ctx* ctx, mlq_field f, const double* host) {
  return guarded(ctx, [&] { ctx->eng->upload(f, host); });
}
int mlq_download(mlq_ctx* ctx, mlq_field f, double* host) {
  return guarded(ctx, [&] { ctx->eng->download(f, host); });
}
int mlq_seeded_field(mlq_ctx* ctx, mlq_field f, unsigned long long seed, int kind, double a) {
  return guarded(ctx, [&] {
    if (ctx->eng->field_ncomp(f) != 3) throw Error(MLQ_EINVAL, "seeded field needs 3 components");
    ctx->eng->seeded(f, seed, kind, a);
  });
}
int mlq_eval_J_batch(mlq_ctx* ctx, mlq_field u, mlq_field d, const double* taus, int ntau,
                     double* J) {
  return guarded(ctx, [&] {
    if (ntau < 0 || (ntau > 0 && (!taus || !J))) throw Error(MLQ_EINVAL, "bad tau list");
    ctx->eng->eval_J_batch(u, d, taus, ntau, J);
    for (int i = 0; i < ntau; ++i)
      if (std::isnan(J[i])) throw Error(MLQ_ENAN, "nan in J");
  });
}
int mlq_rcg_path(const mlq_ctx* ctx, int* resident_state) {
  return guarded(const_cast<mlq_ctx*>(ctx), [&] {
    if (!resident_state) throw Error(MLQ_EINVAL, "null output");
    *resident_state = ctx->eng->rcg_path();
  });
}
int mlq_time_linesearch(mlq_ctx* ctx, mlq_field u, mlq_field d, const double* taus, int ntau,
                        int batch, double* J, double* t_ms) {
  return guarded(ctx, [&] {
    if (ntau < 1 || !taus || !J || !t_ms || batch < -1) throw Error(MLQ_EINVAL, "bad arguments");
    ctx->eng->time_linesearch(u, d, taus, ntau, batch, J, t_ms);
  });
}
int mlq_kappa_test(mlq_ctx* ctx, mlq_field phi, mlq_field grad, mlq_field pert, int mysystem,
                   int inner_kind, double* table, double* frac, double* best_kappa, int* best_row) {
  return guarded(ctx, [&] {
    if (!table || !frac || !best_kappa || !best_row) throw Error(MLQ_EINVAL, "null output");
    if (mysystem != MLQ_J_MAXDLQDT && mysystem != MLQ_J_LINEMIN)
      throw Error(MLQ_EBADKIND, "kappa test system must be maxdLqdt or LineMin");
    if (inner_kind < MLQ_IP_L2 || inner_kind > MLQ_IP_H) throw Error(MLQ_EBADKIND, "bad inner product");
    ctx->eng->kappa_test(phi, grad, pert, mysystem, inner_kind, table, frac, best_kappa, best_row);
  });
}
int mlq_rcg_solve(mlq_ctx* ctx, mlq_field u, const mlq_rcg_cfg* cfg, mlq_rcg_result* res,
                  mlq_log_cb cb, void* user) {
  return guarded(ctx, [&] {
    if (!cfg || !res) throw Error(MLQ_EINVAL, "null cfg/result");
    LogSink log{cb, user};
    std::memset(res, 0, sizeof(*res));
    ctx->eng->rcg_solve(u, *cfg, *res, log);
  });
}

// ---- measurement ----
int mlq_time_fft(mlq_ctx* ctx, int batch, int reps, double* t_ms) {
  return guarded(ctx, [&] { ctx->eng->time_fft(batch, reps, t_ms); });
}
int mlq_autotune(mlq_ctx* ctx, int batch, double* times, int* choice) {
  return guarded(ctx, [&] { ctx->eng->autotune(batch, times, choice); });
}
int mlq_set_tiles(mlq_ctx* ctx, const int* choice) {
  return guarded(ctx, [&] {
    if (!choice) throw Error(MLQ_EINVAL, "null argument");
    for (int p = 0; p < 8; ++p)
      if (choice[p] < 0 || choice[p] > 7) throw Error(MLQ_EINVAL, "tile choice out of range 0..7");
    ctx->eng->set_tiles(choice);
  });
}
int mlq_time_copy(mlq_ctx* ctx, int nfields, int reps, double* t_ms) {
  return guarded(ctx, [&] { ctx->eng->time_copy(nfields, reps, t_ms); });
}
int mlq_time_column_copy(mlq_ctx* ctx, int axis, int tile, int nfields, int reps, double* t_ms,
                         double* bytes) {
  return guarded(ctx, [&] { ctx->eng->time_column_copy(axis, tile, nfields, reps, t_ms, bytes); });
}
int mlq_sizes(const mlq_ctx* ctx, double* rb, double* sb) {
  if (!ctx || !rb || !sb) return MLQ_EINVAL;
  *rb = ctx->eng->real_bytes();
  *sb = ctx->eng->spec_bytes();
  return MLQ_OK;
}
int mlq_prof_enable(mlq_ctx* ctx, int on) {
  return guarded(ctx, [&] { ctx->eng->prof_enable(on != 0); });
}
int mlq_prof_get(mlq_ctx* ctx, int cls, long long* count, double* ms, double* bytes) {
  return guarded(ctx, [&] { ctx->eng->prof_get(cls, count, ms, bytes); });
}
const char* mlq_prof_name(int cls) {
  static const char* names[PC_COUNT] = {"fft_x_fwd", "fft_y_fwd", "fft_z_fwd", "fft_z_inv",
                                        "fft_y_inv", "fft_x_inv", "pointwise", "spectral",
                                        "reduce", "linesearch_pointwise", "fft_z_roundtrip"};
  return (cls >= 0 && cls < PC_COUNT) ? names[cls] : "?";
}
int mlq_mem_live(mlq_ctx* ctx, double* live, double* peak) {
  return guarded(ctx, [&] { ctx->eng->mem_live(live, peak); });
}
int mlq_mem_info(mlq_ctx* ctx, double* used, double* peak, double* dfree, double* dtotal) {
  return guarded(ctx, [&] {
    ctx->eng->mem_info(used, peak);
    size_t f = 0, t = 0;
    MLQ_HIP(hipMemGetInfo(&f, &t));
    *dfree = (double)f;
    *dtotal = (double)t;
  });
}

int mlq_transform_counts(const mlq_ctx* ctx, long long* fwd_fields, long long* inv_fields) {
  if (!ctx || !fwd_fields || !inv_fields) return MLQ_EINVAL;
  ctx->eng->transform_counts(fwd_fields, inv_fields);
  return MLQ_OK;
}

}  // extern "C"
