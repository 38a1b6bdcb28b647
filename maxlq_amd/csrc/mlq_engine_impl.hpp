// Engine<T>: device memory, batched 3-D transforms, primitives and the
// reference field operators for one grid, in precision T (double: the FP64
// parity path; float: the mixed path, whose reductions and scalars stay FP64).
#pragma once
#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <vector>

#include "mlq_engine.hpp"
#include "mlq_pointwise.hpp"
#include "mlq_slab.hpp"

namespace mlq {

// MLQ_ALLOC_LOG=1: one stderr line per device allocation, pool flush and free,
// with the host milliseconds each took; 2 adds each temporary taken from and
// returned to the pool
static inline double alloc_log_ms() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch()).count();
}
static inline int alloc_log_level() {
  static const int level = [] {
    const char* s = std::getenv("MLQ_ALLOC_LOG");
    return s ? std::max(1, std::atoi(s)) : 0;
  }();
  return level;
}

template <typename T> class Engine;
template <typename T> class FastRcg;
template <typename T>
void rcg_solve_fast(Engine<T>&, int, const mlq_rcg_cfg&, mlq_rcg_result&, const LogSink&);

// RAII handle of a pooled temporary: k contiguous real fields or half spectra.
template <typename T, typename P> class Tmp {
 public:
  Tmp() = default;
  Tmp(Engine<T>* e, P* p, size_t bytes) : e_(e), p_(p), bytes_(bytes) {}
  Tmp(Tmp&& o) noexcept : e_(o.e_), p_(o.p_), bytes_(o.bytes_) { o.p_ = nullptr; }
  Tmp& operator=(Tmp&& o) noexcept {
    reset();
    e_ = o.e_; p_ = o.p_; bytes_ = o.bytes_; o.p_ = nullptr;
    return *this;
  }
  Tmp(const Tmp&) = delete;
  Tmp& operator=(const Tmp&) = delete;
  ~Tmp() { reset(); }
  void reset();
  P* get() const { return p_; }
  operator P*() const { return p_; }

 private:
  Engine<T>* e_ = nullptr;
  P* p_ = nullptr;
  size_t bytes_ = 0;
};

// Coefficient fields of the generic line search for phi = u + tau d1, d1 =
// d + gam u: D(phi.phi) = S0 + tau S1 + tau^2 S2, the pressure P0 + tau P1 +
// tau^2 P2 (P: 3 contiguous fields), |grad phi|^2 = Q0 + tau Q1 + tau^2 Q2,
// w_j = D(sum_i phi_i d_j phi_i) = a_j + tau b_j + tau^2 e_j, u and d (3
// contiguous fields each).
template <typename T> struct LsGen {
  const T *S0, *S1, *S2, *P, *Q0, *Q1, *Q2, *a, *b, *e, *u, *d;
  double gam;
};

template <typename T> class Engine final : public IEngine {
 public:
  using C = c2<T>;
  using RT = Tmp<T, T>;
  using ST = Tmp<T, C>;

  Engine(int n, const mlq_params& p);
  ~Engine() override;

  int n() const override { return n_; }
  const mlq_params& params() const override { return p_; }
  void set_params(const mlq_params& p) override { p_ = p; resolve_params(); }
  void sync() override { MLQ_HIP(hipStreamSynchronize(st_)); }
  // wait for the stream, polling (see the definition)
  void wait_spin();
  double real_bytes() const override { return (double)nr_ * sizeof(T); }
  // algorithmic bytes of one half spectrum (the row padding is never touched)
  double spec_bytes() const override { return (double)n_ * n_ * px_ * sizeof(C); }
  // algorithmic bytes of the transform passes under a band max|m_j| <= keep
  // (keep < 0: none): kx columns and ky (or kz) rows a pruned pass touches
  double ab_cols(int keep) const { return keep >= 0 ? keep + 1.0 : (double)px_; }
  double ab_rows(int keep) const { return keep >= 0 ? 2.0 * keep + 1.0 : (double)n_; }
  // one x-y plane row of columns, the kept (kx, ky) columns, the kept cube
  double ab_xcols(int keep) const { return ab_cols(keep) * n_ * n_ * sizeof(C); }
  double ab_kcols(int keep) const { return ab_cols(keep) * ab_rows(keep) * n_ * sizeof(C); }
  double ab_cube(int keep) const {
    return ab_cols(keep) * ab_rows(keep) * ab_rows(keep) * sizeof(C);
  }

  int field_alloc(int ncomp) override;
  void field_free(int h) override;
  void upload(int h, const double* host) override;
  void download(int h, double* host) override;
  int field_ncomp(int h) const override { return fld(h).ncomp; }

  void fft_c2c(const double* in, double* out, bool inverse, int ncomp) override;
  void fft_r2c_host(const double* in, double* out) override;
  void fft_c2r_host(const double* in, double* out) override;

  void gradient(int u_sc, int out_vec) override;
  void divergence(int u_vec, int out_sc) override;
  void laplacian(int f_vec) override;
  void solve_poisson(int f_sc, double l, int out_sc) override;
  void dealias_scalar(int g_sc, double order) override;
  void divAvg_free(int f_vec) override;
  void HilbertGradient(int f_vec, double order) override;
  void calc0thFourierModes(int v, double out[3]) override;
  void calc_uk(int u_vec, double k, int out_sc) override;
  void calc_gk_order2(int g_sc, double k, int out_sc) override;
  double Lq_norm(int u_vec) override;
  void rescale(int u_vec, double B) override;
  void constraint_derivative(int u_vec, int out_vec) override;
  void vector_transport(int u, double q, double B, int eta, int xi, int out) override;
  double inner(int f, int g, int kind) override;
  void dLqdt_parts(int u_vec, double q, double out[3]) override;
  void spectral_data(int u_vec, int normalize, double* out) override;
  double enstrophy(int u_vec) override;
  double dEnstrophydt(int u_vec) override;
  double enstrophy_dev(const T* u);
  double dEnstrophydt_dev(const T* u);
  void spectral_data_dev(const T* u, bool normalize, double* out);
  // spectra records of the solve (save_spectraEveryXiteration)
  static bool spectra_due(const mlq_rcg_cfg& cfg, int iter) {
    return cfg.spectra_every > 0 && iter % cfg.spectra_every == 0;
  }
  void spectrum_of(const mlq_rcg_cfg& cfg, const T* u, std::vector<double>& out) {
    out.assign(2 * (size_t)n_, 0.0);
    spectral_data_dev(u, !(cfg.flags & MLQ_RCG_SPECTRA_RAW), out.data());
  }
  void emit_spectrum(const LogSink& log, const mlq_rcg_cfg& cfg, int iter, int which, const T* u) {
    std::vector<double> s;
    spectrum_of(cfg, u, s);
    log.emit(MLQ_LOG_SPECTRUM, iter, which, 0, s.data());
  }
  // the iterate as an MLQ_LOG_FIELD record (save_field, optimization.f90:388-390)
  void emit_field(const LogSink& log, int iter, int u) {
    field_host_.resize(3 * (size_t)nr_);
    download(u, field_host_.data());
    log.emit(MLQ_LOG_FIELD, iter, iter, 0, field_host_.data());
  }
  std::vector<double> field_host_;
  void grad_L2(int u_vec, int out_vec) override;
  double eval_J(int f_vec, int system) override;
  void eval_J_batch(int u, int d, const double* taus, int ntau, double* J) override;
  void time_linesearch(int u, int d, const double* taus, int ntau, int batch, double* J,
                       double* t_ms) override;
  void kappa_test(int phi, int grad, int pert, int system, int kind, double* table, double* frac,
                  double* best_kappa, int* best_row) override;
  void seeded(int f_vec, unsigned long long seed, int kind, double a) override;
  void rcg_solve(int u, const mlq_rcg_cfg& cfg, mlq_rcg_result& res, const LogSink& log) override;
  void time_fft(int batch, int reps, double* t_ms) override;
  void time_copy(int nfields, int reps, double* t_ms) override;
  void time_column_copy(int axis, int var, int nfields, int reps, double* t_ms,
                        double* bytes) override;
  void autotune(int batch, double* times, int* choice) override;
  void set_tiles(const int* choice) override {
    for (int p = 0; p < 8; ++p) svar_[p] = choice[p];
  }
  void prof_enable(bool on) override;
  void prof_get(int cls, long long* count, double* ms, double* bytes) override;
  void mem_info(double* used, double* peak) override { *used = (double)used_; *peak = (double)peak_; }
  void mem_live(double* live, double* peak) override {
    *live = (double)(used_ - pooled_);
    *peak = (double)live_peak_;
  }
  hipStream_t stream() const override { return st_; }
  const double2* twiddles() const override { return tw_; }
  void transform_counts(long long* fwd, long long* inv) const override {
    *fwd = nfwd_;
    *inv = ninv_;
  }
  long long nfwd_ = 0, ninv_ = 0;   // fields transformed since creation
  // producer x passes (k_xfwd_p) enabled; MLQ_FUSED=0 at context creation
  // selects the pointwise-kernel + transform form instead (same results up to
  // rounding; kept for tests of that form)
  bool fused_ = true;

  // diagnosticScalars (function_ops.f90:2503-2602): out[21] in the mlq_diag
  // layout; fields (4 n^3: |u|, |w|, u.w, Q / max Q) and the ring points
  // (up to cap, 3 coordinates each; *nring counts them all) when not null
  void diag_scalars(const T* u, double* out, double* fields, double* ring, long long cap,
                    long long* nring);
  void diagnostic_scalars(int u_vec, double* out, double* fields, double* ring, long long cap,
                          long long* nring) override {
    if (fld(u_vec).ncomp != 3) throw Error(MLQ_EINVAL, "diagnosticScalars of a vector field");
    diag_scalars(fphys(u_vec), out, fields, ring, cap, nring);
  }
  // kappa_test_pert (function_ops.f90:613-977) into the vector field `out`
  void kappa_pert(int out, int kind, double m1, double m2, double m3, unsigned long long seed,
                  const double* loaded);
  void kappa_test_pert(int out, int kind, double m1, double m2, double m3, unsigned long long seed,
                       const double* loaded) override {
    kappa_pert(out, kind, m1, m2, m3, seed, loaded);
  }
  // The sums of R(phi) at one step size from the generic line-search fields
  // (the |phi|^k chain of calc_uk on D(phi.phi)): out = {sum |phi|^(q-2)
  // |grad phi|^2, sum D(|phi|^(q-4)) w.w, sum D(p |phi|^(q-4)) phi.w,
  // sum (|phi|^(q/2))^2}, unnormalised (function_ops.f90:1304-1358, 1675-1692)
  // the chain of |phi|^k is that of sqrt(s2) phi (D(phi.phi) scaled by s2);
  // the other factors of the sums are those of phi
  void ls_generic(const LsGen<T>& f, double t, double out[4], double s2 = 1.0);
  // q = 6: the polynomial coefficients of the R sums, and the norm sum at t
  void ls6_coefs(const LsGen<T>& f, double r1[7], double r2[7], double r3[8]);
  double ls6_norm(const LsGen<T>& f, double t);
  // an FP64 engine on the same grid and parameters (created on first use):
  // the mixed path evaluates its line-search J(tau) chain there, so that J is
  // a smooth function of tau (brent resolves tau to TOL = 1e-6, J to 1e-12)
  Engine<double>* fp64_twin();

  // device temporaries: k real fields (rtmp) or k spectra (stmp), carved
  // out of slabs and given back to them when the Tmp goes
  RT rtmp(int k = 1);
  ST stmp(int k = 1);
  void pool_put(void* p, size_t bytes);

 private:
  // ---- set-up ----
  void resolve_params();
  void* dmalloc(size_t bytes);
  void dfree(void* p, size_t bytes);
  DevField& fld(int h);
  const DevField& fld(int h) const;
  T* fphys(int h, int c = 0) { return (T*)fld(h).phys + (long long)c * nr_; }

  // ---- launch helpers ----
  template <class F> void map(const F& f, long long n, int cls = PC_POINTWISE);
  template <class F> void smap(const F& f) { map(f, ns_, PC_SPECTRAL); }
  template <int K, class F> void reduce(const F& f, long long n, double* out);
  // the complex values at device addresses p[0..n) (n <= FETCH_MAX), read in
  // one transfer to pinned host memory and one synchronisation
  void fetch(int n, const C* const* p, std::complex<double>* out);
  void prof_begin(int cls);
  void prof_end(int cls, double bytes);

  // ---- transforms ----
  // forward: src[b] (physical) -> dst[b] (spectrum) = op[b] * FFT(src[b])
  void fwd(int nb, const T* const* src, C* const* dst, const SpecOp* ops);
  // inverse: dst[b] = REAL(IFFT(op[b] * sum_i coef[b][i] * src[b][i])); keep >= 0
  // declares every input zero outside the cube max|m_j| <= keep (pruned passes)
  struct InvSrc { const C* s[3]; double c[3]; int ns; };
  void inv(int nb, const InvSrc* src, const SpecOp* ops, T* const* dst, int keep = -1);
  template <int N> void fwd_n(int nb, const T* const* src, C* const* dst, const SpecOp* ops);
  template <int N> void fwd_yz_n(int nb, C* const* dst, const SpecOp* ops, int keep,
                                 bool kept_only = false);
  int fwd_keep(int nb, const SpecOp* ops) const;
  int inv_keep(int nb, const SpecOp* ops, int keep) const;
  // forward transforms of the P::NB fields that the producer P computes per
  // point (no intermediate fields); false when it cannot run (the caller
  // then falls back to a pointwise kernel and fwd)
  template <class P> bool fwd_prod(const P& prod, C* const* dst, const SpecOp* ops);
  template <int N, class P> void fwd_prod_n(const P& prod, C* const* dst, const SpecOp* ops, bool& done);
  template <int N> void inv_n(int nb, const InvSrc* src, const SpecOp* ops, T* const* dst, int keep);
  template <int N>
  XArgs<T> inv_zy_n(int nb, const InvSrc* src, const SpecOp* ops, T* const* dst, int keep);
  // inverse transforms of the P::NI inputs (written to idst[b] where set) and
  // forward transforms of the P::NB fields P computes from them per point,
  // with one x pass for both; false (nothing launched) when the batch cannot
  // run that way (the caller then runs inv and fwd_prod)
  template <class P>
  bool inv_fwd_prod(const InvSrc* src, const SpecOp* iops, T* const* idst, int ikeep, const P& prod,
                    C* const* dst, const SpecOp* ops);
  template <int N, class P>
  void inv_fwd_prod_n(const InvSrc* src, const SpecOp* iops, T* const* idst, int ikeep,
                      const P& prod, C* const* dst, const SpecOp* ops, bool& done);

  // convenience
  void fwd1(const T* src, C* dst, SpecOp op = sp_none()) { fwd(1, &src, &dst, &op); }
  void inv1(const C* src, T* dst, SpecOp op = sp_none()) {
    InvSrc s{{src, nullptr, nullptr}, {1.0, 0.0, 0.0}, 1};
    inv(1, &s, &op, &dst);
  }
  void dealias_dev(T* const* f, int nb, double order);  // in place, batch
  // out[i] = a[i] b[i], dealiased (order `order`) when `dealias`: the product
  // is formed by the forward x pass of the round trip instead of being stored
  // first.  out[i] may be a[i] or b[i].
  void mul_dealias(int nb, const T* const* a, const T* const* b, T* const* out, bool dealias,
                   double order = 2.0);
  void mul_dealias(const T* a, const T* b, T* out, bool dealias) {
    mul_dealias(1, &a, &b, &out, dealias);
  }
  // Round trips dst = REAL(IFFT(op * FFT(src[in]))) of physical fields, nout
  // outputs of nin inputs; the z transforms are fused (k_zrt), so no
  // spectrum is stored.  An output may overwrite its own input only when it
  // is that input's only output (dealias in place).  mul: when given, input
  // b is the pointwise product src[b] * mul[b] (mul[b] may be null).
  struct RtOut { int in; SpecOp op; T* dst; };
  void rtrip(int nin, const T* const* src, int nout, const RtOut* outs,
             const T* const* mul = nullptr);
  template <int N>
  void rtrip_n(int nin, const T* const* src, int nout, const RtOut* outs, int keep,
               const T* const* mul);
  // slot plan of one rtrip launch group; returns the scratch slots it needs
  struct RtPlan {
    int nz = 0, zin[MAXB], zout[MAXB];
    SpecOp zop[MAXB];
    int ny = 0, yz[MAXB], yk[MAXB], yslot[MAXB];
    int oy[MAXB], oxik[MAXB];
    int need = 0;
  };
  int rtrip_plan(int nin, int nout, const RtOut* outs, RtPlan& pl) const;
  // fused q = 4 line-search sums for a batch of steps (k_xfwd_ls, pruned y
  // pass, k_zls): out[2b] = sum_kept m |A_b|^2, out[2b+1] = sum_kept m
  // Re(conj(A_b) Q(tau_b)); requires 0 <= keep < n/2 and nb <= MAXB
  void ls_q4_sums(const T* S0, const T* S1, const T* S2, const C* Q0, const C* Q1, const C* Q2,
                  const double* taus, int nb, int keep, double* out);
  // the points where max(S0 + t S1 + t^2 S2, 0) can clip for some t
  // (k_ls_risky): count, and the (S0, S1, S2) list when it fits its capacity
  void ls_risky(const T* S0, const T* S1, const T* S2, std::vector<double>& list,
                long long& count);
  template <int N>
  void ls_q4_sums_n(const T* S0, const T* S1, const T* S2, const C* Q0, const C* Q1, const C* Q2,
                    const double* taus, int nb, int keep, double* out);
  int mcut_of(double order) const { return dealias_mcut(n_, order); }
  bool mask_is_identity(int mcut) const { return mcut >= n_ / 2; }

  // ---- operators on device pointers (vector = 3 components, stride nr_) ----
  struct Chain {            // |u|^k with the calc_uk / calc_gk_order2 recursion
    RT g, G2;
    std::map<long long, RT> memo;
    bool haveG2 = false;
  };
  void chain_init(Chain& ch, const T* u, const T* s_pre = nullptr);
  const T* chain_pow(Chain& ch, double k);
  RT chain_take(Chain& ch, double k);   // |u|^k, ownership moved out of the chain
  void spec3(const T* u, C* const* uh);                      // 3 forward transforms
  void grad9(C* const* uh, T* const* G);                     // G[i*3+j] = d_j u_i
  void grad9_rt(const T* u, T* G);                           // the same by round trips
  void dLqdt_dev(const T* u, double q, double out[3]);
  void gradL2_dev(const T* u, T* const* V);
  double lq_norm_dev(const T* u);
  double spec_dot(const C* a, const C* b, int kind);         // weighted Parseval sum
  double dot(const T* a, const T* b);
  double sum(const T* a);
  bool has_nan(const T* a, long long n);

  // ---- state ----
  int n_, px_, pp_;   // grid size, n/2+1 modes along x, spectral row pitch
  int sb_ = MAXB;     // spectra in the transform scratch (largest transform batch)
  // tile variant of the strided passes {y fwd, z fwd, z inv, y inv} (autotuned)
  int svar_[8] = {0, 0, 0, 0, 0, 0, 0, 0};   // {y fwd, z fwd, z inv, y inv} x {full, band-limited}
  template <int N> double time_pass_n(int pass, int var, int batch, int reps, int keep = -1,
                                      C* ext = nullptr);
  template <int N>
  double time_colcopy_n(int axis, int var, const C* src, C* dst, int nfields, int reps);
  long long nr_, ns_;
  mlq_params p_;
  bool use_e_u_ = false;
  hipStream_t st_ = nullptr;
  double2* tw_ = nullptr;   // FP64 twiddles for every field precision
  C* scratch_ = nullptr;
  double* part_ = nullptr;
  double* lspart_ = nullptr;   // partial sums of the fused line-search z pass
  size_t lspart_bytes_ = 0;
  double* lsrisky_ = nullptr;  // clipping-candidate list of the line search
  size_t lsrisky_bytes_ = 0;
  double* red_dev_ = nullptr;
  hipEvent_t done_ev_ = nullptr;     // recorded and polled by wait_spin
  double* red_host_ = nullptr;
  std::vector<DevField> fields_;
  std::unique_ptr<IEngine> twin_;   // fp64_twin()
  // Slabs of device memory that the temporaries are carved from (SlabSet: best
  // fit, neighbouring free ranges merged when a block comes back); a new slab
  // of at least 9 real fields only when no free range is large enough.  A
  // pool of blocks keyed by size runs the device out of memory when the
  // shapes change from one phase of an iteration to the next, gives its
  // blocks back and allocates them again - and device memory freed and
  // allocated again is cleared before it is handed out, ~28 ms per GB
  // (1024^3: 6 s of a 10 s iteration).  Slabs are freed only when an
  // allocation fails otherwise.
  SlabSet slabs_;
  void* carve(size_t bytes);
  size_t release_free_slabs();
  size_t used_ = 0, peak_ = 0;
  // bytes of the slabs not carved out (allocated, not in use) and the
  // high-water mark of the bytes in use (allocated minus pooled)
  size_t pooled_ = 0, live_peak_ = 0;
  void note_live() { live_peak_ = std::max(live_peak_, used_ - pooled_); }
  // profiling
  bool prof_ = false;
  struct ProfRec { int cls; hipEvent_t a, b; double bytes; };
  std::vector<ProfRec> prof_recs_;
  std::vector<hipEvent_t> ev_pool_;
  long long prof_count_[PC_COUNT] = {};
  double prof_ms_[PC_COUNT] = {};
  double prof_bytes_[PC_COUNT] = {};
  void prof_flush();

  // the reference-structured solve (every operator recomputed from the field)
  void rcg_solve_ref(int u, const mlq_rcg_cfg& cfg, mlq_rcg_result& res, const LogSink& log);
  void eval_J_batch_ref(int u, int d, const double* taus, int ntau, double* J);
  // the resident-state iteration applies (q >= 4 without e_u, dealiased) and
  // its buffers fit in the device memory left (free plus pooled temporaries)
  bool fast_path_ok() const { return fast_path_applies() && fast_path_fits(); }
  bool fast_path_applies() const { return !use_e_u_ && p_.q >= 4.0 - 1e-12 && p_.toDealias; }
  bool fast_path_fits() const;
  // device bytes of the resident-state iteration at its peak (inside the
  // line-search precompute: 46 real fields and 27 half spectra at q = 4,
  // a few more real fields for the |phi|^k chain at other q)
  double fast_path_bytes() const {
    const bool q4 = std::fabs(p_.q - 4.0) < 1e-14;
    // + 9 fields: the velocity-gradient operand of the deferred state update
    // (q = 6: 2 more at the polynomial precompute, its 10 product fields
    // against the 8 chain temporaries of a J evaluation)
    const bool q6 = std::fabs(p_.q - 6.0) < 1e-14;
    double b = (q4 ? 57.0 : (q6 ? 67.0 : 65.0)) * real_bytes() + 28.0 * (double)sizeof(C) * (double)ns_;
    b += 3.0 * kappa_fields_ * real_bytes();
    // FP32 fields, q != 4: the FP64 twin's copy of the 24 line-search fields,
    // its chain temporaries (8 fields) and its transform scratch (12 spectra)
    if (!q4 && sizeof(T) == 4)
      b += (q6 ? 34.0 : 32.0) * 8.0 * (double)nr_ + 12.0 * 16.0 * (double)n_ * n_ * (n_ / 2 + 8);
    return b;
  }
  // the reference-structured iteration takes its line search from
  // coefficient fields (LeanLs) when the resident state applies but does not
  // fit, and those fields fit in the device memory left beside `held` more
  // real fields
  bool lean_ls_applies() const { return fast_path_applies() && !fast_path_fits(); }
  bool lean_ls_fits(int held) const;
 public:
  bool uses_fast_path() const override { return fast_path_ok(); }
  // (before a solve: beside its 19 component buffers)
  int rcg_path() const override {
    if (fast_path_ok()) return 1;
    return lean_ls_applies() && lean_ls_fits(19) ? 2 : 0;
  }
 private:

  template <typename U> friend class FastRcg;
  template <typename U> friend struct LeanLs;
  template <typename U> friend struct KappaInSolve;
  int kappa_fields_ = 0;   // real vector fields the solve's kappa tests hold (0 or 4)
  template <typename U>
  friend void rcg_solve_fast(Engine<U>&, int, const mlq_rcg_cfg&, mlq_rcg_result&, const LogSink&);
};

template <typename T, typename P> void Tmp<T, P>::reset() {
  if (p_ && e_) e_->pool_put(p_, bytes_);
  p_ = nullptr;
}

}  // namespace mlq
