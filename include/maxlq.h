/*
 * maxlq.h -- C ABI of the MI355X-native maxLq RCG hot path (libmaxlq.so).
 *
 * Every host array on this boundary uses the reference layout: FP64,
 * Fortran column-major with x fastest; a scalar field is n1*n2*n3 doubles
 * (reference dims (1:n(1),1:n(2),1:local_N), local_N = n3 on one GPU), a
 * vector field is 3 such blocks (dims (...,1:3)).  Complex arrays are
 * interleaved (re, im) doubles.  Fields named by an ``mlq_field`` handle live
 * in the device's HBM for the lifetime of the context; host arrays are only
 * read or written inside the call that receives them.
 *
 * Threading: one host thread per context, one HIP stream per context, not
 * re-entrant.  Scalars are returned synchronously.  All functions return an
 * MLQ_* status; algorithm outcomes (mnbrak flags, stop reasons) are carried
 * in result structs.  Each declaration cites the reference routine it
 * replaces (file:line in uniFabiB/maxLq).
 */
#ifndef MAXLQ_H
#define MAXLQ_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define MLQ_OK 0
#define MLQ_ENAN 1       /* NaN in a transform / unit vector: the reference does `stop 1`
                            (fftwfunction.f90:78-89, function_ops.f90:1843-1854) */
#define MLQ_EBADKIND 2   /* unknown/old inner-product kind: reference `call exit`
                            (function_ops.f90:3485-3495) */
#define MLQ_EHIP 3       /* HIP runtime error */
#define MLQ_EOOM 4       /* device allocation failed */
#define MLQ_EINVAL 5     /* bad argument (size not supported, bad handle, ...) */
#define MLQ_EUNSUP 6     /* configuration outside the implemented path */

/* precision modes */
#define MLQ_FP64 0       /* all field arithmetic in FP64 (parity path) */
#define MLQ_MIXED 1      /* FP32 transforms and pointwise work, FP64 reductions/scalars */

/* inner-product kinds (the reference's `mytype` strings, function_ops.f90:3426-3496) */
#define MLQ_IP_L2 0      /* "L2" */
#define MLQ_IP_HL 1      /* "H_l^(3/2-1/q)" */
#define MLQ_IP_H 2       /* "H^(3/2-1/q)" */
#define MLQ_IP_OLD_H 3   /* "H^((3q-1)/(2q))"   -> MLQ_EBADKIND */
#define MLQ_IP_OLD_HL 4  /* "H_l^((3q-1)/(2q))" -> MLQ_EBADKIND */

/* eval_J systems (optimization.f90:533-545) */
#define MLQ_J_MAXDLQDT 0 /* "maxdLqdt": R(u) */
#define MLQ_J_LINEMIN 1  /* "LineMin":  -R(B u/||u||_q) */
#define MLQ_J_LQNORM 2   /* "||u||_q" */

/* The hot-path subset of global_variables.f90 mirrored into one struct. */
typedef struct mlq_params {
  double q;              /* lebesgueQ */
  double B;              /* constraintB (global_variables.f90:115) */
  double visc;           /* visc = 1 (:8) */
  double lambda1;        /* lambda1 = 0.1 (:31) */
  double viscCoefficient;      /* 1 (:113) */
  double pressureCoefficient;  /* 1 (:114) */
  int toDealias;         /* .true. (:68) */
  int use_e_u;           /* -1: automatic as set_e_u_or_just_u (initialize.f90:75-86) */
  int dealiase_if_mult_by_e_u; /* .true. (:40) */
  int precision;         /* MLQ_FP64 or MLQ_MIXED */
} mlq_params;

typedef struct mlq_ctx mlq_ctx;
typedef int mlq_field;   /* handle of a device-resident vector (3 comp) or scalar field */

/* Fill *p with the reference defaults for exponent q (global_variables.f90). */
void mlq_default_params(mlq_params* p, double q);

/* ---- context (replaces init_fft / fft_deallocate, fftwfunction.f90:169-212,
 *      and the grid set-up of initialize, initialize.f90:131-203) ---- */
/* n1 = n2 = n3 (the reference's RESOL, initialize.f90:141-146), a power of
 * two from 8 to 1024; other grids return MLQ_EUNSUP before any device call. */
int mlq_init(mlq_ctx** ctx, int n1, int n2, int n3, const mlq_params* p);
int mlq_finalize(mlq_ctx* ctx);
int mlq_set_params(mlq_ctx* ctx, const mlq_params* p);
const char* mlq_last_error(const mlq_ctx* ctx);
/* single-rank slab: local_N = n3, offset 0 (replaces fftw_mpi_local_size_3d,
 * initialize.f90:149) */
int mlq_local_size_3d(const mlq_ctx* ctx, int* local_n, int* local_k_offset);
/* fftw_mpi_local_size_3d(n0, n1, n2, comm, local_n0, local_0_start) of FFTW's
 * MPI interface on one rank, which initialize.f90:149 calls through
 * maxlq_amd/fortran/fftw3-mpi.f03: local_n0 = n0, local_0_start = 0, returns
 * n0 n1 n2 (no context needed) */
intptr_t mlq_fftw_mpi_local_size_3d(intptr_t n0, intptr_t n1, intptr_t n2, int32_t comm,
                                    intptr_t* local_n0, intptr_t* local_0_start);
int mlq_sync(mlq_ctx* ctx);

/* ---- fftwfunction (fftwfunction.f90:62-165) ---- */
/* fftfwd(u, fu): unnormalised complex forward transform, :62-92 */
int mlq_fftfwd(mlq_ctx* ctx, const double* u_cplx, double* fu_cplx);
/* fftbwd(fu, u): complex backward transform divided by n1*n2*n3, :97-113 */
int mlq_fftbwd(mlq_ctx* ctx, const double* fu_cplx, double* u_cplx);
/* fftfwdr(u, fu): real -> (n1/2+1, n2, n3) half spectrum, :118-134 */
int mlq_fftfwdr(mlq_ctx* ctx, const double* u_real, double* fu_half);
/* mlq_fftfwdr on device arrays, allocating no device memory.  No counterpart
 * in the reference: for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  u_dev: n^3
 * FP64 values, x fastest; fu_dev: (n/2+1, n, n) complex FP64, interleaved,
 * the layout mlq_fftfwdr returns.  FP64 whatever the context's precision.
 * The intermediate spectrum lives in a per-device static array of 128 x 128
 * x 72 complex doubles (18.9 MB, held by every GPU the library runs on until
 * the process ends), so n <= 128; larger grids return MLQ_EINVAL.  Calls on
 * one GPU share that array: calls that overlap on one GPU (from any context)
 * race.  Synchronous on the context's stream. */
int mlq_fftfwdr_noalloc(mlq_ctx* ctx, const double* u_dev, double* fu_dev);
/* fftbwdr(fu, u): half spectrum -> real, divided by n1*n2*n3, :139-165 */
int mlq_fftbwdr(mlq_ctx* ctx, const double* fu_half, double* u_real);
/* fftfwdV / fftbwdV: the vector forms, :24-57 */
int mlq_fftfwdV(mlq_ctx* ctx, const double* vec_cplx, double* fvec_cplx);
int mlq_fftbwdV(mlq_ctx* ctx, const double* fvec_cplx, double* vec_cplx);

/* ---- function_ops field routines, host arrays (function_ops.f90) ---- */
int mlq_gradient(mlq_ctx* ctx, const double* u, double* grad_u);          /* :3147-3195 */
int mlq_divergence(mlq_ctx* ctx, const double* u, double* divU);          /* :3200-3243 */
int mlq_laplacian(mlq_ctx* ctx, double* f);                               /* :3248-3282 */
int mlq_solve_poisson(mlq_ctx* ctx, const double* f, double l, double* phi); /* :1951-1996 */
int mlq_dealias_scalar(mlq_ctx* ctx, double* g, double nonlinOrder);      /* :3332-3350 */
int mlq_divAvg_free(mlq_ctx* ctx, double* myfield);                       /* :2843-2918 */
int mlq_HilbertGradient(mlq_ctx* ctx, double* grad, double order);        /* :3583-3627 */
int mlq_calc0thFourierModes(mlq_ctx* ctx, const double* vec, double out[3]); /* :3048-3094 */
int mlq_calc_uk(mlq_ctx* ctx, const double* u, double k, double* uk);     /* :2324-2343 */
int mlq_calc_gk_order2(mlq_ctx* ctx, const double* g, double k, double* gk); /* :2281-2319 */
int mlq_calc_global_Lq_norm(mlq_ctx* ctx, const double* u, double* norm); /* :1675-1692 */
/* calculate_spectral_data(u, spectral_data), function_ops.f90:2614-2677:
 * the energy spectrum in spherical shells.  spectral_data is the reference's
 * (1:n1, 1:2) array in Fortran order: spectral_data[i] = 2 pi (i+1) (the
 * shell's upper bound), spectral_data[n1+i] = sum of |fftfwd(u_c)|^2 over the
 * three components and the modes with 2 pi i < |K| <= 2 pi (i+1); with
 * normalize (normalizeSpectrumByL2Norm, global_variables.f90:76, default
 * true) column 2 is scaled by ||u||_2^2 / (2 pi sum(column 2)) (:2667-2671). */
int mlq_calculate_spectral_data(mlq_ctx* ctx, const double* u, int normalize,
                                double* spectral_data);
/* the same on a device-resident vector field (no counterpart in the
 * reference: its fields are host arrays) */
int mlq_spectral_data_field(mlq_ctx* ctx, mlq_field u, int normalize, double* spectral_data);
/* calc_global_summed_enstrophy(u), function_ops.f90:2455-2467 (Enstrophy
 * :2422-2442, vel2vort :2982-3042): 1/2 ||curl u||_2^2 */
int mlq_calc_global_summed_enstrophy(mlq_ctx* ctx, const double* u, double* enstrophy);
/* calc_global_dEnstrophydt(u), function_ops.f90:1417-1481:
 * -viscCoefficient visc ||Lap u||_2^2 + <D((u.grad) u), Lap u>_L2 */
int mlq_calc_global_dEnstrophydt(mlq_ctx* ctx, const double* u, double* dEdt);
int mlq_rescaleLqNorm(mlq_ctx* ctx, double* u, double B);                 /* :1699-1709 */
int mlq_calcConstraintDerivativeL2(mlq_ctx* ctx, const double* u, double* out); /* :984-1006 */
int mlq_vectorTransport(mlq_ctx* ctx, const double* u, double q, double B,
                        const double* eta, const double* xi, double* out); /* :1716-1768 */
int mlq_inner_product(mlq_ctx* ctx, const double* f, const double* g, int kind,
                      double* out);                                       /* :3404-3503 */
int mlq_global_inner_product(mlq_ctx* ctx, const double* f, const double* g, int kind,
                             double* out);                                /* :3508-3521 */
int mlq_global_summed_field_inner_product(mlq_ctx* ctx, const double* f, const double* g,
                                          int kind, double* out);         /* :3557-3577 */
int mlq_calc_local_dLqdt_inclParts(mlq_ctx* ctx, const double* u, double q,
                                   double out[3]);                        /* :1304-1358 */
int mlq_calc_global_dLqdt_inclParts(mlq_ctx* ctx, const double* u, double q,
                                    double out[3]);                       /* :1392-1405 */
int mlq_GradL2ForLq(mlq_ctx* ctx, const double* U, double* V);            /* :1066-1211 */

/* ---- optimization (optimization.f90) ---- */
/* eval_J(myfield, mysystem), :520-547 */
int mlq_eval_J(mlq_ctx* ctx, const double* field, int mysystem, double* J);

/* ---- device-resident fields ---- */
int mlq_field_alloc(mlq_ctx* ctx, int ncomp, mlq_field* f);
int mlq_field_free(mlq_ctx* ctx, mlq_field f);
int mlq_upload(mlq_ctx* ctx, mlq_field f, const double* host);
int mlq_download(mlq_ctx* ctx, mlq_field f, double* host);
/* the project's seeded divergence-free field, written into f and rescaled to
 * ||u||_q = B (kind 0: exp(-|m|^2/a), kind 1: |m|^2 exp(-|m|^2/a)) */
int mlq_seeded_field(mlq_ctx* ctx, mlq_field f, unsigned long long seed, int kind, double a);

/* J_LineMin(tau_i) = eval_J(u + tau_i d, "LineMin") for ntau step sizes in one
 * batched evaluation (the J evaluations of mnbrak/brent, optimization.f90:666-752,
 * 833-908) */
int mlq_eval_J_batch(mlq_ctx* ctx, mlq_field u, mlq_field d, const double* taus, int ntau,
                     double* J);

/* Measurement of the batched line search: one coefficient precompute for
 * phi = u + tau d, then J_LineMin at the ntau step sizes in launches of
 * `batch` step sizes each through the fused batched transform (1: one after
 * the other; ntau: all in one batch); -1: the solver's own evaluation (the
 * exact tau-polynomials wherever no grid point clips, the batched transform
 * elsewhere); 0: the reference-structured evaluation, every J recomputed from
 * the field as eval_J does.  t_ms[0] = precompute, t_ms[1] = evaluations (ms).
 * No counterpart in the reference (its mnbrak/brent evaluate J one at a
 * time, optimization.f90:666-752). */
int mlq_time_linesearch(mlq_ctx* ctx, mlq_field u, mlq_field d, const double* taus, int ntau,
                        int batch, double* J, double* t_ms);

/* diagnosticScalars(U), function_ops.f90:2503-2602: the values it passes to
 * save_diagnosticFields_global (data_ops.f90:125-155) with w = vel2vort(u):
 * Energy(U) and Energy(W) (:2394-2412), the largest value of each component,
 * vorCore = calculate_geometric_data(U, W, aux, "Q") (:2681-2832: volume
 * where the Q criterion is positive, its maximum, and a third entry the
 * reference leaves unset, 0 here), max |u|, max |w|, Helicity (:2479-2497)
 * and the extrema of u.w; divU_L2 = inner_product(div u, div u, "L2") */
typedef struct mlq_diag {
  double K[3], E[3], Umax[3], Wmax[3], vorCore[3];
  double magUmax, magWmax, H, maxHel, minHel, divU_L2;
} mlq_diag;
/* fields (may be null): 4 n^3 doubles |u|, |w|, u.w and Q / max Q
 * (save_diagnosticScalars' magU, magW, Helicity, VortexCore); ring_xyz (may
 * be null): the coordinates ((i-1)/n, (j-1)/n, (k-1)/n) of the points with
 * |w| / max |w| > 0.95 in the reference's loop order, at most ring_cap of
 * them (calculate_ring_data :2837-2900); *ring_points counts them all */
int mlq_diagnosticScalars(mlq_ctx* ctx, const double* u, mlq_diag* out, double* fields,
                          double* ring_xyz, long long ring_cap, long long* ring_points);

/* kappa_test_pert(phi_pert, mytype, m1, m2, m3), function_ops.f90:613-977, into
 * the device field `out`.  The "load-*" kinds take the loaded field (3 n^3
 * doubles, e.g. from mlq_read_field_nc) where the reference reads its input
 * file; MLQ_PERT_RANDOM_POLY ("create-random-poly-b", which kappa_test falls
 * back to, optimization.f90:1024-1035) draws its uniform values from the
 * project's seeded generator (SplitMix64 keyed by seed, component and point)
 * where the reference calls random_number. */
#define MLQ_PERT_SINE 0             /* "sine" (:660-673) */
#define MLQ_PERT_DIVFREE_SINE 1     /* "divfree-sine" (:675-689) */
#define MLQ_PERT_LOAD_RAW 2         /* "load-te0080": the loaded field as is (:655-658) */
#define MLQ_PERT_LOAD 3             /* "load-random-a/b" (:692-702) */
#define MLQ_PERT_LOAD_SMOOTH 4      /* "load-random-smooth-a" (:704-713) */
#define MLQ_PERT_LOAD_SMOOTH_RAW 5  /* "load-random-smooth-b", no projection (:927-935) */
#define MLQ_PERT_LOAD_K 6           /* "load-k-random-a/b", factor m1 (:715-767) */
#define MLQ_PERT_LOAD_POLY 7        /* "load-random-poly-a/b" (:769-795, 846-871) */
#define MLQ_PERT_LOAD_EXP 8         /* "load-random-exp-a/b" (:873-925) */
#define MLQ_PERT_RANDOM_POLY 9      /* "create-random-poly-b" (:797-844) */
#define MLQ_PERT_RANDOM 10          /* "save-random": the field it saves (:937-972); the
                                       caller writes the file */
int mlq_kappa_test_pert(mlq_ctx* ctx, mlq_field out, int kind, double m1, double m2, double m3,
                        unsigned long long seed, const double* loaded);

/* kappa_test (optimization.f90:957-1129): J0 = eval_J(phi), then for
 * ii = 1..16, eps = 10^(ii-15): kappa = (eval_J(phi + eps pert) - J0) /
 * (eps <grad, pert>_kind).  The perturbation is the caller's (the reference
 * loads or draws it in kappa_test_pert).  table: 16 rows {eps, <grad, pert>,
 * J1 - J0, kappa, kappa adjusted} as save_kappa_test writes them; *frac the
 * adjustment factor; best kappa and its 1-based row by the reference's
 * criterion.  mysystem: MLQ_J_MAXDLQDT or MLQ_J_LINEMIN. */
int mlq_kappa_test(mlq_ctx* ctx, mlq_field phi, mlq_field grad, mlq_field pert, int mysystem,
                   int inner_kind, double* table, double* frac, double* best_kappa, int* best_row);

/* Which iteration mlq_rcg_solve will run with the current parameters: 1 the
 * resident-state iteration (q >= 4 without the e_u variant, dealiased, and
 * its buffers fit in the device memory left), 2 the reference-structured
 * gradient and direction with the line search from coefficient fields
 * (MLQ_RCG_LEAN: where the resident state applies but does not fit), 0 the
 * reference-structured one (every operator recomputed from the field, the
 * smallest footprint).  mlq_eval_J_batch takes the resident state for 1,
 * else the reference-structured evaluation.  No counterpart in the
 * reference. */
int mlq_rcg_path(const mlq_ctx* ctx, int* resident_state);

/* ---- the whole RCG solve for one B: maxdLqdt, optimization.f90:7-514 ---- */
typedef struct mlq_rcg_cfg {
  int max_iter;          /* MAX_ITER */
  double optim_tol;      /* OPTIM_TOL (initialize.f90:95-100) */
  int start_iter;        /* optimizationIterOffset when resuming (:81-86) */
  int batch_linesearch;  /* evaluate speculative mnbrak/brent step sizes in batches */
  int reset_every;       /* resetMomentumTermEveryXiterations (25) */
  int rolling;           /* toleranceRollingAverageSize (10) */
  int flags;             /* MLQ_RCG_* bits */
  int spectra_every;     /* save_spectraEveryXiteration (global_variables.f90:74): every
                            this many iterations the solve delivers MLQ_LOG_SPECTRUM
                            records of uvec and d (optimization.f90:119-121, 377-381);
                            < 1 never */
  int save_every;        /* save_uvecEveryXiteration (global_variables.f90:70): every this
                            many iterations the solve delivers the iterate as an
                            MLQ_LOG_FIELD record (optimization.f90:388-390); < 1 never */
  int kappa_every;       /* kappaTestEveryXiteration (global_variables.f90:75): every this
                            many iterations a kappa test of the L2 gradient
                            (optimization.f90:147-149); < 1 never */
  unsigned long long kappa_seed;  /* the kappa tests' perturbation when kappa_field is null:
                            "create-random-poly-b" (m1 = -5) drawn with this seed */
  const double* kappa_field;      /* the contents of n<N>_random_b.nc (3 n^3 doubles) for
                            "load-random-poly-b" (m1 = -5, optimization.f90:1019-1031),
                            or null */
} mlq_rcg_cfg;
/* cfg.flags: run every operator from the field as the reference orders it
 * (no resident state, no tau-polynomial line search); validation path */
#define MLQ_RCG_REFERENCE 1
/* cfg.flags: spectra without the L2 normalisation (normalizeSpectrumByL2Norm
 * = .false., global_variables.f90:76) */
#define MLQ_RCG_SPECTRA_RAW 2
/* cfg.flags: kappaTest (global_variables.f90:67): kappa tests of the L2
 * gradient at the start of the solve (optimization.f90:69-72) and, at its
 * end, of the H_l gradient (and, if the solve did not converge, of the
 * direction d1) at uvec0 (:457-469); results as MLQ_LOG_KAPPA records */
#define MLQ_RCG_KAPPA 4
/* cfg.flags: the reference-structured gradient and direction with the line
 * search from coefficient fields (the 18 tau-coefficient fields of
 * D(phi.phi), w_j, the pressure and |grad phi|^2 for phi = u + tau d1,
 * formed once per iteration; each J(tau) then evaluates the |phi|^k chain and
 * the R sums from them, the retraction by homogeneity).  The solve takes this
 * line search by itself where the resident state applies but does not fit
 * (mlq_rcg_path 2); the flag asks for it at any size (q >= 4 without the e_u
 * variant, dealiased; MLQ_RCG_REFERENCE takes precedence).  No counterpart
 * in the reference. */
#define MLQ_RCG_LEAN 8

typedef struct mlq_rcg_result {
  double J1;             /* final R(u) */
  int iters;             /* final iteration counter */
  int successful;        /* optimizationSuccessful */
  int while_flag;
  int stop_reason;       /* 0 converged, 1 max iter, 2 mnbrak flag 1, 3 mnbrak flag 2,
                            4 tau > TAU_MAX, 5 cost not increasing */
  int n_J;               /* J evaluations requested by mnbrak+brent (all iterations) */
  int n_J_computed;      /* J evaluations computed (>= n_J with speculation) */
  double seconds;
  int n_J_poly;          /* of n_J_computed, those given by the exact line-search
                            polynomials (q = 4 and q = 6, no clipping of D(phi.phi)) */
  int pad_;
  double enstrophy;      /* calc_global_summed_enstrophy(Uvec) of the result and */
  double dEnstrophydt;   /* calc_global_dEnstrophydt(Uvec), for
                            save_to_optimizationResultList (optimization.f90:491-493) */
} mlq_rcg_result;

/* log record kinds delivered to the callback (data_ops.f90 writers) */
#define MLQ_LOG_ITER 1     /* save_diagnostics_optim: i0=iter, d[0..2]=tau,beta,J (:764-799);
                              d[3..4] = tau_brack (the t_bra(a), t_bra(b) of the iteration's
                              terminal line, optimization.f90:420), d[5] = J requests of its
                              mnbrak and brent */
#define MLQ_LOG_LINEMIN 2  /* save_linemin_data: i0=brakIter, i1=0 replace/1 append,
                              d[0..5]=tA,tB,tC,FA,FB,FC (:1403-1442) */
#define MLQ_LOG_BRENT 3    /* brent-info: i0=j, d[0..1]=X,FX (optimization.f90:836-946) */
#define MLQ_LOG_MSG 4      /* optim_msg_handle: i0=message code (data_ops.f90:1571-1632) */
#define MLQ_LOG_ERR 5      /* optim_error_handle: i0=error code (:1527-1566) */
#define MLQ_LOG_SPECTRUM 6 /* calculateSaveSpectrum (function_ops.f90:1011-1023): i0=0 uvec,
                              1 d; d[0..2n1-1] = spectral_data(1:n1,1:2) as
                              mlq_calculate_spectral_data returns it (save_spectral_data,
                              data_ops.f90:227-261, writes it) */
#define MLQ_LOG_FIELD 7    /* save_field(Uvec, "u_result") every save_every iterations
                              (optimization.f90:388-390): i0 = iter, d = the iterate, 3 n^3
                              doubles in the reference layout (valid during the call) */
#define MLQ_LOG_KAPPA 8    /* a kappa test (optimization.f90:957-1129, written by
                              save_kappa_test, data_ops.f90:1448-1491): i0 = 0
                              "initial_grad_L2", 1 "gradL2_iter<iter>", 2 "end_gradJ",
                              3 "end_d1"; i1 = useAdjustedKappaTest; d[0..79] the table
                              mlq_kappa_test returns, d[80] frac, d[81] best kappa,
                              d[82] best row */
typedef void (*mlq_log_cb)(void* user, int kind, int iter, int i0, int i1, const double* d);

int mlq_rcg_solve(mlq_ctx* ctx, mlq_field u, const mlq_rcg_cfg* cfg, mlq_rcg_result* res,
                  mlq_log_cb cb, void* user);

/* ---- field files and result-file names (host only, no context) ---- */
/* save_field / save_field_R3toR3_ncdf (data_ops.f90:9-41, 390-447): the vector
 * field (3 component blocks of n1*n2*n3 doubles, x fastest) as a NetCDF classic
 * file with dimensions x, y, z and double variables Ux, Uy, Uz; CDF-1 while
 * its offsets fit 32 bits, else the 64-bit-offset format, with z the record
 * dimension when one variable exceeds 4 GiB (1024^3) */
int mlq_save_field_nc(const char* path, int n1, int n2, int n3, const double* field);
/* read_field_R3toR3_ncdf (data_ops.f90:1198-1397) for the names Ux, Uy, Uz:
 * either format and layout; MLQ_EINVAL when the file, its variables or its
 * dimensions do not match the grid */
int mlq_read_field_nc(const char* path, int n1, int n2, int n3, double* field);
/* the reference's field-file name (data_ops.f90:31):
 * dir + field_name + "_q" + lebesgueQTxt + "_n" + n + "_B" + i3.3(b_iter) +
 * "_iter" + (i5.5(iter), or "end" for iter < 0) + ".nc"; lebesgueQTxt as
 * setLebesgueQtext writes it (initialize.f90:49-70: i1 or i2 for integer q
 * below 100, else ES9.2) */
int mlq_result_filename(char* buf, int buflen, const char* dir, const char* field_name, double q,
                        int n, int b_iter, int iter);
int mlq_q_text(double q, char* buf, int buflen);
/* set_q_resol_bIterOffset_optimIterOffsets (function_ops.f90:4635-4674, with
 * getQfromFileName / getBiterfromFileName :4599-4632): q after the first
 * "_q", resol after the first "_n", the B index after the first "_B", and the
 * iteration after the first "_iter"; a name ending in "iterend" resumes after
 * that B solve (b_iter_offset = B index, optim_iter_offset = 0), any other
 * inside it (b_iter_offset = B index - 1, optim_iter_offset = the iteration) */
int mlq_parse_result_filename(const char* name, double* q, int* resol, int* b_iter_offset,
                              int* optim_iter_offset);

/* ---- measurement helpers ---- */
/* time `reps` launches of the 3-D real transform pair of a batch of `batch`
 * fields on the context stream with HIP events; per-pass average kernel
 * durations (ms) are returned in t_ms[0..5] = {x fwd, y fwd, z fwd, z inv, y inv, x inv}
 * and the batch timed in t_ms[6] (`batch`, at most the transform scratch:
 * 12 spectra, fewer where they would take a large share of the device) */
int mlq_time_fft(mlq_ctx* ctx, int batch, int reps, double* t_ms);
/* time `reps` launches of a plain device copy of `nfields` real fields (the
 * calibration kernel for HBM byte counters); average ms per launch in *t_ms */
int mlq_time_copy(mlq_ctx* ctx, int nfields, int reps, double* t_ms);
/* time `reps` launches of a copy of `nfields` half spectra with the memory
 * access pattern of a strided transform pass and none of its arithmetic:
 * axis 1 the y pattern (columns inside one z plane), 2 the z pattern (columns
 * across the planes); `tile` = tile shape + 4*order as in mlq_autotune.  The
 * rate is the ceiling that access pattern sets for the pass.  *t_ms: ms per
 * launch; *bytes: algorithmic bytes per launch (read + write, padding
 * excluded).  No counterpart in the reference (measurement support). */
int mlq_time_column_copy(mlq_ctx* ctx, int axis, int tile, int nfields, int reps, double* t_ms,
                         double* bytes);
/* tile autotuner of the strided transform passes: times each of 8 tile
 * candidates (4 shapes of elements per thread x columns per workgroup, each
 * in dispatch order and in XCD-contiguous order: candidate = shape + 4*order)
 * of the {y fwd, z fwd, z inv, y inv} passes on `batch` fields, for full
 * transforms (passes 0-3) and for band-limited ones at the 2/3 dealias cube
 * (passes 4-7), and keeps the fastest of each.  times[pass*8+cand] (64 values,
 * ms per launch) and choice[pass] (8 values) may be null.  mlq_init runs it
 * for n >= 128 unless the environment sets MLQ_AUTOTUNE=0. */
int mlq_autotune(mlq_ctx* ctx, int batch, double* times, int* choice);
/* set the tile of each of the 8 strided passes (choice[pass] = shape +
 * 4*order, as mlq_autotune reports it), e.g. to repeat a measurement with the
 * tiles of an earlier run.  No counterpart in the reference. */
int mlq_set_tiles(mlq_ctx* ctx, const int* choice);
/* bytes of one real field and of one half spectrum in the active precision */
int mlq_sizes(const mlq_ctx* ctx, double* real_bytes, double* spec_bytes);
/* per-kernel-class timing accumulated while profiling is on (HIP events on the
 * context stream); class ids are listed by mlq_prof_name */
int mlq_prof_enable(mlq_ctx* ctx, int on);
int mlq_prof_get(mlq_ctx* ctx, int cls, long long* count, double* total_ms, double* bytes);
const char* mlq_prof_name(int cls);
/* device bytes in use by the context (its allocations minus the free space of
 * the slabs its temporaries are carved from) and their high-water mark since
 * mlq_init: the memory the work needed, where mlq_mem_info's peak also counts
 * the slabs' free space.  No counterpart in the reference. */
int mlq_mem_live(mlq_ctx* ctx, double* live_bytes, double* peak_live_bytes);
int mlq_mem_info(mlq_ctx* ctx, double* used_bytes, double* peak_bytes, double* device_free,
                 double* device_total);
/* fields transformed since the context was created (3-D real transforms,
 * forward and inverse; a vector field counts 3).  No counterpart in the
 * reference: it reports the transform count per iteration that the
 * reference's C2C count (548 + 94 n_J per iteration at q = 4) compares to. */
int mlq_transform_counts(const mlq_ctx* ctx, long long* fwd_fields, long long* inv_fields);

#ifdef __cplusplus
}
#endif
#endif /* MAXLQ_H */
