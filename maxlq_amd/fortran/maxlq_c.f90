! iso_c_binding interface to libmaxlq.so (include/maxlq.h).
!
! Host arrays are passed by reference in the reference layout: real(c_double)
! fields dimensioned (1:n1,1:n2,1:n3[,1:3]) and complex(c_double_complex)
! spectra, exactly as the reference's fftwfunction and function_ops routines
! receive them.  Every function returns an MLQ_* status (0 = MLQ_OK).
module maxlq_c
   use, intrinsic :: iso_c_binding
   implicit none

   integer(c_int), parameter :: MLQ_OK = 0, MLQ_ENAN = 1, MLQ_EBADKIND = 2, MLQ_EHIP = 3, &
                                MLQ_EOOM = 4, MLQ_EINVAL = 5, MLQ_EUNSUP = 6
   integer(c_int), parameter :: MLQ_FP64 = 0, MLQ_MIXED = 1
   integer(c_int), parameter :: MLQ_IP_L2 = 0, MLQ_IP_HL = 1, MLQ_IP_H = 2, MLQ_IP_OLD_H = 3, &
                                MLQ_IP_OLD_HL = 4
   integer(c_int), parameter :: MLQ_J_MAXDLQDT = 0, MLQ_J_LINEMIN = 1, MLQ_J_LQNORM = 2
   integer(c_int), parameter :: MLQ_LOG_ITER = 1, MLQ_LOG_LINEMIN = 2, MLQ_LOG_BRENT = 3, &
                                MLQ_LOG_MSG = 4, MLQ_LOG_ERR = 5, MLQ_LOG_SPECTRUM = 6, &
                                MLQ_LOG_FIELD = 7, MLQ_LOG_KAPPA = 8
   integer(c_int), parameter :: MLQ_RCG_REFERENCE = 1, MLQ_RCG_SPECTRA_RAW = 2, MLQ_RCG_KAPPA = 4, &
                                 MLQ_RCG_LEAN = 8
   integer(c_int), parameter :: MLQ_PERT_SINE = 0, MLQ_PERT_DIVFREE_SINE = 1, MLQ_PERT_LOAD_RAW = 2, &
                                MLQ_PERT_LOAD = 3, MLQ_PERT_LOAD_SMOOTH = 4, &
                                MLQ_PERT_LOAD_SMOOTH_RAW = 5, MLQ_PERT_LOAD_K = 6, &
                                MLQ_PERT_LOAD_POLY = 7, MLQ_PERT_LOAD_EXP = 8, &
                                MLQ_PERT_RANDOM_POLY = 9, MLQ_PERT_RANDOM = 10

   type, bind(C) :: mlq_params
      real(c_double) :: q, B, visc, lambda1, viscCoefficient, pressureCoefficient
      integer(c_int) :: toDealias, use_e_u, dealiase_if_mult_by_e_u, precision
   end type mlq_params

   type, bind(C) :: mlq_rcg_cfg
      integer(c_int) :: max_iter
      real(c_double) :: optim_tol
      integer(c_int) :: start_iter, batch_linesearch, reset_every, rolling, flags, spectra_every
      integer(c_int) :: save_every, kappa_every
      integer(c_long_long) :: kappa_seed
      type(c_ptr) :: kappa_field
   end type mlq_rcg_cfg

   type, bind(C) :: mlq_rcg_result
      real(c_double) :: J1
      integer(c_int) :: iters, successful, while_flag, stop_reason, n_J, n_J_computed
      real(c_double) :: seconds
      integer(c_int) :: n_J_poly, pad_
      real(c_double) :: enstrophy, dEnstrophydt
   end type mlq_rcg_result

   interface
      subroutine mlq_default_params(p, q) bind(C, name="mlq_default_params")
         import :: mlq_params, c_double
         type(mlq_params), intent(out) :: p
         real(c_double), value :: q
      end subroutine
      integer(c_int) function mlq_init(ctx, n1, n2, n3, p) bind(C, name="mlq_init")
         import :: c_ptr, c_int, mlq_params
         type(c_ptr), intent(out) :: ctx
         integer(c_int), value :: n1, n2, n3
         type(mlq_params), intent(in) :: p
      end function
      integer(c_int) function mlq_finalize(ctx) bind(C, name="mlq_finalize")
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function mlq_set_params(ctx, p) bind(C, name="mlq_set_params")
         import :: c_ptr, c_int, mlq_params
         type(c_ptr), value :: ctx
         type(mlq_params), intent(in) :: p
      end function
      integer(c_int) function mlq_local_size_3d(ctx, local_n, local_k_offset) &
            bind(C, name="mlq_local_size_3d")
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx
         integer(c_int), intent(out) :: local_n, local_k_offset
      end function
      ! fftwfunction
      integer(c_int) function mlq_fftfwd(ctx, u, fu) bind(C, name="mlq_fftfwd")
         import :: c_ptr, c_int, c_double_complex
         type(c_ptr), value :: ctx
         complex(c_double_complex), intent(in) :: u(*)
         complex(c_double_complex), intent(out) :: fu(*)
      end function
      integer(c_int) function mlq_fftbwd(ctx, fu, u) bind(C, name="mlq_fftbwd")
         import :: c_ptr, c_int, c_double_complex
         type(c_ptr), value :: ctx
         complex(c_double_complex), intent(in) :: fu(*)
         complex(c_double_complex), intent(out) :: u(*)
      end function
      integer(c_int) function mlq_fftfwdV(ctx, v, fv) bind(C, name="mlq_fftfwdV")
         import :: c_ptr, c_int, c_double_complex
         type(c_ptr), value :: ctx
         complex(c_double_complex), intent(in) :: v(*)
         complex(c_double_complex), intent(out) :: fv(*)
      end function
      integer(c_int) function mlq_fftbwdV(ctx, fv, v) bind(C, name="mlq_fftbwdV")
         import :: c_ptr, c_int, c_double_complex
         type(c_ptr), value :: ctx
         complex(c_double_complex), intent(in) :: fv(*)
         complex(c_double_complex), intent(out) :: v(*)
      end function
      integer(c_int) function mlq_fftfwdr(ctx, u, fu) bind(C, name="mlq_fftfwdr")
         import :: c_ptr, c_int, c_double, c_double_complex
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: u(*)
         complex(c_double_complex), intent(out) :: fu(*)
      end function
      integer(c_int) function mlq_fftbwdr(ctx, fu, u) bind(C, name="mlq_fftbwdr")
         import :: c_ptr, c_int, c_double, c_double_complex
         type(c_ptr), value :: ctx
         complex(c_double_complex), intent(in) :: fu(*)
         real(c_double), intent(out) :: u(*)
      end function
      ! function_ops field routines
      integer(c_int) function mlq_gradient(ctx, u, g) bind(C, name="mlq_gradient")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: u(*)
         real(c_double), intent(out) :: g(*)
      end function
      integer(c_int) function mlq_divergence(ctx, u, d) bind(C, name="mlq_divergence")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: u(*)
         real(c_double), intent(out) :: d(*)
      end function
      integer(c_int) function mlq_laplacian(ctx, f) bind(C, name="mlq_laplacian")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(inout) :: f(*)
      end function
      integer(c_int) function mlq_solve_poisson(ctx, f, l, phi) bind(C, name="mlq_solve_poisson")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: f(*)
         real(c_double), value :: l
         real(c_double), intent(out) :: phi(*)
      end function
      integer(c_int) function mlq_dealias_scalar(ctx, g, order) bind(C, name="mlq_dealias_scalar")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(inout) :: g(*)
         real(c_double), value :: order
      end function
      integer(c_int) function mlq_divAvg_free(ctx, f) bind(C, name="mlq_divAvg_free")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(inout) :: f(*)
      end function
      integer(c_int) function mlq_HilbertGradient(ctx, g, order) bind(C, name="mlq_HilbertGradient")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(inout) :: g(*)
         real(c_double), value :: order
      end function
      integer(c_int) function mlq_calc0thFourierModes(ctx, v, out) &
            bind(C, name="mlq_calc0thFourierModes")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: v(*)
         real(c_double), intent(out) :: out(3)
      end function
      integer(c_int) function mlq_calc_uk(ctx, u, k, uk) bind(C, name="mlq_calc_uk")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: u(*)
         real(c_double), value :: k
         real(c_double), intent(out) :: uk(*)
      end function
      integer(c_int) function mlq_calc_global_Lq_norm(ctx, u, nrm) &
            bind(C, name="mlq_calc_global_Lq_norm")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: u(*)
         real(c_double), intent(out) :: nrm
      end function
      integer(c_int) function mlq_rescaleLqNorm(ctx, u, B) bind(C, name="mlq_rescaleLqNorm")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(inout) :: u(*)
         real(c_double), value :: B
      end function
      integer(c_int) function mlq_calcConstraintDerivativeL2(ctx, u, o) &
            bind(C, name="mlq_calcConstraintDerivativeL2")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: u(*)
         real(c_double), intent(out) :: o(*)
      end function
      integer(c_int) function mlq_vectorTransport(ctx, u, q, B, eta, xi, o) &
            bind(C, name="mlq_vectorTransport")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: u(*), eta(*), xi(*)
         real(c_double), value :: q, B
         real(c_double), intent(out) :: o(*)
      end function
      integer(c_int) function mlq_global_inner_product(ctx, f, g, kind, o) &
            bind(C, name="mlq_global_inner_product")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: f(*), g(*)
         integer(c_int), value :: kind
         real(c_double), intent(out) :: o
      end function
      integer(c_int) function mlq_global_summed_field_inner_product(ctx, f, g, kind, o) &
            bind(C, name="mlq_global_summed_field_inner_product")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: f(*), g(*)
         integer(c_int), value :: kind
         real(c_double), intent(out) :: o
      end function
      integer(c_int) function mlq_calc_global_dLqdt_inclParts(ctx, u, q, o) &
            bind(C, name="mlq_calc_global_dLqdt_inclParts")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: u(*)
         real(c_double), value :: q
         real(c_double), intent(out) :: o(3)
      end function
      integer(c_int) function mlq_GradL2ForLq(ctx, U, V) bind(C, name="mlq_GradL2ForLq")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: U(*)
         real(c_double), intent(out) :: V(*)
      end function
      integer(c_int) function mlq_eval_J(ctx, f, sys, J) bind(C, name="mlq_eval_J")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: f(*)
         integer(c_int), value :: sys
         real(c_double), intent(out) :: J
      end function
      ! device-resident fields and the whole solve
      integer(c_int) function mlq_field_alloc(ctx, ncomp, h) bind(C, name="mlq_field_alloc")
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx
         integer(c_int), value :: ncomp
         integer(c_int), intent(out) :: h
      end function
      integer(c_int) function mlq_field_free(ctx, h) bind(C, name="mlq_field_free")
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx
         integer(c_int), value :: h
      end function
      integer(c_int) function mlq_upload(ctx, h, a) bind(C, name="mlq_upload")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: h
         real(c_double), intent(in) :: a(*)
      end function
      integer(c_int) function mlq_download(ctx, h, a) bind(C, name="mlq_download")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: h
         real(c_double), intent(out) :: a(*)
      end function
      integer(c_int) function mlq_rcg_solve(ctx, h, cfg, res, cb, user) bind(C, name="mlq_rcg_solve")
         import :: c_ptr, c_int, c_funptr, mlq_rcg_cfg, mlq_rcg_result
         type(c_ptr), value :: ctx
         integer(c_int), value :: h
         type(mlq_rcg_cfg), intent(in) :: cfg
         type(mlq_rcg_result), intent(out) :: res
         type(c_funptr), value :: cb
         type(c_ptr), value :: user
      end function
      type(c_ptr) function mlq_last_error(ctx) bind(C, name="mlq_last_error")
         import :: c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function mlq_sync(ctx) bind(C, name="mlq_sync")
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx
      end function
      integer(c_int) function mlq_calc_gk_order2(ctx, g, k, gk) bind(C, name="mlq_calc_gk_order2")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: g(*)
         real(c_double), value :: k
         real(c_double), intent(out) :: gk(*)
      end function
      integer(c_int) function mlq_inner_product(ctx, f, g, kind, res) bind(C, name="mlq_inner_product")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: f(*), g(*)
         integer(c_int), value :: kind
         real(c_double), intent(out) :: res
      end function
      integer(c_int) function mlq_calc_local_dLqdt_inclParts(ctx, u, q, res) &
            bind(C, name="mlq_calc_local_dLqdt_inclParts")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: u(*)
         real(c_double), value :: q
         real(c_double), intent(out) :: res(3)
      end function
      integer(c_int) function mlq_seeded_field(ctx, h, seed, kind, a) bind(C, name="mlq_seeded_field")
         import :: c_ptr, c_int, c_double, c_long_long
         type(c_ptr), value :: ctx
         integer(c_int), value :: h
         integer(c_long_long), value :: seed
         integer(c_int), value :: kind
         real(c_double), value :: a
      end function
      integer(c_int) function mlq_eval_J_batch(ctx, u, d, taus, ntau, J) bind(C, name="mlq_eval_J_batch")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: u, d, ntau
         real(c_double), intent(in) :: taus(*)
         real(c_double), intent(out) :: J(*)
      end function
      integer(c_int) function mlq_time_linesearch(ctx, u, d, taus, ntau, batch, J, t_ms) &
            bind(C, name="mlq_time_linesearch")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: u, d, ntau, batch
         real(c_double), intent(in) :: taus(*)
         real(c_double), intent(out) :: J(*), t_ms(2)
      end function
      integer(c_int) function mlq_kappa_test(ctx, phi, grad, pert, mysystem, inner_kind, table, &
            frac, best_kappa, best_row) bind(C, name="mlq_kappa_test")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: phi, grad, pert, mysystem, inner_kind
         real(c_double), intent(out) :: table(5, 16), frac, best_kappa
         integer(c_int), intent(out) :: best_row
      end function
      ! diagnosticScalars (function_ops.f90:2503-2602): out(21) = K(3), E(3),
      ! Umax(3), Wmax(3), vorCore(3), magUmax, magWmax, H, maxHel, minHel, divU_L2
      integer(c_int) function mlq_diagnosticScalars(ctx, u, out, fields, ring_xyz, ring_cap, &
            ring_points) bind(C, name="mlq_diagnosticScalars")
         import :: c_ptr, c_int, c_double, c_long_long
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: u(*)
         real(c_double), intent(out) :: out(21)
         type(c_ptr), value :: fields, ring_xyz
         integer(c_long_long), value :: ring_cap
         integer(c_long_long), intent(out) :: ring_points
      end function
      ! kappa_test_pert (function_ops.f90:613-977); loaded may be c_null_ptr
      integer(c_int) function mlq_kappa_test_pert(ctx, out, kind, m1, m2, m3, seed, loaded) &
            bind(C, name="mlq_kappa_test_pert")
         import :: c_ptr, c_int, c_double, c_long_long
         type(c_ptr), value :: ctx
         integer(c_int), value :: out, kind
         real(c_double), value :: m1, m2, m3
         integer(c_long_long), value :: seed
         type(c_ptr), value :: loaded
      end function
      integer(c_int) function mlq_rcg_path(ctx, resident_state) bind(C, name="mlq_rcg_path")
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx
         integer(c_int), intent(out) :: resident_state
      end function
      integer(c_int) function mlq_time_fft(ctx, batch, reps, t_ms) bind(C, name="mlq_time_fft")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: batch, reps
         real(c_double), intent(out) :: t_ms(7)
      end function
      integer(c_int) function mlq_calculate_spectral_data(ctx, u, normalize, spectral_data) &
            bind(C, name="mlq_calculate_spectral_data")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: u(*)
         integer(c_int), value :: normalize
         real(c_double), intent(out) :: spectral_data(*)
      end function
      integer(c_int) function mlq_spectral_data_field(ctx, u, normalize, spectral_data) &
            bind(C, name="mlq_spectral_data_field")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: u, normalize
         real(c_double), intent(out) :: spectral_data(*)
      end function
      integer(c_int) function mlq_calc_global_summed_enstrophy(ctx, u, enstrophy) &
            bind(C, name="mlq_calc_global_summed_enstrophy")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: u(*)
         real(c_double), intent(out) :: enstrophy
      end function
      integer(c_int) function mlq_calc_global_dEnstrophydt(ctx, u, dEdt) &
            bind(C, name="mlq_calc_global_dEnstrophydt")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: u(*)
         real(c_double), intent(out) :: dEdt
      end function
      integer(c_int) function mlq_time_copy(ctx, nfields, reps, t_ms) bind(C, name="mlq_time_copy")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: nfields, reps
         real(c_double), intent(out) :: t_ms
      end function
      integer(c_int) function mlq_time_column_copy(ctx, axis, tile, nfields, reps, t_ms, bytes) &
            bind(C, name="mlq_time_column_copy")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: axis, tile, nfields, reps
         real(c_double), intent(out) :: t_ms, bytes
      end function
      integer(c_int) function mlq_autotune(ctx, batch, times, choice) bind(C, name="mlq_autotune")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         integer(c_int), value :: batch
         real(c_double), intent(out) :: times(64)
         integer(c_int), intent(out) :: choice(8)
      end function
      integer(c_int) function mlq_sizes(ctx, real_bytes, spec_bytes) bind(C, name="mlq_sizes")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: real_bytes, spec_bytes
      end function
      integer(c_int) function mlq_prof_enable(ctx, on) bind(C, name="mlq_prof_enable")
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx
         integer(c_int), value :: on
      end function
      integer(c_int) function mlq_prof_get(ctx, cls, count, total_ms, bytes) bind(C, name="mlq_prof_get")
         import :: c_ptr, c_int, c_double, c_long_long
         type(c_ptr), value :: ctx
         integer(c_int), value :: cls
         integer(c_long_long), intent(out) :: count
         real(c_double), intent(out) :: total_ms, bytes
      end function
      type(c_ptr) function mlq_prof_name(cls) bind(C, name="mlq_prof_name")
         import :: c_ptr, c_int
         integer(c_int), value :: cls
      end function
      integer(c_int) function mlq_mem_info(ctx, used_bytes, peak_bytes, device_free, device_total) &
            bind(C, name="mlq_mem_info")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: used_bytes, peak_bytes, device_free, device_total
      end function
      ! field files and result-file names (host only): strings are
      ! NUL-terminated character(kind=c_char) arrays
      integer(c_int) function mlq_save_field_nc(path, n1, n2, n3, field) &
            bind(C, name="mlq_save_field_nc")
         import :: c_int, c_double, c_char
         character(kind=c_char), intent(in) :: path(*)
         integer(c_int), value :: n1, n2, n3
         real(c_double), intent(in) :: field(*)
      end function
      integer(c_int) function mlq_read_field_nc(path, n1, n2, n3, field) &
            bind(C, name="mlq_read_field_nc")
         import :: c_int, c_double, c_char
         character(kind=c_char), intent(in) :: path(*)
         integer(c_int), value :: n1, n2, n3
         real(c_double), intent(out) :: field(*)
      end function
      integer(c_int) function mlq_result_filename(buf, buflen, dir, field_name, q, n, b_iter, iter) &
            bind(C, name="mlq_result_filename")
         import :: c_int, c_double, c_char
         character(kind=c_char), intent(out) :: buf(*)
         integer(c_int), value :: buflen, n, b_iter, iter
         character(kind=c_char), intent(in) :: dir(*), field_name(*)
         real(c_double), value :: q
      end function
      integer(c_int) function mlq_q_text(q, buf, buflen) bind(C, name="mlq_q_text")
         import :: c_int, c_double, c_char
         real(c_double), value :: q
         character(kind=c_char), intent(out) :: buf(*)
         integer(c_int), value :: buflen
      end function
      integer(c_int) function mlq_parse_result_filename(name, q, resol, b_iter_offset, &
            optim_iter_offset) bind(C, name="mlq_parse_result_filename")
         import :: c_int, c_double, c_char
         character(kind=c_char), intent(in) :: name(*)
         real(c_double), intent(out) :: q
         integer(c_int), intent(out) :: resol, b_iter_offset, optim_iter_offset
      end function
      integer(c_int) function mlq_transform_counts(ctx, fwd_fields, inv_fields) &
            bind(C, name="mlq_transform_counts")
         import :: c_ptr, c_int, c_long_long
         type(c_ptr), value :: ctx
         integer(c_long_long), intent(out) :: fwd_fields, inv_fields
      end function
   end interface

contains

   ! a Fortran string as a NUL-terminated C string
   function cstr(s) result(c)
      character(len=*), intent(in) :: s
      character(kind=c_char) :: c(len_trim(s) + 1)
      integer :: i
      do i = 1, len_trim(s)
         c(i) = s(i:i)
      end do
      c(len_trim(s) + 1) = c_null_char
   end function cstr

   ! a NUL-terminated C string as a Fortran string
   function fstr(c) result(s)
      character(kind=c_char), intent(in) :: c(:)
      character(len=:), allocatable :: s
      integer :: i, n
      n = 0
      do i = 1, size(c)
         if (c(i) == c_null_char) exit
         n = i
      end do
      allocate(character(len=n) :: s)
      do i = 1, n
         s(i:i) = c(i)
      end do
   end function fstr

end module maxlq_c
