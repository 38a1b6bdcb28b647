! Drop-in replacement of the reference's module fftwfunction (fftwfunction.f90)
! backed by libmaxlq.so on one MI355X.  Same module name, same public
! procedures and dummies; the device context lives in this module.
!
! Needs the reference's module global_variables (grid size n, local_N, the
! physical and numerical parameters) and the interface module maxlq_c.
module fftwfunction
   use, intrinsic :: iso_c_binding
   use global_variables
   use maxlq_c
   implicit none

   type(c_ptr), save :: mlq_ctx_ptr = c_null_ptr
   ! precision of the device context: MLQ_FP64 (the parity path) or MLQ_MIXED
   ! (FP32 fields and transforms, FP64 reductions and scalars); the caller may
   ! set it before init_fft
   integer(c_int), save :: mlq_precision = MLQ_FP64

contains

   ! map a status to the reference's behaviour: NaN -> stop 1
   subroutine mlq_check(rc, where)
      integer(c_int), intent(in) :: rc
      character(len=*), intent(in) :: where
      if (rc == MLQ_OK) return
      if (rc == MLQ_ENAN) then
         print*, "nan in ", where
         stop 1
      end if
      print*, "maxlq error ", rc, " in ", where
      stop 2
   end subroutine mlq_check

   ! mirror the module globals the hot path reads into the device context
   subroutine mlq_sync_params()
      type(mlq_params) :: p
      if (.not. c_associated(mlq_ctx_ptr)) return
      call fill_params(p)
      call mlq_check(mlq_set_params(mlq_ctx_ptr, p), "set_params")
   end subroutine mlq_sync_params

   subroutine fill_params(p)
      type(mlq_params), intent(out) :: p
      call mlq_default_params(p, real(lebesgueQ, c_double))
      p%B = constraintB
      p%visc = visc
      p%lambda1 = lambda1
      p%viscCoefficient = viscCoefficient
      p%pressureCoefficient = pressureCoefficient
      p%toDealias = merge(1_c_int, 0_c_int, toDealias)
      p%use_e_u = merge(1_c_int, 0_c_int, use_e_u_instead_of_uqMinus4)
      p%dealiase_if_mult_by_e_u = merge(1_c_int, 0_c_int, dealiase_if_mult_by_e_u)
      p%precision = mlq_precision
   end subroutine fill_params

   subroutine init_fft()                               ! fftwfunction.f90:169-185
      type(mlq_params) :: p
      integer(c_int) :: ln, off
      call fill_params(p)
      call mlq_check(mlq_init(mlq_ctx_ptr, n(1), n(2), n(3), p), "init_fft")
      call mlq_check(mlq_local_size_3d(mlq_ctx_ptr, ln, off), "init_fft")
      local_N = ln
      local_k_offset = off
   end subroutine init_fft

   subroutine fft_deallocate()                         ! fftwfunction.f90:191-212
      if (c_associated(mlq_ctx_ptr)) call mlq_check(mlq_finalize(mlq_ctx_ptr), "fft_deallocate")
      mlq_ctx_ptr = c_null_ptr
   end subroutine fft_deallocate

   subroutine fftfwdV(vec, fvec)                       ! fftwfunction.f90:24-40
      complex(pr), dimension(1:n(1),1:n(2),1:local_N,1:3), intent(in) :: vec
      complex(pr), dimension(1:n(1),1:n(2),1:local_N,1:3), intent(out) :: fvec
      call mlq_check(mlq_fftfwdV(mlq_ctx_ptr, vec, fvec), "fftfwdV")
   end subroutine fftfwdV

   subroutine fftbwdV(fvec, vec)                       ! fftwfunction.f90:45-57
      complex(pr), dimension(1:n(1),1:n(2),1:local_N,1:3), intent(in) :: fvec
      complex(pr), dimension(1:n(1),1:n(2),1:local_N,1:3), intent(out) :: vec
      call mlq_check(mlq_fftbwdV(mlq_ctx_ptr, fvec, vec), "fftbwdV")
   end subroutine fftbwdV

   subroutine fftfwd(u, fu)                            ! fftwfunction.f90:62-92
      complex(pr), dimension(1:n(1),1:n(2),1:local_N), intent(in) :: u
      complex(pr), dimension(1:n(1),1:n(2),1:local_N), intent(out) :: fu
      call mlq_check(mlq_fftfwd(mlq_ctx_ptr, u, fu), "fftfwd")
   end subroutine fftfwd

   subroutine fftbwd(fu, u)                            ! fftwfunction.f90:97-113
      complex(pr), dimension(1:n(1),1:n(2),1:local_N), intent(in) :: fu
      complex(pr), dimension(1:n(1),1:n(2),1:local_N), intent(out) :: u
      call mlq_check(mlq_fftbwd(mlq_ctx_ptr, fu, u), "fftbwd")
   end subroutine fftbwd

   subroutine fftfwdr(u, fu)                           ! fftwfunction.f90:118-134
      real(pr), dimension(1:n(1),1:n(2),1:local_N), intent(in) :: u
      complex(pr), dimension(1:n(1)/2+1,1:n(2),1:local_N), intent(out) :: fu
      call mlq_check(mlq_fftfwdr(mlq_ctx_ptr, u, fu), "fftfwdr")
   end subroutine fftfwdr

   subroutine fftbwdr(fu, u)                           ! fftwfunction.f90:139-165
      complex(pr), dimension(1:n(1)/2+1,1:n(2),1:local_N), intent(in) :: fu
      real(pr), dimension(1:n(1),1:n(2),1:local_N), intent(out) :: u
      call mlq_check(mlq_fftbwdr(mlq_ctx_ptr, fu, u), "fftbwdr")
   end subroutine fftbwdr

   ! single-rank forms of the FFTW-MPI calls the driver makes
   ! (maxdLpdt.f90:36, 518): nothing to initialise or clean up
   subroutine fftw_mpi_init()
   end subroutine fftw_mpi_init

   subroutine fftw_mpi_cleanup()
   end subroutine fftw_mpi_cleanup

end module fftwfunction
