! The function_ops routines of the RCG hot path on the GPU, with the
! reference's dummies (function_ops.f90); function_ops.f90 forwards to these
! from its own routines of the same names.  Every call mirrors the current module globals
! into the device context first, so q- and B-continuation keep working.
module mlq_field_ops
   use, intrinsic :: iso_c_binding
   use global_variables
   use maxlq_c
   use fftwfunction, only: mlq_ctx_ptr, mlq_check, mlq_sync_params
   implicit none

contains

   integer(c_int) function ip_kind(mytype)
      character(len=*), intent(in) :: mytype
      select case (mytype)
      case ("L2")
         ip_kind = MLQ_IP_L2
      case ("H_l^(3/2-1/q)")
         ip_kind = MLQ_IP_HL
      case ("H^(3/2-1/q)")
         ip_kind = MLQ_IP_H
      case ("H^((3q-1)/(2q))")
         ip_kind = MLQ_IP_OLD_H
      case ("H_l^((3q-1)/(2q))")
         ip_kind = MLQ_IP_OLD_HL
      case default
         ip_kind = -1_c_int
      end select
   end function ip_kind

   subroutine gpu_GradL2ForLq(U, V)                            ! :1066-1211
      real(pr), dimension(1:n(1),1:n(2),1:local_N,1:3), intent(in) :: U
      real(pr), dimension(1:n(1),1:n(2),1:local_N,1:3), intent(out) :: V
      call mlq_sync_params()
      call mlq_check(mlq_GradL2ForLq(mlq_ctx_ptr, U, V), "GradL2ForLq")
   end subroutine gpu_GradL2ForLq

   subroutine gpu_HilbertGradient(grad, order)                 ! :3583-3627
      real(pr), dimension(1:n(1),1:n(2),1:local_N,1:3), intent(inout) :: grad
      real(pr), intent(in) :: order
      call mlq_sync_params()
      call mlq_check(mlq_HilbertGradient(mlq_ctx_ptr, grad, order), "HilbertGradient")
   end subroutine gpu_HilbertGradient

   subroutine gpu_divAvg_free(myfield)                         ! :2843-2918
      real(pr), dimension(1:n(1),1:n(2),1:local_N,1:3), intent(inout) :: myfield
      call mlq_check(mlq_divAvg_free(mlq_ctx_ptr, myfield), "divAvg_free")
   end subroutine gpu_divAvg_free

   subroutine gpu_calcConstraintDerivativeL2(u, resultVec)     ! :984-1006
      real(pr), dimension(1:n(1),1:n(2),1:local_N,1:3), intent(in) :: u
      real(pr), dimension(1:n(1),1:n(2),1:local_N,1:3), intent(out) :: resultVec
      call mlq_sync_params()
      call mlq_check(mlq_calcConstraintDerivativeL2(mlq_ctx_ptr, u, resultVec), &
                     "calcConstraintDerivativeL2")
   end subroutine gpu_calcConstraintDerivativeL2

   function gpu_calc_global_Lq_norm(u) result(norm)            ! :1675-1692
      real(pr), dimension(1:n(1),1:n(2),1:local_N,1:3), intent(in) :: u
      real(pr) :: norm
      call mlq_sync_params()
      call mlq_check(mlq_calc_global_Lq_norm(mlq_ctx_ptr, u, norm), "calc_global_Lq_norm")
   end function gpu_calc_global_Lq_norm

   subroutine gpu_calculate_spectral_data(u, spectral_data)    ! :2614-2677
      real(pr), dimension(1:n(1),1:n(2),1:local_N,1:3), intent(in) :: u
      real(pr), dimension(1:n(1),1:2), intent(out) :: spectral_data
      integer(c_int) :: norm_flag
      call mlq_sync_params()
      norm_flag = 0
      if (normalizeSpectrumByL2Norm) norm_flag = 1
      call mlq_check(mlq_calculate_spectral_data(mlq_ctx_ptr, u, norm_flag, spectral_data), &
                     "calculate_spectral_data")
   end subroutine gpu_calculate_spectral_data

   function gpu_calc_global_summed_enstrophy(u) result(res)    ! :2455-2467
      real(pr), dimension(1:n(1),1:n(2),1:local_N,1:3), intent(in) :: u
      real(pr) :: res
      call mlq_sync_params()
      call mlq_check(mlq_calc_global_summed_enstrophy(mlq_ctx_ptr, u, res), &
                     "calc_global_summed_enstrophy")
   end function gpu_calc_global_summed_enstrophy

   function gpu_calc_global_dEnstrophydt(u) result(res)        ! :1417-1481
      real(pr), dimension(1:n(1),1:n(2),1:local_N,1:3), intent(in) :: u
      real(pr) :: res
      call mlq_sync_params()
      call mlq_check(mlq_calc_global_dEnstrophydt(mlq_ctx_ptr, u, res), "calc_global_dEnstrophydt")
   end function gpu_calc_global_dEnstrophydt

   subroutine gpu_rescaleLqNorm(u, B)                          ! :1699-1709
      real(pr), dimension(1:n(1),1:n(2),1:local_N,1:3), intent(inout) :: u
      real(pr), intent(in) :: B
      call mlq_sync_params()
      call mlq_check(mlq_rescaleLqNorm(mlq_ctx_ptr, u, B), "rescaleLqNorm")
   end subroutine gpu_rescaleLqNorm

   function gpu_vectorTransport(u, q, B, eta, xi) result(resultVec)   ! :1716-1768
      real(pr), dimension(1:n(1),1:n(2),1:local_N,1:3), intent(in) :: u, eta, xi
      real(pr), intent(in) :: q, B
      real(pr), dimension(1:n(1),1:n(2),1:local_N,1:3) :: resultVec
      call mlq_sync_params()
      call mlq_check(mlq_vectorTransport(mlq_ctx_ptr, u, q, B, eta, xi, resultVec), &
                     "vectorTransport")
   end function gpu_vectorTransport

   function gpu_global_inner_product(f, g, mytype) result(inn_prod)   ! :3508-3521
      real(pr), dimension(1:n(1),1:n(2),1:local_N), intent(in) :: f, g
      character(len=*), intent(in) :: mytype
      real(pr) :: inn_prod
      integer(c_int) :: rc
      call mlq_sync_params()
      rc = mlq_global_inner_product(mlq_ctx_ptr, f, g, ip_kind(mytype), inn_prod)
      if (rc == MLQ_EBADKIND) then
         print*, "ERROR IN function_ops.f90 inner_product: wrong exponent ", mytype
         call exit
      end if
      call mlq_check(rc, "global_inner_product")
   end function gpu_global_inner_product

   function gpu_global_summed_field_inner_product(f, g, mytype) result(inn_prod)  ! :3557-3577
      real(pr), dimension(1:n(1),1:n(2),1:local_N,1:3), intent(in) :: f, g
      character(len=*), intent(in) :: mytype
      real(pr) :: inn_prod
      integer(c_int) :: rc
      call mlq_sync_params()
      rc = mlq_global_summed_field_inner_product(mlq_ctx_ptr, f, g, ip_kind(mytype), inn_prod)
      if (rc == MLQ_EBADKIND) then
         print*, "ERROR IN function_ops.f90 inner_product: wrong exponent ", mytype
         call exit
      end if
      call mlq_check(rc, "global_summed_field_inner_product")
   end function gpu_global_summed_field_inner_product

   function gpu_calc_global_dLqdt_inclParts(u, q) result(r)    ! :1392-1405
      real(pr), dimension(1:n(1),1:n(2),1:local_N,1:3), intent(in) :: u
      real(pr), intent(in) :: q
      real(pr), dimension(3) :: r
      call mlq_sync_params()
      call mlq_check(mlq_calc_global_dLqdt_inclParts(mlq_ctx_ptr, u, q, r), "dLqdt")
   end function gpu_calc_global_dLqdt_inclParts

   subroutine gpu_gradient(u, grad_u)                          ! :3147-3195
      real(pr), dimension(1:n(1),1:n(2),1:local_N), intent(in) :: u
      real(pr), dimension(1:n(1),1:n(2),1:local_N,1:3), intent(out) :: grad_u
      call mlq_check(mlq_gradient(mlq_ctx_ptr, u, grad_u), "gradient")
   end subroutine gpu_gradient

   subroutine gpu_divergence(u, divU)                          ! :3200-3243
      real(pr), dimension(1:n(1),1:n(2),1:local_N,1:3), intent(in) :: u
      real(pr), dimension(1:n(1),1:n(2),1:local_N), intent(out) :: divU
      call mlq_check(mlq_divergence(mlq_ctx_ptr, u, divU), "divergence")
   end subroutine gpu_divergence

   subroutine gpu_laplacian(f)                                 ! :3248-3282
      real(pr), dimension(1:n(1),1:n(2),1:local_N,1:3), intent(inout) :: f
      call mlq_check(mlq_laplacian(mlq_ctx_ptr, f), "laplacian")
   end subroutine gpu_laplacian

   subroutine gpu_solve_poisson(f, l, phi)                     ! :1951-1996
      real(pr), dimension(1:n(1),1:n(2),1:local_N), intent(in) :: f
      real(pr), intent(in) :: l
      real(pr), dimension(1:n(1),1:n(2),1:local_N), intent(out) :: phi
      call mlq_check(mlq_solve_poisson(mlq_ctx_ptr, f, l, phi), "solve_poisson")
   end subroutine gpu_solve_poisson

   subroutine gpu_dealias_scalar(g, nonlinOrder)               ! :3332-3350
      real(pr), dimension(1:n(1),1:n(2),1:local_N), intent(inout) :: g
      real(pr), intent(in) :: nonlinOrder
      call mlq_sync_params()
      call mlq_check(mlq_dealias_scalar(mlq_ctx_ptr, g, nonlinOrder), "dealias_scalar")
   end subroutine gpu_dealias_scalar

   subroutine gpu_calc_uk(U, k, uk)                            ! :2324-2343
      real(pr), dimension(1:n(1),1:n(2),1:local_N,1:3), intent(in) :: U
      real(pr), intent(in) :: k
      real(pr), dimension(1:n(1),1:n(2),1:local_N), intent(out) :: uk
      call mlq_sync_params()
      call mlq_check(mlq_calc_uk(mlq_ctx_ptr, U, k, uk), "calc_uk")
   end subroutine gpu_calc_uk

   function gpu_eval_J(myfield, mysystem) result(J)            ! optimization.f90:520-547
      real(pr), dimension(1:n(1),1:n(2),1:local_N,1:3), intent(in) :: myfield
      character(len=*), intent(in) :: mysystem
      real(pr) :: J
      integer(c_int) :: sys
      select case (mysystem)
      case ("maxdLqdt")
         sys = MLQ_J_MAXDLQDT
      case ("LineMin")
         sys = MLQ_J_LINEMIN
      case ("||u||_q")
         sys = MLQ_J_LQNORM
      case default
         J = 0.0_pr
         return
      end select
      call mlq_sync_params()
      call mlq_check(mlq_eval_J(mlq_ctx_ptr, myfield, sys, J), "eval_J")
   end function gpu_eval_J

end module mlq_field_ops
