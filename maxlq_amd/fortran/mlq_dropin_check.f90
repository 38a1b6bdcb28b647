! The reference's hot-path routines called the way its driver calls them:
! through the drop-in modules fftwfunction (fftwfunction_gpu.f90) and
! mlq_field_ops, against the reference's own module global_variables, on the
! GPU.  Closed forms on the ABC flow u = (cos 2 pi z + cos 2 pi y,
! cos 2 pi x + cos 2 pi z, cos 2 pi y + cos 2 pi x) at 32^3, q = 4 (a Beltrami
! field: R1 + R2 = -4 pi^2 ||u||_4^4, R3 = 0, and R is homogeneous of degree q
! in u, so <GradL2ForLq(u), u>_L2 = q R):
!   * fftfwd of cos(2 pi z) is n^3/2 at m_z = +-1; fftbwd(fftfwd(f)) = f
!   * calc_global_Lq_norm, calc_global_dLqdt_inclParts, GradL2ForLq
!   * the same R through a mixed-precision context (mlq_precision before init_fft)
program mlq_dropin_check
   use, intrinsic :: iso_c_binding
   use global_variables
   use fftwfunction
   use mlq_field_ops
   implicit none
   real(pr), parameter :: tp = 2.0_pr * 3.14159265358979323846264338327950288_pr
   complex(pr), allocatable :: c(:,:,:), fc(:,:,:), c2(:,:,:)
   real(pr), allocatable :: V(:,:,:,:), s(:,:,:)
   real(pr) :: x, y, z, lqq, r(3), rm(3), nrm, ip
   integer :: i, j, k, fails

   fails = 0
   resol = 32
   n = resol
   lebesgueQ = 4.0_pr
   constraintB = 1.0_pr
   use_e_u_instead_of_uqMinus4 = .false.        ! set_e_u_or_just_u for q >= 4 (initialize.f90:75-86)
   call init_fft()
   print*, "local_N", local_N, "local_k_offset", local_k_offset
   allocate(Uvec(n(1), n(2), local_N, 3), V(n(1), n(2), local_N, 3), s(n(1), n(2), local_N))
   allocate(c(n(1), n(2), local_N), fc(n(1), n(2), local_N), c2(n(1), n(2), local_N))
   do k = 1, local_N
      do j = 1, n(2)
         do i = 1, n(1)
            x = real(i - 1, pr) / n(1)
            y = real(j - 1, pr) / n(2)
            z = real(k - 1, pr) / n(3)
            Uvec(i, j, k, 1) = cos(tp * z) + cos(tp * y)
            Uvec(i, j, k, 2) = cos(tp * x) + cos(tp * z)
            Uvec(i, j, k, 3) = cos(tp * y) + cos(tp * x)
            c(i, j, k) = cmplx(cos(tp * z), sin(tp * x), pr)
         end do
      end do
   end do

   call fftfwd(c, fc)
   if (abs(real(fc(1, 1, 2)) - 0.5_pr * n(1)**3) > 1.0e-9_pr * n(1)**3 .or. &
       abs(real(fc(1, 1, n(3))) - 0.5_pr * n(1)**3) > 1.0e-9_pr * n(1)**3) then
      print*, "fftfwd of cos(2 pi z):", fc(1, 1, 2), fc(1, 1, n(3))
      fails = fails + 1
   end if
   call fftbwd(fc, c2)
   if (maxval(abs(c2 - c)) > 1.0e-13_pr) then
      print*, "fftbwd(fftfwd(f)) - f:", maxval(abs(c2 - c))
      fails = fails + 1
   end if

   s = sum(Uvec * Uvec, dim=4)
   lqq = sum(s * s) / real(n(1), pr)**3
   nrm = gpu_calc_global_Lq_norm(Uvec)
   if (abs(nrm - lqq**0.25_pr) > 1.0e-13_pr * nrm) then
      print*, "calc_global_Lq_norm", nrm, "expected", lqq**0.25_pr
      fails = fails + 1
   end if
   r = gpu_calc_global_dLqdt_inclParts(Uvec, lebesgueQ)
   if (abs(r(2) + tp**2 * lqq) > 1.0e-10_pr * tp**2 * lqq .or. &
       abs(r(3)) > 1.0e-9_pr * abs(r(2))) then
      print*, "calc_global_dLqdt_inclParts", r, "expected R1+R2 =", -tp**2 * lqq
      fails = fails + 1
   end if
   call gpu_GradL2ForLq(Uvec, V)
   ip = gpu_global_summed_field_inner_product(V, Uvec, "L2")
   if (abs(ip - lebesgueQ * r(1)) > 1.0e-10_pr * abs(lebesgueQ * r(1))) then
      print*, "<GradL2ForLq(u), u>", ip, "expected q R =", lebesgueQ * r(1)
      fails = fails + 1
   end if

   ! the mixed-precision path selected through the module variable
   call fft_deallocate()
   mlq_precision = MLQ_MIXED
   call init_fft()
   rm = gpu_calc_global_dLqdt_inclParts(Uvec, lebesgueQ)
   if (abs(rm(1) - r(1)) > 1.0e-6_pr * abs(r(1))) then
      print*, "mixed R", rm(1), "fp64 R", r(1)
      fails = fails + 1
   end if
   call fft_deallocate()

   if (fails /= 0) then
      print*, "FAIL", fails
      stop 1
   end if
   print*, "PASS mlq_dropin_check", nrm, r(1), ip, rm(1)
end program mlq_dropin_check
