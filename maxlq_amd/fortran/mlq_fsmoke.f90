! Fortran caller of libmaxlq.so through the maxlq_c interface: checks, on the
! GPU, two closed forms of the reference's operators at 32^3:
!   * ||u||_q = c for u = c (sin 2 pi z, cos 2 pi z, 0)      (|u| = c everywhere)
!   * R(u) = -4 pi^2 ||u||_q^q, R3 = 0 for the ABC flow     (Beltrami field)
! and a device-resident RCG solve of 2 iterations (J increases).
program mlq_fsmoke
   use, intrinsic :: iso_c_binding
   use maxlq_c
   implicit none
   integer, parameter :: n = 32
   real(c_double), parameter :: pi = 3.14159265358979323846264338327950288_c_double
   real(c_double), allocatable :: u(:,:,:,:), s(:,:,:)
   type(c_ptr) :: ctx
   type(mlq_params) :: p
   type(mlq_rcg_cfg) :: cfg
   type(mlq_rcg_result) :: res
   real(c_double) :: nrm, r(3), lqq, x, y, z, c, ens
   integer(c_int) :: rc, h
   integer :: i, j, k, fails

   fails = 0
   allocate(u(n,n,n,3), s(n,n,n))
   call mlq_default_params(p, 4.0_c_double)
   rc = mlq_init(ctx, n, n, n, p)
   if (rc /= MLQ_OK) stop 3

   c = 0.75_c_double
   do k = 1, n
      do j = 1, n
         do i = 1, n
            z = real(k - 1, c_double) / n
            u(i, j, k, 1) = c * sin(2 * pi * z)
            u(i, j, k, 2) = c * cos(2 * pi * z)
            u(i, j, k, 3) = 0.0_c_double
         end do
      end do
   end do
   rc = mlq_calc_global_Lq_norm(ctx, u, nrm)
   if (rc /= MLQ_OK .or. abs(nrm - c) > 1e-13_c_double) then
      print*, "Lq norm", nrm, "expected", c
      fails = fails + 1
   end if

   do k = 1, n
      do j = 1, n
         do i = 1, n
            x = real(i - 1, c_double) / n
            y = real(j - 1, c_double) / n
            z = real(k - 1, c_double) / n
            u(i, j, k, 1) = cos(2 * pi * z) + cos(2 * pi * y)
            u(i, j, k, 2) = cos(2 * pi * x) + cos(2 * pi * z)
            u(i, j, k, 3) = cos(2 * pi * y) + cos(2 * pi * x)
         end do
      end do
   end do
   s = sum(u * u, dim=4)
   lqq = sum(s * s) / real(n, c_double)**3
   rc = mlq_calc_global_dLqdt_inclParts(ctx, u, 4.0_c_double, r)
   if (rc /= MLQ_OK .or. abs(r(2) + 4 * pi**2 * lqq) > 1e-10_c_double * 4 * pi**2 * lqq &
       .or. abs(r(3)) > 1e-9_c_double * abs(r(2))) then
      print*, "R parts", r, "expected R1+R2 =", -4 * pi**2 * lqq
      fails = fails + 1
   end if

   ! two RCG iterations on a perturbed ABC field, resident on the device
   do k = 1, n
      do j = 1, n
         do i = 1, n
            x = real(i - 1, c_double) / n
            u(i, j, k, 2) = u(i, j, k, 2) + 0.3_c_double * sin(4 * pi * x)
         end do
      end do
   end do
   rc = mlq_field_alloc(ctx, 3, h)
   rc = mlq_upload(ctx, h, u)
   cfg = mlq_rcg_cfg(2, 1.0e-5_c_double, 0, 0, 25, 10, 0, 0, 0, 0, 0_c_long_long, c_null_ptr)
   rc = mlq_rcg_solve(ctx, h, cfg, res, c_null_funptr, c_null_ptr)
   if (rc /= MLQ_OK .or. res%iters < 1) then
      print*, "rcg solve rc", rc, "iters", res%iters
      fails = fails + 1
   end if
   ! the result's enstrophy equals calc_global_summed_enstrophy of the result
   rc = mlq_download(ctx, h, u)
   rc = mlq_calc_global_summed_enstrophy(ctx, u, ens)
   if (rc /= MLQ_OK .or. .not. (res%enstrophy > 0.0_c_double) .or. &
       abs(res%enstrophy - ens) > 1.0e-10_c_double * ens) then
      print*, "enstrophy", rc, res%enstrophy, ens
      fails = fails + 1
   end if
   rc = mlq_finalize(ctx)
   if (fails /= 0) then
      print*, "FAIL", fails
      stop 1
   end if
   print*, "PASS mlq_fsmoke", nrm, r(1), res%J1
end program mlq_fsmoke
