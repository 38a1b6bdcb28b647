! B-continuation on one GPU: the constraint loop of the reference's driver
! (maxdLpdt.f90:458-512) over the C ABI, writing the reference's files:
!   <home>results_q<q>.dat               save_to_optimizationResultList (data_ops.f90:728-759)
!   <home>constraintDirs/B<bbb>-<B>/     initializeConstraintDirectory (data_ops.f90:804-837)
!      iteration-info-B<bbb>.dat         save_diagnostics_optim (data_ops.f90:764-799)
!   <home>ncFiles/u_result_q<q>_n<n>_B<bbb>_iter<iiiii|end>.nc
!                                        save_field (data_ops.f90:9-41; optimization.f90:388-390, 477)
!   <home>constraintDirs/B<bbb>-<B>/kappaTest/kappa_B<bbb>_<name>.dat
!                                        save_kappa_test (data_ops.f90:1448-1491), with kappa=
! Usage:
!   mlq_continuation <home/> <n> <q> <nB> <max_iter> <optim_tol> <fp64|mixed> <save_every>
!                    [kappa=<every>] [<file>]
! kappa=<every> turns kappaTest on (tests at the start and end of every solve,
! optimization.f90:69-72, 457-469) with kappaTestEveryXiteration = <every>
! (< 1 never, :147-149); the perturbation is create-random-poly-b with seed 1.
! Without <file> the first B solve starts from the seeded divergence-free field
! (seed 1, exp(-|m|^2/16)).  With <file> (a u_result name) q, n, bIterOffset
! and optimizationIterOffset come from its name (function_ops.f90:4635-4674)
! and the initial field from the file; the loop runs B numbers
! bIterOffset+1 .. bIterOffset+nB, the first of them from iteration
! optimizationIterOffset (optimization.f90:81-86).
module mlq_cont_io
   use, intrinsic :: iso_c_binding
   use maxlq_c
   implicit none
   character(len=:), allocatable :: home_dir, nc_dir, constraint_dir, bIterTxt
   integer(c_int) :: grid_n, b_iter
   real(c_double) :: q_value
contains
   subroutine make_dir(dir)                     ! createDirectoryIfNonExistent (data_ops.f90:842-870)
      character(len=*), intent(in) :: dir
      logical :: ex
      inquire(file=trim(dir), exist=ex)
      if (.not. ex) call execute_command_line("mkdir -p " // trim(dir))
   end subroutine make_dir

   function field_file(iter) result(name)
      integer(c_int), intent(in) :: iter
      character(len=:), allocatable :: name
      character(kind=c_char) :: buf(4096)
      if (mlq_result_filename(buf, 4096, cstr(nc_dir), cstr("u_result"), q_value, grid_n, b_iter, &
                              iter) /= MLQ_OK) stop 4
      name = fstr(buf)
   end function field_file

   ! the solve's records: iteration-info lines and the periodic field files
   subroutine on_log(user, kind, iter, i0, i1, d) bind(C)
      type(c_ptr), value :: user
      integer(c_int), value :: kind, iter, i0, i1
      type(c_ptr), value :: d
      real(c_double), pointer :: dd(:)
      character(len=:), allocatable :: fname
      select case (kind)
      case (MLQ_LOG_ITER)
         call c_f_pointer(d, dd, [3])
         fname = constraint_dir // "iteration-info" // "-B" // trim(bIterTxt) // ".dat"
         if (i0 == 0) then
            open(10, file=fname, form="FORMATTED", status="REPLACE")
            write(10, *) "Iter  Tau  Beta  J"
            close(10)
         end if
         open(10, file=fname, form="FORMATTED", position="APPEND")
         write(10, "(I5.4, 3 ES20.12)") i0, dd(1), dd(2), dd(3)
         close(10)
      case (MLQ_LOG_KAPPA)
         call c_f_pointer(d, dd, [83])
         call save_kappa(iter, i0, i1 /= 0, dd)
      case (MLQ_LOG_FIELD)
         call c_f_pointer(d, dd, [3 * grid_n**3])
         fname = field_file(i0)
         if (mlq_save_field_nc(cstr(fname), grid_n, grid_n, grid_n, dd) /= MLQ_OK) then
            print*, "cannot write ", fname
            stop 5
         end if
      end select
   end subroutine on_log

   ! save_kappa_test for the 16 rows of one kappa test (data_ops.f90:1448-1491;
   ! without the adjusted columns the first row is written twice, as there)
   subroutine save_kappa(iter, id, adjusted, d)
      integer(c_int), intent(in) :: iter, id
      logical, intent(in) :: adjusted
      real(c_double), intent(in) :: d(83)
      character(len=:), allocatable :: kdir, fname, ident
      character(len=5) :: itxt
      integer :: ii
      real(c_double) :: eps, ip, dJ, kap, kadj
      select case (id)
      case (0); ident = "initial_grad_L2"
      case (1)
         write(itxt, "(i5.5)") iter
         ident = "gradL2_iter" // itxt
      case (2); ident = "end_gradJ"
      case default; ident = "end_d1"
      end select
      kdir = constraint_dir // "kappaTest/"
      call make_dir(kdir)
      fname = kdir // "kappa" // "_B" // trim(bIterTxt) // "_" // ident // ".dat"
      do ii = 1, 16
         eps = d((ii - 1) * 5 + 1); ip = d((ii - 1) * 5 + 2); dJ = d((ii - 1) * 5 + 3)
         kap = d((ii - 1) * 5 + 4); kadj = d((ii - 1) * 5 + 5)
         if (adjusted) then
            if (ii == 1) then
               open(10, file=fname, form="FORMATTED", status="REPLACE")
               write(10, "(9 G20.12)") "eps", "deltaJ", "deltaJ/eps", "inner_prod", "kappa_adj", &
                  "LOG10|kap_adj-1|", "kappa", "LOG10|kap-1|", "adj_factor"
               close(10)
            end if
            open(10, file=fname, form="FORMATTED", status="OLD", position="APPEND")
            write(10, "(9 ES20.12)") eps, dJ, dJ / eps, ip, kadj, log10(abs(kadj - 1.0_c_double)), &
               kap, log10(abs(kap - 1.0_c_double)), d(81)
            close(10)
         else
            if (ii == 1) then
               open(10, file=fname, form="FORMATTED", status="REPLACE")
               write(10, "(6 G20.12)") "eps", "deltaJ", "deltaJ/eps", "inner_prod", "kappa", &
                  "LOG10|kap-1|"
               write(10, "(6 ES20.12)") eps, dJ, dJ / eps, ip, kap, log10(abs(kap - 1.0_c_double))
               close(10)
            end if
            open(10, file=fname, form="FORMATTED", status="OLD", position="APPEND")
            write(10, "(6 ES20.12)") eps, dJ, dJ / eps, ip, kap, log10(abs(kap - 1.0_c_double))
            close(10)
         end if
      end do
   end subroutine save_kappa
end module mlq_cont_io

program mlq_continuation
   use, intrinsic :: iso_c_binding
   use maxlq_c
   use mlq_cont_io
   implicit none
   character(len=4096) :: arg, resume
   integer(c_int) :: n, nB, max_iter, save_every, prec, rc, h, bIterOffset, optimIterOffset, j, jend
   integer(c_int) :: resol, iterOut, kappa_every, ia
   logical :: kappa_on
   real(c_double) :: q, optim_tol, B
   real(c_double), allocatable :: B_list(:), u(:,:,:,:)
   type(c_ptr) :: ctx
   type(mlq_params) :: p
   type(mlq_rcg_cfg) :: cfg
   type(mlq_rcg_result) :: res
   character(len=20) :: bTxtBuf
   character(len=:), allocatable :: results_file
   logical :: ex

   if (command_argument_count() < 8) then
      print*, "usage: mlq_continuation <home/> <n> <q> <nB> <max_iter> <optim_tol> <fp64|mixed>", &
              " <save_every> [kappa=<every>] [<u_result file>]"
      stop 2
   end if
   call get_command_argument(1, arg); home_dir = trim(arg)
   call get_command_argument(2, arg); read(arg, *) n
   call get_command_argument(3, arg); read(arg, *) q
   call get_command_argument(4, arg); read(arg, *) nB
   call get_command_argument(5, arg); read(arg, *) max_iter
   call get_command_argument(6, arg); read(arg, *) optim_tol
   call get_command_argument(7, arg); prec = merge(MLQ_MIXED, MLQ_FP64, trim(arg) == "mixed")
   call get_command_argument(8, arg); read(arg, *) save_every
   bIterOffset = 0
   optimIterOffset = 0
   resume = ""
   kappa_on = .false.
   kappa_every = 0
   do ia = 9, command_argument_count()
      call get_command_argument(ia, arg)
      if (arg(1:6) == "kappa=") then
         kappa_on = .true.
         read(arg(7:), *) kappa_every
      else
         resume = arg
      end if
   end do
   if (len_trim(resume) > 0) then
      rc = mlq_parse_result_filename(cstr(resume), q, resol, bIterOffset, optimIterOffset)
      if (rc /= MLQ_OK) stop 3
      n = resol
      print*, "assigned lebesgueQ = ", q, " resol = ", resol, " bIterOffset = ", bIterOffset, &
              " optimizationIterOffset = ", optimIterOffset
   end if
   grid_n = n
   q_value = q
   nc_dir = home_dir // "ncFiles/"
   call make_dir(home_dir)                                     ! maxdLpdt.f90:253-254
   call make_dir(nc_dir)

   ! B values (maxdLpdt.f90:465-481)
   allocate(B_list(32 + bIterOffset))
   B_list(1) = 1.0_c_double
   do j = 2, size(B_list)
      if (abs(q - 3.0_c_double) < 2.0e-16_c_double) then
         B = B_list(j - 1) * 10.0_c_double**(1.0_c_double / 4.0_c_double)
         if (B > 100.0_c_double) B = B_list(j - 1) * 10.0_c_double**(1.0_c_double / 16.0_c_double)
         if (B > 110.0_c_double) B = B_list(j - 1) * 10.0_c_double**(1.0_c_double / 32.0_c_double)
      else
         B = B_list(j - 1) * 10.0_c_double**(1.0_c_double / 8.0_c_double)
      end if
      B_list(j) = B
   end do

   call mlq_default_params(p, q)
   p%B = B_list(1 + bIterOffset)
   p%precision = prec
   if (mlq_init(ctx, n, n, n, p) /= MLQ_OK) stop 6
   if (mlq_field_alloc(ctx, 3, h) /= MLQ_OK) stop 6
   allocate(u(n, n, n, 3))
   if (len_trim(resume) > 0) then
      if (mlq_read_field_nc(cstr(resume), n, n, n, u) /= MLQ_OK) then
         print*, "cannot read ", trim(resume)
         stop 7
      end if
      if (mlq_upload(ctx, h, u) /= MLQ_OK) stop 6
   else
      if (mlq_seeded_field(ctx, h, 1_c_long_long, 0, 16.0_c_double) /= MLQ_OK) stop 6
   end if

   results_file = home_dir // "results_q" // qtxt(q) // ".dat"
   jend = min(size(B_list), bIterOffset + nB)
   do j = 1 + bIterOffset, jend                                ! maxdLpdt.f90:498-512
      B = B_list(j)
      b_iter = j
      write(bTxtBuf, "(i3.3)") j
      bIterTxt = trim(bTxtBuf)
      write(bTxtBuf, "(ES7.1)") B
      print*, "constraint B" // bIterTxt, B
      call make_dir(home_dir // "constraintDirs/")            ! data_ops.f90:804-837
      constraint_dir = home_dir // "constraintDirs/" // "B" // bIterTxt // "-" // trim(bTxtBuf) // "/"
      call make_dir(constraint_dir)
      p%B = B
      if (mlq_set_params(ctx, p) /= MLQ_OK) stop 6
      cfg = mlq_rcg_cfg(max_iter, optim_tol, 0, merge(3, 0, prec == MLQ_MIXED), 25, 10, &
                        merge(MLQ_RCG_KAPPA, 0, kappa_on), 0, save_every, kappa_every, &
                        1_c_long_long, c_null_ptr)
      if (j == bIterOffset + 1) cfg%start_iter = optimIterOffset    ! optimization.f90:81-86
      rc = mlq_rcg_solve(ctx, h, cfg, res, c_funloc(on_log), c_null_ptr)
      if (rc /= MLQ_OK) then
         print*, "maxdLqdt failed", rc
         stop 8
      end if
      ! save_data_Optim: the result field with optimizationIterationTxt = "end"
      if (mlq_download(ctx, h, u) /= MLQ_OK) stop 6
      if (mlq_save_field_nc(cstr(field_file(-1_c_int)), n, n, n, u) /= MLQ_OK) stop 5
      ! save_to_optimizationResultList (data_ops.f90:728-759); the iteration
      ! count is negative for an unsuccessful solve
      iterOut = res%iters
      if (res%successful == 0) iterOut = -res%iters
      inquire(file=results_file, exist=ex)
      if (ex) then
         open(10, file=results_file, form="FORMATTED", status="OLD", position="APPEND")
      else
         open(10, file=results_file, form="FORMATTED", status="REPLACE")
         write(10, "(7 G20.12)") "B#", "B", "J1", "iter", "enstrophy", "enstrophy_lhs", "q"
      end if
      write(10, "(A20, 2 ES20.12, I20, 3 ES20.12)") bIterTxt, B, res%J1, iterOut, res%enstrophy, &
                                                     res%dEnstrophydt, q
      close(10)
   end do
   rc = mlq_finalize(ctx)
   print*, "end"

contains
   function qtxt(qq) result(s)
      real(c_double), intent(in) :: qq
      character(len=:), allocatable :: s
      character(kind=c_char) :: buf(64)
      if (mlq_q_text(qq, buf, 64) /= MLQ_OK) stop 4
      s = fstr(buf)
   end function qtxt
end program mlq_continuation
