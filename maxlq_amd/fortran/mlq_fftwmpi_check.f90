! The fftw3-mpi.f03 stand-in used the way initialize.f90:131-153 uses FFTW's
! (INCLUDE inside the subroutine, fftw_mpi_local_size_3d on the C_INTPTR_T
! grid sizes): one rank holds the whole slab.
subroutine slab_sizes(resol, comm, local_N, local_k_offset, alloc)
   use, intrinsic :: iso_c_binding
   implicit none
   INCLUDE "fftw3-mpi.f03"
   integer, intent(in) :: resol, comm
   integer, intent(out) :: local_N, local_k_offset
   integer(C_INTPTR_T), intent(out) :: alloc
   integer(C_INTPTR_T) :: C_n(3), C_local_N, C_local_k_offset
   C_n = resol
   alloc = fftw_mpi_local_size_3d(C_n(3), C_n(2), C_n(1), int(comm, C_INT32_T), C_local_N, &
                                  C_local_k_offset)
   local_N = int(C_local_N)
   local_k_offset = int(C_local_k_offset)
end subroutine slab_sizes

program mlq_fftwmpi_check
   use, intrinsic :: iso_c_binding
   implicit none
   integer :: local_N, local_k_offset
   integer(C_INTPTR_T) :: alloc
   call slab_sizes(64, 91, local_N, local_k_offset, alloc)
   print*, local_N, local_k_offset, alloc
   if (local_N /= 64 .or. local_k_offset /= 0 .or. alloc /= 64_C_INTPTR_T**3) stop 1
end program mlq_fftwmpi_check
