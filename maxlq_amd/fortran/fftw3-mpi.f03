! Stand-in for FFTW's fftw3-mpi.f03 on one GPU: the declarations that
! initialize.f90:136-153 needs (INCLUDE "fftw3-mpi.f03" inside SUBROUTINE
! initialize), for the reference's single-rank slab: local_N = n3, offset 0,
! as fftw_mpi_local_size_3d returns on one rank (mlq_local_size_3d in the
! library).  The reference's driver keeps its own MPI calls (np = 1).
      interface
        integer(C_INTPTR_T) function fftw_mpi_local_size_3d(n0, n1, n2, comm, local_n0, &
              local_0_start) bind(C, name="mlq_fftw_mpi_local_size_3d")
          import
          integer(C_INTPTR_T), value :: n0, n1, n2
          integer(C_INT32_T), value :: comm
          integer(C_INTPTR_T), intent(out) :: local_n0, local_0_start
        end function fftw_mpi_local_size_3d
      end interface
