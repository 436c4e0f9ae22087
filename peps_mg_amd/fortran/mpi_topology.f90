!> One-process Cartesian topology (module mpi_topology, src/mpi_topology.f90).
!> One GPU owns the whole grid, so every communicator has one rank and no
!> neighbours; the variables keep the reference's names for callers.
module mpi_topology

    implicit none

    integer(kind=4), parameter :: MPI_PROC_NULL_1P = -2

    integer(kind=4) :: mpi_world_cart = 0
    integer(kind=4) :: nprocs = 1       !< number of processes
    integer(kind=4) :: myrank = 0       !< rank
    integer(kind=4) :: np_dim(0:2) = 1  !< processes per direction (must be 1)
    logical         :: period(0:2) = .false.

    !> type cart_comm_1d (src/mpi_topology.f90:32-38)
    type, public :: cart_comm_1d
        integer(kind=4) :: myrank = 0
        integer(kind=4) :: nprocs = 1
        integer(kind=4) :: west_rank = MPI_PROC_NULL_1P
        integer(kind=4) :: east_rank = MPI_PROC_NULL_1P
        integer(kind=4) :: mpi_comm = 0
    end type cart_comm_1d

    type(cart_comm_1d) :: comm_1d_x, comm_1d_y, comm_1d_z

contains

    !> mpi_topology_create (src/mpi_topology.f90:53-99) for one process.
    subroutine mpi_topology_create()
        if (any(np_dim /= 1)) then
            print '(a,3i4)', '[Error] this build runs on one process (one GPU); npx, npy, npz = ', np_dim
            stop 301
        endif
        if (any(period)) then
            print '(a)', '[Error] periodic boundaries are not supported (homogeneous Dirichlet only)'
            stop 301
        endif
    end subroutine mpi_topology_create

    !> mpi_topology_destroy (src/mpi_topology.f90:104-111)
    subroutine mpi_topology_destroy()
    end subroutine mpi_topology_destroy

end module mpi_topology
