!> Heptadiagonal 7-point matrix (module matrix, src/matrix.f90).  The GPU
!> path is matrix-free; the AoS array is kept for callers that use it, and
!> the library checks it against the grid metrics on every solve.
module matrix

    implicit none

    !> type matrix_heptadiagonal (src/matrix.f90:27-30)
    type, public :: matrix_heptadiagonal
        integer(kind=4) :: dof
        real(kind=8), allocatable :: coeff(:,:,:,:)
    end type matrix_heptadiagonal

contains

    !> matrix_heptadiagonal_create (src/matrix.f90:39-94); index 0 centre,
    !> 1 (i-1), 2 (i+1), 3 (j-1), 4 (j+1), 5 (k-1), 6 (k+1).
    subroutine matrix_heptadiagonal_create(a_poisson, sdm)
        use geometry, only : subdomain
        type(matrix_heptadiagonal), intent(inout) :: a_poisson
        type(subdomain), intent(in)               :: sdm
        integer(kind=4) :: i, j, k
        real(kind=8) :: cxw, cxe, cyw, cye, czw, cze
        a_poisson%dof = sdm%nx * sdm%ny * sdm%nz
        allocate(a_poisson%coeff(0:6, sdm%nx, sdm%ny, sdm%nz))
        do k = 1, sdm%nz
            czw = 1.0d0 / (sdm%dzm(k) * sdm%dzg(k))
            cze = 1.0d0 / (sdm%dzm(k) * sdm%dzg(k+1))
            do j = 1, sdm%ny
                cyw = 1.0d0 / (sdm%dym(j) * sdm%dyg(j))
                cye = 1.0d0 / (sdm%dym(j) * sdm%dyg(j+1))
                do i = 1, sdm%nx
                    cxw = 1.0d0 / (sdm%dxm(i) * sdm%dxg(i))
                    cxe = 1.0d0 / (sdm%dxm(i) * sdm%dxg(i+1))
                    a_poisson%coeff(0,i,j,k) = -(cxw + cxe + cyw + cye + czw + cze)
                    a_poisson%coeff(1,i,j,k) = cxw
                    a_poisson%coeff(2,i,j,k) = cxe
                    a_poisson%coeff(3,i,j,k) = cyw
                    a_poisson%coeff(4,i,j,k) = cye
                    a_poisson%coeff(5,i,j,k) = czw
                    a_poisson%coeff(6,i,j,k) = cze
                enddo
            enddo
        enddo
    end subroutine matrix_heptadiagonal_create

    !> matrix_heptadiagonal_destroy (src/matrix.f90:100-109)
    subroutine matrix_heptadiagonal_destroy(a_poisson)
        type(matrix_heptadiagonal), intent(inout) :: a_poisson
        a_poisson%dof = 0
        if (allocated(a_poisson%coeff)) deallocate(a_poisson%coeff)
    end subroutine matrix_heptadiagonal_destroy

end module matrix
