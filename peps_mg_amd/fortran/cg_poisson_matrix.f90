!> Conjugate gradient (module cg_poisson_matrix, src/cg_poisson_matrix.f90)
!> on the GPU through libpepsmg.so.
module cg_poisson_matrix

    use, intrinsic :: iso_c_binding
    use pepsmg_c
    use pepsmg_shim
    implicit none

    private
    public :: cg_solver_poisson_matrix
    public :: cg_iterator_poisson_matrix

contains

    !> cg_solver_poisson_matrix (src/cg_poisson_matrix.f90:47-146)
    subroutine cg_solver_poisson_matrix(sol, a_poisson, rhs, dm, maxiteration, tolerance, is_aggregated)
        use matrix, only : matrix_heptadiagonal
        use geometry, only : subdomain
        use mpi_topology, only : myrank
        real(kind=8), intent(inout), target            :: sol(0:,0:,0:)
        type(matrix_heptadiagonal), intent(in), target :: a_poisson
        real(kind=8), intent(in), target               :: rhs(0:,0:,0:)
        type(subdomain), intent(in), target            :: dm
        integer(kind=4), intent(in)                    :: maxiteration
        real(kind=8), intent(in)                       :: tolerance
        logical, intent(in)                            :: is_aggregated(0:2)
        type(pepsmg_grid_t) :: g
        integer(c_int32_t) :: it
        real(c_double) :: rel
        g = pepsmg_grid_of(dm)
        call pepsmg_check_padded(sol, dm%nx, dm%ny, dm%nz, 'sol', 'cg_solver_poisson_matrix')
        call pepsmg_check_padded(rhs, dm%nx, dm%ny, dm%nz, 'rhs', 'cg_solver_poisson_matrix')
        call pepsmg_check_coeff(a_poisson, dm%nx, dm%ny, dm%nz, 'cg_solver_poisson_matrix')
        call pepsmg_check(pepsmg_cg_solver(g, pepsmg_ptr3(sol), pepsmg_coeff_ptr(a_poisson), pepsmg_ptr3(rhs), &
                          maxiteration, tolerance, it, rel), 'cg_solver_poisson_matrix')
        if (myrank == 0) print '(a,i5,a,e15.7)', '[CG] Finished with total iteration = ', it, ', r_mse = ', rel
    end subroutine cg_solver_poisson_matrix

    !> cg_iterator_poisson_matrix (src/cg_poisson_matrix.f90:157-241)
    subroutine cg_iterator_poisson_matrix(sol, a_poisson, rhs, dm, maxiteration, is_aggregated)
        use matrix, only : matrix_heptadiagonal
        use geometry, only : subdomain
        real(kind=8), intent(inout), target            :: sol(0:,0:,0:)
        type(matrix_heptadiagonal), intent(in), target :: a_poisson
        real(kind=8), intent(in), target               :: rhs(0:,0:,0:)
        type(subdomain), intent(in), target            :: dm
        integer(kind=4), intent(in)                    :: maxiteration
        logical, intent(in)                            :: is_aggregated(0:2)
        type(pepsmg_grid_t) :: g
        g = pepsmg_grid_of(dm)
        call pepsmg_check_padded(sol, dm%nx, dm%ny, dm%nz, 'sol', 'cg_iterator_poisson_matrix')
        call pepsmg_check_padded(rhs, dm%nx, dm%ny, dm%nz, 'rhs', 'cg_iterator_poisson_matrix')
        call pepsmg_check_coeff(a_poisson, dm%nx, dm%ny, dm%nz, 'cg_iterator_poisson_matrix')
        call pepsmg_check(pepsmg_cg_iterator(g, pepsmg_ptr3(sol), pepsmg_coeff_ptr(a_poisson), pepsmg_ptr3(rhs), &
                          maxiteration), 'cg_iterator_poisson_matrix')
    end subroutine cg_iterator_poisson_matrix

end module cg_poisson_matrix
