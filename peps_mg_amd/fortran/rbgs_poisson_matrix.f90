!> Red/black Gauss-Seidel/SOR (module rbgs_poisson_matrix,
!> src/rbgs_poisson_matrix.f90) on the GPU through libpepsmg.so.
module rbgs_poisson_matrix

    use, intrinsic :: iso_c_binding
    use pepsmg_c
    use pepsmg_shim
    implicit none

    private
    public :: rbgs_solver_poisson_matrix
    public :: rbgs_iterator_poisson_matrix

contains

    !> rbgs_solver_poisson_matrix (src/rbgs_poisson_matrix.f90:47-179)
    subroutine rbgs_solver_poisson_matrix(sol, a_poisson, rhs, dm, maxiteration, tolerance, omega, is_aggregated)
        use matrix, only : matrix_heptadiagonal
        use geometry, only : subdomain
        use mpi_topology, only : myrank
        real(kind=8), intent(inout), target            :: sol(0:,0:,0:)
        type(matrix_heptadiagonal), intent(in), target :: a_poisson
        real(kind=8), intent(in), target               :: rhs(0:,0:,0:)
        type(subdomain), intent(in), target            :: dm
        integer(kind=4), intent(in)                    :: maxiteration
        real(kind=8), intent(in)                       :: tolerance, omega
        logical, intent(in)                            :: is_aggregated(0:2)
        type(pepsmg_grid_t) :: g
        integer(c_int32_t) :: it
        real(c_double) :: rel
        g = pepsmg_grid_of(dm)
        call pepsmg_check_padded(sol, dm%nx, dm%ny, dm%nz, 'sol', 'rbgs_solver_poisson_matrix')
        call pepsmg_check_padded(rhs, dm%nx, dm%ny, dm%nz, 'rhs', 'rbgs_solver_poisson_matrix')
        call pepsmg_check_coeff(a_poisson, dm%nx, dm%ny, dm%nz, 'rbgs_solver_poisson_matrix')
        call pepsmg_check(pepsmg_rbgs_solver(g, pepsmg_ptr3(sol), pepsmg_coeff_ptr(a_poisson), pepsmg_ptr3(rhs), &
                          maxiteration, tolerance, omega, it, rel), 'rbgs_solver_poisson_matrix')
        if (myrank == 0) print '(a,i5,a,e15.7)', '   [RBGS solver] Solution obtained. Iter. = ', it, ', r_mse = ', rel
    end subroutine rbgs_solver_poisson_matrix

    !> rbgs_iterator_poisson_matrix (src/rbgs_poisson_matrix.f90:191-279)
    subroutine rbgs_iterator_poisson_matrix(sol, a_poisson, rhs, dm, maxiteration, omega, is_aggregated)
        use matrix, only : matrix_heptadiagonal
        use geometry, only : subdomain
        real(kind=8), intent(inout), target            :: sol(0:,0:,0:)
        type(matrix_heptadiagonal), intent(in), target :: a_poisson
        real(kind=8), intent(in), target               :: rhs(0:,0:,0:)
        type(subdomain), intent(in), target            :: dm
        integer(kind=4), intent(in)                    :: maxiteration
        real(kind=8), intent(in)                       :: omega
        logical, intent(in)                            :: is_aggregated(0:2)
        type(pepsmg_grid_t) :: g
        g = pepsmg_grid_of(dm)
        call pepsmg_check_padded(sol, dm%nx, dm%ny, dm%nz, 'sol', 'rbgs_iterator_poisson_matrix')
        call pepsmg_check_padded(rhs, dm%nx, dm%ny, dm%nz, 'rhs', 'rbgs_iterator_poisson_matrix')
        call pepsmg_check_coeff(a_poisson, dm%nx, dm%ny, dm%nz, 'rbgs_iterator_poisson_matrix')
        call pepsmg_check(pepsmg_rbgs_iterator(g, pepsmg_ptr3(sol), pepsmg_coeff_ptr(a_poisson), pepsmg_ptr3(rhs), &
                          maxiteration, omega), 'rbgs_iterator_poisson_matrix')
    end subroutine rbgs_iterator_poisson_matrix

end module rbgs_poisson_matrix
