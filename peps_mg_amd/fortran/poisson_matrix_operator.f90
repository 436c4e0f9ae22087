!> Matrix-vector and dot products (module poisson_matrix_operator,
!> src/poisson_matrix_operator.f90) on the GPU through libpepsmg.so.
module poisson_matrix_operator

    use, intrinsic :: iso_c_binding
    use pepsmg_c
    use pepsmg_shim
    implicit none

    private
    public :: vv_dot_3d_matrix
    public :: mv_mul_poisson_matrix

contains

    !> vv_dot_3d_matrix (src/poisson_matrix_operator.f90:40-99)
    subroutine vv_dot_3d_matrix(result, x, y, nx, ny, nz, is_serial)
        real(kind=8), intent(out)             :: result
        real(kind=8), intent(in), target      :: x(0:,0:,0:), y(0:,0:,0:)
        integer(kind=4), intent(in)           :: nx, ny, nz
        logical, intent(in)                   :: is_serial(0:2)
        real(c_double) :: r
        call pepsmg_check_padded(x, nx, ny, nz, 'x', 'vv_dot_3d_matrix')
        call pepsmg_check_padded(y, nx, ny, nz, 'y', 'vv_dot_3d_matrix')
        call pepsmg_check(pepsmg_vv_dot(nx, ny, nz, pepsmg_ptr3(x), pepsmg_ptr3(y), r), 'vv_dot_3d_matrix')
        result = r
    end subroutine vv_dot_3d_matrix

    !> mv_mul_poisson_matrix (src/poisson_matrix_operator.f90:110-161)
    subroutine mv_mul_poisson_matrix(y, a_poisson, x, dm, is_serial)
        use geometry, only : subdomain
        use matrix, only : matrix_heptadiagonal
        real(kind=8), intent(out), target                 :: y(0:,0:,0:)
        type(matrix_heptadiagonal), intent(in), target    :: a_poisson
        real(kind=8), intent(inout), target               :: x(0:,0:,0:)
        type(subdomain), intent(in), target               :: dm
        logical, intent(in)                               :: is_serial(0:2)
        type(pepsmg_grid_t) :: g
        g = pepsmg_grid_of(dm)
        call pepsmg_check_padded(y, dm%nx, dm%ny, dm%nz, 'y', 'mv_mul_poisson_matrix')
        call pepsmg_check_padded(x, dm%nx, dm%ny, dm%nz, 'x', 'mv_mul_poisson_matrix')
        call pepsmg_check_coeff(a_poisson, dm%nx, dm%ny, dm%nz, 'mv_mul_poisson_matrix')
        call pepsmg_check(pepsmg_mv_mul(g, pepsmg_ptr3(y), pepsmg_coeff_ptr(a_poisson), pepsmg_ptr3(x)), &
                          'mv_mul_poisson_matrix')
    end subroutine mv_mul_poisson_matrix

end module poisson_matrix_operator
