!> Helpers shared by the drop-in modules: C descriptors of Fortran arrays.
module pepsmg_shim

    use, intrinsic :: iso_c_binding
    use pepsmg_c
    use geometry, only : subdomain
    use matrix, only : matrix_heptadiagonal
    implicit none

contains

    !> pepsmg_grid from a subdomain's metric arrays (valid while sdm lives).
    function pepsmg_grid_of(sdm) result(g)
        type(subdomain), intent(in), target :: sdm
        type(pepsmg_grid_t) :: g
        g%nx = sdm%nx
        g%ny = sdm%ny
        g%nz = sdm%nz
        g%dxm = c_loc(sdm%dxm(lbound(sdm%dxm, 1)))
        g%dxg = c_loc(sdm%dxg(lbound(sdm%dxg, 1)))
        g%dym = c_loc(sdm%dym(lbound(sdm%dym, 1)))
        g%dyg = c_loc(sdm%dyg(lbound(sdm%dyg, 1)))
        g%dzm = c_loc(sdm%dzm(lbound(sdm%dzm, 1)))
        g%dzg = c_loc(sdm%dzg(lbound(sdm%dzg, 1)))
    end function pepsmg_grid_of

    !> Address of a padded (0:nx+1,0:ny+1,0:nz+1) contiguous array.
    function pepsmg_ptr3(a) result(p)
        real(kind=8), intent(in), target, contiguous :: a(0:,0:,0:)
        type(c_ptr) :: p
        p = c_loc(a(0,0,0))
    end function pepsmg_ptr3

    !> Address of a matrix's coefficient array (NULL when unallocated).
    function pepsmg_coeff_ptr(a_poisson) result(p)
        type(matrix_heptadiagonal), intent(in), target :: a_poisson
        type(c_ptr) :: p
        if (allocated(a_poisson%coeff)) then
            p = c_loc(a_poisson%coeff(0,1,1,1))
        else
            p = c_null_ptr
        endif
    end function pepsmg_coeff_ptr

end module pepsmg_shim
