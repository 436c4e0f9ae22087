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

    !> Stops with an [Error] line unless a has the padded extents
    !> (0:nx+1,0:ny+1,0:nz+1) of the grid: the library copies that many
    !> values in and out through the array's address.
    subroutine pepsmg_check_padded(a, nx, ny, nz, name, where)
        real(kind=8), intent(in), contiguous :: a(0:,0:,0:)
        integer(kind=4), intent(in) :: nx, ny, nz
        character(*), intent(in)    :: name, where
        if (size(a, 1) /= nx + 2 .or. size(a, 2) /= ny + 2 .or. size(a, 3) /= nz + 2) then
            print '(a,a,a,a,a,3i7,a,3i7)', '[Error] ', where, ': ', name, ' has extents', size(a, 1), size(a, 2), &
                size(a, 3), ', expected', nx + 2, ny + 2, nz + 2
            stop 301
        endif
    end subroutine pepsmg_check_padded

    !> Stops with an [Error] line unless an allocated coefficient array is
    !> coeff(0:6,nx,ny,nz).
    subroutine pepsmg_check_coeff(a_poisson, nx, ny, nz, where)
        type(matrix_heptadiagonal), intent(in) :: a_poisson
        integer(kind=4), intent(in) :: nx, ny, nz
        character(*), intent(in)    :: where
        if (.not. allocated(a_poisson%coeff)) return
        if (size(a_poisson%coeff, 1) /= 7 .or. size(a_poisson%coeff, 2) /= nx .or. &
            size(a_poisson%coeff, 3) /= ny .or. size(a_poisson%coeff, 4) /= nz) then
            print '(a,a,a,4i7)', '[Error] ', where, ': coeff has extents', size(a_poisson%coeff, 1), &
                size(a_poisson%coeff, 2), size(a_poisson%coeff, 3), size(a_poisson%coeff, 4)
            stop 301
        endif
    end subroutine pepsmg_check_coeff

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
