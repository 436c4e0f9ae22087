!> iso_c_binding interface to libpepsmg.so (C declarations in include/pepsmg.h).
module pepsmg_c

    use, intrinsic :: iso_c_binding
    implicit none

    integer(c_int), parameter :: PEPSMG_OK = 0

    !> struct pepsmg_grid
    type, bind(C) :: pepsmg_grid_t
        integer(c_int32_t) :: nx, ny, nz
        type(c_ptr)        :: dxm, dxg, dym, dyg, dzm, dzg
    end type pepsmg_grid_t

    !> struct pepsmg_options
    type, bind(C) :: pepsmg_options_t
        integer(c_int32_t) :: smoother, precision, coarse_switch, inner_cycles
        integer(c_int32_t) :: max_refine, fused_smoother, tile_y, verbose
        integer(c_int32_t) :: tile_x, inner_nu
        real(c_double)     :: inner_omega
        integer(c_int32_t) :: inner_nu_post
    end type pepsmg_options_t

    !> struct pepsmg_stats
    type, bind(C) :: pepsmg_stats_t
        integer(c_int32_t) :: cycles, refine_steps, converged, history_cap
        real(c_double)     :: res0, rho
        type(c_ptr)        :: history, coarse_iters
        real(c_double)     :: device_ms
    end type pepsmg_stats_t

    interface
        function pepsmg_last_error() bind(C, name="pepsmg_last_error") result(p)
            import :: c_ptr
            type(c_ptr) :: p
        end function

        subroutine pepsmg_default_options(opt) bind(C, name="pepsmg_default_options")
            import :: pepsmg_options_t
            type(pepsmg_options_t), intent(out) :: opt
        end subroutine

        function pepsmg_multigrid_create(grid, nlevel, ncycle, aggr, cga, cgpsa, ctx) &
                bind(C, name="pepsmg_multigrid_create") result(rc)
            import :: c_int, c_int32_t, c_ptr, pepsmg_grid_t
            type(pepsmg_grid_t), intent(in) :: grid
            integer(c_int32_t), value       :: nlevel, ncycle, aggr, cga
            integer(c_int32_t), intent(in)  :: cgpsa(3)
            type(c_ptr), intent(out)        :: ctx
            integer(c_int)                  :: rc
        end function

        function pepsmg_set_options(ctx, opt) bind(C, name="pepsmg_set_options") result(rc)
            import :: c_int, c_ptr, pepsmg_options_t
            type(c_ptr), value                 :: ctx
            type(pepsmg_options_t), intent(in) :: opt
            integer(c_int)                     :: rc
        end function

        function pepsmg_multigrid_solve_vcycle(ctx, sol, rsd, coeff, rhs, maxit, tol, omega, stats) &
                bind(C, name="pepsmg_multigrid_solve_vcycle") result(rc)
            import :: c_int, c_int32_t, c_double, c_ptr, pepsmg_stats_t
            type(c_ptr), value         :: ctx, sol, rsd, coeff, rhs
            integer(c_int32_t), value  :: maxit
            real(c_double), value      :: tol, omega
            type(pepsmg_stats_t)       :: stats
            integer(c_int)             :: rc
        end function

        subroutine pepsmg_multigrid_destroy(ctx) bind(C, name="pepsmg_multigrid_destroy")
            import :: c_ptr
            type(c_ptr), value :: ctx
        end subroutine

        function pepsmg_rbgs_iterator(grid, sol, coeff, rhs, maxit, omega) &
                bind(C, name="pepsmg_rbgs_iterator") result(rc)
            import :: c_int, c_int32_t, c_double, c_ptr, pepsmg_grid_t
            type(pepsmg_grid_t), intent(in) :: grid
            type(c_ptr), value              :: sol, coeff, rhs
            integer(c_int32_t), value       :: maxit
            real(c_double), value           :: omega
            integer(c_int)                  :: rc
        end function

        function pepsmg_rbgs_solver(grid, sol, coeff, rhs, maxit, tol, omega, iters, rel) &
                bind(C, name="pepsmg_rbgs_solver") result(rc)
            import :: c_int, c_int32_t, c_double, c_ptr, pepsmg_grid_t
            type(pepsmg_grid_t), intent(in) :: grid
            type(c_ptr), value              :: sol, coeff, rhs
            integer(c_int32_t), value       :: maxit
            real(c_double), value           :: tol, omega
            integer(c_int32_t), intent(out) :: iters
            real(c_double), intent(out)     :: rel
            integer(c_int)                  :: rc
        end function

        function pepsmg_cg_iterator(grid, sol, coeff, rhs, maxit) &
                bind(C, name="pepsmg_cg_iterator") result(rc)
            import :: c_int, c_int32_t, c_ptr, pepsmg_grid_t
            type(pepsmg_grid_t), intent(in) :: grid
            type(c_ptr), value              :: sol, coeff, rhs
            integer(c_int32_t), value       :: maxit
            integer(c_int)                  :: rc
        end function

        function pepsmg_cg_solver(grid, sol, coeff, rhs, maxit, tol, iters, rel) &
                bind(C, name="pepsmg_cg_solver") result(rc)
            import :: c_int, c_int32_t, c_double, c_ptr, pepsmg_grid_t
            type(pepsmg_grid_t), intent(in) :: grid
            type(c_ptr), value              :: sol, coeff, rhs
            integer(c_int32_t), value       :: maxit
            real(c_double), value           :: tol
            integer(c_int32_t), intent(out) :: iters
            real(c_double), intent(out)     :: rel
            integer(c_int)                  :: rc
        end function

        function pepsmg_mv_mul(grid, y, coeff, x) bind(C, name="pepsmg_mv_mul") result(rc)
            import :: c_int, c_ptr, pepsmg_grid_t
            type(pepsmg_grid_t), intent(in) :: grid
            type(c_ptr), value              :: y, coeff, x
            integer(c_int)                  :: rc
        end function

        function pepsmg_fill_rhs_host(nx, ny, nz, out, seed, scale_rand, scale_sin, xg, yg, zg) &
                bind(C, name="pepsmg_fill_rhs_host") result(rc)
            import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
            integer(c_int32_t), value :: nx, ny, nz
            type(c_ptr), value        :: out, xg, yg, zg
            integer(c_int64_t), value :: seed
            real(c_double), value     :: scale_rand, scale_sin
            integer(c_int)            :: rc
        end function

        function pepsmg_vv_dot(nx, ny, nz, x, y, res) bind(C, name="pepsmg_vv_dot") result(rc)
            import :: c_int, c_int32_t, c_double, c_ptr
            integer(c_int32_t), value   :: nx, ny, nz
            type(c_ptr), value          :: x, y
            real(c_double), intent(out) :: res
            integer(c_int)              :: rc
        end function
    end interface

contains

    !> Abort with the library's message when rc /= PEPSMG_OK (the reference
    !> prints and stops on every error condition, src/multigrid.f90:101-124).
    subroutine pepsmg_check(rc, where)
        integer(c_int), intent(in)   :: rc
        character(*), intent(in)     :: where
        character(kind=c_char), pointer :: msg(:)
        character(len=512) :: text
        integer :: i
        if (rc == PEPSMG_OK) return
        text = ''
        call c_f_pointer(pepsmg_last_error(), msg, [512])
        do i = 1, 512
            if (msg(i) == c_null_char) exit
            text(i:i) = msg(i)
        enddo
        print '(a,a,a,i3,a,a)', '[Error] ', where, ' failed (', rc, '): ', trim(text)
        stop 301
    end subroutine pepsmg_check

end module pepsmg_c
