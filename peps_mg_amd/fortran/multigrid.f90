!> Geometric multigrid V-cycle (module multigrid, src/multigrid.f90) on the
!> GPU through libpepsmg.so.  Like the reference the module holds one
!> hierarchy (src/multigrid.f90:40-55); here it is the library's context.
module multigrid

    use, intrinsic :: iso_c_binding
    use pepsmg_c
    use pepsmg_shim
    use geometry, only : subdomain
    use matrix, only : matrix_heptadiagonal
    implicit none

    private

    type(c_ptr), save            :: mg_ctx = c_null_ptr
    integer(kind=4), save        :: n_vcycles = 0
    type(pepsmg_options_t), save :: mg_opt
    logical, save                :: mg_opt_set = .false.
    type(pepsmg_stats_t), save, public :: mg_last_stats   !< statistics of the last solve

    public :: multigrid_create
    public :: multigrid_solve_vcycle
    public :: multigrid_destroy
    public :: multigrid_set_options

contains

    !> multigrid_create (src/multigrid.f90:75-159): the library builds the
    !> level hierarchy and applies the reference's checks on nlevel,
    !> aggregation method and levels (it stops with the message on error).
    subroutine multigrid_create(sdm, nlevel, ncycle, aggr_method, single_aggr_level, adaptive_aggr_level)
        type(subdomain), intent(in), target :: sdm
        integer(kind=4), intent(in)         :: nlevel, ncycle, aggr_method, single_aggr_level
        integer(kind=4), intent(in)         :: adaptive_aggr_level(3)
        type(pepsmg_grid_t) :: g
        if (c_associated(mg_ctx)) call multigrid_destroy()
        g = pepsmg_grid_of(sdm)
        call pepsmg_check(pepsmg_multigrid_create(g, nlevel, ncycle, aggr_method, single_aggr_level, &
                          adaptive_aggr_level, mg_ctx), 'multigrid_create')
        n_vcycles = ncycle
        if (mg_opt_set) call pepsmg_check(pepsmg_set_options(mg_ctx, mg_opt), 'multigrid_set_options')
        print '(a)', '[MG] Multigrid geometry constructed.'
    end subroutine multigrid_create

    !> Options with no counterpart in the reference (smoother, precision,
    !> persistent coarse-level switch, refinement schedule, verbosity; the
    !> FP32 inner V-cycles' pre- and post-smoothing steps and SOR factor if
    !> given); they apply to the current hierarchy and to hierarchies
    !> created later.
    subroutine multigrid_set_options(smoother, precision, coarse_switch, inner_cycles, max_refine, verbose, &
                                     inner_nu, inner_nu_post, inner_omega)
        integer(kind=4), intent(in) :: smoother, precision, coarse_switch, inner_cycles, max_refine, verbose
        integer(kind=4), intent(in), optional :: inner_nu, inner_nu_post
        real(kind=8), intent(in), optional :: inner_omega
        if (.not. mg_opt_set) call pepsmg_default_options(mg_opt)
        mg_opt%smoother = smoother
        mg_opt%precision = precision
        mg_opt%coarse_switch = coarse_switch
        mg_opt%inner_cycles = inner_cycles
        mg_opt%max_refine = max_refine
        mg_opt%verbose = verbose
        if (present(inner_nu)) mg_opt%inner_nu = inner_nu
        if (present(inner_nu_post)) mg_opt%inner_nu_post = inner_nu_post
        if (present(inner_omega)) mg_opt%inner_omega = inner_omega
        mg_opt_set = .true.
        if (c_associated(mg_ctx)) call pepsmg_check(pepsmg_set_options(mg_ctx, mg_opt), 'multigrid_set_options')
    end subroutine multigrid_set_options

    !> multigrid_solve_vcycle (src/multigrid.f90:1864-1894): V-cycles until
    !> sqrt(|r|^2/|r0|^2) < tolerance or number_of_vcycles; sol returns the
    !> solution, rsd the final residual.  Prints the reference's per-cycle line.
    subroutine multigrid_solve_vcycle(sol, rsd, a_poisson, rhs, sdm, maxiteration, tolerance, omega_sor)
        use mpi_topology, only : myrank
        real(kind=8), intent(inout), target            :: sol(0:,0:,0:), rsd(0:,0:,0:)
        type(matrix_heptadiagonal), intent(in), target :: a_poisson
        real(kind=8), intent(in), target               :: rhs(0:,0:,0:)
        type(subdomain), intent(in), target            :: sdm
        integer(kind=4), intent(in)                    :: maxiteration
        real(kind=8), intent(in)                       :: tolerance, omega_sor
        real(c_double), allocatable, target   :: hist(:)
        integer(c_int32_t), allocatable, target :: cit(:)
        integer :: c, cap, n
        if (.not. c_associated(mg_ctx)) then
            print '(a)', '[Error] multigrid_create has not been called'
            stop 301
        endif
        call pepsmg_check_padded(sol, sdm%nx, sdm%ny, sdm%nz, 'sol', 'multigrid_solve_vcycle')
        call pepsmg_check_padded(rsd, sdm%nx, sdm%ny, sdm%nz, 'rsd', 'multigrid_solve_vcycle')
        call pepsmg_check_padded(rhs, sdm%nx, sdm%ny, sdm%nz, 'rhs', 'multigrid_solve_vcycle')
        call pepsmg_check_coeff(a_poisson, sdm%nx, sdm%ny, sdm%nz, 'multigrid_solve_vcycle')
        cap = max(n_vcycles, 1) + 64
        allocate(hist(cap), cit(cap))
        mg_last_stats%history_cap = cap
        mg_last_stats%history = c_loc(hist(1))
        mg_last_stats%coarse_iters = c_loc(cit(1))
        call pepsmg_check(pepsmg_multigrid_solve_vcycle(mg_ctx, pepsmg_ptr3(sol), pepsmg_ptr3(rsd), &
                          pepsmg_coeff_ptr(a_poisson), pepsmg_ptr3(rhs), maxiteration, tolerance, omega_sor, &
                          mg_last_stats), 'multigrid_solve_vcycle')
        if (myrank == 0) then
            if (mg_last_stats%refine_steps > 0) then
                n = mg_last_stats%refine_steps
                do c = 1, n
                    print '(a,i4,a,3(e15.7,x))', '[MG] refinement step = ', c, ', Error:', &
                        hist(c) * sqrt(mg_last_stats%res0), sqrt(mg_last_stats%res0), hist(c)
                enddo
            else
                n = mg_last_stats%cycles
                do c = 1, n
                    print '(a,i4,a,3(e15.7,x))', '[MG] cycle = ', c, ', Error:', &
                        hist(c) * sqrt(mg_last_stats%res0), sqrt(mg_last_stats%res0), hist(c)
                enddo
            endif
            ! The reference stops at number_of_vcycles without a word
            ! (src/multigrid.f90:1928-1974); its own sample input caps the
            ! solve at 3 cycles on purpose, so this is reported, not fatal.
            if (mg_last_stats%converged == 0) print '(a,i5,a,e15.7,a,e15.7)', &
                '[MG] Not converged: stopped at the cycle cap after ', mg_last_stats%cycles, &
                ' V-cycles, rho = ', mg_last_stats%rho, ' >= tolerance = ', tolerance
        endif
        mg_last_stats%history = c_null_ptr
        mg_last_stats%coarse_iters = c_null_ptr
        deallocate(hist, cit)
    end subroutine multigrid_solve_vcycle

    !> multigrid_destroy (src/multigrid.f90:1320-1334)
    subroutine multigrid_destroy()
        if (c_associated(mg_ctx)) call pepsmg_multigrid_destroy(mg_ctx)
        mg_ctx = c_null_ptr
    end subroutine multigrid_destroy

end module multigrid
