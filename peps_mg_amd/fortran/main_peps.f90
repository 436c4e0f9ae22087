!> Driver of the Poisson solve, reading the reference's PARA_INPUT.inp
!> (src/main_peps.f90:44-60: namelists meshes, origin, length, mesh_stretch,
!> procs, control, coefficients) plus one optional namelist of this build:
!>
!>   &gpu smoother = 0|1 (RBGS | CG), precision = 0|1 (FP64 | FP32 V-cycles in
!>        FP64 refinement), coarse_switch, inner_cycles, max_refine, verbose,
!>        inner_nu, inner_nu_post, inner_omega (the FP32 V-cycles' pre- and
!>        post-smoothing steps and SOR factor), rhs_mode = 0|1, seed,
!>        output_file, rhs_file, write_solution /
!>
!> rhs_mode 0 is the reference's "Problem 1" cosine right-hand side with its
!> RMS check (src/main_peps.f90:78-90, :172-182); rhs_mode 1 is the synthetic
!> workload b = -(U[-1,1] + 3 pi^2 sin(pi x) sin(pi y) sin(pi z)) with U the
!> SplitMix64 field of the given seed.  output_file, if set, receives the
!> padded solution array as raw float64 (stream access), rhs_file the padded
!> right-hand side the same way.  write_solution (default .false.) writes the
!> reference's seven solution files (file_write, src/main_peps.f90:204-446,
!> on one process): solution_all.dat, solution_all_gatherv.dat and
!> solution_all_gatherv_halo.dat (x, y, z, u per interior cell, format
!> (3D20.10,1E30.20)), solution_all.bin, solution_all_gatherv.bin and
!> solution_all_gatherv_halo.bin (one unformatted record of
!> u(1:nx,1:ny,1:nz)) and solution_all_MPI_IO.bin (the same values as raw
!> float64), then checks that solution_all.bin and the MPI-IO file agree
!> (binary_file_check, :448-484).
program main_peps

    use, intrinsic :: iso_c_binding
    use mpi_topology
    use geometry
    use matrix
    use multigrid
    use pepsmg_c
    use pepsmg_shim

    implicit none

    real(kind=8), parameter :: PI = 3.1415926535897932384626d0

    integer(kind=4) :: npx, npy, npz
    integer(kind=4) :: nx, ny, nz
    real(kind=8)    :: ox, oy, oz, lx, ly, lz, ax, ay, az
    real(kind=8)    :: alpha_x, alpha_y, alpha_z
    integer(kind=4) :: maxiteration, number_of_vcycles, number_of_levels, aggregation_method
    integer(kind=4) :: CGA_level, CGPSA_level(3)
    real(kind=8)    :: tolerance, omega_sor
    integer(kind=4) :: smoother, precision, coarse_switch, inner_cycles, max_refine, verbose, rhs_mode
    integer(kind=4) :: inner_nu, inner_nu_post
    real(kind=8)    :: inner_omega
    integer(kind=8) :: seed
    character(len=256) :: output_file, rhs_file
    logical :: write_solution

    type(domain), target     :: g_domain
    type(subdomain), target  :: s_domain
    type(matrix_heptadiagonal), target :: a_poisson
    real(kind=8), allocatable :: ref_sub(:,:,:)
    real(kind=8)    :: rms
    integer(kind=4) :: i, j, k, ios, u
    integer(kind=8) :: t0, t1, rate

    namelist /meshes/ nx, ny, nz
    namelist /origin/ ox, oy, oz
    namelist /length/ lx, ly, lz
    namelist /mesh_stretch/ ax, ay, az
    namelist /procs/ npx, npy, npz
    namelist /control/ maxiteration, tolerance, number_of_vcycles, number_of_levels, aggregation_method, &
                       CGA_level, CGPSA_level, omega_sor
    namelist /coefficients/ alpha_x, alpha_y, alpha_z
    namelist /gpu/ smoother, precision, coarse_switch, inner_cycles, max_refine, verbose, rhs_mode, seed, output_file, &
                   write_solution, inner_nu, inner_nu_post, inner_omega, rhs_file

    smoother = 0; precision = 0; coarse_switch = 4096; inner_cycles = 2; max_refine = 30; verbose = 0
    inner_nu = 2; inner_nu_post = 0; inner_omega = 1.35d0
    rhs_mode = 0; seed = 20210501_8; output_file = ''; rhs_file = ''; write_solution = .false.
    alpha_x = 1.0d0; alpha_y = 1.0d0; alpha_z = 1.0d0

    open(unit=1, file="PARA_INPUT.inp", status='old')
    read(1, meshes)
    read(1, origin)
    read(1, length)
    read(1, mesh_stretch)
    read(1, procs)
    read(1, control)
    read(1, coefficients, iostat=ios)
    rewind(1)
    read(1, gpu, iostat=ios)
    close(1)

    np_dim(0) = npx; np_dim(1) = npy; np_dim(2) = npz
    period(0:2) = .false.
    call mpi_topology_create()

    call geometry_domain_create(g_domain, nx, ny, nz, ox, oy, oz, lx, ly, lz, ax, ay, az, period)
    print '(a)', '[Poisson] Geometry and matrix size initialized.'
    call geometry_subdomain_create(s_domain, g_domain)
    call geometry_subdomain_ddt_create(s_domain)

    allocate(ref_sub(0:s_domain%nx+1, 0:s_domain%ny+1, 0:s_domain%nz+1))
    ref_sub = 0.0d0
    if (rhs_mode == 0) then
        ! Problem 1: cosine over -l/2 <= x,y,z <= l/2 (src/main_peps.f90:78-90)
        do k = 1, s_domain%nz
            do j = 1, s_domain%ny
                do i = 1, s_domain%nx
                    s_domain%x(i, j, k) = 0.0d0
                    s_domain%b(i, j, k) = -cos(s_domain%xg(i)/lx*PI) * cos(s_domain%yg(j)/ly*PI) &
                                          * cos(s_domain%zg(k)/lz*PI) * PI*PI*(1.0d0/(lx*lx)+1.0d0/(ly*ly)+1.0d0/(lz*lz))
                    ref_sub(i, j, k) = cos(s_domain%xg(i)/lx*PI) * cos(s_domain%yg(j)/ly*PI) * cos(s_domain%zg(k)/lz*PI)
                enddo
            enddo
        enddo
    else
        call pepsmg_check(pepsmg_fill_rhs_host(nx, ny, nz, pepsmg_ptr3(s_domain%b), seed, -1.0d0, -3.0d0*PI*PI, &
                          c_loc(s_domain%xg(0)), c_loc(s_domain%yg(0)), c_loc(s_domain%zg(0))), 'fill_rhs')
    endif
    print '(a)', '[Poisson] Geometry and rhs constructed.'
    if (len_trim(rhs_file) > 0) then
        open(newunit=u, file=trim(rhs_file), access='stream', form='unformatted', status='replace')
        write(u) s_domain%b
        close(u)
    endif

    call matrix_heptadiagonal_create(a_poisson, s_domain)
    print '(a)', '[Poisson] Poisson matrix constructed.'
    print '(a)', '[Poisson] Start solving equations.'

    call multigrid_set_options(smoother, precision, coarse_switch, inner_cycles, max_refine, verbose, &
                               inner_nu, inner_nu_post, inner_omega)
    call system_clock(t0, rate)
    call multigrid_create(s_domain, number_of_levels, number_of_vcycles, aggregation_method, CGA_level, CGPSA_level)
    call multigrid_solve_vcycle(s_domain%x, s_domain%r, a_poisson, s_domain%b, s_domain, maxiteration, &
                                tolerance, omega_sor)
    call multigrid_destroy
    call system_clock(t1)

    if (rhs_mode == 0) then
        rms = 0.0d0
        do k = 1, s_domain%nz
            do j = 1, s_domain%ny
                do i = 1, s_domain%nx
                    rms = rms + (s_domain%x(i,j,k) - ref_sub(i,j,k))**2
                enddo
            enddo
        enddo
        print '(a,e20.10)', '[Poisson] RMS = ', rms/g_domain%nx/g_domain%ny/g_domain%nz
    endif
    print '(a,i4,a,i3,a,e15.7,a,f12.6,a)', '[Poisson] V-cycles = ', mg_last_stats%cycles, &
        ', converged = ', mg_last_stats%converged, ', rho = ', mg_last_stats%rho, &
        ', device time = ', mg_last_stats%device_ms * 1.0d-3, ' s'
    print '(a,f12.6)', '[Poisson] Solution obtained. Execution time = ', dble(t1 - t0) / dble(rate)

    if (len_trim(output_file) > 0) then
        open(newunit=u, file=trim(output_file), access='stream', form='unformatted', status='replace')
        write(u) s_domain%x
        close(u)
        print '(a,a)', '[Poisson] Solution written to ', trim(output_file)
    endif

    if (write_solution) then
        call file_write
        call binary_file_check
        print '(a)', '[Poisson] Solution printed.'
    endif

    deallocate(ref_sub)
    call matrix_heptadiagonal_destroy(a_poisson)
    call mpi_topology_destroy
    call geometry_subdomain_ddt_destroy(s_domain)
    call geometry_subdomain_destroy(s_domain)
    call geometry_domain_destroy(g_domain)
    print '(a)', '[Poisson] Memory deallocated.'

contains

    !> The reference's seven solution files for one process
    !> (src/main_peps.f90:204-446).  Its point-to-point, gatherv and
    !> gatherv-with-halo collections all reduce to the subdomain's interior on
    !> one process, so each formatted file holds the same lines and each
    !> unformatted file the same record (the reference's
    !> solution_all_gatherv_halo.bin also writes the gatherv array, :397).
    subroutine file_write
        integer(kind=4) :: uf
        call write_formatted("solution_all.dat")              ! :304-313
        call write_record("solution_all.bin")                 ! :315-317
        call write_formatted("solution_all_gatherv.dat")      ! :346-355
        call write_record("solution_all_gatherv.bin")         ! :357-359
        call write_formatted("solution_all_gatherv_halo.dat") ! :385-394
        call write_record("solution_all_gatherv_halo.bin")    ! :396-398
        open(newunit=uf, form="unformatted", access='stream', file="solution_all_MPI_IO.bin", status='replace')
        write(uf) s_domain%x(1:s_domain%nx, 1:s_domain%ny, 1:s_domain%nz)   ! :413-419
        close(uf)
    end subroutine file_write

    !> One line (x, y, z, u) per interior cell, x fastest, format (3D20.10,1E30.20).
    subroutine write_formatted(fname)
        character(*), intent(in) :: fname
        integer(kind=4) :: i, j, k, uf
        open(newunit=uf, form="formatted", file=fname, status='replace')
        do k = 1, g_domain%nz
            do j = 1, g_domain%ny
                do i = 1, g_domain%nx
                    write(uf, '(3D20.10,1E30.20)') g_domain%xg(i), g_domain%yg(j), g_domain%zg(k), s_domain%x(i, j, k)
                enddo
            enddo
        enddo
        close(uf)
    end subroutine write_formatted

    !> One unformatted sequential record of u(1:nx,1:ny,1:nz).
    subroutine write_record(fname)
        character(*), intent(in) :: fname
        integer(kind=4) :: uf
        open(newunit=uf, form="unformatted", file=fname, status='replace')
        write(uf) s_domain%x(1:s_domain%nx, 1:s_domain%ny, 1:s_domain%nz)
        close(uf)
    end subroutine write_record

    !> Reads both binary files back and prints every cell where they differ
    !> (src/main_peps.f90:448-484).
    subroutine binary_file_check
        real(kind=8), allocatable :: temp1(:,:,:), temp2(:,:,:)
        integer(kind=4) :: i, j, k, uf
        allocate(temp1(1:g_domain%nx, 1:g_domain%ny, 1:g_domain%nz))
        allocate(temp2(1:g_domain%nx, 1:g_domain%ny, 1:g_domain%nz))
        open(newunit=uf, form='unformatted', file="solution_all.bin", status='old')
        read(uf) temp1
        close(uf)
        open(newunit=uf, form='unformatted', access='stream', file="solution_all_MPI_IO.bin", status='old')
        read(uf) temp2
        close(uf)
        do k = 1, g_domain%nz
            do j = 1, g_domain%ny
                do i = 1, g_domain%nx
                    if (temp1(i, j, k) /= temp2(i, j, k)) print *, i, j, k, temp1(i, j, k), temp2(i, j, k)
                enddo
            enddo
        enddo
        deallocate(temp1, temp2)
    end subroutine binary_file_check

end program main_peps
