!> Domain and subdomain types and grid metrics (module geometry,
!> src/geometry.f90), for one process owning the whole grid.  The metric
!> formulas are the reference's (they define the discrete problem); ghost
!> exchange is not needed because the ghost layer holds the homogeneous
!> Dirichlet value 0 and is never written.
module geometry

    use mpi_topology, only : comm_1d_x, comm_1d_y, comm_1d_z
    implicit none

    !> type domain (src/geometry.f90:30-53)
    type, public :: domain
        integer(kind=4) :: nx, ny, nz
        real(kind=8)    :: lx, ly, lz
        real(kind=8)    :: ox, oy, oz
        real(kind=8), allocatable, dimension(:) :: dxm, dym, dzm
        real(kind=8), allocatable, dimension(:) :: dxg, dyg, dzg
        real(kind=8), allocatable, dimension(:) :: xg, yg, zg
        logical :: is_periodic(0:2)
    end type domain

    !> type subdomain (src/geometry.f90:59-118)
    type, public :: subdomain
        integer(kind=4) :: nx, ny, nz
        real(kind=8)    :: lx, ly, lz
        real(kind=8)    :: ox, oy, oz
        real(kind=8), allocatable, dimension(:) :: dxm, dym, dzm
        real(kind=8), allocatable, dimension(:) :: dxg, dyg, dzg
        real(kind=8), allocatable, dimension(:) :: xg, yg, zg
        logical :: is_periodic(0:2)
        real(kind=8), allocatable, dimension(:,:,:) :: x
        real(kind=8), allocatable, dimension(:,:,:) :: b
        real(kind=8), allocatable, dimension(:,:,:) :: r
        logical :: is_aggregated(0:2)
        integer(kind=4), public :: ista, iend
        integer(kind=4), public :: jsta, jend
        integer(kind=4), public :: ksta, kend
        integer(kind=4), public :: ddt_yz_plane_x0, ddt_yz_plane_x1, ddt_yz_plane_xn, ddt_yz_plane_xn1
        integer(kind=4), public :: ddt_xz_plane_y0, ddt_xz_plane_y1, ddt_xz_plane_yn, ddt_xz_plane_yn1
        integer(kind=4), public :: ddt_xy_plane_z0, ddt_xy_plane_z1, ddt_xy_plane_zn, ddt_xy_plane_zn1
        integer(kind=4), public :: ddt_inner_domain
        logical :: is_x0_boundary, is_x1_boundary
        logical :: is_y0_boundary, is_y1_boundary
        logical :: is_z0_boundary, is_z1_boundary
    end type subdomain

contains

    !> One axis of geometry_domain_create (src/geometry.f90:176-209): uniform
    !> (a = 1) or symmetric geometric stretch with ratio a.
    subroutine geometry_axis(n, o, l, a, dm, dg, g)
        integer(kind=4), intent(in) :: n
        real(kind=8), intent(in)    :: o, l, a
        real(kind=8), intent(out)   :: dm(0:), dg(0:), g(0:)
        integer(kind=4) :: i
        dm(0) = 0.0d0
        dg(0) = 0.0d0
        g(0) = o
        if (a == 1.0d0) then
            do i = 1, n
                dm(i) = l / dble(n)
            enddo
        else
            dm(1) = 0.5d0 * l * (a - 1.0d0) / (a**(n/2) - 1.0d0)
            do i = 2, n/2
                dm(i) = dm(i-1) * a
                dm(n-i+1) = dm(i)
            enddo
            dm(n) = dm(1)
        endif
        dm(n+1) = 0.0d0
        do i = 1, n
            dg(i) = 0.5d0 * (dm(i) + dm(i-1))
            g(i) = g(i-1) + dg(i)
        enddo
        dg(n+1) = 0.5d0 * (dm(n+1) + dm(n))
        g(n+1) = g(n) + dg(n+1)
    end subroutine geometry_axis

    !> geometry_domain_create (src/geometry.f90:139-281)
    subroutine geometry_domain_create(gdm, nx, ny, nz, ox, oy, oz, lx, ly, lz, ax, ay, az, period)
        type(domain), intent(inout)  :: gdm
        integer(kind=4), intent(in)  :: nx, ny, nz
        real(kind=8), intent(in)     :: ox, oy, oz, lx, ly, lz, ax, ay, az
        logical, intent(in)          :: period(0:2)
        gdm%nx = nx; gdm%ny = ny; gdm%nz = nz
        gdm%ox = ox; gdm%oy = oy; gdm%oz = oz
        gdm%lx = lx; gdm%ly = ly; gdm%lz = lz
        gdm%is_periodic(0:2) = period(0:2)
        allocate(gdm%dxm(0:nx+1), gdm%dxg(0:nx+1), gdm%xg(0:nx+1))
        allocate(gdm%dym(0:ny+1), gdm%dyg(0:ny+1), gdm%yg(0:ny+1))
        allocate(gdm%dzm(0:nz+1), gdm%dzg(0:nz+1), gdm%zg(0:nz+1))
        call geometry_axis(nx, ox, lx, ax, gdm%dxm, gdm%dxg, gdm%xg)
        call geometry_axis(ny, oy, ly, ay, gdm%dym, gdm%dyg, gdm%yg)
        call geometry_axis(nz, oz, lz, az, gdm%dzm, gdm%dzg, gdm%zg)
    end subroutine geometry_domain_create

    !> geometry_domain_destroy (src/geometry.f90:287-297)
    subroutine geometry_domain_destroy(gdm)
        type(domain), intent(inout) :: gdm
        deallocate(gdm%dxm, gdm%dym, gdm%dzm, gdm%dxg, gdm%dyg, gdm%dzg, gdm%xg, gdm%yg, gdm%zg)
    end subroutine geometry_domain_destroy

    !> geometry_subdomain_create (src/geometry.f90:304-423) on one process:
    !> the subdomain is the whole domain, aggregated in every direction.
    subroutine geometry_subdomain_create(sdm, gdm)
        type(subdomain), intent(inout) :: sdm
        type(domain), intent(in)       :: gdm
        sdm%is_periodic(0:2) = gdm%is_periodic(0:2)
        sdm%is_aggregated(0:2) = .true.
        sdm%ista = 1; sdm%iend = gdm%nx
        sdm%jsta = 1; sdm%jend = gdm%ny
        sdm%ksta = 1; sdm%kend = gdm%nz
        sdm%nx = gdm%nx; sdm%ny = gdm%ny; sdm%nz = gdm%nz
        sdm%lx = gdm%lx; sdm%ly = gdm%ly; sdm%lz = gdm%lz
        allocate(sdm%dxm(0:sdm%nx+1), sdm%dxg(0:sdm%nx+1), sdm%xg(0:sdm%nx+1))
        allocate(sdm%dym(0:sdm%ny+1), sdm%dyg(0:sdm%ny+1), sdm%yg(0:sdm%ny+1))
        allocate(sdm%dzm(0:sdm%nz+1), sdm%dzg(0:sdm%nz+1), sdm%zg(0:sdm%nz+1))
        sdm%dxm = gdm%dxm; sdm%dxg = gdm%dxg; sdm%xg = gdm%xg
        sdm%dym = gdm%dym; sdm%dyg = gdm%dyg; sdm%yg = gdm%yg
        sdm%dzm = gdm%dzm; sdm%dzg = gdm%dzg; sdm%zg = gdm%zg
        sdm%ox = sdm%xg(1) - 0.5d0 * sdm%dxm(1)
        sdm%oy = sdm%yg(1) - 0.5d0 * sdm%dym(1)
        sdm%oz = sdm%zg(1) - 0.5d0 * sdm%dzm(1)
        sdm%is_x0_boundary = .true.; sdm%is_x1_boundary = .true.
        sdm%is_y0_boundary = .true.; sdm%is_y1_boundary = .true.
        sdm%is_z0_boundary = .true.; sdm%is_z1_boundary = .true.
        allocate(sdm%x(0:sdm%nx+1, 0:sdm%ny+1, 0:sdm%nz+1))
        allocate(sdm%b(0:sdm%nx+1, 0:sdm%ny+1, 0:sdm%nz+1))
        allocate(sdm%r(0:sdm%nx+1, 0:sdm%ny+1, 0:sdm%nz+1))
        sdm%x = 0.0d0
        sdm%b = 0.0d0
        sdm%r = 0.0d0
        comm_1d_x%nprocs = 1; comm_1d_y%nprocs = 1; comm_1d_z%nprocs = 1
    end subroutine geometry_subdomain_create

    !> geometry_subdomain_destroy (src/geometry.f90)
    subroutine geometry_subdomain_destroy(sdm)
        type(subdomain), intent(inout) :: sdm
        if (allocated(sdm%dxm)) deallocate(sdm%dxm, sdm%dym, sdm%dzm, sdm%dxg, sdm%dyg, sdm%dzg)
        if (allocated(sdm%xg)) deallocate(sdm%xg, sdm%yg, sdm%zg)
        if (allocated(sdm%x)) deallocate(sdm%x)
        if (allocated(sdm%b)) deallocate(sdm%b)
        if (allocated(sdm%r)) deallocate(sdm%r)
    end subroutine geometry_subdomain_destroy

    !> geometry_subdomain_ddt_create (src/geometry.f90:445-541): no derived
    !> datatypes are needed on one process.
    subroutine geometry_subdomain_ddt_create(sdm)
        type(subdomain), intent(inout) :: sdm
        sdm%ddt_inner_domain = 0
    end subroutine geometry_subdomain_ddt_create

    subroutine geometry_subdomain_ddt_destroy(sdm)
        type(subdomain), intent(inout) :: sdm
        sdm%ddt_inner_domain = 0
    end subroutine geometry_subdomain_ddt_destroy

    !> geometry_halocell_update_selectively (src/geometry.f90:621-660): every
    !> direction is aggregated on one process, so there is nothing to exchange.
    subroutine geometry_halocell_update_selectively(val, sdm, is_serial)
        real(kind=8), intent(inout) :: val(0:,0:,0:)
        type(subdomain), intent(in) :: sdm
        logical, intent(in)         :: is_serial(0:2)
    end subroutine geometry_halocell_update_selectively

end module geometry
