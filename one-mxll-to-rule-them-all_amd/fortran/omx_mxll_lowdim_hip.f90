! omx_mxll_lowdim_hip.f90 -- the 2D and 1D GPU solvers made by
! maxwell_factory(2) and maxwell_factory(1).
!
! omx_mxll_lowdim_mod declares TMxll_2D and TMxll_1D, the shapes of the
! reference's 2D and 1D solver types (src/FDTD/mxll_2D_mod.f90:9-50,
! mxll_1D_mod.f90:9-46) as callers see them through `class is (TMxll_2D)`:
! grid size, boundaries, 2D mode, and host mirrors of the fields with the
! reference's bounds (2D: Ex(0:nx,ny+1), Ey(nx+1,0:ny), Hz(0:nx,0:ny) for TEz,
! Ez(nx+1,ny+1), Hx(nx,0:ny), Hy(0:nx,ny) for TMz, both for FULL; 1D: Ex(nz),
! Hy(nz-1)), filled by sync_to_host and written back by sync_to_device.
!
! omx_mxll_lowdim_hip_mod implements them over the engine's omxl_* entry
! points (FP64).  init keeps the reference's signature
! (mxll_base_mod.f90:36-51) and reads the reference's medium_XXXX.in files
! with "x y" (2D) or "x" (1D) points (classical_medium_mod.f90:203-337,
! 341-483); the 2D mode comes in init's `mode` argument (TMZ_2D_MODE,
! TEZ_2D_MODE or FULL_2D_MODE, input_mod.f90:163-177).  Every engine status
! is checked; a failure stops with the engine's message.
module omx_mxll_lowdim_mod

    use omx_units_mod
    use omx_mxll_base_mod
    implicit none

    integer, parameter :: TMZ_2D_MODE = 1, TEZ_2D_MODE = 2, FULL_2D_MODE = 3

    type, abstract, extends(TMxll) :: TMxll_2D
        integer :: nx = 0, ny = 0
        integer :: boundaries(2) = 0
        real(dp), allocatable :: Ex(:, :), Ey(:, :), Ez(:, :)
        real(dp), allocatable :: Hx(:, :), Hy(:, :), Hz(:, :)
    contains
        procedure(sync2_interface), deferred :: sync_to_host
        procedure(sync2_interface), deferred :: sync_to_device
        procedure :: alloc_mirrors => alloc_mirrors_2D
        procedure :: free_mirrors => free_mirrors_2D
    end type TMxll_2D

    type, abstract, extends(TMxll) :: TMxll_1D
        integer :: nz = 0
        integer :: boundaries = 0
        real(dp), allocatable :: Ex(:), Hy(:)
    contains
        procedure(sync1_interface), deferred :: sync_to_host
        procedure(sync1_interface), deferred :: sync_to_device
    end type TMxll_1D

    abstract interface
        subroutine sync2_interface(this)
            import :: TMxll_2D
            class(TMxll_2D), intent(inout) :: this
        end subroutine sync2_interface
        subroutine sync1_interface(this)
            import :: TMxll_1D
            class(TMxll_1D), intent(inout) :: this
        end subroutine sync1_interface
    end interface

contains

    subroutine alloc_mirrors_2D(this)
        class(TMxll_2D), intent(inout) :: this
        integer :: nx, ny
        nx = this%nx; ny = this%ny
        if (this%mode /= TMZ_2D_MODE .and. .not. allocated(this%Ex)) then
            allocate(this%Ex(0:nx, ny + 1), this%Ey(nx + 1, 0:ny), this%Hz(0:nx, 0:ny))
            this%Ex = 0.0d0; this%Ey = 0.0d0; this%Hz = 0.0d0
        end if
        if (this%mode /= TEZ_2D_MODE .and. .not. allocated(this%Ez)) then
            allocate(this%Ez(nx + 1, ny + 1), this%Hx(nx, 0:ny), this%Hy(0:nx, ny))
            this%Ez = 0.0d0; this%Hx = 0.0d0; this%Hy = 0.0d0
        end if
    end subroutine alloc_mirrors_2D

    subroutine free_mirrors_2D(this)
        class(TMxll_2D), intent(inout) :: this
        if (allocated(this%Ex)) deallocate(this%Ex, this%Ey, this%Hz)
        if (allocated(this%Ez)) deallocate(this%Ez, this%Hx, this%Hy)
    end subroutine free_mirrors_2D

end module omx_mxll_lowdim_mod


module omx_mxll_lowdim_hip_mod

    use iso_c_binding
    use omx_units_mod
    use omx_mxll_base_mod
    use omx_mxll_lowdim_mod
    use omx_fdtd_cabi
    implicit none

    type, extends(TMxll_2D) :: TMxll_2D_hip
        type(c_ptr) :: ctx = c_null_ptr
        integer :: n_samples = 0
    contains
        procedure :: init => init_2D_hip
        procedure :: kill => kill_2D_hip
        procedure :: td_propagate_H_field => td_propagate_H_2D_hip
        procedure :: td_propagate_E_field => td_propagate_E_2D_hip
        procedure :: add_source => add_source_2D_hip
        procedure :: inject_sources => inject_sources_2D_hip
        procedure :: add_detector => add_detector_2D_hip
        procedure :: sample_detectors => sample_2D_hip
        procedure :: sync_to_host => sync_to_host_2D_hip
        procedure :: sync_to_device => sync_to_device_2D_hip
        procedure :: synchronize => synchronize_2D_hip
    end type TMxll_2D_hip

    type, extends(TMxll_1D) :: TMxll_1D_hip
        type(c_ptr) :: ctx = c_null_ptr
        integer :: n_samples = 0
    contains
        procedure :: init => init_1D_hip
        procedure :: kill => kill_1D_hip
        procedure :: td_propagate_H_field => td_propagate_H_1D_hip
        procedure :: td_propagate_E_field => td_propagate_E_1D_hip
        procedure :: add_source => add_source_1D_hip
        procedure :: inject_sources => inject_sources_1D_hip
        procedure :: add_detector => add_detector_1D_hip
        procedure :: sample_detectors => sample_1D_hip
        procedure :: sync_to_host => sync_to_host_1D_hip
        procedure :: sync_to_device => sync_to_device_1D_hip
        procedure :: synchronize => synchronize_1D_hip
    end type TMxll_1D_hip

contains

    ! medium_XXXX.in with dims coordinates per point (nm) to EOF; a point marks
    ! cell int(x/dr) + int(N/2) per axis.  1D: a dielectric of eps_r = 1 leaves
    ! its cells unclaimed (classical_medium_mod.f90:259-268).
    subroutine read_media_lowdim(dims, n_media, grid, dr, table, aux)
        integer,  intent(in) :: dims, n_media, grid(2)
        real(dp), intent(in) :: dr
        type(omx_medium), allocatable, intent(out) :: table(:)
        integer(c_int8_t), allocatable, intent(out) :: aux(:, :)
        character(len=20) :: name, kind_ch
        character(len=10) :: material
        integer :: n, funit, ierr, a, ij(2)
        real(dp) :: x(2), eps_read, omegaD, GammaD
        logical :: claims

        allocate(table(n_media), aux(grid(1), grid(2)))
        aux = 0_c_int8_t
        do n = 1, n_media
            write (name, '("medium_", I4.4, ".in")') n
            open (newunit=funit, file=trim(name), action="read", iostat=ierr)
            if (ierr /= 0) then
                write (*, '("Error: medium file ",A," could not be opened")') trim(name)
                error stop 1
            end if
            read (funit, *, iostat=ierr) kind_ch
            if (ierr /= 0) then
                write (*, '("Error: medium file ",A," has invalid format")') trim(name)
                error stop 1
            end if
            claims = .true.
            select case (trim(kind_ch))
            case ("dielectric")
                read (funit, *) eps_read
                table(n)%kind = DIELECTRIC_MEDIUM
                table(n)%eps_r = eps_read
                if (dims == 1 .and. eps_read == 1.0d0) claims = .false.
            case ("drude")
                read (funit, *) omegaD, GammaD, eps_read
                table(n)%kind = DRUDE_MEDIUM
                table(n)%omega_d = omegaD * ev_to_au
                table(n)%gamma_d = GammaD * ev_to_au
                table(n)%eps_r = eps_read
            case ("drude-lorentz")
                read (funit, *) material, eps_read
                call omx_check(omx_medium_library(trim(material) // c_null_char, eps_read, table(n)), &
                               "omx_medium_library")
            case default
                write (*, '("Error: medium type ",A," not recognized in",A)') trim(kind_ch), trim(name)
                error stop 1
            end select
            do
                x = 0.0d0
                read (funit, *, iostat=ierr) x(1:dims)
                if (ierr /= 0) exit
                ij = 1
                do a = 1, dims
                    ij(a) = int(x(a) * nm_to_au / dr) + int(grid(a) / 2)
                    if (ij(a) < 1 .or. ij(a) > grid(a)) then
                        write (*, '("Error: medium ",I4.4," has a point outside the grid")') n
                        error stop 1
                    end if
                end do
                if (.not. claims) cycle
                if (aux(ij(1), ij(2)) /= 0) then
                    write (*, '("Error: medium ",I4.4," overlaps medium ",I4.4)') int(aux(ij(1), ij(2))), n
                else
                    aux(ij(1), ij(2)) = int(n, c_int8_t)
                end if
            end do
            close (funit)
        end do
    end subroutine read_media_lowdim

    ! compute_kernel and determine_indx_and_ranks, 1D and 2D
    ! (source_mod.f90:251-300, 336-346): Blackman-Harris window over x (1D,
    ! not clipped) or r (2D, clipped at 0), unit sum; kernel point i at
    ! i + int(r0/dr) + int(N/2) (1D: the x coordinate).
    subroutine lowdim_kernel(dims, dr, grid, r0, radius, n_ker, ind, ker)
        integer,  intent(in) :: dims, grid(2)
        real(dp), intent(in) :: dr, r0(3), radius
        integer,  intent(out) :: n_ker
        integer(c_int32_t), allocatable, intent(out) :: ind(:)
        real(c_double), allocatable, intent(out) :: ker(:)
        integer :: w, np, i, j, p, jlo
        real(dp) :: r_max, rc, r, norm, v

        n_ker = int(radius / dr)
        w = 2 * n_ker + 1
        np = w**dims
        allocate(ker(np), ind(dims * np))
        r_max = (n_ker * 2 + 2) * dr
        rc = r_max / 2.0d0
        norm = 0.0d0
        jlo = -n_ker
        if (dims == 1) jlo = 0
        do j = jlo, -jlo
        do i = -n_ker, n_ker
            p = (i + n_ker) + w * (j - jlo) + 1
            if (dims == 1) then
                r = i * dr
            else
                r = sqrt((i * dr)**2 + (j * dr)**2)
            end if
            v = a_bh(1) + a_bh(2) * cos(2.0d0 * pi0 * (r - rc) / r_max) + &
                a_bh(3) * cos(2.0d0 * pi0 * 2.0d0 * (r - rc) / r_max) + &
                a_bh(4) * cos(2.0d0 * pi0 * 3.0d0 * (r - rc) / r_max)
            if (dims == 2 .and. v < 0.0d0) v = 0.0d0
            ker(p) = v
            norm = norm + v
            ind(p) = i + int(r0(1) / dr) + int(grid(1) / 2)
            if (dims == 2) ind(p + np) = j + int(r0(2) / dr) + int(grid(2) / 2)
        end do
        end do
        ker = ker / norm
    end subroutine lowdim_kernel

    integer function dir_code(dir, dims)
        character(len=*), intent(in) :: dir
        integer, intent(in) :: dims
        select case (dir(1:1))
        case ('x'); dir_code = 1
        case ('y'); dir_code = 2
        case ('z'); dir_code = 3
        case default; dir_code = 0
        end select
        if (dir_code == 0 .or. (dims == 1 .and. dir_code /= 1)) then
            write (*, '("Error: Unknown source direction in ", I0, "D mxll.")') dims
            error stop 1
        end if
    end function dir_code

    ! ---- 2D -----------------------------------------------------------------
    subroutine init_2D_hip(this, grid_Ndims, npml, boundaries, dt, dr, mode, n_media, mpi_coords, mpi_dims)
        class(TMxll_2D_hip), intent(inout) :: this
        integer,  intent(in) :: grid_Ndims(3), npml, boundaries(3), mode, n_media
        integer,  intent(in) :: mpi_coords(3), mpi_dims(3)
        real(dp), intent(in) :: dt, dr
        type(omxl_desc) :: d
        type(omx_medium), allocatable :: table(:)
        integer(c_int8_t), allocatable :: aux(:, :)

        if (any(mpi_dims /= 1) .or. any(mpi_coords /= 0)) then
            write (*, '("Error: TMxll_2D_hip runs one domain per GPU (mpi_dims must be 1)")')
            error stop 1
        end if
        this%dimensions = 2
        this%chunk_coor = mpi_coords
        this%nx = grid_Ndims(1); this%ny = grid_Ndims(2)
        this%npml = npml
        this%boundaries = boundaries(1:2)
        this%mode = mode
        this%n_media = n_media
        this%dt = dt
        this%dr = dr
        this%dt_eps0 = dt / eps0
        this%dt_mu0 = dt / mu0
        this%time = 0.0d0
        this%t_skip = 0.0d0
        d%dims = 2
        d%mode = mode
        d%n1 = this%nx; d%n2 = this%ny
        d%npml = npml
        d%boundaries = boundaries(1:2)
        d%precision = 8
        d%dt = dt
        d%dr = dr
        call omx_check(omxl_create(d, this%ctx), "omxl_create")
        if (n_media > 0) then
            call read_media_lowdim(2, n_media, (/this%nx, this%ny/), dr, table, aux)
            call omx_check(omxl_set_media(this%ctx, int(n_media, c_int32_t), table, aux), "omxl_set_media")
        end if
    end subroutine init_2D_hip

    subroutine kill_2D_hip(this)
        class(TMxll_2D_hip), intent(inout) :: this
        integer(c_int) :: rc
        if (c_associated(this%ctx)) then
            rc = omxl_destroy(this%ctx)
            this%ctx = c_null_ptr
        end if
        call this%free_mirrors()
    end subroutine kill_2D_hip

    subroutine td_propagate_H_2D_hip(this)
        class(TMxll_2D_hip), intent(inout) :: this
        call omx_check(omxl_step_h(this%ctx), "omxl_step_h")
    end subroutine td_propagate_H_2D_hip

    subroutine td_propagate_E_2D_hip(this, t)
        class(TMxll_2D_hip), intent(inout) :: this
        integer, intent(in) :: t
        call omx_check(omxl_step_e(this%ctx), "omxl_step_e")
    end subroutine td_propagate_E_2D_hip

    ! A soft current source as TSource%init_source builds it (au, dir 'x', 'y' or 'z').
    subroutine add_source_2D_hip(this, dir, j_amp, w0, t0, tau, t_init, t_final, phase, r0, radius)
        class(TMxll_2D_hip), intent(inout) :: this
        character(len=*), intent(in) :: dir
        real(dp), intent(in) :: j_amp, w0, t0, tau, t_init, t_final, phase, r0(3), radius
        integer(c_int32_t), allocatable :: ind(:)
        real(c_double), allocatable :: ker(:)
        integer :: n_ker
        call lowdim_kernel(2, this%dr, (/this%nx, this%ny/), r0, radius, n_ker, ind, ker)
        call omx_check(omxl_add_source(this%ctx, dir_code(dir, 2), j_amp, w0, t0, tau, t_init, t_final, phase, &
                                       n_ker, ind, ker), "omxl_add_source")
    end subroutine add_source_2D_hip

    subroutine inject_sources_2D_hip(this, time)
        class(TMxll_2D_hip), intent(inout) :: this
        real(dp), intent(in) :: time
        call omx_check(omxl_inject_sources(this%ctx, time), "omxl_inject_sources")
    end subroutine inject_sources_2D_hip

    ! indx_list(nd,2) of init_detector (detector_mod.f90:239-371).
    subroutine add_detector_2D_hip(this, field, indx_list)
        class(TMxll_2D_hip), intent(inout) :: this
        integer, intent(in) :: field
        integer, intent(in) :: indx_list(:, :)
        integer(c_int32_t), allocatable :: idx(:)
        integer :: nd
        nd = size(indx_list, 1)
        allocate(idx(2 * nd))
        idx = reshape(indx_list, (/2 * nd/))
        call omx_check(omxl_add_detector(this%ctx, field, nd, idx), "omxl_add_detector")
        this%n_samples = this%n_samples + nd
    end subroutine add_detector_2D_hip

    ! All detector samples (write_2D_field stencils), in the order added.
    subroutine sample_2D_hip(this, values)
        class(TMxll_2D_hip), intent(inout) :: this
        real(dp), intent(out) :: values(:)
        call omx_check(omxl_sample(this%ctx, values), "omxl_sample")
    end subroutine sample_2D_hip

    subroutine sync_to_host_2D_hip(this)
        class(TMxll_2D_hip), intent(inout) :: this
        call this%alloc_mirrors()
        if (allocated(this%Ex)) then
            call omx_check(omxl_copy_field(this%ctx, Ex_FIELD, this%Ex, 0), "omxl_copy_field")
            call omx_check(omxl_copy_field(this%ctx, Ey_FIELD, this%Ey, 0), "omxl_copy_field")
            call omx_check(omxl_copy_field(this%ctx, Hz_FIELD, this%Hz, 0), "omxl_copy_field")
        end if
        if (allocated(this%Ez)) then
            call omx_check(omxl_copy_field(this%ctx, Ez_FIELD, this%Ez, 0), "omxl_copy_field")
            call omx_check(omxl_copy_field(this%ctx, Hx_FIELD, this%Hx, 0), "omxl_copy_field")
            call omx_check(omxl_copy_field(this%ctx, Hy_FIELD, this%Hy, 0), "omxl_copy_field")
        end if
    end subroutine sync_to_host_2D_hip

    subroutine sync_to_device_2D_hip(this)
        class(TMxll_2D_hip), intent(inout) :: this
        if (allocated(this%Ex)) then
            call omx_check(omxl_copy_field(this%ctx, Ex_FIELD, this%Ex, 1), "omxl_copy_field")
            call omx_check(omxl_copy_field(this%ctx, Ey_FIELD, this%Ey, 1), "omxl_copy_field")
            call omx_check(omxl_copy_field(this%ctx, Hz_FIELD, this%Hz, 1), "omxl_copy_field")
        end if
        if (allocated(this%Ez)) then
            call omx_check(omxl_copy_field(this%ctx, Ez_FIELD, this%Ez, 1), "omxl_copy_field")
            call omx_check(omxl_copy_field(this%ctx, Hx_FIELD, this%Hx, 1), "omxl_copy_field")
            call omx_check(omxl_copy_field(this%ctx, Hy_FIELD, this%Hy, 1), "omxl_copy_field")
        end if
    end subroutine sync_to_device_2D_hip

    subroutine synchronize_2D_hip(this)
        class(TMxll_2D_hip), intent(inout) :: this
        call omx_check(omxl_synchronize(this%ctx), "omxl_synchronize")
    end subroutine synchronize_2D_hip

    ! ---- 1D -----------------------------------------------------------------
    subroutine init_1D_hip(this, grid_Ndims, npml, boundaries, dt, dr, mode, n_media, mpi_coords, mpi_dims)
        class(TMxll_1D_hip), intent(inout) :: this
        integer,  intent(in) :: grid_Ndims(3), npml, boundaries(3), mode, n_media
        integer,  intent(in) :: mpi_coords(3), mpi_dims(3)
        real(dp), intent(in) :: dt, dr
        type(omxl_desc) :: d
        type(omx_medium), allocatable :: table(:)
        integer(c_int8_t), allocatable :: aux(:, :)

        this%dimensions = 1
        this%chunk_coor = mpi_coords
        this%nz = grid_Ndims(1)
        this%npml = npml
        this%boundaries = boundaries(1)
        this%mode = mode
        this%n_media = n_media
        this%dt = dt
        this%dr = dr
        this%dt_eps0 = dt / eps0
        this%dt_mu0 = dt / mu0
        this%time = 0.0d0
        this%t_skip = 0.0d0
        d%dims = 1
        d%mode = 0
        d%n1 = this%nz; d%n2 = 1
        d%npml = npml
        d%boundaries = (/boundaries(1), 1/)
        d%precision = 8
        d%dt = dt
        d%dr = dr
        call omx_check(omxl_create(d, this%ctx), "omxl_create")
        if (n_media > 0) then
            call read_media_lowdim(1, n_media, (/this%nz, 1/), dr, table, aux)
            call omx_check(omxl_set_media(this%ctx, int(n_media, c_int32_t), table, aux), "omxl_set_media")
        end if
    end subroutine init_1D_hip

    subroutine kill_1D_hip(this)
        class(TMxll_1D_hip), intent(inout) :: this
        integer(c_int) :: rc
        if (c_associated(this%ctx)) then
            rc = omxl_destroy(this%ctx)
            this%ctx = c_null_ptr
        end if
        if (allocated(this%Ex)) deallocate(this%Ex, this%Hy)
    end subroutine kill_1D_hip

    subroutine td_propagate_H_1D_hip(this)
        class(TMxll_1D_hip), intent(inout) :: this
        call omx_check(omxl_step_h(this%ctx), "omxl_step_h")
    end subroutine td_propagate_H_1D_hip

    subroutine td_propagate_E_1D_hip(this, t)
        class(TMxll_1D_hip), intent(inout) :: this
        integer, intent(in) :: t
        call omx_check(omxl_step_e(this%ctx), "omxl_step_e")
    end subroutine td_propagate_E_1D_hip

    subroutine add_source_1D_hip(this, dir, j_amp, w0, t0, tau, t_init, t_final, phase, r0, radius)
        class(TMxll_1D_hip), intent(inout) :: this
        character(len=*), intent(in) :: dir
        real(dp), intent(in) :: j_amp, w0, t0, tau, t_init, t_final, phase, r0(3), radius
        integer(c_int32_t), allocatable :: ind(:)
        real(c_double), allocatable :: ker(:)
        integer :: n_ker
        call lowdim_kernel(1, this%dr, (/this%nz, 1/), r0, radius, n_ker, ind, ker)
        call omx_check(omxl_add_source(this%ctx, dir_code(dir, 1), j_amp, w0, t0, tau, t_init, t_final, phase, &
                                       n_ker, ind, ker), "omxl_add_source")
    end subroutine add_source_1D_hip

    subroutine inject_sources_1D_hip(this, time)
        class(TMxll_1D_hip), intent(inout) :: this
        real(dp), intent(in) :: time
        call omx_check(omxl_inject_sources(this%ctx, time), "omxl_inject_sources")
    end subroutine inject_sources_1D_hip

    ! indx_list(nd,1) of init_detector (detector_mod.f90:207-238).
    subroutine add_detector_1D_hip(this, field, indx_list)
        class(TMxll_1D_hip), intent(inout) :: this
        integer, intent(in) :: field
        integer, intent(in) :: indx_list(:, :)
        integer(c_int32_t), allocatable :: idx(:)
        integer :: nd
        nd = size(indx_list, 1)
        allocate(idx(nd))
        idx = indx_list(:, 1)
        call omx_check(omxl_add_detector(this%ctx, field, nd, idx), "omxl_add_detector")
        this%n_samples = this%n_samples + nd
    end subroutine add_detector_1D_hip

    subroutine sample_1D_hip(this, values)
        class(TMxll_1D_hip), intent(inout) :: this
        real(dp), intent(out) :: values(:)
        call omx_check(omxl_sample(this%ctx, values), "omxl_sample")
    end subroutine sample_1D_hip

    subroutine sync_to_host_1D_hip(this)
        class(TMxll_1D_hip), intent(inout) :: this
        if (.not. allocated(this%Ex)) then
            allocate(this%Ex(this%nz), this%Hy(this%nz - 1))
            this%Ex = 0.0d0; this%Hy = 0.0d0
        end if
        call omx_check(omxl_copy_field(this%ctx, Ex_FIELD, this%Ex, 0), "omxl_copy_field")
        call omx_check(omxl_copy_field(this%ctx, Hy_FIELD, this%Hy, 0), "omxl_copy_field")
    end subroutine sync_to_host_1D_hip

    subroutine sync_to_device_1D_hip(this)
        class(TMxll_1D_hip), intent(inout) :: this
        if (.not. allocated(this%Ex)) return
        call omx_check(omxl_copy_field(this%ctx, Ex_FIELD, this%Ex, 1), "omxl_copy_field")
        call omx_check(omxl_copy_field(this%ctx, Hy_FIELD, this%Hy, 1), "omxl_copy_field")
    end subroutine sync_to_device_1D_hip

    subroutine synchronize_1D_hip(this)
        class(TMxll_1D_hip), intent(inout) :: this
        call omx_check(omxl_synchronize(this%ctx), "omxl_synchronize")
    end subroutine synchronize_1D_hip

end module omx_mxll_lowdim_hip_mod
