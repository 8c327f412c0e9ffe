! omx_mxll_3d_hip.f90 -- TMxll_3D_hip, the GPU 3D solver: a TMxll_3D (so
! callers that branch on `class is (TMxll_3D)` reach it), made by
! maxwell_factory(3).  Host side of the drop-in: the four TMxll bindings keep
! the reference's signatures (src/FDTD/mxll_base_mod.f90:27-30,36-67); init
! also reads the reference's medium_XXXX.in files
! (classical_medium_mod.f90:603-690) and hands the cell medium map to the
! engine, which derives the staggered per-component map and the update
! coefficients on the device.  sync_to_host / sync_to_device copy E and H
! between the device and TMxll_3D's host mirrors (reference bounds).
!
! The couplings the reference performs by touching the solver's arrays
! (sources, E at quantum sites, J scatter, detectors, field dumps) are type-
! bound calls into the engine.  Every engine status is checked; a failure
! stops with the engine's message, as the reference stops on fatal input.
! Precision: `precision` set before init, else the environment variable
! OMX_FDTD_PRECISION (4 or 8), else 8.
module omx_mxll_3d_hip_mod

    use iso_c_binding
    use omx_units_mod
    use omx_mxll_base_mod
    use omx_mxll_3d_mod
    use omx_fdtd_cabi
    implicit none

    type, extends(TMxll_3D) :: TMxll_3D_hip
        type(c_ptr) :: ctx = c_null_ptr
        integer :: precision = 0
        integer :: n_samples = 0
        integer :: n_sites = 0
    contains
        procedure :: init => init_3D_hip
        procedure :: kill => kill_3D_hip
        procedure :: td_propagate_H_field => td_propagate_H_3D_hip
        procedure :: td_propagate_E_field => td_propagate_E_3D_hip
        procedure :: add_source => add_source_3D_hip
        procedure :: inject_sources => inject_sources_3D_hip
        procedure :: add_detector => add_detector_3D_hip
        procedure :: sample_detectors => sample_3D_hip
        procedure :: set_q_sites => set_q_sites_3D_hip
        procedure :: set_q_groups => set_q_groups_3D_hip
        procedure :: set_two_level => set_two_level_3D_hip
        procedure :: two_level_step => two_level_step_3D_hip
        procedure :: get_two_level => get_two_level_3D_hip
        procedure :: gather_E => gather_E_3D_hip
        procedure :: scatter_J => scatter_J_3D_hip
        procedure :: energy => energy_3D_hip
        procedure :: get_field => get_field_3D_hip
        procedure :: set_field => set_field_3D_hip
        procedure :: sync_to_host => sync_to_host_3D_hip
        procedure :: sync_to_device => sync_to_device_3D_hip
        procedure :: synchronize => synchronize_3D_hip
    end type TMxll_3D_hip

contains

    subroutine init_3D_hip(this, grid_Ndims, npml, boundaries, dt, dr, mode, n_media, mpi_coords, mpi_dims)
        class(TMxll_3D_hip), intent(inout) :: this
        integer,  intent(in) :: grid_Ndims(3), npml, boundaries(3), mode, n_media
        integer,  intent(in) :: mpi_coords(3), mpi_dims(3)
        real(dp), intent(in) :: dt, dr
        type(omx3d_desc) :: d
        type(omx_medium), allocatable :: table(:)
        integer(c_int8_t), allocatable :: aux(:, :, :)
        character(len=16) :: env
        integer :: stat, p

        if (any(mpi_dims /= 1) .or. any(mpi_coords /= 0)) then
            write (*, '("Error: TMxll_3D_hip runs one domain per GPU (mpi_dims must be 1)")')
            error stop 1
        end if
        this%dimensions = 3
        this%chunk_coor = mpi_coords
        this%nx = grid_Ndims(1); this%ny = grid_Ndims(2); this%nz = grid_Ndims(3)
        this%npml = npml
        this%boundaries = boundaries
        this%mode = mode
        this%n_media = n_media
        this%dt = dt
        this%dr = dr
        this%dt_eps0 = dt / eps0
        this%dt_mu0 = dt / mu0
        this%time = 0.0d0
        this%t_skip = 0.0d0
        p = this%precision
        if (p == 0) then
            call get_environment_variable("OMX_FDTD_PRECISION", env, status=stat)
            p = 8
            if (stat == 0) read (env, *) p
        end if
        this%precision = p
        d%nx = this%nx; d%ny = this%ny; d%nz = this%nz
        d%npml = npml
        d%boundaries = boundaries
        d%precision = p
        d%dt = dt
        d%dr = dr
        call omx_check(omx3d_create(d, this%ctx), "omx3d_create")
        if (n_media > 0) then
            call read_media_files(n_media, this%nx, this%ny, this%nz, dr, table, aux)
            call omx_check(omx3d_set_media(this%ctx, int(n_media, c_int32_t), table, aux), "omx3d_set_media")
        end if
    end subroutine init_3D_hip

    ! medium_XXXX.in: type line, parameter line, then "x y z" (nm) points to
    ! EOF; a point marks cell int(x/dr) + int(N/2) (classical_medium_mod.f90:603-690).
    subroutine read_media_files(n_media, nx, ny, nz, dr, table, aux)
        integer,  intent(in) :: n_media, nx, ny, nz
        real(dp), intent(in) :: dr
        type(omx_medium), allocatable, intent(out) :: table(:)
        integer(c_int8_t), allocatable, intent(out) :: aux(:, :, :)
        character(len=20) :: name, kind_ch
        character(len=10) :: material
        integer :: n, funit, ierr, ii, jj, kk
        real(dp) :: x, y, z, eps_read, omegaD, GammaD

        allocate(table(n_media), aux(nx, ny, nz))
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
            select case (trim(kind_ch))
            case ("dielectric")
                read (funit, *) eps_read
                table(n)%kind = DIELECTRIC_MEDIUM
                table(n)%eps_r = eps_read
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
                read (funit, *, iostat=ierr) x, y, z
                if (ierr /= 0) exit
                ii = int(x * nm_to_au / dr) + int(nx / 2)
                jj = int(y * nm_to_au / dr) + int(ny / 2)
                kk = int(z * nm_to_au / dr) + int(nz / 2)
                if (ii < 1 .or. ii > nx .or. jj < 1 .or. jj > ny .or. kk < 1 .or. kk > nz) then
                    write (*, '("Error: medium ",I4.4," has a point outside the grid")') n
                    error stop 1
                end if
                if (aux(ii, jj, kk) /= 0) then
                    write (*, '("Error: medium ",I4.4," overlaps medium ",I4.4)') int(aux(ii, jj, kk)), n
                else
                    aux(ii, jj, kk) = int(n, c_int8_t)
                end if
            end do
            close (funit)
        end do
    end subroutine read_media_files

    subroutine kill_3D_hip(this)
        class(TMxll_3D_hip), intent(inout) :: this
        integer(c_int) :: rc
        if (c_associated(this%ctx)) then
            rc = omx3d_destroy(this%ctx)
            this%ctx = c_null_ptr
        end if
        call this%free_mirrors()
    end subroutine kill_3D_hip

    subroutine td_propagate_H_3D_hip(this)
        class(TMxll_3D_hip), intent(inout) :: this
        call omx_check(omx3d_step_h(this%ctx), "omx3d_step_h")
    end subroutine td_propagate_H_3D_hip

    subroutine td_propagate_E_3D_hip(this, t)
        class(TMxll_3D_hip), intent(inout) :: this
        integer, intent(in) :: t
        call omx_check(omx3d_step_e(this%ctx, this%time, this%t_skip), "omx3d_step_e")
    end subroutine td_propagate_E_3D_hip

    ! A soft current source as TSource%init_source builds it: Blackman-Harris
    ! kernel (source_mod.f90:251-327) at indices i + int(r0/dr) + int(N/2)
    ! (source_mod.f90:347-356).  Arguments in au, dir = 'x', 'y' or 'z'.
    subroutine add_source_3D_hip(this, dir, j_amp, w0, t0, tau, t_init, t_final, phase, r0, radius)
        class(TMxll_3D_hip), intent(inout) :: this
        character(len=*), intent(in) :: dir
        real(dp), intent(in) :: j_amp, w0, t0, tau, t_init, t_final, phase, r0(3), radius
        integer :: n_ker, w, i, j, k, p, np, idir
        real(dp) :: r_max, rc, r, norm, v
        real(c_double), allocatable :: ker(:)
        integer(c_int32_t), allocatable :: ind(:)

        n_ker = int(radius / this%dr)
        w = 2 * n_ker + 1
        np = w**3
        allocate(ker(np), ind(3 * np))
        r_max = (n_ker * 2 + 2) * this%dr
        rc = r_max / 2.0d0
        norm = 0.0d0
        do k = -n_ker, n_ker
        do j = -n_ker, n_ker
        do i = -n_ker, n_ker
            p = (i + n_ker) + w * ((j + n_ker) + w * (k + n_ker)) + 1
            r = sqrt((i * this%dr)**2 + (j * this%dr)**2 + (k * this%dr)**2)
            v = a_bh(1) + a_bh(2) * cos(2.0d0 * pi0 * (r - rc) / r_max) + &
                a_bh(3) * cos(2.0d0 * pi0 * 2.0d0 * (r - rc) / r_max) + &
                a_bh(4) * cos(2.0d0 * pi0 * 3.0d0 * (r - rc) / r_max)
            if (v < 0.0d0) v = 0.0d0
            ker(p) = v
            norm = norm + v
            ind(p) = i + int(r0(1) / this%dr) + int(this%nx / 2)
            ind(p + np) = j + int(r0(2) / this%dr) + int(this%ny / 2)
            ind(p + 2 * np) = k + int(r0(3) / this%dr) + int(this%nz / 2)
        end do
        end do
        end do
        ker = ker / norm
        select case (dir(1:1))
        case ('x'); idir = 1
        case ('y'); idir = 2
        case ('z'); idir = 3
        case default
            write (*, '("Error: Unknown source direction in 3D mxll.")')
            error stop 1
        end select
        call omx_check(omx3d_add_source(this%ctx, idir, j_amp, w0, t0, tau, t_init, t_final, phase, &
                                        n_ker, ind, ker), "omx3d_add_source")
    end subroutine add_source_3D_hip

    subroutine inject_sources_3D_hip(this, time)
        class(TMxll_3D_hip), intent(inout) :: this
        real(dp), intent(in) :: time
        call omx_check(omx3d_inject_sources(this%ctx, time), "omx3d_inject_sources")
    end subroutine inject_sources_3D_hip

    subroutine add_detector_3D_hip(this, det_type, field, indx_list)
        class(TMxll_3D_hip), intent(inout) :: this
        integer, intent(in) :: det_type, field
        integer, intent(in) :: indx_list(:, :)
        integer(c_int32_t), allocatable :: idx(:)
        integer :: nd
        nd = size(indx_list, 1)
        allocate(idx(3 * nd))
        idx = reshape(indx_list, (/3 * nd/))
        call omx_check(omx3d_add_detector(this%ctx, det_type, field, nd, idx), "omx3d_add_detector")
        this%n_samples = this%n_samples + nd
    end subroutine add_detector_3D_hip

    ! All detector samples, in the order the detectors were added.
    subroutine sample_3D_hip(this, values)
        class(TMxll_3D_hip), intent(inout) :: this
        real(dp), intent(out) :: values(:)
        call omx_check(omx3d_sample(this%ctx, values), "omx3d_sample")
    end subroutine sample_3D_hip

    subroutine set_q_sites_3D_hip(this, ijk, density, dt_q)
        class(TMxll_3D_hip), intent(inout) :: this
        integer, intent(in) :: ijk(:, :)
        real(dp), intent(in) :: density, dt_q
        integer(c_int32_t), allocatable :: idx(:)
        this%n_sites = size(ijk, 1)
        allocate(idx(3 * this%n_sites))
        idx = reshape(ijk, (/3 * this%n_sites/))
        call omx_check(omx3d_set_q_sites(this%ctx, this%n_sites, idx, density, dt_q), "omx3d_set_q_sites")
    end subroutine set_q_sites_3D_hip

    ! Systems of q-groups (omx3d_set_q_groups): group_type(g) Q_MATERIAL or
    ! Q_SINGLE, group_of(q) the 1-based group of system q (systems in group
    ! order), ijk(q,:) its cell (a single group's kernel centre).
    subroutine set_q_groups_3D_hip(this, group_type, group_of, ijk, density, dt_q)
        class(TMxll_3D_hip), intent(inout) :: this
        integer, intent(in) :: group_type(:), group_of(:), ijk(:, :)
        real(dp), intent(in) :: density, dt_q
        integer(c_int32_t), allocatable :: idx(:), gt(:), gof(:)
        this%n_sites = size(ijk, 1)
        allocate(idx(3 * this%n_sites), gt(size(group_type)), gof(this%n_sites))
        idx = reshape(ijk, (/3 * this%n_sites/))
        gt = group_type
        gof = group_of - 1
        call omx_check(omx3d_set_q_groups(this%ctx, size(group_type), gt, this%n_sites, gof, idx, density, dt_q), &
                       "omx3d_set_q_groups")
    end subroutine set_q_groups_3D_hip

    ! Two-level emitters at the sites (au; pol(n,3)).
    subroutine set_two_level_3D_hip(this, omega0, mu, pol, gamma2)
        class(TMxll_3D_hip), intent(inout) :: this
        real(dp), intent(in) :: omega0(:), mu(:), pol(:, :), gamma2(:)
        call omx_check(omx3d_set_two_level(this%ctx, omega0, mu, pol, gamma2), "omx3d_set_two_level")
    end subroutine set_two_level_3D_hip

    ! One quantum step of every emitter: gather E, RK4 over dt_q, scatter J.
    subroutine two_level_step_3D_hip(this)
        class(TMxll_3D_hip), intent(inout) :: this
        call omx_check(omx3d_two_level_step(this%ctx), "omx3d_two_level_step")
    end subroutine two_level_step_3D_hip

    ! dipole(n,3) and density matrix (rho_gg, rho_ee, Re rho_eg, Im rho_eg)(n,4).
    subroutine get_two_level_3D_hip(this, dipole, rho)
        class(TMxll_3D_hip), intent(inout) :: this
        real(dp), intent(out) :: dipole(:, :), rho(:, :)
        call omx_check(omx3d_get_two_level(this%ctx, dipole, rho), "omx3d_get_two_level")
    end subroutine get_two_level_3D_hip

    ! E_field_list(n,3) of a q-group (interactions_mod.f90:656-691)
    subroutine gather_E_3D_hip(this, E_field_list)
        class(TMxll_3D_hip), intent(inout) :: this
        real(dp), intent(out) :: E_field_list(:, :)
        call omx_check(omx3d_gather_e(this%ctx, E_field_list), "omx3d_gather_e")
    end subroutine gather_E_3D_hip

    ! J_old <- J; J(site) = density*dPt_dt; dJ = (J - J_old)/dt_q
    ! (interactions_mod.f90:1067-1069,1184-1190,1216-1218)
    subroutine scatter_J_3D_hip(this, dPt_dt)
        class(TMxll_3D_hip), intent(inout) :: this
        real(dp), intent(in) :: dPt_dt(:, :)
        call omx_check(omx3d_scatter_j(this%ctx, dPt_dt), "omx3d_scatter_j")
    end subroutine scatter_J_3D_hip

    subroutine energy_3D_hip(this, w64, w32c)
        class(TMxll_3D_hip), intent(inout) :: this
        real(dp), intent(out) :: w64, w32c
        call omx_check(omx3d_energy(this%ctx, w64, w32c), "omx3d_energy")
    end subroutine energy_3D_hip

    ! Host mirror of one field with the reference's bounds, e.g. Ex(0:nx,1:ny+1,1:nz+1).
    subroutine get_field_3D_hip(this, field, host)
        class(TMxll_3D_hip), intent(inout) :: this
        integer, intent(in) :: field
        real(dp), intent(inout) :: host(:, :, :)
        call omx_check(omx3d_copy_field(this%ctx, field, host, 0), "omx3d_copy_field")
    end subroutine get_field_3D_hip

    ! Host array of one field with the reference's bounds -> the device.
    subroutine set_field_3D_hip(this, field, host)
        class(TMxll_3D_hip), intent(inout) :: this
        integer, intent(in) :: field
        real(dp), intent(inout) :: host(:, :, :)
        call omx_check(omx3d_copy_field(this%ctx, field, host, 1), "omx3d_copy_field")
    end subroutine set_field_3D_hip

    ! E and H of the device into the mirrors this%Ex .. this%Hz (waits for
    ! the queued steps; FP32 grids are widened to FP64).
    subroutine sync_to_host_3D_hip(this)
        class(TMxll_3D_hip), intent(inout) :: this
        call this%alloc_mirrors()
        call omx_check(omx3d_copy_field(this%ctx, Ex_FIELD, this%Ex, 0), "omx3d_copy_field")
        call omx_check(omx3d_copy_field(this%ctx, Ey_FIELD, this%Ey, 0), "omx3d_copy_field")
        call omx_check(omx3d_copy_field(this%ctx, Ez_FIELD, this%Ez, 0), "omx3d_copy_field")
        call omx_check(omx3d_copy_field(this%ctx, Hx_FIELD, this%Hx, 0), "omx3d_copy_field")
        call omx_check(omx3d_copy_field(this%ctx, Hy_FIELD, this%Hy, 0), "omx3d_copy_field")
        call omx_check(omx3d_copy_field(this%ctx, Hz_FIELD, this%Hz, 0), "omx3d_copy_field")
    end subroutine sync_to_host_3D_hip

    ! The mirrors (as the caller left them) onto the device: a caller that
    ! edits fields on the host, as the reference's couplings do, hands them
    ! back before the next step.
    subroutine sync_to_device_3D_hip(this)
        class(TMxll_3D_hip), intent(inout) :: this
        if (.not. allocated(this%Ex)) return
        call omx_check(omx3d_copy_field(this%ctx, Ex_FIELD, this%Ex, 1), "omx3d_copy_field")
        call omx_check(omx3d_copy_field(this%ctx, Ey_FIELD, this%Ey, 1), "omx3d_copy_field")
        call omx_check(omx3d_copy_field(this%ctx, Ez_FIELD, this%Ez, 1), "omx3d_copy_field")
        call omx_check(omx3d_copy_field(this%ctx, Hx_FIELD, this%Hx, 1), "omx3d_copy_field")
        call omx_check(omx3d_copy_field(this%ctx, Hy_FIELD, this%Hy, 1), "omx3d_copy_field")
        call omx_check(omx3d_copy_field(this%ctx, Hz_FIELD, this%Hz, 1), "omx3d_copy_field")
    end subroutine sync_to_device_3D_hip

    subroutine synchronize_3D_hip(this)
        class(TMxll_3D_hip), intent(inout) :: this
        call omx_check(omx3d_synchronize(this%ctx), "omx3d_synchronize")
    end subroutine synchronize_3D_hip

end module omx_mxll_3d_hip_mod
