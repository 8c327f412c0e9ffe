! omx_fdtd_main.f90 -- omx_fdtd_3d, the driver that reads the reference's
! own input files and writes its detector files:
!   inp            namelist &OMxRTA (input_mod.f90:7-42,71-75); mxll_dimensions
!                  3, 2 (mxll_2D_mode "TMz", "TEz" or "full") or 1
!   sources.in     one soft source per line (source_mod.f90:194-209)
!   medium_XXXX.in read by TMxll_3D_hip%init
!   detectors.in   "type field xmin xmax ymin ymax zmin zmax" [nm] (output_mod.f90:64)
!   mol_group_XXXXXXX.in  quantum groups (omx_qgroup_io: two-level systems)
! The solver comes from maxwell_factory (factory_mod.f90:15-30) and the time
! loop is that of mxim_mxll.f90:66-132, on the GPU, quantum sub-steps
! included (3D).  Output: the reference's output_detector_XXXXXXX/ files
! (omx_detector_io) and output_q_group_XXXX/energy_and_dipole.dat.  2D and 1D
! decks run the same loop without quantum groups (FP64).
! Detector rows come from the engine's samples; with OMX_FDTD_DET_MIRROR=1
! they come instead from the TMxll_3D host mirrors (sync_to_host at every
! print step), the way the reference's writers read its arrays.
program omx_fdtd_3d

    use iso_c_binding
    use omx_units_mod
    use omx_mxll_base_mod
    use omx_mxll_3d_mod
    use omx_mxll_3d_hip_mod
    use omx_mxll_lowdim_mod
    use omx_mxll_lowdim_hip_mod
    use omx_factory_mod
    use omx_detector_io_mod
    use omx_qgroup_io_mod
    implicit none

    character(len=10) :: mxll_boundaries(3) = "none", mxll_2D_mode = "none"
    integer  :: mxll_dimensions = 1, mxll_npml = 19, mxll_n_src = 0, mxll_n_media = 0
    integer  :: mxll_n_q_groups = 0, mxll_n_detectors = 0, mpi_procs_per_axis(3) = 1
    real(dp) :: mxll_box_size(3) = 0.0d0, mxll_total_time = 1.0d0, mxll_dr = 1.0d0
    real(dp) :: mxll_dt = 10.0E10_dp, mxll_dt_q = 0.01d0, mxll_dt_det_print = 0.0d0
    real(dp) :: mxll_dt_q_print = 0.0d0, mxll_density_factor = 1.0d-3, mxll_eps_r = 1.0d0
    namelist /OMxRTA/ mxll_boundaries, mxll_2D_mode, mxll_dimensions, mxll_npml, mxll_n_src, &
                      mxll_box_size, mxll_total_time, mxll_dr, mxll_dt, mxll_density_factor, mxll_eps_r, &
                      mxll_n_media, mxll_n_q_groups, mxll_dt_q, mpi_procs_per_axis, mxll_n_detectors, &
                      mxll_dt_det_print, mxll_dt_q_print

    class(TMxll), allocatable :: mxll
    type(TDetectorOut), allocatable :: dets(:)
    type(TQGroupIn), allocatable :: qgroups(:)
    integer  :: grid(3), bc(3), Nt, t_det_print, t_q_print, i, funit, ierr, stat, mode_2D
    real(dp) :: dr, dt, dt_q, dt_q_print, density
    character(len=8) :: env
    logical :: use_mirror

    open (newunit=funit, file="inp", action="read", iostat=ierr)
    if (ierr /= 0) then
        write (*, '("Error: input file inp does not exist")')
        error stop 1
    end if
    read (funit, nml=OMxRTA, iostat=ierr)
    if (ierr /= 0) write (*, '("Error: invalid Namelist format")')
    close (funit)
    do i = 1, 3
        select case (mxll_boundaries(i))
        case ("close");    bc(i) = CLOSE_BOUNDARIES
        case ("periodic"); bc(i) = PERIODIC_BOUNDARIES
        case ("cpml");     bc(i) = CPML_BOUNDARIES
        case default
            write (*, '("Error: invalid boundary condition in axis ", I0)') i
            error stop 1
        end select
    end do
    grid = int(mxll_box_size / mxll_dr)                          ! input_mod.f90:110-112
    select case (mxll_2D_mode)                                   ! input_mod.f90:163-177
    case ("TMz");  mode_2D = TMZ_2D_MODE
    case ("TEz");  mode_2D = TEZ_2D_MODE
    case ("full"); mode_2D = FULL_2D_MODE
    case ("none"); mode_2D = 0
    case default
        write (*, '("mxll_2D_mode: ", A)') mxll_2D_mode
        write (*, '("Error: invalid 2D mode")')
        error stop 1
    end select
    if (mxll_dimensions == 2 .and. mode_2D == 0) then
        write (*, '("Error: a 2D grid needs mxll_2D_mode TMz, TEz or full")')
        error stop 1
    end if
    if (mxll_dimensions /= 3 .and. mxll_n_q_groups > 0) then
        write (*, '("Error: quantum groups run on 3D grids in this build")')
        error stop 1
    end if
    dr = mxll_dr * nm_to_au                                      ! input_mod.f90:121
    dt = MIN(dr / (2.0 * c0), mxll_dt)                           ! input_mod.f90:122
    Nt = INT((DBLE(mxll_total_time) * fs_to_au) / dt)            ! input_mod.f90:124
    t_det_print = 1
    if (mxll_dt_det_print > 0.0d0) t_det_print = int(mxll_dt_det_print * fs_to_au / dt)
    if (t_det_print <= 0) then
        write (*, *) "Error: dt_det_print must be greater than Maxwell dt."
        error stop 1
    end if
    ! quantum step and print stride as input_mod.f90:123,125,149-160 set them
    ! (dt_q is taken as given, not converted from fs)
    dt_q = mxll_dt_q
    density = mxll_density_factor / (nm_to_au)**3
    t_q_print = 1
    if (mxll_n_q_groups > 0) then
        if (mxll_dt_q_print == 0.0d0) then
            dt_q_print = dt
        else if (mxll_dt_q_print < dt_q * au_to_fs) then
            write (*, '("Error: mxll_dt_q_print must be greater than or equal to the quantum system time step ' // &
                      'dt_q = ", E12.4, "fs")') dt_q * au_to_fs
            error stop 1
        else
            dt_q_print = mxll_dt_q_print * fs_to_au
        end if
        t_q_print = int(dt_q_print / dt)
        if (t_q_print <= 0) t_q_print = 1
    end if

    call get_environment_variable("OMX_FDTD_DET_MIRROR", env, status=stat)
    use_mirror = stat == 0 .and. trim(env) == "1"

    allocate(mxll, source=maxwell_factory(mxll_dimensions))
    call mxll%init(grid, mxll_npml, bc, dt, dr, mode_2D, mxll_n_media, (/0, 0, 0/), (/1, 1, 1/))
    select type (mxll)
    class is (TMxll_3D_hip)
        call run_3d(mxll)
    class is (TMxll_2D_hip)
        call run_lowdim(mxll)
    class is (TMxll_1D_hip)
        call run_lowdim(mxll)
    class default
        write (*, '("Error: no GPU solver for this grid")')
        error stop 1
    end select
    call mxll%kill()

contains

    subroutine run_3d(m)
        class(TMxll_3D_hip), intent(inout) :: m
        character(len=2)  :: dir
        real(dp) :: j_amp, freq, t0, tau, r0(3), radius, t_init, t_final, phase, time
        real(dp), allocatable :: values(:), omega0(:), mu(:), pol(:, :), gamma2(:), dip(:, :), rho(:, :), et(:)
        integer, allocatable :: idx(:, :), qijk(:, :), gtype(:), gof(:)
        integer :: tt, n, g, nq, print_step, print_q_step, tt_0
        logical :: move_q_system

        if (mxll_n_src > 0) then
            open (newunit=funit, file="sources.in", action="read", iostat=ierr)
            if (ierr /= 0) then
                write (*, '("Error: could not open source input file sources.in")')
                error stop 1
            end if
            do i = 1, mxll_n_src
                read (funit, *) dir, j_amp, freq, t0, tau, r0(1), r0(2), r0(3), radius, t_init, t_final, phase
                call m%add_source(dir, j_amp, freq * ev_to_au, t0 * fs_to_au, tau * fs_to_au, &
                                  t_init * fs_to_au, t_final * fs_to_au, phase, r0 * nm_to_au, radius * nm_to_au)
            end do
            close (funit)
        end if

        if (mxll_n_detectors > 0) then
            call read_detectors(mxll_n_detectors, grid, dr, dets)
            do n = 1, mxll_n_detectors
                dets(n)%first = m%n_samples + 1
                if (dets(n)%nd == 0) cycle
                idx = dets(n)%indx_list
                call m%add_detector(dets(n)%detector_type, dets(n)%field, idx)
            end do
            call init_detector_files(dets, grid, dr)
        end if
        allocate(values(max(m%n_samples, 1)))

        nq = 0
        if (mxll_n_q_groups > 0) then
            call read_q_groups(mxll_n_q_groups, grid, dr, qgroups)
            do g = 1, mxll_n_q_groups
                nq = nq + qgroups(g)%n_systems
            end do
            if (nq > 0) then
                allocate(gtype(mxll_n_q_groups), gof(nq), qijk(nq, 3))
                do g = 1, mxll_n_q_groups
                    gtype(g) = qgroups(g)%group_type
                    do i = 1, qgroups(g)%n_systems
                        gof(qgroups(g)%first + i - 1) = g
                        qijk(qgroups(g)%first + i - 1, :) = qgroups(g)%ijk(i, :)
                    end do
                end do
                call m%set_q_groups(gtype, gof, qijk, density, dt_q)
                call read_two_level(qgroups, omega0, mu, pol, gamma2)
                call m%set_two_level(omega0, mu, pol, gamma2)
                allocate(dip(nq, 3), rho(nq, 4), et(nq))
            end if
            call init_q_group_outputs(qgroups)
        end if

        ! quantum sub-steps (mxim_mxll.f90:73-79): the systems move every
        ! n_skip_steps = NINT(dt_q/dt) steps when dt_q > dt, else every step
        m%n_skip_steps = 1
        if (dt_q > dt) m%n_skip_steps = NINT(dt_q / dt)
        tt_0 = 0
        if (m%n_skip_steps == 1) tt_0 = 1
        m%tq_step_old = 0
        print_step = 0
        print_q_step = 0
        do tt = 1, Nt
            time = dt * (DBLE(tt) - 0.5d0)
            m%time = time
            m%tq_step = 1 + int((tt - tt_0) / m%n_skip_steps)
            move_q_system = .false.
            if (m%tq_step > m%tq_step_old) then
                m%t_skip = time
                m%tq_step_old = m%tq_step
                move_q_system = .true.
                if (nq == 0) move_q_system = .false.
            end if
            call m%td_propagate_H_field()
            call m%inject_sources(time)
            call m%td_propagate_E_field(tt)
            if (move_q_system) call m%two_level_step()
            if (mxll_n_detectors > 0 .and. mod(tt, t_det_print) == 0) then
                call write_detector_headers(dets, print_step, time, grid, dr)
                if (use_mirror) then
                    call m%sync_to_host()
                    call write_3D_field_mirror(dets, print_step, time, grid, dr, m)
                else
                    call m%sample_detectors(values)
                    call write_detector_rows(dets, print_step, time, grid, dr, values)
                end if
                print_step = print_step + 1
            end if
            if (mxll_n_q_groups > 0 .and. mod(tt, t_q_print) == 0) then
                if (nq > 0) then
                    call m%get_two_level(dip, rho)
                    et = 0.5d0 * omega0 * (rho(:, 2) - rho(:, 1))
                    call write_q_group_outputs(qgroups, print_q_step, time, et, dip)
                end if
                print_q_step = print_q_step + 1
            end if
        end do
        call m%synchronize()
        write (*, '("omx_fdtd_3d: ", I0, " steps on a ", I0, "x", I0, "x", I0, " grid, precision ", I0)') &
            Nt, grid(1), grid(2), grid(3), m%precision
    end subroutine run_3d

    ! A 2D or 1D deck: sources, detectors and the driver loop's step order
    ! (mxim_mxll.f90:92-132), detector files as write_2D_field /
    ! write_1D_field write them.
    subroutine run_lowdim(m)
        class(TMxll), intent(inout) :: m
        character(len=2)  :: dir
        real(dp) :: j_amp, freq, t0, tau, r0(3), radius, t_init, t_final, phase, time
        real(dp), allocatable :: values(:)
        integer :: tt, n, dims, print_step, ns

        dims = mxll_dimensions
        ns = 0
        if (mxll_n_src > 0) then
            open (newunit=funit, file="sources.in", action="read", iostat=ierr)
            if (ierr /= 0) then
                write (*, '("Error: could not open source input file sources.in")')
                error stop 1
            end if
            do i = 1, mxll_n_src
                read (funit, *) dir, j_amp, freq, t0, tau, r0(1), r0(2), r0(3), radius, t_init, t_final, phase
                select type (m)
                class is (TMxll_2D_hip)
                    call m%add_source(dir, j_amp, freq * ev_to_au, t0 * fs_to_au, tau * fs_to_au, &
                                      t_init * fs_to_au, t_final * fs_to_au, phase, r0 * nm_to_au, radius * nm_to_au)
                class is (TMxll_1D_hip)
                    call m%add_source(dir, j_amp, freq * ev_to_au, t0 * fs_to_au, tau * fs_to_au, &
                                      t_init * fs_to_au, t_final * fs_to_au, phase, r0 * nm_to_au, radius * nm_to_au)
                end select
            end do
            close (funit)
        end if
        if (mxll_n_detectors > 0) then
            call read_detectors(mxll_n_detectors, grid, dr, dets, dims)
            do n = 1, mxll_n_detectors
                call check_2D_mode(dets(n))
                dets(n)%first = ns + 1
                if (dets(n)%nd == 0) cycle
                select type (m)
                class is (TMxll_2D_hip)
                    call m%add_detector(dets(n)%field, dets(n)%indx_list(:, 1:2))
                    ns = m%n_samples
                class is (TMxll_1D_hip)
                    call m%add_detector(dets(n)%field, dets(n)%indx_list(:, 1:1))
                    ns = m%n_samples
                end select
            end do
            call init_detector_files(dets, grid, dr)
        end if
        allocate(values(max(ns, 1)))
        print_step = 0
        do tt = 1, Nt
            time = dt * (DBLE(tt) - 0.5d0)
            m%time = time
            m%t_skip = time
            call m%td_propagate_H_field()
            select type (m)
            class is (TMxll_2D_hip)
                call m%inject_sources(time)
            class is (TMxll_1D_hip)
                call m%inject_sources(time)
            end select
            call m%td_propagate_E_field(tt)
            if (mxll_n_detectors > 0 .and. mod(tt, t_det_print) == 0) then
                call write_detector_headers(dets, print_step, time, grid, dr)
                if (ns > 0) then
                    select type (m)
                    class is (TMxll_2D_hip)
                        call m%sample_detectors(values)
                    class is (TMxll_1D_hip)
                        call m%sample_detectors(values)
                    end select
                end if
                call write_detector_rows(dets, print_step, time, grid, dr, values)
                print_step = print_step + 1
            end if
        end do
        select type (m)
        class is (TMxll_2D_hip)
            call m%synchronize()
        class is (TMxll_1D_hip)
            call m%synchronize()
        end select
        write (*, '("omx_fdtd_3d: ", I0, " steps on a ", I0, "D grid of ", I0, "x", I0, ", precision 8")') &
            Nt, dims, grid(1), merge(grid(2), 1, dims == 2)
    end subroutine run_lowdim

    ! Fields a 2D mode cannot detect (output_mod.f90:81-103).
    subroutine check_2D_mode(det)
        type(TDetectorOut), intent(in) :: det
        if (mode_2D == TMZ_2D_MODE .and. (det%field == Ex_FIELD .or. det%field == Ey_FIELD .or. &
                                          det%field == Hz_FIELD)) then
            write (*, *) "Error: " // det%f_ch // " field cannot be detected in TMZ 2D mode."
            error stop 1
        end if
        if (mode_2D == TEZ_2D_MODE .and. (det%field == Ez_FIELD .or. det%field == Hx_FIELD .or. &
                                          det%field == Hy_FIELD)) then
            write (*, *) "Error: " // det%f_ch // " field cannot be detected in TEZ 2D mode."
            error stop 1
        end if
    end subroutine check_2D_mode

end program omx_fdtd_3d
