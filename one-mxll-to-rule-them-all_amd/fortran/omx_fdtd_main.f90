! omx_fdtd_main.f90 -- omx_fdtd_3d, a 3D driver over TMxll_3D_hip that reads
! the reference's own input files and writes its detector files:
!   inp            namelist &OMxRTA (input_mod.f90:7-42,71-75), 3D only
!   sources.in     one soft source per line (source_mod.f90:194-209)
!   medium_XXXX.in read by TMxll_3D_hip%init
!   detectors.in   "type field xmin xmax ymin ymax zmin zmax" [nm] (output_mod.f90:64)
! and runs the time loop of mxim_mxll.f90:92-132 on the GPU.  Quantum groups
! (DFTB+) are not part of this driver; a run that asks for them stops.
! Output: output_detector_XXXXXXX/point.dat ("time value" rows) and one file
! per print step for line/plane/volume detectors ("coordinates value" rows).
program omx_fdtd_3d

    use iso_c_binding
    use omx_units_mod
    use omx_mxll_3d_hip_mod
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

    type(TMxll_3D_hip) :: mxll
    integer  :: grid(3), bc(3), Nt, t_det_print, tt, i, funit, ierr, print_step
    integer  :: det_type(999), det_field(999), det_nd(999), det_first(999)
    integer, allocatable :: idx(:, :)
    real(dp), allocatable :: values(:)
    real(dp) :: dr, dt, time
    character(len=2)  :: dir
    real(dp) :: j_amp, freq, t0, tau, r0(3), radius, t_init, t_final, phase

    open (newunit=funit, file="inp", action="read", iostat=ierr)
    if (ierr /= 0) then
        write (*, '("Error: input file inp does not exist")')
        error stop 1
    end if
    read (funit, nml=OMxRTA, iostat=ierr)
    if (ierr /= 0) write (*, '("Error: invalid Namelist format")')
    close (funit)
    if (mxll_dimensions /= 3) then
        write (*, '("Error: omx_fdtd_3d runs 3D grids only")')
        error stop 1
    end if
    if (mxll_n_q_groups > 0) then
        write (*, '("Error: quantum groups need the DFTB+ host path; not available in omx_fdtd_3d")')
        error stop 1
    end if
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
    dr = mxll_dr * nm_to_au                                      ! input_mod.f90:121
    dt = MIN(dr / (2.0 * c0), mxll_dt)                           ! input_mod.f90:122
    Nt = INT((DBLE(mxll_total_time) * fs_to_au) / dt)            ! input_mod.f90:124
    t_det_print = 1
    if (mxll_dt_det_print > 0.0d0) t_det_print = int(mxll_dt_det_print * fs_to_au / dt)
    if (t_det_print <= 0) then
        write (*, *) "Error: dt_det_print must be greater than Maxwell dt."
        error stop 1
    end if

    call mxll%init(grid, mxll_npml, bc, dt, dr, 0, mxll_n_media, (/0, 0, 0/), (/1, 1, 1/))

    if (mxll_n_src > 0) then
        open (newunit=funit, file="sources.in", action="read", iostat=ierr)
        if (ierr /= 0) then
            write (*, '("Error: could not open source input file sources.in")')
            error stop 1
        end if
        do i = 1, mxll_n_src
            read (funit, *) dir, j_amp, freq, t0, tau, r0(1), r0(2), r0(3), radius, t_init, t_final, phase
            call mxll%add_source(dir, j_amp, freq * ev_to_au, t0 * fs_to_au, tau * fs_to_au, &
                                 t_init * fs_to_au, t_final * fs_to_au, phase, r0 * nm_to_au, radius * nm_to_au)
        end do
        close (funit)
    end if

    call init_detectors()
    allocate(values(max(mxll%n_samples, 1)))

    print_step = 0
    mxll%n_skip_steps = 1
    do tt = 1, Nt
        time = dt * (DBLE(tt) - 0.5d0)
        mxll%time = time
        mxll%t_skip = time
        call mxll%td_propagate_H_field()
        call mxll%inject_sources(time)
        call mxll%td_propagate_E_field(tt)
        if (mxll_n_detectors > 0 .and. mod(tt, t_det_print) == 0) then
            call mxll%sample_detectors(values)
            call write_detectors(time)
            print_step = print_step + 1
        end if
    end do
    call mxll%synchronize()
    write (*, '("omx_fdtd_3d: ", I0, " steps on a ", I0, "x", I0, "x", I0, " grid, precision ", I0)') &
        Nt, grid(1), grid(2), grid(3), mxll%precision
    call mxll%kill()

contains

    ! init_detector for one domain (detector_mod.f90:373-707): index box
    ! int(x/dr) + int(N/2), cells listed with i fastest.
    subroutine init_detectors()
        character(len=10) :: type_ch
        character(len=4)  :: field_ch
        character(len=20) :: number
        real(dp) :: xmn, xmx, ymn, ymx, zmn, zmx
        integer  :: n, lo(3), hi(3), a, nd, ii, jj, kk, q
        if (mxll_n_detectors <= 0) return
        open (newunit=funit, file="detectors.in", action="read", iostat=ierr)
        if (ierr /= 0) then
            write (*, *) "Error: Could not open file detectors.in"
            error stop 1
        end if
        do n = 1, mxll_n_detectors
            read (funit, *) type_ch, field_ch, xmn, xmx, ymn, ymx, zmn, zmx
            select case (trim(type_ch))
            case ("point", "POINT", "Point");          det_type(n) = POINT_DETECTOR
            case ("line_x", "LINE_X", "Line_x");       det_type(n) = LINE_X_DETECTOR
            case ("line_y", "LINE_Y", "Line_y");       det_type(n) = LINE_Y_DETECTOR
            case ("line_z", "LINE_Z", "Line_z");       det_type(n) = LINE_Z_DETECTOR
            case ("plane_xy", "PLANE_XY", "Plane_xy"); det_type(n) = PLANE_XY_DETECTOR
            case ("plane_yz", "PLANE_YZ", "Plane_yz"); det_type(n) = PLANE_YZ_DETECTOR
            case ("plane_zx", "PLANE_ZX", "Plane_zx"); det_type(n) = PLANE_ZX_DETECTOR
            case ("volume", "VOLUME", "Volume");       det_type(n) = VOLUME_DETECTOR
            case default
                write (*, '("Error: invalid detector type")')
                error stop 1
            end select
            select case (trim(field_ch))
            case ("Ex", "ex", "EX"); det_field(n) = Ex_FIELD
            case ("Ey", "ey", "EY"); det_field(n) = Ey_FIELD
            case ("Ez", "ez", "EZ"); det_field(n) = Ez_FIELD
            case ("Hx", "hx", "HX"); det_field(n) = Hx_FIELD
            case ("Hy", "hy", "HY"); det_field(n) = Hy_FIELD
            case ("Hz", "hz", "HZ"); det_field(n) = Hz_FIELD
            case default
                write (*, '("Error: invalid field type for detector")')
                error stop 1
            end select
            lo = (/int(xmn * nm_to_au / dr), int(ymn * nm_to_au / dr), int(zmn * nm_to_au / dr)/) + grid / 2
            hi = (/int(xmx * nm_to_au / dr), int(ymx * nm_to_au / dr), int(zmx * nm_to_au / dr)/) + grid / 2
            select case (det_type(n))
            case (POINT_DETECTOR);    hi = lo
            case (LINE_X_DETECTOR);   hi(2:3) = lo(2:3)
            case (LINE_Y_DETECTOR);   hi(1) = lo(1); hi(3) = lo(3)
            case (LINE_Z_DETECTOR);   hi(1:2) = lo(1:2)
            case (PLANE_XY_DETECTOR); hi(3) = lo(3)
            case (PLANE_YZ_DETECTOR); hi(1) = lo(1)
            case (PLANE_ZX_DETECTOR); hi(2) = lo(2)
            end select
            do a = 1, 3
                lo(a) = max(lo(a), 1)
                hi(a) = min(hi(a), grid(a))
            end do
            nd = max(0, hi(1) - lo(1) + 1) * max(0, hi(2) - lo(2) + 1) * max(0, hi(3) - lo(3) + 1)
            det_nd(n) = nd
            det_first(n) = mxll%n_samples + 1
            if (nd == 0) cycle
            allocate(idx(nd, 3))
            q = 0
            do kk = lo(3), hi(3)
            do jj = lo(2), hi(2)
            do ii = lo(1), hi(1)
                q = q + 1
                idx(q, :) = (/ii, jj, kk/)
            end do
            end do
            end do
            call mxll%add_detector(det_type(n), det_field(n), idx)
            deallocate(idx)
            write (number, '(I7.7)') n
            call execute_command_line("mkdir -p ./output_detector_" // trim(number))
            if (det_type(n) == POINT_DETECTOR) then
                open (newunit=q, file="./output_detector_" // trim(number) // "/point.dat", status="replace")
                write (q, '("# Time (a.u.)   ", A, " (a.u.)")') trim(field_ch)
                close (q)
            end if
        end do
        close (funit)
    end subroutine init_detectors

    subroutine write_detectors(time)
        real(dp), intent(in) :: time
        character(len=20) :: number, pstep
        character(len=2)  :: fname(6) = (/"Ex", "Ey", "Ez", "Hx", "Hy", "Hz"/)
        character(len=8)  :: kind_ch
        integer :: n, q, u, a
        real(dp) :: c(3)
        integer, allocatable :: ijk(:, :)
        do n = 1, mxll_n_detectors
            if (det_nd(n) == 0) cycle
            write (number, '(I7.7)') n
            if (det_type(n) == POINT_DETECTOR) then
                open (newunit=u, file="./output_detector_" // trim(number) // "/point.dat", &
                      status="old", position="append", action="write")
                write (u, *) time, values(det_first(n))
                close (u)
                cycle
            end if
            write (pstep, '(I9.9)') print_step
            kind_ch = "_line_"
            if (det_type(n) >= PLANE_XY_DETECTOR) kind_ch = "_plane_"
            if (det_type(n) == VOLUME_DETECTOR) kind_ch = "_volume_"
            open (newunit=u, file="./output_detector_" // trim(number) // "/" // fname(det_field(n)) // &
                  trim(kind_ch) // trim(pstep) // ".dat", status="replace", action="write")
            write (u, '("# Time = ", ES18.8, " (a.u.)")') time
            call detector_cells(n, ijk)
            do q = 1, det_nd(n)
                do a = 1, 3
                    c(a) = (ijk(q, a) * dr - int(grid(a) / 2) * dr) * au_to_nm
                end do
                select case (det_type(n))
                case (LINE_X_DETECTOR);   write (u, *) c(1), values(det_first(n) + q - 1)
                case (LINE_Y_DETECTOR);   write (u, *) c(2), values(det_first(n) + q - 1)
                case (LINE_Z_DETECTOR);   write (u, *) c(3), values(det_first(n) + q - 1)
                case (PLANE_XY_DETECTOR); write (u, *) c(1), c(2), values(det_first(n) + q - 1)
                case (PLANE_YZ_DETECTOR); write (u, *) c(2), c(3), values(det_first(n) + q - 1)
                case (PLANE_ZX_DETECTOR); write (u, *) c(3), c(1), values(det_first(n) + q - 1)
                case default;             write (u, *) c(1), c(2), c(3), values(det_first(n) + q - 1)
                end select
            end do
            close (u)
        end do
    end subroutine write_detectors

    ! the cells of detector n, rebuilt in the order they were registered
    subroutine detector_cells(n, ijk)
        integer, intent(in) :: n
        integer, allocatable, intent(out) :: ijk(:, :)
        character(len=10) :: type_ch
        character(len=4)  :: field_ch
        real(dp) :: v(6)
        integer :: lo(3), hi(3), a, q, ii, jj, kk, m, fu
        open (newunit=fu, file="detectors.in", action="read")
        do m = 1, n
            read (fu, *) type_ch, field_ch, v
        end do
        close (fu)
        lo = (/int(v(1) * nm_to_au / dr), int(v(3) * nm_to_au / dr), int(v(5) * nm_to_au / dr)/) + grid / 2
        hi = (/int(v(2) * nm_to_au / dr), int(v(4) * nm_to_au / dr), int(v(6) * nm_to_au / dr)/) + grid / 2
        select case (det_type(n))
        case (LINE_X_DETECTOR);   hi(2:3) = lo(2:3)
        case (LINE_Y_DETECTOR);   hi(1) = lo(1); hi(3) = lo(3)
        case (LINE_Z_DETECTOR);   hi(1:2) = lo(1:2)
        case (PLANE_XY_DETECTOR); hi(3) = lo(3)
        case (PLANE_YZ_DETECTOR); hi(1) = lo(1)
        case (PLANE_ZX_DETECTOR); hi(2) = lo(2)
        end select
        do a = 1, 3
            lo(a) = max(lo(a), 1)
            hi(a) = min(hi(a), grid(a))
        end do
        allocate(ijk(det_nd(n), 3))
        q = 0
        do kk = lo(3), hi(3)
        do jj = lo(2), hi(2)
        do ii = lo(1), hi(1)
            q = q + 1
            ijk(q, :) = (/ii, jj, kk/)
        end do
        end do
        end do
    end subroutine detector_cells

end program omx_fdtd_3d
