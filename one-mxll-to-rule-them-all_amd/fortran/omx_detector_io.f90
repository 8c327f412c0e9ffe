! omx_detector_io.f90 -- the reference's detector input and output files.
!
!   detectors.in  one detector per line, "type field x_min x_max y_min y_max
!                 z_min z_max" in nm (output_mod.f90:56-76); cell box
!                 int(x*nm_to_au/dr) + int(N/2) per axis
!                 (detector_mod.f90:373-707), cells listed i fastest and
!                 clipped to the grid.
!   output_detector_XXXXXXX/point.dat
!                 a header with the point's position (output_mod.f90:105-165),
!                 then one list-directed "time value" row per print step.
!   output_detector_XXXXXXX/<field>_line|plane|volume_XXXXXXXXX.dat
!                 one file per print step: two header lines
!                 (write_fields_subs_mod.f90:631-748), then list-directed
!                 "coordinates value" rows (:751-1509).
!
! 2D grids (dims 2) take point, line_x, line_y and plane_xy detectors at
! (x, y) (detector_mod.f90:239-371) and write <field>_line_ / <field>_plane_xy_
! files (write_fields_subs_mod.f90:197-628); 1D grids (dims 1) take point and
! line_z detectors at z along their one axis (:207-238) and write Ex_line_ /
! Hy_line_z_ files (:16-193).  A 2D point.dat header prints z = 0 (the
! reference prints a k index init_detector never sets, output_mod.f90:133-137).
!
! Rows come either from the engine's detector samples (write_detector_rows)
! or, for a caller that holds a TMxll_3D mirror, from the mirror with the
! reference's averaging stencils (write_3D_field_mirror).  One domain (no
! MPI split): the reference's mpi_dims are 1 and mpi_coords 0.
module omx_detector_io_mod

    use omx_units_mod
    use omx_mxll_3d_mod
    implicit none
    private

    public :: TDetectorOut, read_detectors, init_detector_files, write_detector_headers, &
              write_detector_rows, write_3D_field_mirror

    type :: TDetectorOut
        integer :: detector_type = 0
        integer :: field = 0
        character(len=2) :: f_ch = "  "
        ! the reference's index box, before clipping to the grid
        integer :: i_min = 0, i_max = 0, j_min = 0, j_max = 0, k_min = 0, k_max = 0
        integer :: nd = 0                         ! cells inside the grid
        integer, allocatable :: indx_list(:, :)   ! (nd, 3), i fastest
        integer :: first = 0                      ! first of its values in the engine's sample vector
        integer :: dims = 3                       ! dimensions of the grid it samples
    end type TDetectorOut


    character(len=*), parameter :: dir0 = "./output_detector"
    character(len=2), parameter :: fnames(6) = (/"Ex", "Ey", "Ez", "Hx", "Hy", "Hz"/)

contains

    pure real(dp) function coord(idx, n, dr)
        integer,  intent(in) :: idx, n
        real(dp), intent(in) :: dr
        coord = idx * dr - int(n / 2) * dr
    end function coord

    subroutine read_detectors(n_det, grid, dr, dets, dims)
        integer,  intent(in) :: n_det, grid(3)
        real(dp), intent(in) :: dr
        type(TDetectorOut), allocatable, intent(out) :: dets(:)
        integer, intent(in), optional :: dims
        character(len=10) :: type_ch
        character(len=4)  :: field_ch
        real(dp) :: v(6)
        integer :: funit, ierr, n, lo(3), hi(3), a, q, ii, jj, kk, nd_axes

        allocate(dets(n_det))
        open (newunit=funit, file="detectors.in", action="read", iostat=ierr)
        if (ierr /= 0) then
            write (*, *) "Error: Could not open file detectors.in"
            error stop 1
        end if
        do n = 1, n_det
            read (funit, *) type_ch, field_ch, v
            v = v * nm_to_au
            select case (trim(type_ch))
            case ("point", "POINT", "Point");          dets(n)%detector_type = POINT_DETECTOR
            case ("line_x", "LINE_X", "Line_x");       dets(n)%detector_type = LINE_X_DETECTOR
            case ("line_y", "LINE_Y", "Line_y");       dets(n)%detector_type = LINE_Y_DETECTOR
            case ("line_z", "LINE_Z", "Line_z");       dets(n)%detector_type = LINE_Z_DETECTOR
            case ("plane_xy", "PLANE_XY", "Plane_xy"); dets(n)%detector_type = PLANE_XY_DETECTOR
            case ("plane_yz", "PLANE_YZ", "Plane_yz"); dets(n)%detector_type = PLANE_YZ_DETECTOR
            case ("plane_zx", "PLANE_ZX", "Plane_zx"); dets(n)%detector_type = PLANE_ZX_DETECTOR
            case ("volume", "VOLUME", "Volume");       dets(n)%detector_type = VOLUME_DETECTOR
            case default
                write (*, '("Error: invalid detector type")')
                error stop 1
            end select
            select case (trim(field_ch))
            case ("Ex", "ex", "EX"); dets(n)%field = Ex_FIELD
            case ("Ey", "ey", "EY"); dets(n)%field = Ey_FIELD
            case ("Ez", "ez", "EZ"); dets(n)%field = Ez_FIELD
            case ("Hx", "hx", "HX"); dets(n)%field = Hx_FIELD
            case ("Hy", "hy", "HY"); dets(n)%field = Hy_FIELD
            case ("Hz", "hz", "HZ"); dets(n)%field = Hz_FIELD
            case default
                write (*, '("Error: invalid field type for detector")')
                error stop 1
            end select
            dets(n)%f_ch = fnames(dets(n)%field)
            nd_axes = 3
            if (present(dims)) nd_axes = dims
            dets(n)%dims = nd_axes
            call check_lowdim(nd_axes, dets(n))
            if (nd_axes == 1) then
                ! the 1D line is the z axis: its cells come from z (detector_mod.f90:207-238)
                lo = 1; hi = 1
                lo(1) = int(v(5) / dr) + int(grid(1) / 2)
                hi(1) = int(v(6) / dr) + int(grid(1) / 2)
            else
                do a = 1, nd_axes
                    lo(a) = int(v(2 * a - 1) / dr) + int(grid(a) / 2)
                    hi(a) = int(v(2 * a) / dr) + int(grid(a) / 2)
                end do
                if (nd_axes == 2) then
                    lo(3) = 1; hi(3) = 1
                end if
            end if
            ! a point, line or plane spans its first corner along the other axes
            select case (dets(n)%detector_type)
            case (POINT_DETECTOR);    hi = lo
            case (LINE_X_DETECTOR);   hi(2:3) = lo(2:3)
            case (LINE_Y_DETECTOR);   hi(1) = lo(1); hi(3) = lo(3)
            case (LINE_Z_DETECTOR)
                if (nd_axes == 3) hi(1:2) = lo(1:2)
            case (PLANE_XY_DETECTOR); hi(3) = lo(3)
            case (PLANE_YZ_DETECTOR); hi(1) = lo(1)
            case (PLANE_ZX_DETECTOR); hi(2) = lo(2)
            end select
            dets(n)%i_min = lo(1); dets(n)%i_max = hi(1)
            dets(n)%j_min = lo(2); dets(n)%j_max = hi(2)
            dets(n)%k_min = lo(3); dets(n)%k_max = hi(3)
            if (nd_axes == 1) then
                dets(n)%k_min = lo(1); dets(n)%k_max = hi(1)
            end if
            do a = 1, 3
                lo(a) = max(lo(a), 1)
                if (a <= nd_axes) hi(a) = min(hi(a), grid(a))
            end do
            dets(n)%nd = max(0, hi(1) - lo(1) + 1) * max(0, hi(2) - lo(2) + 1) * max(0, hi(3) - lo(3) + 1)
            allocate(dets(n)%indx_list(dets(n)%nd, 3))
            q = 0
            do kk = lo(3), hi(3)
                do jj = lo(2), hi(2)
                    do ii = lo(1), hi(1)
                        q = q + 1
                        dets(n)%indx_list(q, :) = (/ii, jj, kk/)
                    end do
                end do
            end do
        end do
        close (funit)
    end subroutine read_detectors

    ! Detector types and fields a 1D or 2D grid cannot take (detector_mod.f90:86-201).
    subroutine check_lowdim(dims, det)
        integer, intent(in) :: dims
        type(TDetectorOut), intent(in) :: det
        if (dims == 1) then
            if (det%field /= Ex_FIELD .and. det%field /= Hy_FIELD) then
                write (*, '("Error: ",A," field cannot be detected in 1D")') det%f_ch
                write (*, '("Use Ex or Hy fields instead")')
                error stop 1
            end if
            if (det%detector_type /= POINT_DETECTOR .and. det%detector_type /= LINE_Z_DETECTOR) then
                write (*, '("Error: invalid combination of detector type and dimensions")')
                error stop 1
            end if
        else if (dims == 2) then
            select case (det%detector_type)
            case (POINT_DETECTOR, LINE_X_DETECTOR, LINE_Y_DETECTOR, PLANE_XY_DETECTOR)
            case default
                write (*, '("Error: invalid combination of detector type and dimensions")')
                error stop 1
            end select
        end if
    end subroutine check_lowdim

    ! The detectors' directories and the point detectors' point.dat headers.
    subroutine init_detector_files(dets, grid, dr)
        type(TDetectorOut), intent(in) :: dets(:)
        integer,  intent(in) :: grid(3)
        real(dp), intent(in) :: dr
        character(len=20) :: number
        character(len=99) :: dname
        integer :: n, funit, ierr
        real(dp) :: x, y, z

        do n = 1, size(dets)
            write (number, '(I7.7)') n
            dname = dir0 // "_" // trim(number)
            call execute_command_line("mkdir -p " // trim(dname), exitstat=ierr)
            if (ierr /= 0) then
                write (*, *) "Error: Could not create directory ", trim(dname)
                error stop 1
            end if
            if (dets(n)%detector_type /= POINT_DETECTOR) cycle
            open (newunit=funit, file=trim(dname) // "/point.dat", status="replace", action="write", iostat=ierr)
            if (ierr /= 0) then
                write (*, *) "Error: Could not open file for writing: ", trim(dname) // "/point.dat"
                error stop 1
            end if
            select case (dets(n)%dims)
            case (1)
                x = 0.0d0; y = 0.0d0
                z = coord(dets(n)%k_min, grid(1), dr) * au_to_nm
            case (2)
                x = coord(dets(n)%i_min, grid(1), dr) * au_to_nm
                y = coord(dets(n)%j_min, grid(2), dr) * au_to_nm
                z = 0.0d0
            case default
                x = coord(dets(n)%i_min, grid(1), dr) * au_to_nm
                y = coord(dets(n)%j_min, grid(2), dr) * au_to_nm
                z = coord(dets(n)%k_min, grid(3), dr) * au_to_nm
            end select
            write (funit, '("# Time (a.u.)              ' // dets(n)%f_ch // &
                          ' (a.u.),      x = ", F10.5, " nm,   y = ", F10.5, " nm,   z = ", F10.5, " nm")') x, y, z
            close (funit)
        end do
    end subroutine init_detector_files

    function step_file(n, det, print_step) result(name)
        integer, intent(in) :: n, print_step
        type(TDetectorOut), intent(in) :: det
        character(len=99) :: name
        character(len=20) :: number, pnum
        character(len=10) :: kind_ch
        write (number, '(I7.7)') n
        write (pnum, '(I9.9)') print_step
        select case (det%detector_type)
        case (LINE_X_DETECTOR, LINE_Y_DETECTOR, LINE_Z_DETECTOR); kind_ch = "_line_"
        case (VOLUME_DETECTOR); kind_ch = "_volume_"
        case default; kind_ch = "_plane_"
        end select
        if (det%dims == 2 .and. det%detector_type == PLANE_XY_DETECTOR) kind_ch = "_plane_xy_"
        if (det%dims == 1 .and. det%field == Hy_FIELD) kind_ch = "_line_z_"
        name = dir0 // "_" // trim(number) // "/" // det%f_ch // trim(kind_ch) // trim(pnum) // ".dat"
    end function step_file

    ! This print step's line / plane / volume files with their two header
    ! lines (write_fields_subs_mod.f90:631-748).  The reference's header
    ! writer prints coordinates that an earlier detector of the same call set:
    ! a line along y shows the z of the detector before it, a line along z
    ! that detector's x and y, a plane the value the last line printed first
    ! (:669-714); and a line's second column name is a character nothing in
    ! the call has set (a NUL byte in its build; after a plane, the plane's).
    ! The same values carry over here from detector to detector within a
    ! call; each call starts from a NUL name and zero coordinates.
    subroutine write_detector_headers(dets, print_step, time, grid, dr)
        type(TDetectorOut), intent(in) :: dets(:)
        integer,  intent(in) :: print_step, grid(3)
        real(dp), intent(in) :: time, dr
        character(len=1) :: h1, h2, w1, hw2
        real(dp) :: hx, hy, hz, hv1, hv2
        integer :: n, funit, ierr

        hx = 0.0d0; hy = 0.0d0; hz = 0.0d0; hv1 = 0.0d0; hv2 = 0.0d0
        hw2 = achar(0)

        if (size(dets) > 0) then
            if (dets(1)%dims /= 3) then
                call write_lowdim_headers(dets, print_step, time, grid, dr)
                return
            end if
        end if
        do n = 1, size(dets)
            select case (dets(n)%detector_type)
            case (LINE_X_DETECTOR, LINE_Y_DETECTOR, LINE_Z_DETECTOR)
                select case (dets(n)%detector_type)
                case (LINE_X_DETECTOR)
                    h1 = "y"; h2 = "z"; w1 = "x"
                    hy = coord(dets(n)%j_min, grid(2), dr)
                    hz = coord(dets(n)%k_min, grid(3), dr)
                    hv1 = hy * au_to_nm; hv2 = hz * au_to_nm
                case (LINE_Y_DETECTOR)
                    h1 = "z"; h2 = "x"; w1 = "y"
                    hx = coord(dets(n)%i_min, grid(1), dr)
                    hv1 = hz * au_to_nm; hv2 = hx * au_to_nm
                case default
                    h1 = "x"; h2 = "y"; w1 = "z"
                    hz = coord(dets(n)%k_min, grid(3), dr)
                    hv1 = hx * au_to_nm; hv2 = hy * au_to_nm
                end select
                open (newunit=funit, file=trim(step_file(n, dets(n), print_step)), status="replace", &
                      action="write", iostat=ierr)
                write (funit, '("# Time = ", ES18.8, " (a.u.), ' // h1 // ' = ", F12.6, " (nm), ' // h2 // &
                              ' = ", F12.6, " (nm)")') time, hv1, hv2
                write (funit, '("# ' // w1 // ' (nm)                   ' // trim(hw2) // &
                              ' (nm)                   ' // dets(n)%f_ch // ' (a.u.)")')
                close (funit)
            case (PLANE_XY_DETECTOR, PLANE_YZ_DETECTOR, PLANE_ZX_DETECTOR)
                select case (dets(n)%detector_type)
                case (PLANE_XY_DETECTOR); h1 = "z"; w1 = "x"; hw2 = "y"
                case (PLANE_YZ_DETECTOR); h1 = "x"; w1 = "y"; hw2 = "z"
                case default;             h1 = "y"; w1 = "z"; hw2 = "x"
                end select
                open (newunit=funit, file=trim(step_file(n, dets(n), print_step)), status="replace", &
                      action="write", iostat=ierr)
                write (funit, '("# Time = ", ES18.8, " (a.u.), ' // h1 // ' = ", F12.6, " (nm)")') time, hv1
                write (funit, '("# ' // w1 // ' (nm)                   ' // trim(hw2) // &
                              ' (nm)                   ' // dets(n)%f_ch // ' (a.u.)")')
                close (funit)
            case (VOLUME_DETECTOR)
                open (newunit=funit, file=trim(step_file(n, dets(n), print_step)), status="replace", &
                      action="write", iostat=ierr)
                write (funit, '("# Time = ", ES18.8, " (a.u.)")') time
                write (funit, '("# x (nm)                   y (nm)                   z (nm)                    ' &
                              // dets(n)%f_ch // ' (a.u.)")')
                close (funit)
            end select
        end do
    end subroutine write_detector_headers

    ! write_2D_headers (write_fields_subs_mod.f90:197-276) and write_1D_header
    ! (:16-85): this print step's line and plane files with their two
    ! header lines.
    subroutine write_lowdim_headers(dets, print_step, time, grid, dr)
        type(TDetectorOut), intent(in) :: dets(:)
        integer,  intent(in) :: print_step, grid(3)
        real(dp), intent(in) :: time, dr
        character(len=1) :: ch, cw
        real(dp) :: cv
        integer :: n, funit, ierr
        do n = 1, size(dets)
            select case (dets(n)%detector_type)
            case (LINE_X_DETECTOR, LINE_Y_DETECTOR)
                if (dets(n)%detector_type == LINE_X_DETECTOR) then
                    ch = "y"; cw = "x"
                    cv = coord(dets(n)%j_min, grid(2), dr) * au_to_nm
                else
                    ch = "x"; cw = "y"
                    cv = coord(dets(n)%i_min, grid(1), dr) * au_to_nm
                end if
                open (newunit=funit, file=trim(step_file(n, dets(n), print_step)), status="replace", &
                      action="write", iostat=ierr)
                write (funit, '("# Time = ", ES18.8, " (a.u.), ' // ch // ' = ", F12.6, " (nm)")') time, cv
                write (funit, '("# ' // cw // ' (nm)                   ' // trim(dets(n)%f_ch) // ' (a.u.)")')
                close (funit)
            case (PLANE_XY_DETECTOR)
                open (newunit=funit, file=trim(step_file(n, dets(n), print_step)), status="replace", &
                      action="write", iostat=ierr)
                write (funit, '("# Time = ", ES18.8, " (a.u.)")') time
                write (funit, '("# x (nm)                   y (nm)                   ' // trim(dets(n)%f_ch) // &
                              ' (a.u.)")')
                close (funit)
            case (LINE_Z_DETECTOR)
                open (newunit=funit, file=trim(step_file(n, dets(n), print_step)), status="replace", &
                      action="write", iostat=ierr)
                write (funit, '("# Time = ", ES18.8, " (a.u.)")') time
                write (funit, '("# z (nm)                   ' // dets(n)%f_ch // ' (a.u.)")')
                close (funit)
            end select
        end do
    end subroutine write_lowdim_headers

    ! One row per cell of the detectors that see the grid, after the
    ! headers of this print step: point.dat gets "time value", the others
    ! "coordinates value" (the coordinates the reference writes for each
    ! detector type, write_fields_subs_mod.f90:751-1509).
    subroutine write_row(funit, det, q, time, grid, dr, val)
        integer,  intent(in) :: funit, q, grid(3)
        type(TDetectorOut), intent(in) :: det
        real(dp), intent(in) :: time, dr, val
        real(dp) :: x, y, z
        x = coord(det%indx_list(q, 1), grid(1), dr) * au_to_nm
        y = coord(det%indx_list(q, 2), grid(2), dr) * au_to_nm
        z = coord(det%indx_list(q, 3), grid(3), dr) * au_to_nm
        if (det%dims == 1) z = x
        select case (det%detector_type)
        case (POINT_DETECTOR);    write (funit, *) time, val
        case (LINE_X_DETECTOR);   write (funit, *) x, val
        case (LINE_Y_DETECTOR);   write (funit, *) y, val
        case (LINE_Z_DETECTOR);   write (funit, *) z, val
        case (PLANE_XY_DETECTOR); write (funit, *) x, y, val
        case (PLANE_YZ_DETECTOR); write (funit, *) y, z, val
        case (PLANE_ZX_DETECTOR); write (funit, *) z, x, val
        case default;             write (funit, *) x, y, z, val
        end select
    end subroutine write_row

    subroutine open_rows(n, det, print_step, funit)
        integer, intent(in) :: n, print_step
        type(TDetectorOut), intent(in) :: det
        integer, intent(out) :: funit
        character(len=20) :: number
        write (number, '(I7.7)') n
        if (det%detector_type == POINT_DETECTOR) then
            open (newunit=funit, file=dir0 // "_" // trim(number) // "/point.dat", status="old", &
                  action="write", position="append")
        else
            open (newunit=funit, file=trim(step_file(n, det, print_step)), status="old", action="write", &
                  position="append")
        end if
    end subroutine open_rows

    ! Rows from the engine's samples: values(dets(n)%first + q - 1) is cell q
    ! of detector n.
    subroutine write_detector_rows(dets, print_step, time, grid, dr, values)
        type(TDetectorOut), intent(in) :: dets(:)
        integer,  intent(in) :: print_step, grid(3)
        real(dp), intent(in) :: time, dr, values(:)
        integer :: n, q, funit
        do n = 1, size(dets)
            if (dets(n)%nd == 0) cycle
            call open_rows(n, dets(n), print_step, funit)
            do q = 1, dets(n)%nd
                call write_row(funit, dets(n), q, time, grid, dr, values(dets(n)%first + q - 1))
            end do
            close (funit)
        end do
    end subroutine write_detector_rows

    ! Rows from a TMxll_3D mirror (sync_to_host first), with the reference's
    ! stencils: E components averaged over their two Yee positions, H over
    ! four, Ez of a point detector unaveraged, and Hz of a zx plane summed
    ! in its own order (write_fields_subs_mod.f90:801-1509).
    subroutine write_3D_field_mirror(dets, print_step, time, grid, dr, mxll)
        type(TDetectorOut), intent(in) :: dets(:)
        integer,  intent(in) :: print_step, grid(3)
        real(dp), intent(in) :: time, dr
        class(TMxll_3D), intent(in) :: mxll
        integer :: n, q, funit, i, j, k
        real(dp) :: val
        do n = 1, size(dets)
            if (dets(n)%nd == 0) cycle
            call open_rows(n, dets(n), print_step, funit)
            do q = 1, dets(n)%nd
                i = dets(n)%indx_list(q, 1); j = dets(n)%indx_list(q, 2); k = dets(n)%indx_list(q, 3)
                select case (dets(n)%field)
                case (Ex_FIELD)
                    val = 0.5 * (mxll%Ex(i - 1, j, k) + mxll%Ex(i, j, k))
                case (Ey_FIELD)
                    val = 0.5 * (mxll%Ey(i, j - 1, k) + mxll%Ey(i, j, k))
                case (Ez_FIELD)
                    if (dets(n)%detector_type == POINT_DETECTOR) then
                        val = mxll%Ez(i, j, k)
                    else
                        val = 0.5 * (mxll%Ez(i, j, k - 1) + mxll%Ez(i, j, k))
                    end if
                case (Hx_FIELD)
                    val = 0.25 * (mxll%Hx(i, j - 1, k) + mxll%Hx(i, j, k) + mxll%Hx(i, j, k - 1) + mxll%Hx(i, j - 1, k - 1))
                case (Hy_FIELD)
                    val = 0.25 * (mxll%Hy(i - 1, j, k) + mxll%Hy(i, j, k) + mxll%Hy(i, j, k - 1) + mxll%Hy(i - 1, j, k - 1))
                case default
                    if (dets(n)%detector_type == PLANE_ZX_DETECTOR) then
                        val = 0.25 * (mxll%Hz(i - 1, j - 1, k) + mxll%Hz(i, j - 1, k) + mxll%Hz(i - 1, j, k) + &
                                      mxll%Hz(i, j, k))
                    else
                        val = 0.25 * (mxll%Hz(i, j, k) + mxll%Hz(i - 1, j, k) + mxll%Hz(i, j - 1, k) + &
                                      mxll%Hz(i - 1, j - 1, k))
                    end if
                end select
                call write_row(funit, dets(n), q, time, grid, dr, val)
            end do
            close (funit)
        end do
    end subroutine write_3D_field_mirror

end module omx_detector_io_mod
