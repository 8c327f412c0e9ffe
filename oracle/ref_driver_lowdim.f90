! ref_driver_lowdim.f90 -- test driver that runs the reference's own 2D and
! 1D solver modules (compiled from the reference tree by oracle/Makefile,
! target `ref`) to produce golden vectors.  TEST INFRASTRUCTURE ONLY.
!
! It follows the driver loop of src/FDTD/mxim_mxll.f90:92-132 with the
! reference's init_2Dgrid / init_1Dgrid, td_propagate_H_*field /
! td_propagate_E_*field, read_init_media_2D / _1D (medium_XXXX.in),
! read_init_sources and update_sources (sources.in), init_detector
! (detectors.in, point detectors) and write_2D_field / write_1D_field.
! Restated here in a few lines each, citing what they follow:
!   source_interactions, 1D / 2D   interactions_mod.f90:32-100 (its module
!                                  needs DFTB+; J_mat past the kernel's edge
!                                  taken as 0)
!   point.dat creation             output_mod.f90:40-168 (header line only)
! The reference's 2D init allocates eps_x, eps_y only in TEz mode and eps_z
! only in TMz mode (mxll_2D_mod.f90:110-163), while read_init_media_2D writes
! all three for every dielectric cell (classical_medium_mod.f90:525-535): the
! driver allocates the missing ones (set to 1) before init so that a grid
! with a dielectric runs at all; the propagation never reads them.  As
! init_classical_medium never stores eps_r of a dielectric
! (classical_medium_mod.f90:62-63), the 2D eps arrays then get 1/eps_r from
! the case file (diel_eps) for the cells of the Ex map, as
! read_init_media_2D assigns them (:525-535).
!
! Input (current directory): case.nml (namelist &case) and the reference's
! own medium_XXXX.in, sources.in and detectors.in.  With timing /= 0 the
! driver writes no files and prints the wall time of its step loop (the CPU
! baseline of the 2D / 1D bench workloads).
! Output: coeffs.bin (be ce bh ch, then den_ex den_ey den_hx den_hy, or 1D
! den_ez den_hz), media.bin (media_map), fields.bin (the mode's field arrays,
! whole, at every step with mod(tt, field_every) == 0), and the reference's
! output_detector_XXXXXXX/point.dat files.
program omx_ref_driver_lowdim

    use constants_mod
    use classical_medium_mod
    use mxll_1D_mod
    use mxll_2D_mod
    use source_mod
    use detector_mod
    use write_fields_subs_mod
    use omp_lib

    implicit none

    integer  :: dims = 2, mode = 1, n1 = 32, n2 = 32, npml = 4, bc(3) = 3
    integer  :: n_steps = 10, record_every = 1, field_every = 0, n_media = 0, n_src = 0, n_det = 0, timing = 0
    real(dp) :: dx_nm = 5.0d0
    real(dp) :: diel_eps(15) = 1.0d0
    namelist /case/ dims, mode, n1, n2, npml, bc, n_steps, record_every, field_every, n_media, n_src, n_det, dx_nm, &
                    diel_eps, timing

    type(TMxll_2D)               :: m2
    type(TMxll_1D)               :: m1
    type(TSource), allocatable   :: source_list(:)
    type(TDetector), allocatable :: detectors(:)
    integer  :: grid_Ndims(3), mpi_coords(3), mpi_dims(3)
    integer  :: tt, n, i, j, funit, ffield, print_det_step
    integer(8) :: c0_clock, c1_clock, rate
    real(dp) :: dr, dt, time
    real(dp) :: x_min, x_max, y_min, y_max, z_min, z_max
    character(len=4)  :: field_name
    character(len=10) :: detector_type
    character(len=20) :: number
    character(len=99) :: dirname

    open(newunit=funit, file="case.nml", action="read")
    read(funit, nml=case)
    close(funit)

    if (dims == 1) then
        grid_Ndims = (/n1, 1, 1/)
    else
        grid_Ndims = (/n1, n2, 1/)
    end if
    mpi_coords = 0
    mpi_dims = 1
    dr = dx_nm * nm_to_au                                    ! input_mod.f90:121
    dt = MIN(dr / (2.0 * c0), 10.0E10_dp)                    ! input_mod.f90:122

    call read_init_sources(source_list, n_src, dims, dr, grid_Ndims, mpi_coords, mpi_dims)
    if (dims == 1) then
        call m1%init(grid_Ndims, npml, bc, dt, dr, 0, n_media, mpi_coords, mpi_dims)
    else
        allocate(m2%eps_x(n1, n2), m2%eps_y(n1, n2), m2%eps_z(n1, n2))
        m2%eps_x = 1.0d0
        m2%eps_y = 1.0d0
        m2%eps_z = 1.0d0
        call m2%init(grid_Ndims, npml, bc, dt, dr, mode, n_media, mpi_coords, mpi_dims)
        do n = 1, n_media
            if (m2%media(n)%medium_type /= DIELECTRIC_MEDIUM) cycle
            do j = 1, n2
            do i = 1, n1
                if (m2%media_map(i, j, 1) == n) then
                    m2%eps_x(i, j) = 1.0d0 / diel_eps(n)
                    m2%eps_y(i, j) = 1.0d0 / diel_eps(n)
                    m2%eps_z(i, j) = 1.0d0 / diel_eps(n)
                end if
            end do
            end do
        end do
    end if

    ! ---- golden tables ----------------------------------------------------
    if (timing /= 0) then
        n_det = 0
        field_every = 0
    end if
    if (timing == 0) then
    open(newunit=funit, file="coeffs.bin", access="stream", form="unformatted", status="replace")
    if (dims == 1) then
        if (allocated(m1%be)) write(funit) m1%be(1:npml), m1%ce(1:npml), m1%bh(1:npml-1), m1%ch(1:npml-1)
        write(funit) m1%den_ez, m1%den_hz
    else
        if (allocated(m2%be)) write(funit) m2%be(1:npml), m2%ce(1:npml), m2%bh(1:npml-1), m2%ch(1:npml-1)
        write(funit) m2%den_ex, m2%den_ey, m2%den_hx, m2%den_hy
    end if
    close(funit)
    open(newunit=funit, file="media.bin", access="stream", form="unformatted", status="replace")
    if (dims == 1) then
        write(funit) m1%media_map
    else
        write(funit) m2%media_map
    end if
    close(funit)
    end if

    ! ---- point detectors (output_mod.f90:40-168) ----------------------------
    if (n_det > 0) then
        allocate(detectors(n_det))
        open(newunit=funit, file="detectors.in", action="read")
        do n = 1, n_det
            read(funit, *) detector_type, field_name, x_min, x_max, y_min, y_max, z_min, z_max
            call detectors(n)%init_detector(field_name, detector_type, dims, grid_Ndims, dr, mpi_dims, &
                 mpi_coords, x_min*nm_to_au, x_max*nm_to_au, y_min*nm_to_au, y_max*nm_to_au, &
                 z_min*nm_to_au, z_max*nm_to_au)
        end do
        close(funit)
        do n = 1, n_det
            write(number, '(I7.7)') n
            dirname = "./output_detector_" // trim(number)
            call execute_command_line("mkdir -p " // trim(dirname))
            open(newunit=funit, file=trim(dirname) // "/point.dat", status="replace", action="write")
            write(funit, '("# Time (a.u.)              ", A)') detectors(n)%f_ch
            close(funit)
        end do
    end if

    if (timing == 0) open(newunit=ffield, file="fields.bin", access="stream", form="unformatted", status="replace")
    print_det_step = 0
    call system_clock(c0_clock, rate)
    do tt = 1, n_steps
        time = dt * (DBLE(tt) - 0.5d0)
        if (dims == 1) then
            m1%time = time
            m1%t_skip = time     ! a new quantum step every step (mxim_mxll.f90:97-104, n_skip_steps = 1)
            call m1%td_propagate_H_field()
            call update_sources(source_list, n_src, time)
            call inject_sources_1d()
            call m1%td_propagate_E_field(tt)
        else
            m2%time = time
            m2%t_skip = time
            call m2%td_propagate_H_field()
            call update_sources(source_list, n_src, time)
            call inject_sources_2d()
            call m2%td_propagate_E_field(tt)
        end if
        if (n_det > 0 .and. record_every > 0) then
            if (mod(tt, record_every) == 0) then
                if (dims == 1) then
                    call write_1D_field(detectors, m1, n_det, print_det_step, time, dr, grid_Ndims, &
                                        mpi_dims, mpi_coords, 0)
                else
                    call write_2D_field(detectors, m2, n_det, print_det_step, time, dr, grid_Ndims, &
                                        mpi_dims, mpi_coords, 0)
                end if
                print_det_step = print_det_step + 1
            end if
        end if
        if (field_every > 0) then
            if (mod(tt, field_every) == 0) then
                if (dims == 1) then
                    write(ffield) m1%Ex, m1%Hy
                else
                    if (mode == TMZ_2D_MODE) write(ffield) m2%Ez, m2%Hx, m2%Hy
                    if (mode == TEZ_2D_MODE) write(ffield) m2%Ex, m2%Ey, m2%Hz
                end if
            end if
        end if
    end do
    call system_clock(c1_clock)
    if (timing == 0) close(ffield)
    write(*, '("REF_LOWDIM dims=", I0, " mode=", I0, " steps=", I0)') dims, mode, n_steps
    write(*, '("REF_TIMING steps=", I0, " seconds=", ES16.8, " threads=", I0, " cells=", I0)') &
        n_steps, dble(c1_clock - c0_clock) / dble(rate), omp_get_max_threads(), n1 * merge(n2, 1, dims == 2)

contains

    ! source_interactions, 1D branch (interactions_mod.f90:38-55)
    subroutine inject_sources_1d()
        integer  :: s, si, i_id
        real(dp) :: c_src
        c_src = m1%dt_eps0 / m1%dr / c0 / 2.0d0
        do s = 1, n_src
            do si = -source_list(s)%n_ker, source_list(s)%n_ker
                i_id = source_list(s)%ind_i(si, 1, 1)
                m1%Ex(i_id) = m1%Ex(i_id) + c_src * source_list(s)%J_mat(si, 1, 1)
            end do
        end do
    end subroutine inject_sources_1d

    ! source_interactions, 2D branch (interactions_mod.f90:57-100)
    subroutine inject_sources_2d()
        integer  :: s, si, sj, nk, i_id, j_id
        real(dp) :: J_av, J_next, c_src
        c_src = m2%dt_eps0 / m2%dr / c0 / 2.0d0
        do s = 1, n_src
            nk = source_list(s)%n_ker
            do si = -nk, nk
            do sj = -nk, nk
                i_id = source_list(s)%ind_i(si, sj, 1)
                j_id = source_list(s)%ind_j(si, sj, 1)
                select case (source_list(s)%dir)
                case ('x')
                    J_next = 0.0d0
                    if (si + 1 <= nk) J_next = source_list(s)%J_mat(si + 1, sj, 1)
                    J_av = 0.5d0 * (source_list(s)%J_mat(si, sj, 1) + J_next)
                    m2%Ex(i_id, j_id) = m2%Ex(i_id, j_id) + c_src * J_av
                case ('y')
                    J_next = 0.0d0
                    if (sj + 1 <= nk) J_next = source_list(s)%J_mat(si, sj + 1, 1)
                    J_av = 0.5d0 * (source_list(s)%J_mat(si, sj, 1) + J_next)
                    m2%Ey(i_id, j_id) = m2%Ey(i_id, j_id) + c_src * J_av
                case ('z')
                    J_av = source_list(s)%J_mat(si, sj, 1)
                    m2%Ez(i_id, j_id) = m2%Ez(i_id, j_id) + c_src * J_av
                end select
            end do
            end do
        end do
    end subroutine inject_sources_2d

end program omx_ref_driver_lowdim
