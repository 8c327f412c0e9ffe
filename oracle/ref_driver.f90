! ref_driver.f90 -- test driver that runs the reference's own 3D solver
! modules (compiled from the reference tree by oracle/Makefile, target `ref`)
! to produce golden vectors and the reference CPU timing.  TEST
! INFRASTRUCTURE ONLY.
!
! It follows the driver loop of src/FDTD/mxim_mxll.f90:92-132 with the
! reference's init_3Dgrid / td_propagate_H_3Dfield / td_propagate_E_3Dfield,
! read_init_media_3D (medium_XXXX.in), read_init_sources (sources.in),
! update_sources, init_detector and write_3D_headers / write_3D_field.  The
! steps whose reference code cannot be compiled here (their modules need
! DFTB+) are restated in a few lines each, citing the lines they follow:
!   source_interactions, 3D       interactions_mod.f90:102-156 (J_mat past the
!                                  kernel edge taken as 0)
!   detector output set-up         output_mod.f90:40-168
!   E gather / J scatter (material, non-MPI)
!                                  interactions_mod.f90:656-672, 1067-1069,
!                                  1184-1190, 1216-1218
! Dielectric media get eps = 1/eps_r from the case file after init (the
! reference never stores eps_r for them, classical_medium_mod.f90:62-63).
!
! Input (current directory): case.nml (namelist &case), plus the reference's
! own medium_XXXX.in, sources.in and detectors.in when used, and qsites.in
! ("i j k a_x a_y a_z" per site, plus its 1-based q-group when n_qg > 1:
! dPt/dt(c, tt) = a_c*sin(0.3*tt + c - 1)); q-group g is "material"
! (qg_type(g) = 1) or "single" (2: J over the 7^3 kernel of
! q_group_mod.f90:216-239 around the site).
! Output: coeffs.bin, media.bin, qsites_e.bin (stream doubles) and the
! reference's output_detector_XXXXXXX/ text files.
program omx_ref_driver

    use constants_mod
    use classical_medium_mod
    use mxll_3D_mod
    use source_mod
    use detector_mod
    use write_fields_subs_mod
    use omp_lib

    implicit none

    integer  :: nx = 16, ny = 16, nz = 16, npml = 4, bc(3) = 3
    integer  :: n_steps = 10, record_every = 1, n_media = 0, n_src = 0, n_det = 0, n_q = 0
    integer  :: timing = 0, n_skip = 1, n_qg = 1, qg_type(8) = 1
    real(dp) :: dx_nm = 5.0d0, q_density_nm3 = 1.0d-3
    real(dp) :: diel_eps(15) = 1.0d0
    namelist /case/ nx, ny, nz, npml, bc, n_steps, record_every, n_media, n_src, n_det, n_q, &
                    timing, dx_nm, q_density_nm3, diel_eps, n_skip, n_qg, qg_type

    type(TMxll_3D)               :: mxll
    type(TSource), allocatable   :: source_list(:)
    type(TDetector), allocatable :: detectors(:)
    integer  :: grid_Ndims(3), mpi_coords(3), mpi_dims(3)
    integer  :: tt, i, j, k, n, s, c, funit, ierr, print_det_step
    real(dp) :: dr, dt, time, density
    real(dp) :: x_min, x_max, y_min, y_max, z_min, z_max
    character(len=4)  :: field_name
    character(len=10) :: detector_type
    character(len=20) :: number
    character(len=99) :: dirname
    integer,  allocatable :: q_ijk(:, :), q_grp(:)
    real(dp) :: q_ker(-3:3, -3:3, -3:3)
    real(dp), allocatable :: q_amp(:, :), e_sites(:, :), dpdt(:, :)
    integer(8) :: c0_clock, c1_clock, rate
    integer  :: tt_0
    logical  :: move_q_system
    real(dp) :: dt_q

    open(newunit=funit, file="case.nml", action="read", iostat=ierr)
    if (ierr /= 0) error stop "case.nml missing"
    read(funit, nml=case)
    close(funit)

    grid_Ndims = (/nx, ny, nz/)
    mpi_coords = 0
    mpi_dims = 1
    dr = dx_nm * nm_to_au                                    ! input_mod.f90:121
    dt = MIN(dr / (2.0 * c0), 10.0E10_dp)                    ! input_mod.f90:122
    density = q_density_nm3 / (nm_to_au)**3                  ! input_mod.f90:125

    call read_init_sources(source_list, n_src, 3, dr, grid_Ndims, mpi_coords, mpi_dims)
    call mxll%init(grid_Ndims, npml, bc, dt, dr, 0, n_media, mpi_coords, mpi_dims)

    do n = 1, n_media
        if (mxll%media(n)%medium_type /= DIELECTRIC_MEDIUM) cycle
        do k = 1, nz
        do j = 1, ny
        do i = 1, nx
            if (mxll%media_map(i, j, k, 1) == n) mxll%eps_x(i, j, k) = 1.0d0 / diel_eps(n)
            if (mxll%media_map(i, j, k, 2) == n) mxll%eps_y(i, j, k) = 1.0d0 / diel_eps(n)
            if (mxll%media_map(i, j, k, 3) == n) mxll%eps_z(i, j, k) = 1.0d0 / diel_eps(n)
        end do
        end do
        end do
    end do

    ! ---- golden tables ----------------------------------------------------
    if (timing == 0) then
        open(newunit=funit, file="coeffs.bin", access="stream", form="unformatted", status="replace")
        if (allocated(mxll%be)) then
            write(funit) mxll%be(1:npml), mxll%ce(1:npml), mxll%bh(1:npml-1), mxll%ch(1:npml-1)
        end if
        write(funit) mxll%den_ex, mxll%den_ey, mxll%den_ez, mxll%den_hx, mxll%den_hy, mxll%den_hz
        close(funit)
        open(newunit=funit, file="media.bin", access="stream", form="unformatted", status="replace")
        do n = 1, n_media
            write(funit) mxll%media(n)%A1, mxll%media(n)%A2, mxll%media(n)%C1, mxll%media(n)%C2, &
                         mxll%media(n)%C3, mxll%media(n)%C4
            if (mxll%media(n)%medium_type == DL_MEDIUM) then
                write(funit) mxll%media(n)%omegaD, mxll%media(n)%GammaD, mxll%media(n)%omegaP_i, &
                             mxll%media(n)%GammaP_i, mxll%media(n)%AP_i, mxll%media(n)%alpha_k, &
                             mxll%media(n)%beta_k, mxll%media(n)%gamma_k, mxll%media(n)%B1_k, &
                             mxll%media(n)%B2_k
            end if
        end do
        write(funit) mxll%media_map
        close(funit)
    end if

    ! ---- detectors (output_mod.f90:40-168) --------------------------------
    if (n_det > 0) then
        allocate(detectors(n_det))
        open(newunit=funit, file="detectors.in", action="read")
        do n = 1, n_det
            read(funit, *) detector_type, field_name, x_min, x_max, y_min, y_max, z_min, z_max
            call detectors(n)%init_detector(field_name, detector_type, 3, grid_Ndims, dr, mpi_dims, &
                 mpi_coords, x_min*nm_to_au, x_max*nm_to_au, y_min*nm_to_au, y_max*nm_to_au, &
                 z_min*nm_to_au, z_max*nm_to_au)
        end do
        close(funit)
        if (timing == 0) then
            do n = 1, n_det
                write(number, '(I7.7)') n
                dirname = "./output_detector_" // trim(number)
                call execute_command_line("mkdir -p " // trim(dirname))
                if (detectors(n)%detector_type == POINT_DETECTOR) then
                    ! header with the point's position (output_mod.f90:131-161)
                    x_min = (detectors(n)%i_min*dr - int(mpi_dims(1)*grid_Ndims(1)/2)*dr)*au_to_nm
                    y_min = (detectors(n)%j_min*dr - int(mpi_dims(2)*grid_Ndims(2)/2)*dr)*au_to_nm
                    z_min = (detectors(n)%k_min*dr - int(mpi_dims(3)*grid_Ndims(3)/2)*dr)*au_to_nm
                    open(newunit=funit, file=trim(dirname) // "/point.dat", status="replace", action="write")
                    write(funit, '("# Time (a.u.)              ' // detectors(n)%f_ch // ' (a.u.),      &
                        &x = ", F10.5, " nm,   y = ", F10.5, " nm,   z = ", F10.5, " nm")') x_min, y_min, z_min
                    close(funit)
                end if
            end do
        end if
    end if

    ! ---- quantum sites -----------------------------------------------------
    if (n_q > 0) then
        allocate(q_ijk(n_q, 3), q_amp(n_q, 3), e_sites(n_q, 3), dpdt(n_q, 3), q_grp(n_q))
        q_grp = 1
        open(newunit=funit, file="qsites.in", action="read")
        do n = 1, n_q
            if (n_qg > 1) then
                read(funit, *) q_ijk(n, :), q_amp(n, :), q_grp(n)
            else
                read(funit, *) q_ijk(n, :), q_amp(n, :)
            end if
        end do
        close(funit)
        call single_kernel()
        if (timing == 0) then
            open(unit=77, file="qsites_e.bin", access="stream", form="unformatted", status="replace")
            open(unit=78, file="qmoves.bin", access="stream", form="unformatted", status="replace")
        end if
    end if

    ! ---- time loop (mxim_mxll.f90:73-132): the quantum sub-step bookkeeping
    ! with n_skip_steps given directly (the reference derives it from dt_q,
    ! mxim_mxll.f90:73-74), the quantum step dt_q = n_skip_steps*dt (:79)
    mxll%n_skip_steps = n_skip
    tt_0 = 0
    if (mxll%n_skip_steps == 1) tt_0 = 1
    mxll%tq_step_old = 0
    dt_q = dble(mxll%n_skip_steps) * dt
    print_det_step = 0
    call system_clock(c0_clock, rate)
    do tt = 1, n_steps
        time = dt * (DBLE(tt) - 0.5d0)
        mxll%time = time
        mxll%tq_step = 1 + int((tt - tt_0) / mxll%n_skip_steps)
        move_q_system = .false.
        if (mxll%tq_step > mxll%tq_step_old) then
            mxll%t_skip = time
            mxll%tq_step_old = mxll%tq_step
            move_q_system = .true.
            if (n_q == 0) move_q_system = .false.
        end if

        call mxll%td_propagate_H_field()
        call update_sources(source_list, n_src, time)
        call inject_sources()
        call mxll%td_propagate_E_field(tt)

        if (move_q_system) then
            call gather_e()
            if (timing == 0) write(77) e_sites
            if (timing == 0) write(78) tt
            do c = 1, 3
                dpdt(:, c) = q_amp(:, c) * sin(0.3d0 * tt + (c - 1))
            end do
            call scatter_j()
        end if

        if (timing == 0 .and. n_det > 0 .and. record_every > 0) then
            if (mod(tt, record_every) == 0) then
                call write_3D_headers(detectors, n_det, print_det_step, time, dr, grid_Ndims, mpi_dims, 0)
                call write_3D_field(detectors, mxll, n_det, print_det_step, time, dr, grid_Ndims, mpi_dims, &
                                    mpi_coords, 0)
                print_det_step = print_det_step + 1
            end if
        end if
    end do
    call system_clock(c1_clock)
    if (n_q > 0 .and. timing == 0) close(77)
    if (n_q > 0 .and. timing == 0) close(78)
    write(*, '("REF_TIMING steps=", I0, " seconds=", ES16.8, " threads=", I0, " cells=", I0)') &
        n_steps, dble(c1_clock - c0_clock) / dble(rate), omp_get_max_threads(), nx * ny * nz
    call mxll%kill()

contains

    ! source_interactions, 3D branch (interactions_mod.f90:102-156)
    subroutine inject_sources()
        integer  :: si, sj, sk, nk, i_id, j_id, k_id
        real(dp) :: J_av, J_next, c_src
        c_src = mxll%dt_eps0 / mxll%dr / c0 / 2.0d0
        do s = 1, n_src
            nk = source_list(s)%n_ker
            do si = -nk, nk
            do sj = -nk, nk
            do sk = -nk, nk
                i_id = source_list(s)%ind_i(si, sj, sk)
                j_id = source_list(s)%ind_j(si, sj, sk)
                k_id = source_list(s)%ind_k(si, sj, sk)
                select case (source_list(s)%dir)
                case ('x')
                    J_next = 0.0d0
                    if (si + 1 <= nk) J_next = source_list(s)%J_mat(si + 1, sj, sk)
                    J_av = 0.5d0 * (source_list(s)%J_mat(si, sj, sk) + J_next)
                    mxll%Ex(i_id, j_id, k_id) = mxll%Ex(i_id, j_id, k_id) + c_src * J_av
                case ('y')
                    J_next = 0.0d0
                    if (sj + 1 <= nk) J_next = source_list(s)%J_mat(si, sj + 1, sk)
                    J_av = 0.5d0 * (source_list(s)%J_mat(si, sj, sk) + J_next)
                    mxll%Ey(i_id, j_id, k_id) = mxll%Ey(i_id, j_id, k_id) + c_src * J_av
                case ('z')
                    J_next = 0.0d0
                    if (sk + 1 <= nk) J_next = source_list(s)%J_mat(si, sj, sk + 1)
                    J_av = 0.5d0 * (source_list(s)%J_mat(si, sj, sk) + J_next)
                    mxll%Ez(i_id, j_id, k_id) = mxll%Ez(i_id, j_id, k_id) + c_src * J_av
                end select
            end do
            end do
            end do
        end do
    end subroutine inject_sources

    ! the 3D "single" group kernel, q_group_mod.f90:216-239 (n_ker = 3; its
    ! norm is summed there but never applied)
    subroutine single_kernel()
        integer  :: i, j, k
        real(dp) :: r_max, r0, x, y, z, r
        r_max = dble(3*2+4)
        r0    = r_max/2.0d0
        do i = -3, 3
        do j = -3, 3
        do k = -3, 3
            x = dble(i)
            y = dble(j)
            z = dble(k)
            r = SQRT(x**2 + y**2 + z**2)
            q_ker(i,j,k) = aBH(1)+ &
                  aBH(2)*DCOS(2.0*pi0*(r-r0)/r_max)+ &
                  aBH(3)*DCOS(2.0*pi0*2.0*(r-r0)/r_max)+ &
                  aBH(4)*DCOS(2.0*pi0*3.0*(r-r0)/r_max)
            if (q_ker(i,j,k) < M_ZERO) q_ker(i,j,k) = M_ZERO
        end do
        end do
        end do
    end subroutine single_kernel

    ! send_E_3D_to_J_ranks, non-MPI (interactions_mod.f90:656-691): at the
    ! system's cell, a single group's kernel centre
    subroutine gather_e()
        integer :: q, ii, jj, kk
        do q = 1, n_q
            ii = q_ijk(q, 1); jj = q_ijk(q, 2); kk = q_ijk(q, 3)
            e_sites(q, 1) = 0.5d0 * (mxll%Ex(ii - 1, jj, kk) + mxll%Ex(ii, jj, kk))
            e_sites(q, 2) = 0.5d0 * (mxll%Ey(ii, jj - 1, kk) + mxll%Ey(ii, jj, kk))
            e_sites(q, 3) = 0.5d0 * (mxll%Ez(ii, jj, kk - 1) + mxll%Ez(ii, jj, kk))
        end do
    end subroutine gather_e

    ! send_J_to_E_3D_ranks, non-MPI, once per q-group in group order
    ! (interactions_mod.f90:1067-1069,1179-1218; mxim_mxll.f90:117-121)
    subroutine scatter_j()
        integer :: q, g, ii, jj, kk, i, j, k
        do g = 1, n_qg
            mxll%Jx_old = mxll%Jx
            mxll%Jy_old = mxll%Jy
            mxll%Jz_old = mxll%Jz
            do q = 1, n_q
                if (q_grp(q) /= g) cycle
                ii = q_ijk(q, 1); jj = q_ijk(q, 2); kk = q_ijk(q, 3)
                if (qg_type(g) == 1) then
                    mxll%Jx(ii, jj, kk) = density * dpdt(q, 1)
                    mxll%Jy(ii, jj, kk) = density * dpdt(q, 2)
                    mxll%Jz(ii, jj, kk) = density * dpdt(q, 3)
                else
                    do k = -3, 3
                    do j = -3, 3
                    do i = -3, 3
                        mxll%Jx(ii + i, jj + j, kk + k) = density*q_ker(i,j,k)*dpdt(q, 1)
                        mxll%Jy(ii + i, jj + j, kk + k) = density*q_ker(i,j,k)*dpdt(q, 2)
                        mxll%Jz(ii + i, jj + j, kk + k) = density*q_ker(i,j,k)*dpdt(q, 3)
                    end do
                    end do
                    end do
                end if
            end do
            mxll%dJx = (mxll%Jx - mxll%Jx_old) / dt_q
            mxll%dJy = (mxll%Jy - mxll%Jy_old) / dt_q
            mxll%dJz = (mxll%Jz - mxll%Jz_old) / dt_q
        end do
    end subroutine scatter_j

end program omx_ref_driver
