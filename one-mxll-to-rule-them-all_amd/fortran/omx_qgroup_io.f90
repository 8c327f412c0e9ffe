! omx_qgroup_io.f90 -- the reference's quantum-group input and output files.
!
!   mol_group_XXXXXXX.in  (one per group, q_group_mod.f90:98-298): a line
!                         "group_type q_sys_type n_systems print_option"
!                         (material | single; print_none | print_all |
!                         print_selected), then per system "mol_id id_file x
!                         y z" in nm, plus "print_on | print_off" with
!                         print_selected.  A system sits in cell
!                         int(x*nm_to_au/dr) + int(N/2) per axis (the kernel
!                         centre of a single group, :482-514, :717-734).
!   output_q_group_XXXX/energy_and_dipole.dat
!                         at every print step a "# Time" line and a column
!                         line, then "mol_id energy(eV) dipole_x y z" per
!                         printed system (output_mod.f90:236-388).
!
! The reference's only quantum system is DFTB+ (q_sys_dftb_mod.f90), which
! this build does not have.  Its systems here are the two-level emitters of
! the engine (the TQ_sys_base contract: a dipole and dPt/dt = (dipole -
! dip_old)/dt_q per quantum step), q_sys_type "two_level", each reading
! two_level_XXXXXXX.in (its id_file): "omega0 mu pol_x pol_y pol_z gamma2",
! omega0 and gamma2 in eV, mu in a.u.  Their energy is <H0> =
! (omega0/2)(rho_ee - rho_gg).
module omx_qgroup_io_mod

    use omx_units_mod
    implicit none
    private

    public :: TQGroupIn, read_q_groups, read_two_level, init_q_group_outputs, write_q_group_outputs

    type :: TQGroupIn
        integer :: group_type = 0
        integer :: n_systems = 0
        integer :: print_option = PRINT_NONE
        integer :: first = 0                      ! its first system in the list of all groups' systems
        integer, allocatable :: mol_id(:), id_file(:), ijk(:, :)
        logical, allocatable :: print_on(:)
    end type TQGroupIn

contains

    subroutine read_q_groups(n_groups, grid, dr, groups)
        integer,  intent(in) :: n_groups, grid(3)
        real(dp), intent(in) :: dr
        type(TQGroupIn), allocatable, intent(out) :: groups(:)
        character(len=20) :: group_type_str, q_sys_type_str, number, print_on_ch
        character(len=50) :: print_option_ch, input_name
        real(dp) :: coord(3)
        integer :: g, i, a, funit, ierr, n, first

        allocate(groups(n_groups))
        first = 1
        do g = 1, n_groups
            write (number, '(I7.7)') g
            input_name = "mol_group_" // trim(number) // ".in"
            open (newunit=funit, file=trim(input_name), action="read", iostat=ierr)
            if (ierr /= 0) then
                write (*, *) "Error: Could not open file ", trim(input_name)
                error stop 1
            end if
            read (funit, *, iostat=ierr) group_type_str, q_sys_type_str, n, print_option_ch
            if (ierr /= 0 .or. n < 0) then
                write (*, *) "Error: invalid first line in ", trim(input_name)
                error stop 1
            end if
            select case (trim(print_option_ch))
            case ("print_none");     groups(g)%print_option = PRINT_NONE
            case ("print_all");      groups(g)%print_option = PRINT_ALL
            case ("print_selected"); groups(g)%print_option = PRINT_SELECTED
            case default
                write (*, *) "Error: Unknown print option for the q_system ", trim(print_option_ch)
                error stop 1
            end select
            select case (trim(group_type_str))
            case ("material"); groups(g)%group_type = Q_MATERIAL
            case ("single");   groups(g)%group_type = Q_SINGLE
            case default
                write (*, *) "Error: Unknown Q_group type ", trim(group_type_str)
                error stop 1
            end select
            select case (trim(q_sys_type_str))
            case ("two_level")
            case ("dftb")
                write (*, *) "Error: DFTB+ quantum systems are not part of this build (use two_level)"
                error stop 1
            case default
                write (*, *) "Error: Unknown Q_sys type ", trim(q_sys_type_str)
                error stop 1
            end select
            groups(g)%n_systems = n
            groups(g)%first = first
            first = first + n
            allocate(groups(g)%mol_id(n), groups(g)%id_file(n), groups(g)%ijk(n, 3), groups(g)%print_on(n))
            groups(g)%print_on = groups(g)%print_option == PRINT_ALL
            do i = 1, n
                if (groups(g)%print_option == PRINT_SELECTED) then
                    read (funit, *, iostat=ierr) groups(g)%mol_id(i), groups(g)%id_file(i), coord, print_on_ch
                    if (trim(print_on_ch) == "print_on") then
                        groups(g)%print_on(i) = .true.
                    else if (trim(print_on_ch) == "print_off") then
                        groups(g)%print_on(i) = .false.
                    else
                        write (*, *) "Error: Unknown print option for the q_system ", i
                        error stop 1
                    end if
                else
                    read (funit, *, iostat=ierr) groups(g)%mol_id(i), groups(g)%id_file(i), coord
                end if
                if (ierr /= 0) then
                    write (*, *) "Error: invalid system line ", i, " in ", trim(input_name)
                    error stop 1
                end if
                coord = coord * nm_to_au
                do a = 1, 3
                    groups(g)%ijk(i, a) = int(coord(a) / dr) + int(grid(a) / 2)
                end do
            end do
            close (funit)
        end do
    end subroutine read_q_groups

    ! Two-level parameters of every system (groups in order), from
    ! two_level_XXXXXXX.in of its id_file; converted to a.u.
    subroutine read_two_level(groups, omega0, mu, pol, gamma2)
        type(TQGroupIn), intent(in) :: groups(:)
        real(dp), allocatable, intent(out) :: omega0(:), mu(:), pol(:, :), gamma2(:)
        character(len=20) :: number
        character(len=60) :: name
        integer :: g, i, q, n, funit, ierr
        real(dp) :: v(6)

        n = 0
        do g = 1, size(groups)
            n = n + groups(g)%n_systems
        end do
        allocate(omega0(n), mu(n), pol(n, 3), gamma2(n))
        do g = 1, size(groups)
            do i = 1, groups(g)%n_systems
                q = groups(g)%first + i - 1
                write (number, '(I7.7)') groups(g)%id_file(i)
                name = "two_level_" // trim(number) // ".in"
                open (newunit=funit, file=trim(name), action="read", iostat=ierr)
                if (ierr /= 0) then
                    write (*, *) "Error: Could not open file ", trim(name)
                    error stop 1
                end if
                read (funit, *, iostat=ierr) v
                if (ierr /= 0) then
                    write (*, *) "Error: invalid two-level parameters in ", trim(name)
                    error stop 1
                end if
                close (funit)
                omega0(q) = v(1) * ev_to_au
                mu(q) = v(2)
                pol(q, :) = v(3:5)
                gamma2(q) = v(6) * ev_to_au
            end do
        end do
    end subroutine read_two_level

    subroutine init_q_group_outputs(groups)
        type(TQGroupIn), intent(in) :: groups(:)
        character(len=4) :: number
        integer :: g, ierr
        do g = 1, size(groups)
            if (groups(g)%print_option /= PRINT_NONE) then
                write (number, '(I4.4)') g
                call execute_command_line("mkdir -p ./output_q_group_" // number, exitstat=ierr)
                if (ierr /= 0) then
                    write (*, *) "Error: Could not create directory ", "./output_q_group_" // number
                    error stop 1
                end if
            else
                write (*, *) "Output for q_group ", g, " is turned off."
            end if
        end do
    end subroutine init_q_group_outputs

    ! One print step of energy_and_dipole.dat for every printed group:
    ! energy(q) in a.u., dipole(q,:) of system q of the list of all systems.
    subroutine write_q_group_outputs(groups, print_q_step, time, energy, dipole)
        type(TQGroupIn), intent(in) :: groups(:)
        integer,  intent(in) :: print_q_step
        real(dp), intent(in) :: time, energy(:), dipole(:, :)
        character(len=4) :: number
        character(len=99) :: name
        integer :: g, i, q, funit, ierr
        do g = 1, size(groups)
            if (groups(g)%print_option == PRINT_NONE) cycle
            write (number, '(I4.4)') g
            name = "./output_q_group_" // number // "/energy_and_dipole.dat"
            if (print_q_step == 0) then
                open (newunit=funit, file=trim(name), status="replace", action="write", iostat=ierr)
            else
                open (newunit=funit, file=trim(name), status="old", action="write", position="append", iostat=ierr)
            end if
            if (ierr /= 0) then
                write (*, *) "Error: Could not open file ", trim(name)
                error stop 1
            end if
            write (funit, '("# Time = ", ES18.8, " (a.u.)")') time
            write (funit, '("# N_mol        Energy (eV)                   Dipole_x (a.u.)             ' // &
                          'Dipole_y (a.u.)             Dipole_z (a.u.)")')
            do i = 1, groups(g)%n_systems
                if (.not. groups(g)%print_on(i)) cycle
                q = groups(g)%first + i - 1
                write (funit, *) groups(g)%mol_id(i), energy(q) * au_to_ev, dipole(q, 1), dipole(q, 2), dipole(q, 3)
            end do
            close (funit)
        end do
    end subroutine write_q_group_outputs

end module omx_qgroup_io_mod
