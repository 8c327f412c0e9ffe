! omx_mxll_base.f90 -- units and the abstract field-propagation interface.
!
! omx_units_mod holds the atomic-unit constants with the values the
! reference's constants module compiles to (src/FDTD/constants_mod.f90:13-105);
! its default-REAL literals are written here as _sp literals so the same
! single-precision roundings result.
!
! omx_mxll_base_mod declares TMxll with the components and deferred bindings
! of the reference's interface (src/FDTD/mxll_base_mod.f90:7-69): same names,
! argument lists and meanings, so that a caller written against the
! reference's TMxll drives the GPU solver unchanged.
module omx_units_mod

    implicit none

    integer, parameter :: dp = kind(1.0d0)
    integer, parameter :: sp = kind(1.0)
    real(dp), parameter :: pi0 = 3.141592653589793d0
    real(dp), parameter :: c0 = 137.0d0
    real(dp), parameter :: eps0 = 1.0d0 / (4.0d0 * pi0)
    real(dp), parameter :: mu0 = 1.0d0 / (eps0 * c0**2)
    real(dp), parameter :: nm_to_au = real(18.897261259077823_sp, dp)
    real(dp), parameter :: au_to_nm = 1.0d0 / nm_to_au
    real(dp), parameter :: ev_to_au = real(1.0_sp / 27.2114_sp, dp)
    real(dp), parameter :: au_to_ev = real(27.2114_sp, dp)
    real(dp), parameter :: fs_to_au = 1.0d0 / 2.4188843265864d-2
    real(dp), parameter :: au_to_fs = 2.4188843265864d-2
    real(dp), parameter :: a_bh(4) = (/0.353222222d0, -0.488d0, 0.145d0, -0.010222222d0/)

    integer, parameter :: CLOSE_BOUNDARIES = 1, PERIODIC_BOUNDARIES = 2, CPML_BOUNDARIES = 3
    integer, parameter :: NONE_MEDIUM = 0, DIELECTRIC_MEDIUM = 1, DRUDE_MEDIUM = 2, DL_MEDIUM = 3
    integer, parameter :: POINT_DETECTOR = 1, LINE_X_DETECTOR = 2, LINE_Y_DETECTOR = 3, &
                          LINE_Z_DETECTOR = 4, PLANE_XY_DETECTOR = 5, PLANE_YZ_DETECTOR = 6, &
                          PLANE_ZX_DETECTOR = 7, VOLUME_DETECTOR = 8
    integer, parameter :: Ex_FIELD = 1, Ey_FIELD = 2, Ez_FIELD = 3, Hx_FIELD = 4, Hy_FIELD = 5, Hz_FIELD = 6
    integer, parameter :: Q_MATERIAL = 1, Q_SINGLE = 2
    integer, parameter :: PRINT_NONE = 0, PRINT_ALL = 1, PRINT_SELECTED = 2

end module omx_units_mod


module omx_mxll_base_mod

    use omx_units_mod
    implicit none

    type, abstract :: TMxll
        integer  :: dimensions
        integer  :: tq_step
        integer  :: tq_step_old
        integer  :: n_skip_steps
        integer  :: n_media
        integer  :: npml
        integer  :: mode
        real(dp) :: time
        real(dp) :: t_skip
        real(dp) :: dt_eps0
        real(dp) :: dt_mu0
        real(dp) :: dt
        real(dp) :: dr
        integer  :: chunk_coor(3) = 0
    contains
        procedure(init_interface),      deferred :: init
        procedure(kill_interface),      deferred :: kill
        procedure(td_prop_H_interface), deferred :: td_propagate_H_field
        procedure(td_prop_E_interface), deferred :: td_propagate_E_field
    end type TMxll

    abstract interface
        subroutine init_interface(this, grid_Ndims, npml, boundaries, dt, dr, mode, n_media, &
                                  mpi_coords, mpi_dims)
            import :: TMxll, dp
            class(TMxll), intent(inout) :: this
            integer,      intent(in)    :: grid_Ndims(3)
            integer,      intent(in)    :: npml
            integer,      intent(in)    :: boundaries(3)
            integer,      intent(in)    :: mode
            integer,      intent(in)    :: n_media
            integer,      intent(in)    :: mpi_coords(3)
            integer,      intent(in)    :: mpi_dims(3)
            real(dp),     intent(in)    :: dt
            real(dp),     intent(in)    :: dr
        end subroutine init_interface

        subroutine kill_interface(this)
            import :: TMxll
            class(TMxll), intent(inout) :: this
        end subroutine kill_interface

        subroutine td_prop_H_interface(this)
            import :: TMxll
            class(TMxll), intent(inout) :: this
        end subroutine td_prop_H_interface

        subroutine td_prop_E_interface(this, t)
            import :: TMxll
            class(TMxll), intent(inout) :: this
            integer,      intent(in)    :: t
        end subroutine td_prop_E_interface
    end interface

end module omx_mxll_base_mod


! omx_mxll_3d_mod declares TMxll_3D, the shape of the reference's 3D solver
! type (src/FDTD/mxll_3D_mod.f90:9-58) as callers see it through
! `class is (TMxll_3D)`: the grid size and boundaries, and host mirrors of E
! and H with the reference's bounds (mxll_3D_mod.f90:123-131), e.g.
! Ex(0:nx, 1:ny+1, 1:nz+1).  An implementation that keeps the fields on a
! device fills the mirrors in sync_to_host and writes them back in
! sync_to_device; the mirrors are allocated (and zeroed) on first use, so a
! run that never syncs holds no host copy of the grid.
module omx_mxll_3d_mod

    use omx_units_mod
    use omx_mxll_base_mod
    implicit none

    type, abstract, extends(TMxll) :: TMxll_3D
        integer :: nx = 0, ny = 0, nz = 0
        integer :: boundaries(3) = 0
        real(dp), allocatable :: Ex(:, :, :), Ey(:, :, :), Ez(:, :, :)
        real(dp), allocatable :: Hx(:, :, :), Hy(:, :, :), Hz(:, :, :)
    contains
        procedure(sync_interface), deferred :: sync_to_host
        procedure(sync_interface), deferred :: sync_to_device
        procedure :: alloc_mirrors => alloc_mirrors_3D
        procedure :: free_mirrors => free_mirrors_3D
    end type TMxll_3D

    abstract interface
        subroutine sync_interface(this)
            import :: TMxll_3D
            class(TMxll_3D), intent(inout) :: this
        end subroutine sync_interface
    end interface

contains

    subroutine alloc_mirrors_3D(this)
        class(TMxll_3D), intent(inout) :: this
        integer :: nx, ny, nz
        nx = this%nx; ny = this%ny; nz = this%nz
        if (.not. allocated(this%Ex)) then
            allocate(this%Ex(0:nx, ny + 1, nz + 1), this%Ey(nx + 1, 0:ny, nz + 1), this%Ez(nx + 1, ny + 1, 0:nz))
            allocate(this%Hx(nx, 0:ny, 0:nz), this%Hy(0:nx, ny, 0:nz), this%Hz(0:nx, 0:ny, nz))
            this%Ex = 0.0d0; this%Ey = 0.0d0; this%Ez = 0.0d0
            this%Hx = 0.0d0; this%Hy = 0.0d0; this%Hz = 0.0d0
        end if
    end subroutine alloc_mirrors_3D

    subroutine free_mirrors_3D(this)
        class(TMxll_3D), intent(inout) :: this
        if (allocated(this%Ex)) deallocate(this%Ex, this%Ey, this%Ez, this%Hx, this%Hy, this%Hz)
    end subroutine free_mirrors_3D

end module omx_mxll_3d_mod
