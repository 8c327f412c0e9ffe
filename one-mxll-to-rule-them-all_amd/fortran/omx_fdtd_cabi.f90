! omx_fdtd_cabi.f90 -- iso_c_binding interfaces of libomx_fdtd_hip.so
! (include/omx_fdtd.h).  A thin layer: one interface per C entry point, the
! two argument structs, and a helper that turns a non-zero status into an
! error stop carrying the engine's message.
module omx_fdtd_cabi

    use iso_c_binding
    implicit none

    integer, parameter :: OMX_MAX_POLES = 8

    type, bind(C) :: omx3d_desc
        integer(c_int32_t) :: nx, ny, nz, npml
        integer(c_int32_t) :: boundaries(3)
        integer(c_int32_t) :: precision
        real(c_double)     :: dt, dr
    end type omx3d_desc

    type, bind(C) :: omxl_desc
        integer(c_int32_t) :: dims, mode, n1, n2, npml
        integer(c_int32_t) :: boundaries(2)
        integer(c_int32_t) :: precision
        real(c_double)     :: dt, dr
    end type omxl_desc

    type, bind(C) :: omx_medium
        integer(c_int32_t) :: kind = 0, n_poles = 0
        real(c_double)     :: eps_r = 1.0d0, omega_d = 0.0d0, gamma_d = 0.0d0
        real(c_double)     :: omega_k(OMX_MAX_POLES) = 0.0d0
        real(c_double)     :: gamma_k(OMX_MAX_POLES) = 0.0d0
        real(c_double)     :: a_k(OMX_MAX_POLES) = 0.0d0
    end type omx_medium

    interface
        function omx_last_error() bind(C, name="omx_last_error") result(p)
            import :: c_ptr
            type(c_ptr) :: p
        end function
        integer(c_int) function omx_medium_library(name, eps_r, m) bind(C, name="omx_medium_library")
            import :: c_char, c_double, c_int, omx_medium
            character(kind=c_char), intent(in) :: name(*)
            real(c_double), value :: eps_r
            type(omx_medium), intent(out) :: m
        end function
        integer(c_int) function omx3d_create(desc, ctx) bind(C, name="omx3d_create")
            import :: c_int, c_ptr, omx3d_desc
            type(omx3d_desc), intent(in) :: desc
            type(c_ptr), intent(out) :: ctx
        end function
        integer(c_int) function omx3d_destroy(ctx) bind(C, name="omx3d_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function omx3d_set_media(ctx, n, table, aux) bind(C, name="omx3d_set_media")
            import :: c_int, c_ptr, c_int32_t, c_int8_t, omx_medium
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: n
            type(omx_medium), intent(in) :: table(*)
            integer(c_int8_t), intent(in) :: aux(*)
        end function
        integer(c_int) function omx3d_step_h(ctx) bind(C, name="omx3d_step_h")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function omx3d_step_e(ctx, time, t_skip) bind(C, name="omx3d_step_e")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: time, t_skip
        end function
        integer(c_int) function omx3d_add_source(ctx, dir, j_amp, w0, t0, tau, t_init, t_final, phase, &
                                                 n_ker, ind, ker) bind(C, name="omx3d_add_source")
            import :: c_int, c_ptr, c_double, c_int32_t
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: dir, n_ker
            real(c_double), value :: j_amp, w0, t0, tau, t_init, t_final, phase
            integer(c_int32_t), intent(in) :: ind(*)
            real(c_double), intent(in) :: ker(*)
        end function
        integer(c_int) function omx3d_inject_sources(ctx, time) bind(C, name="omx3d_inject_sources")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: time
        end function
        integer(c_int) function omx3d_add_detector(ctx, det_type, field, nd, idx) bind(C, name="omx3d_add_detector")
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: det_type, field, nd
            integer(c_int32_t), intent(in) :: idx(*)
        end function
        integer(c_int) function omx3d_sample(ctx, out) bind(C, name="omx3d_sample")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: out(*)
        end function
        integer(c_int) function omx3d_set_q_sites(ctx, n, ijk, density, dt_q) bind(C, name="omx3d_set_q_sites")
            import :: c_int, c_ptr, c_double, c_int32_t
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: n
            integer(c_int32_t), intent(in) :: ijk(*)
            real(c_double), value :: density, dt_q
        end function
        integer(c_int) function omx3d_set_q_groups(ctx, n_groups, group_type, n, group_of, ijk, density, dt_q) &
                bind(C, name="omx3d_set_q_groups")
            import :: c_int, c_ptr, c_double, c_int32_t
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: n_groups, n
            integer(c_int32_t), intent(in) :: group_type(*), group_of(*), ijk(*)
            real(c_double), value :: density, dt_q
        end function
        integer(c_int) function omx3d_set_two_level(ctx, omega0, mu, pol, gamma2) bind(C, name="omx3d_set_two_level")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(in) :: omega0(*), mu(*), pol(*), gamma2(*)
        end function
        integer(c_int) function omx3d_two_level_step(ctx) bind(C, name="omx3d_two_level_step")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function omx3d_get_two_level(ctx, dipole, rho) bind(C, name="omx3d_get_two_level")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: dipole(*), rho(*)
        end function
        integer(c_int) function omx3d_gather_e(ctx, e) bind(C, name="omx3d_gather_e")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: e(*)
        end function
        integer(c_int) function omx3d_scatter_j(ctx, dpdt) bind(C, name="omx3d_scatter_j")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(in) :: dpdt(*)
        end function
        integer(c_int) function omx3d_copy_field(ctx, field, host, to_device) bind(C, name="omx3d_copy_field")
            import :: c_int, c_ptr, c_double, c_int32_t
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: field, to_device
            real(c_double), intent(inout) :: host(*)
        end function
        integer(c_int) function omx3d_energy(ctx, w64, w32c) bind(C, name="omx3d_energy")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: w64, w32c
        end function
        integer(c_int) function omx3d_synchronize(ctx) bind(C, name="omx3d_synchronize")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        function c_strlen(s) bind(C, name="strlen") result(n)
            import :: c_ptr, c_size_t
            type(c_ptr), value :: s
            integer(c_size_t) :: n
        end function
        ! 2D / 1D grids
        integer(c_int) function omxl_create(desc, ctx) bind(C, name="omxl_create")
            import :: c_int, c_ptr, omxl_desc
            type(omxl_desc), intent(in) :: desc
            type(c_ptr), intent(out) :: ctx
        end function
        integer(c_int) function omxl_destroy(ctx) bind(C, name="omxl_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function omxl_set_media(ctx, n, table, aux) bind(C, name="omxl_set_media")
            import :: c_int, c_ptr, c_int32_t, c_int8_t, omx_medium
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: n
            type(omx_medium), intent(in) :: table(*)
            integer(c_int8_t), intent(in) :: aux(*)
        end function
        integer(c_int) function omxl_step_h(ctx) bind(C, name="omxl_step_h")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function omxl_step_e(ctx) bind(C, name="omxl_step_e")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function omxl_add_source(ctx, dir, j_amp, w0, t0, tau, t_init, t_final, phase, &
                                                n_ker, ind, ker) bind(C, name="omxl_add_source")
            import :: c_int, c_ptr, c_double, c_int32_t
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: dir, n_ker
            real(c_double), value :: j_amp, w0, t0, tau, t_init, t_final, phase
            integer(c_int32_t), intent(in) :: ind(*)
            real(c_double), intent(in) :: ker(*)
        end function
        integer(c_int) function omxl_inject_sources(ctx, time) bind(C, name="omxl_inject_sources")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: time
        end function
        integer(c_int) function omxl_add_detector(ctx, field, nd, idx) bind(C, name="omxl_add_detector")
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: field, nd
            integer(c_int32_t), intent(in) :: idx(*)
        end function
        integer(c_int) function omxl_sample(ctx, out) bind(C, name="omxl_sample")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: out(*)
        end function
        integer(c_int) function omxl_copy_field(ctx, field, host, to_device) bind(C, name="omxl_copy_field")
            import :: c_int, c_ptr, c_double, c_int32_t
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: field, to_device
            real(c_double), intent(inout) :: host(*)
        end function
        integer(c_int) function omxl_synchronize(ctx) bind(C, name="omxl_synchronize")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
    end interface

contains

    ! Non-zero status -> error stop with the engine's message (the reference
    ! stops with a message on every fatal input error).
    subroutine omx_check(rc, where)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: where
        character(kind=c_char), pointer :: msg(:)
        type(c_ptr) :: p
        integer :: n, i
        character(len=:), allocatable :: text
        if (rc == 0) return
        p = omx_last_error()
        n = int(c_strlen(p))
        call c_f_pointer(p, msg, [n])
        allocate(character(len=n) :: text)
        do i = 1, n
            text(i:i) = msg(i)
        end do
        write (*, '("Error: ", A, " failed (status ", I0, "): ", A)') where, rc, text
        error stop 1
    end subroutine omx_check

end module omx_fdtd_cabi
