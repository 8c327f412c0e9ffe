! omx_factory.f90 -- maxwell_factory, the solver selection of the reference
! (src/FDTD/factory_mod.f90:15-30), returning the GPU solvers of this
! package: dim 3 gives TMxll_3D_hip (a TMxll_3D).
module omx_factory_mod

    use omx_mxll_base_mod
    use omx_mxll_3d_hip_mod
    implicit none

contains

    function maxwell_factory(dim) result(maxwell_obj)
        integer, intent(in) :: dim
        class(TMxll), allocatable :: maxwell_obj

        select case (dim)
        case (3)
            allocate(TMxll_3D_hip :: maxwell_obj)
        case (1, 2)
            write (*, '("Error: this build has no GPU solver for dim = ", I0)') dim
            error stop 1
        case default
            error stop "The selected dim, must be 1, 2, or 3."
        end select
    end function maxwell_factory

end module omx_factory_mod
