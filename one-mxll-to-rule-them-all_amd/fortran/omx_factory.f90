! omx_factory.f90 -- maxwell_factory, the solver selection of the reference
! (src/FDTD/factory_mod.f90:15-30), returning the GPU solvers of this
! package: dim 3 gives TMxll_3D_hip (a TMxll_3D), dim 2 TMxll_2D_hip (a
! TMxll_2D) and dim 1 TMxll_1D_hip (a TMxll_1D).
module omx_factory_mod

    use omx_mxll_base_mod
    use omx_mxll_3d_hip_mod
    use omx_mxll_lowdim_hip_mod
    implicit none

contains

    function maxwell_factory(dim) result(maxwell_obj)
        integer, intent(in) :: dim
        class(TMxll), allocatable :: maxwell_obj

        select case (dim)
        case (1)
            allocate(TMxll_1D_hip :: maxwell_obj)
        case (2)
            allocate(TMxll_2D_hip :: maxwell_obj)
        case (3)
            allocate(TMxll_3D_hip :: maxwell_obj)
        case default
            error stop "The selected dim, must be 1, 2, or 3."
        end select
    end function maxwell_factory

end module omx_factory_mod
