! header_harness.f90 -- test harness (CPU only): writes the detector files'
! headers with omx_detector_io_mod for the deck in the current directory
! (detectors.in plus harness.nml: grid, cell size in nm, record stride and
! step count), for the first and the last print step, so that a test can
! compare them byte for byte with the reference's own files.
program header_harness

    use omx_units_mod
    use omx_detector_io_mod
    implicit none

    integer  :: nx = 0, ny = 0, nz = 0, n_det = 0, record_every = 1, steps = 1
    real(dp) :: dx_nm = 1.0d0
    namelist /harness/ nx, ny, nz, n_det, record_every, steps, dx_nm

    type(TDetectorOut), allocatable :: dets(:)
    integer  :: funit, grid(3), nrec, tt_last
    real(dp) :: dr, dt

    open (newunit=funit, file="harness.nml", action="read")
    read (funit, nml=harness)
    close (funit)
    grid = (/nx, ny, nz/)
    dr = dx_nm * nm_to_au
    dt = dr / (2.0 * c0)
    nrec = steps / record_every
    tt_last = nrec * record_every
    call read_detectors(n_det, grid, dr, dets)
    call init_detector_files(dets, grid, dr)
    call write_detector_headers(dets, 0, dt * (DBLE(record_every) - 0.5d0), grid, dr)
    call write_detector_headers(dets, nrec - 1, dt * (DBLE(tt_last) - 0.5d0), grid, dr)

end program header_harness
