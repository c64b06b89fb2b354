! ref_ranu_main.f90 -- test harness (repo-written) that drives the
! reference's own, unmodified Nusselt / Reynolds statistics
! (MPI/Buoyancy_driven_cavity/fortran/3d/mpi_blocked/RaNu.F90, calNuRe, with
! that directory's module.F90) over a sequence of given fields, one call per
! time unit as the reference's main loop makes it.
!
! TEST INFRASTRUCTURE ONLY: it produces golden fixtures; nothing in the
! product calls it.  diffusivity, viscosity and timeUnit are the reference's
! compile-time parameters (module.F90:18-31); only the grid and the fields vary.
!
! usage: ref_ranu nx ny nz ntimes itc_step fields.bin out.bin
!   fields.bin: ntimes records of u, v, w, T (each nx*ny*nz FP64, x fastest);
!   before call it, itc = it * itc_step.  Nu_Bulk.dat, Re_Bulk.dat and
!   statistically.txt are written in the working directory by calNuRe.
!   out.bin: diffusivity, viscosity, timeUnit, then per call it = 1..ntimes:
!   NuVolAvg(it), ReVolAvg(it), NuVolAvg_mean(it), ReVolAvg_mean(it) and, after
!   the call, statisticallyStationaryState, statisticallyStationaryTime,
!   statisticallyConvergeState (as FP64).
program ref_ranu_main
    use commondata
    implicit none
    integer :: ntimes, itc_step, it
    character(len=256) :: arg, fields_file, out_file
    real(8), allocatable :: rec(:, :)

    call get_command_argument(1, arg); read(arg, *) nx
    call get_command_argument(2, arg); read(arg, *) ny
    call get_command_argument(3, arg); read(arg, *) nz
    call get_command_argument(4, arg); read(arg, *) ntimes
    call get_command_argument(5, arg); read(arg, *) itc_step
    call get_command_argument(6, fields_file)
    call get_command_argument(7, out_file)

    allocate(u(nx, ny, nz), v(nx, ny, nz), w(nx, ny, nz), T(nx, ny, nz))
    allocate(rec(7, ntimes))
    dimensionlessTime = 0
    statisticallyStationaryState = 0
    statisticallyStationaryTime = 0
    statisticallyConvergeState = 0
    NuVolAvg = 0.0d0
    ReVolAvg = 0.0d0
    NuVolAvg_mean = 0.0d0
    ReVolAvg_mean = 0.0d0

    open(10, file=trim(fields_file), form='unformatted', access='stream', status='old')
    do it = 1, ntimes
        read(10) u, v, w, T
        itc = it * itc_step
        call calNuRe()
        rec(1, it) = NuVolAvg(it)
        rec(2, it) = ReVolAvg(it)
        rec(3, it) = NuVolAvg_mean(it)
        rec(4, it) = ReVolAvg_mean(it)
        rec(5, it) = dble(statisticallyStationaryState)
        rec(6, it) = dble(statisticallyStationaryTime)
        rec(7, it) = dble(statisticallyConvergeState)
    enddo
    close(10)

    open(11, file=trim(out_file), form='unformatted', access='stream', status='replace')
    write(11) diffusivity, viscosity, timeUnit
    write(11) rec
    close(11)
end program ref_ranu_main
