! ref_thermal_main.f90 -- test harness (repo-written) that drives the
! reference's own, unmodified 3-D buoyancy-cavity subroutines
! (MPI/Buoyancy_driven_cavity/fortran/3d/mpi_blocked/{module,initial,
! evolution_f,evolution_g}.F90, built with the reference Makefile's macros)
! on one subdomain that is the whole cavity, in the step order of
! bouyancy3d_mpi.F90:222-248, and writes the parameters and rho, u, v, w, T
! as raw FP64 (x fastest).
!
! TEST INFRASTRUCTURE ONLY: it produces golden fixtures.  Rayleigh, Prandtl,
! Mach, Ekman and the wall temperatures are the reference's compile-time
! parameters (module.F90:11-33); only nx, ny, nz and the step count vary.
!
! Prints 'ref_loop_seconds <s>', the wall time of the step loop.
!
! usage: ref_thermal nx ny nz nsteps outfile
program ref_thermal_main
    use commondata
    implicit none
    integer :: nsteps, s
    integer(8) :: t0, t1, rate
    character(len=256) :: arg, outfile

    call get_command_argument(1, arg); read(arg, *) nx
    call get_command_argument(2, arg); read(arg, *) ny
    call get_command_argument(3, arg); read(arg, *) nz
    call get_command_argument(4, arg); read(arg, *) nsteps
    call get_command_argument(5, outfile)

    rank = 0
    dims = 1
    coords = 0

    call initial()
    call system_clock(t0, rate)
    do s = 1, nsteps
        itc = itc + 1
        call collision()
        call streaming()
        call bounceback()
        call collisionT()
        call streamingT()
        call bouncebackT()
        call macro()
        call macroT()
    enddo
    call system_clock(t1)
    ! the loop's wall time, for the CPU baseline (initial and output excluded)
    write(*, '(a, es24.16)') 'ref_loop_seconds ', dble(t1 - t0) / dble(rate)

    open(10, file=trim(outfile), form='unformatted', access='stream', status='replace')
    write(10) tauf, paraA, gBeta, omegaRatating, Thot, Tcold, Tref, diffusivity
    write(10) rho, u, v, w, T
    close(10)
end program ref_thermal_main
