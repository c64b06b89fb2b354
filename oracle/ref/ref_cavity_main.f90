! ref_cavity_main.f90 -- test harness (repo-written) that drives the
! reference's own, unmodified 3-D cavity subroutines
! (MPI/Lid_driven_cavity/fortran/3d/mpi_3d_blocked/{commondata,initial,
! collision,streaming,bounce_back,macro}.f90) on one subdomain that is the
! whole cavity, in the loop order of main.f90:85-103, and writes the
! resulting rho, u, v, w as raw FP64 (x fastest) to a file.
!
! TEST INFRASTRUCTURE ONLY: it produces golden fixtures; nothing in the
! product calls it.  tauf, U0 and Reynolds are the reference's compile-time
! parameters (commondata.f90:4-9); only nx, ny, nz and the step count vary.
!
! Prints 'ref_loop_seconds <s>', the wall time of the step loop.
!
! usage: ref_cavity nx ny nz nsteps outfile
program ref_cavity_main
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

    ! one rank owning every wall (bounce_back.f90 tests coords against dims)
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
        call macro()
    enddo
    call system_clock(t1)
    ! the loop's wall time, for the CPU baseline (initial and output excluded)
    write(*, '(a, es24.16)') 'ref_loop_seconds ', dble(t1 - t0) / dble(rate)

    open(10, file=trim(outfile), form='unformatted', access='stream', status='replace')
    write(10) tauf, U0, rho0
    write(10) rho, u, v, w
    close(10)
end program ref_cavity_main
