! ref_dkt_main.f90 -- test harness (repo-written) that drives the
! reference's own, unmodified sequential particle program
! (MPI/Micro_particles/fortran/case4/seq/CaseIV_DKT.F90: two discs settling
! side by side -- drafting, kissing, tumbling -- in a 201 x 801 D2Q9 channel
! whose window follows them) for nsteps steps in the loop order of its main
! program (CaseIV_DKT.F90:94-115), and writes the state as raw FP64 to
! dkt_state.bin in the working directory:
!   itc, shiftTimes, rhoAvg,
!   xCenter(2), yCenter(2), Uc(2), Vc(2), rationalOmega(2),
!   wallTotalForceX(2), wallTotalForceY(2), totalTorque(2),
!   rho(nx,ny), u(nx,ny), v(nx,ny), obst(1:nx,1:ny), f(0:8,1:nx,1:ny)
! (Fortran order).  The reference's Tecplot output every 500 steps is left
! out (it changes no state).
!
! TEST INFRASTRUCTURE ONLY: it produces golden fixtures; nothing in the
! product calls it.  Entered from ref_dkt_entry.c.
subroutine ref_dkt_run(nsteps) bind(C, name='ref_dkt_run')
    use iso_c_binding
    use commondata
    implicit none
    integer(c_int), value :: nsteps
    integer :: s, i, j

    call initial()
    do s = 1, nsteps
        call collision()
        call streaming()
        call macro()
        call calForce()
        itc = itc + 1
        if (mod(itc, 2000) == 0) call check()
        call updateCenter()
    enddo

    open(10, file='dkt_state.bin', form='unformatted', access='stream', status='replace')
    write(10) dble(itc), dble(shiftTimes), rhoAvg
    write(10) xCenter, yCenter, Uc, Vc, rationalOmega, wallTotalForceX, wallTotalForceY, totalTorque
    write(10) rho, u, v
    write(10) ((dble(obst(i, j)), i = 1, nx), j = 1, ny)
    write(10) f(0:8, 1:nx, 1:ny)
    close(10)
end subroutine ref_dkt_run
