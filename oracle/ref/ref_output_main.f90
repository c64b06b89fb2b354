! ref_output_main.f90 -- test harness (repo-written) that drives the output
! writers of the reference's own, unmodified sequential 3-D cavity program
! (MPI/Lid_driven_cavity/fortran/3d/seq/lid_driven_cavity_3d.f90:583-723:
! output_binary, output_Tecplot, getVelocity; 65^3, its compile-time grid) on
! given fields: its initial() sets the node coordinates, then rho, u, v, w
! are read from fields.bin (raw FP64, x fastest, in that order) in the
! working directory and the writers run with itc = the argument, writing
! their files there.
!
! TEST INFRASTRUCTURE ONLY: it produces golden fixtures; nothing in the
! product calls it.  Entered from ref_output_entry.c.
subroutine ref_output_run(step) bind(C, name='ref_output_run')
    use iso_c_binding
    use commondata
    implicit none
    integer(c_int), value :: step

    call initial()
    open(10, file='fields.bin', form='unformatted', access='stream', status='old')
    read(10) rho, u, v, w
    close(10)
    itc = step
    call output_binary()
    call output_Tecplot()
    call getVelocity()
end subroutine ref_output_run
