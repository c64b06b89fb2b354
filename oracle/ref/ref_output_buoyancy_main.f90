! ref_output_buoyancy_main.f90 -- test harness (repo-written) that drives the
! output writers of the reference's own, unmodified sequential 3-D buoyancy
! program (MPI/Buoyancy_driven_cavity/fortran/3d/seq/bouyancy3d.F90:984-1164:
! output_binary, output_Tecplot, with the benchmarkCavity naming its file
! defines; 51^3, its compile-time grid) on given fields: its initial() sets
! the node coordinates, then u, v, w, T are read from fields.bin (raw FP64,
! x fastest, in that order) in the working directory and the writers run
! with itc = the argument, writing their files there.
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
    read(10) u, v, w, T
    close(10)
    itc = step
    call output_binary()
    call output_Tecplot()
end subroutine ref_output_run
