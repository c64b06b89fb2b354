! commondata.f90 -- run parameters of the lid-driven cavity driver, under the
! names of the reference's module (MPI/Lid_driven_cavity/fortran/3d/
! mpi_3d_blocked/commondata.f90:4-42).  The reference fixes them as
! compile-time parameters; here they keep the reference's values as defaults
! and can be overridden from a namelist file (&cavity ... /) named on the
! command line, so a new configuration needs no rebuild.
module commondata
    implicit none

    integer :: total_nx = 65, total_ny = 65, total_nz = 65
    integer :: nx, ny, nz
    real(8) :: Reynolds = 1000.0d0
    real(8) :: rho0 = 1.0d0
    real(8) :: U0 = 0.1d0
    real(8) :: tauf                         ! commondata.f90:9, set in read_parameters

    integer :: itc
    integer :: itc_max = 50000000
    real(8) :: errorU
    real(8) :: eps = 1e-6
    integer :: check_every = 2000           ! main.f90:99 checks every 2000 steps

    character(len=8) :: collision_model = 'MRT'   ! MRT (reference) or BGK
    character(len=8) :: storage = 'F64'           ! F64, F32 or F16 populations
    integer :: block_x = 0, block_y = 0           ! launch shape, 0 = library default
    logical :: halo_loopback = .false.            ! run the six-face exchange every step, looped back
    character(len=256) :: output_dir = '.'
    logical :: write_tecplot = .true.
    logical :: write_centerlines = .true.
    logical :: write_binary = .false.
    logical :: write_raw = .false.                ! rho,u,v,w as raw FP64 stream (x fastest)
    character(len=256) :: restart_file = ''       ! start from this saved state instead of initial
    character(len=256) :: backup_file = ''        ! save the final state here (exact restart)

    real(8), allocatable :: xp(:), yp(:), zp(:)
    real(8), allocatable :: u(:, :, :), v(:, :, :), w(:, :, :), rho(:, :, :)

    namelist /cavity/ total_nx, total_ny, total_nz, Reynolds, rho0, U0, itc_max, eps, check_every, &
                      collision_model, storage, block_x, block_y, halo_loopback, output_dir, write_tecplot, &
                      write_centerlines, write_binary, write_raw, restart_file, backup_file

contains

    subroutine read_parameters()
        character(len=512) :: path
        integer :: ios, unit
        if (command_argument_count() >= 1) then
            call get_command_argument(1, path)
            open(newunit=unit, file=trim(path), status='old', action='read', iostat=ios)
            if (ios /= 0) then
                write(*,*) 'error: cannot open parameter file ', trim(path)
                error stop 1
            endif
            read(unit, nml=cavity, iostat=ios)
            if (ios /= 0) then
                write(*,*) 'error: bad &cavity namelist in ', trim(path)
                error stop 1
            endif
            close(unit)
        endif
        ! one GPU holds the whole cavity: the subdomain is the domain
        nx = total_nx
        ny = total_ny
        nz = total_nz
        tauf = U0 * dble(total_nx) / Reynolds * 3.0d0 + 0.5d0
    end subroutine read_parameters

end module commondata
