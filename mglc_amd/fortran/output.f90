! output.f90 -- result files of the cavity driver in the reference's formats
! (MPI/Lid_driven_cavity/fortran/3d/mpi_3d_blocked/output.f90):
!   MRTcavity-<itc:9>.plt   Tecplot binary #!TDV101, point packed, float32
!                           X, Y, Z, U, V, W, Pressure = rho/3  (output.f90:173-300)
!   u-z_<itc:9>.dat         u(mid, mid, z)/U0 against z/nz       (output.f90:318-340)
!   x-w_<itc:9>.dat         x/nx against w(x, mid, mid)/U0
!   MRTcavity-<itc>.bin     Fortran sequential records u, v, rho (output.f90:350-366)
!   cavity-raw-<itc:9>.bin  raw FP64 stream rho, u, v, w, x fastest (this driver's
!                           own, for exact comparisons)
! and of the buoyancy driver (Buoyancy_driven_cavity/fortran/3d/mpi_blocked/
! output.F90):
!   buoyancyCavity-<itc>.plt  Tecplot binary, X, Y, Z, U, V, W, T   (output.F90:151-283)
!   buoyancyCavity-<itc>.bin  Fortran sequential records u, v, w, T (output.F90:121-146)
!   buoyancy-raw-<itc:9>.bin  raw FP64 stream rho, u, v, w, T
! One GPU owns the whole cavity, so the rank-0 gather of output.f90:1-133
! is a single download.
module cavity_output
    implicit none
    private
    public :: output_Tecplot, getVelocity, output_binary, output_raw
    public :: output_Tecplot_buoyancy, output_binary_buoyancy, output_raw_buoyancy

contains

    function step_tag(itc) result(tag)
        integer, intent(in) :: itc
        character(len=9) :: tag
        write(tag, '(i9.9)') itc
    end function step_tag

    subroutine put_string(unit, s)
        integer, intent(in) :: unit
        character(len=*), intent(in) :: s
        integer :: q
        do q = 1, len_trim(s)
            write(unit) ichar(s(q:q))
        enddo
        write(unit) 0
    end subroutine put_string

    ! Tecplot binary #!TDV101, one ordered zone, point packing, float32: the
    ! coordinates, the velocity and a seventh variable s
    subroutine tecplot_point(path, name7, xp, yp, zp, u, v, w, s, nx, ny, nz)
        character(len=*), intent(in) :: path, name7
        integer, intent(in) :: nx, ny, nz
        real(8), intent(in) :: xp(0:nx+1), yp(0:ny+1), zp(0:nz+1)
        real(8), intent(in) :: u(nx, ny, nz), v(nx, ny, nz), w(nx, ny, nz), s(nx, ny, nz)
        integer :: i, j, k, unit, q
        real(4), parameter :: zone_marker = 299.0, eoh_marker = 357.0
        character(len=8) :: names(7)
        names = ['X       ', 'Y       ', 'Z       ', 'U       ', 'V       ', 'W       ', '        ']
        names(7) = name7

        open(newunit=unit, file=path, access='stream', form='unformatted', status='replace')
        write(unit) '#!TDV101'
        write(unit) 1                       ! byte order marker
        call put_string(unit, 'MyFirst')    ! title
        write(unit) 7                       ! number of variables
        do q = 1, 7
            call put_string(unit, names(q))
        enddo
        write(unit) zone_marker
        call put_string(unit, 'ZONE 001')
        write(unit) -1                      ! zone colour
        write(unit) 0                       ! zone type: ordered
        write(unit) 1                       ! data packing: point
        write(unit) 0                       ! variable location not specified
        write(unit) 0                       ! no face neighbour connections
        write(unit) nx
        write(unit) ny
        write(unit) nz
        write(unit) 0                       ! no auxiliary data
        write(unit) eoh_marker
        write(unit) zone_marker
        do q = 1, 7
            write(unit) 1                   ! float
        enddo
        write(unit) 0                       ! no variable sharing
        write(unit) -1                      ! no connectivity sharing
        do k = 1, nz
            do j = 1, ny
                do i = 1, nx
                    write(unit) real(xp(i)), real(yp(j)), real(zp(k)), real(u(i,j,k)), &
                                real(v(i,j,k)), real(w(i,j,k)), real(s(i,j,k))
                enddo
            enddo
        enddo
        close(unit)
    end subroutine tecplot_point

    subroutine output_Tecplot(dir, xp, yp, zp, u, v, w, rho, nx, ny, nz, itc)
        character(len=*), intent(in) :: dir
        integer, intent(in) :: nx, ny, nz, itc
        real(8), intent(in) :: xp(0:nx+1), yp(0:ny+1), zp(0:nz+1)
        real(8), intent(in) :: u(nx, ny, nz), v(nx, ny, nz), w(nx, ny, nz), rho(nx, ny, nz)
        call tecplot_point(trim(dir)//'/MRTcavity-'//step_tag(itc)//'.plt', 'Pressure', xp, yp, zp, u, v, w, &
                           rho / 3.0d0, nx, ny, nz)
    end subroutine output_Tecplot

    subroutine output_Tecplot_buoyancy(dir, xp, yp, zp, u, v, w, T, nx, ny, nz, itc)
        character(len=*), intent(in) :: dir
        integer, intent(in) :: nx, ny, nz, itc
        real(8), intent(in) :: xp(0:nx+1), yp(0:ny+1), zp(0:nz+1)
        real(8), intent(in) :: u(nx, ny, nz), v(nx, ny, nz), w(nx, ny, nz), T(nx, ny, nz)
        character(len=32) :: tag
        write(tag, *) itc
        call tecplot_point(trim(dir)//'/buoyancyCavity-'//trim(adjustl(tag))//'.plt', 'T', xp, yp, zp, u, v, w, &
                           T, nx, ny, nz)
    end subroutine output_Tecplot_buoyancy

    subroutine getVelocity(dir, xp, zp, u, w, nx, ny, nz, U0, itc)
        character(len=*), intent(in) :: dir
        integer, intent(in) :: nx, ny, nz, itc
        real(8), intent(in) :: xp(0:nx+1), zp(0:nz+1)
        real(8), intent(in) :: u(nx, ny, nz), w(nx, ny, nz), U0
        integer :: i, k, unit, xh, yh, zh
        xh = (nx - 1) / 2 + 1
        yh = (ny - 1) / 2 + 1
        zh = (nz - 1) / 2 + 1
        open(newunit=unit, file=trim(dir)//'/u-z_'//step_tag(itc)//'.dat', status='replace')
        do k = 1, nz
            write(unit, *) u(xh, yh, k) / U0, zp(k) / dble(nz)
        enddo
        close(unit)
        open(newunit=unit, file=trim(dir)//'/x-w_'//step_tag(itc)//'.dat', status='replace')
        do i = 1, nx
            write(unit, *) xp(i) / dble(nx), w(i, yh, zh) / U0
        enddo
        close(unit)
    end subroutine getVelocity

    subroutine output_binary(dir, u, v, rho, nx, ny, nz, itc)
        character(len=*), intent(in) :: dir
        integer, intent(in) :: nx, ny, nz, itc
        real(8), intent(in) :: u(nx, ny, nz), v(nx, ny, nz), rho(nx, ny, nz)
        integer :: unit
        character(len=32) :: tag
        write(tag, *) itc
        open(newunit=unit, file=trim(dir)//'/MRTcavity-'//trim(adjustl(tag))//'.bin', form='unformatted', &
             access='sequential', status='replace')
        write(unit) u
        write(unit) v
        write(unit) rho
        close(unit)
    end subroutine output_binary

    subroutine output_raw(dir, rho, u, v, w, nx, ny, nz, itc)
        character(len=*), intent(in) :: dir
        integer, intent(in) :: nx, ny, nz, itc
        real(8), intent(in) :: rho(nx, ny, nz), u(nx, ny, nz), v(nx, ny, nz), w(nx, ny, nz)
        integer :: unit
        open(newunit=unit, file=trim(dir)//'/cavity-raw-'//step_tag(itc)//'.bin', form='unformatted', &
             access='stream', status='replace')
        write(unit) rho, u, v, w
        close(unit)
    end subroutine output_raw

    subroutine output_binary_buoyancy(dir, u, v, w, T, nx, ny, nz, itc)
        character(len=*), intent(in) :: dir
        integer, intent(in) :: nx, ny, nz, itc
        real(8), intent(in) :: u(nx, ny, nz), v(nx, ny, nz), w(nx, ny, nz), T(nx, ny, nz)
        integer :: unit
        character(len=32) :: tag
        write(tag, *) itc
        open(newunit=unit, file=trim(dir)//'/buoyancyCavity-'//trim(adjustl(tag))//'.bin', form='unformatted', &
             access='sequential', status='replace')
        write(unit) u
        write(unit) v
        write(unit) w
        write(unit) T
        close(unit)
    end subroutine output_binary_buoyancy

    subroutine output_raw_buoyancy(dir, rho, u, v, w, T, nx, ny, nz, itc)
        character(len=*), intent(in) :: dir
        integer, intent(in) :: nx, ny, nz, itc
        real(8), intent(in) :: rho(nx, ny, nz), u(nx, ny, nz), v(nx, ny, nz), w(nx, ny, nz), T(nx, ny, nz)
        integer :: unit
        open(newunit=unit, file=trim(dir)//'/buoyancy-raw-'//step_tag(itc)//'.bin', form='unformatted', &
             access='stream', status='replace')
        write(unit) rho, u, v, w, T
        close(unit)
    end subroutine output_raw_buoyancy

end module cavity_output
