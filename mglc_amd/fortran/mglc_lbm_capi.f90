! mglc_lbm_capi.f90 -- iso_c_binding interfaces to libmglc_lbm.so
! (include/mglc_lbm.h).  Each interface replaces one of the reference's
! no-argument subroutines of MPI/Lid_driven_cavity/fortran/3d/mpi_3d_blocked/:
!   lbm_initial        initial.f90:1-76
!   lbm_step(n)        n x (collision.f90 + streaming.f90 + bounce_back.f90 + macro.f90)
!   lbm_check          check.f90:9-34 (returns errorU)
!   lbm_download_macro the gather half of output.f90:1-133 (one GPU = one rank)
!   lbm_save_state / lbm_load_state   backupData / loadInitField of the
!                      buoyancy codes (seq/bouyancy3d.F90:1011-1030)
!   lbm_halo_pack / lbm_halo_unpack   message_passing_sendrecv (ex_sendrecv.f90)
!   lbm_step_begin / lbm_step_end     collision_with_message_exchange
!                      (mpi_3d_nonblocked/lid3_mpi_nonblock.f90:1108-1231)
!   lbm_check_thermal, lbm_nusselt, lbm_reynolds   buoyancy check.F90, RaNu.F90
!   lbm_particles_*    the moving-sphere chain of Micro_particles case4
module mglc_lbm_capi
    use iso_c_binding
    implicit none

    integer(c_int), parameter :: LBM_OK = 0, LBM_ERR_ARG = 1, LBM_ERR_HIP = 2, LBM_ERR_NAN = 3
    integer(c_int), parameter :: LBM_ERR_TAU = 4, LBM_ERR_NOMEM = 5, LBM_ERR_STATE = 6
    integer(c_int), parameter :: LBM_MRT = 0, LBM_BGK = 1
    integer(c_int), parameter :: LBM_STORE_F64 = 0, LBM_STORE_F32 = 1, LBM_STORE_F16 = 2

    ! mirrors struct lbm_params
    type, bind(C) :: lbm_params
        integer(c_int) :: nx, ny, nz
        integer(c_int) :: collision
        integer(c_int) :: storage
        integer(c_int) :: meq12_rho
        integer(c_int) :: f16_scale_log2
        integer(c_int) :: thermal
        integer(c_int) :: block_x, block_y
        integer(c_int) :: halo_faces
        integer(c_int) :: halo_loopback
        integer(c_int) :: thermal_axis
        integer(c_int) :: halo_overlap
        integer(c_int) :: waves_per_cu
        integer(c_int) :: transform
        real(c_double) :: tau
        real(c_double) :: u_lid
        real(c_double) :: rho0
        real(c_double) :: para_a, g_beta, omega_rot
        real(c_double) :: t_hot, t_cold, t_ref
        real(c_double) :: reserved_d(2)
    end type lbm_params

    ! mirrors struct lbm_particle_params
    type, bind(C) :: lbm_particle_params
        integer(c_int) :: np, reserved_i
        real(c_double) :: rho_solid, gravity
        real(c_double) :: threshold_wall, stiff_wall
        real(c_double) :: threshold_particle, stiff_particle
        real(c_double) :: reserved_d(4)
    end type lbm_particle_params

    integer(c_int), parameter :: LBM_FACE_XM = 0, LBM_FACE_XP = 1, LBM_FACE_YM = 2, LBM_FACE_YP = 3
    integer(c_int), parameter :: LBM_FACE_ZM = 4, LBM_FACE_ZP = 5, LBM_NUM_REGIONS = 18
    integer(c_int), parameter :: LBM_ERR_UNSUPPORTED = 7, LBM_ERR_PARTICLE = 8

    ! mirrors struct lbm_info
    type, bind(C) :: lbm_info
        integer(c_long_long) :: nx, ny, nz, pitch_x, pop_stride, steps_done, device_bytes
        real(c_double) :: bytes_per_lup
        integer(c_int) :: storage, collision
    end type lbm_info

    interface
        integer(c_int) function lbm_create(p, ctx) bind(C, name='lbm_create')
            import :: c_int, c_ptr, lbm_params
            type(lbm_params), intent(in) :: p
            type(c_ptr), intent(out) :: ctx
        end function lbm_create

        integer(c_int) function lbm_destroy(ctx) bind(C, name='lbm_destroy')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function lbm_destroy

        integer(c_int) function lbm_initial(ctx) bind(C, name='lbm_initial')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function lbm_initial

        integer(c_int) function lbm_load_fields(ctx, rho, u, v, w, T) bind(C, name='lbm_load_fields')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(in) :: rho(*), u(*), v(*), w(*), T(*)
        end function lbm_load_fields

        integer(c_int) function lbm_check_thermal(ctx, error_u, error_t) bind(C, name='lbm_check_thermal')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: error_u, error_t
        end function lbm_check_thermal

        integer(c_int) function lbm_nusselt(ctx, nu_vol, nu_wall) bind(C, name='lbm_nusselt')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: nu_vol, nu_wall
        end function lbm_nusselt

        integer(c_int) function lbm_reynolds(ctx, re) bind(C, name='lbm_reynolds')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: re
        end function lbm_reynolds

        integer(c_int) function lbm_download_temperature(ctx, T) bind(C, name='lbm_download_temperature')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: T(*)
        end function lbm_download_temperature

        integer(c_int) function lbm_step(ctx, nsteps) bind(C, name='lbm_step')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: nsteps
        end function lbm_step

        integer(c_int) function lbm_check(ctx, error_u) bind(C, name='lbm_check')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: error_u
        end function lbm_check

        integer(c_int) function lbm_download_macro(ctx, rho, u, v, w) bind(C, name='lbm_download_macro')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: rho(*), u(*), v(*), w(*)
        end function lbm_download_macro

        integer(c_int) function lbm_synchronize(ctx) bind(C, name='lbm_synchronize')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function lbm_synchronize

        integer(c_int) function lbm_state_bytes(ctx, bytes) bind(C, name='lbm_state_bytes')
            import :: c_int, c_ptr, c_size_t
            type(c_ptr), value :: ctx
            integer(c_size_t), intent(out) :: bytes
        end function lbm_state_bytes

        integer(c_int) function lbm_save_state(ctx, blob, bytes) bind(C, name='lbm_save_state')
            import :: c_int, c_ptr, c_size_t, c_int8_t
            type(c_ptr), value :: ctx
            integer(c_int8_t), intent(out) :: blob(*)
            integer(c_size_t), value :: bytes
        end function lbm_save_state

        integer(c_int) function lbm_load_state(ctx, blob, bytes) bind(C, name='lbm_load_state')
            import :: c_int, c_ptr, c_size_t, c_int8_t
            type(c_ptr), value :: ctx
            integer(c_int8_t), intent(in) :: blob(*)
            integer(c_size_t), value :: bytes
        end function lbm_load_state

        integer(c_int) function lbm_get_info(ctx, info) bind(C, name='lbm_get_info')
            import :: c_int, c_ptr, lbm_info
            type(c_ptr), value :: ctx
            type(lbm_info), intent(out) :: info
        end function lbm_get_info

        integer(c_int) function lbm_set_stream(ctx, hip_stream) bind(C, name='lbm_set_stream')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, hip_stream
        end function lbm_set_stream

        integer(c_int) function lbm_check_sums(ctx, e1, e2) bind(C, name='lbm_check_sums')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: e1, e2
        end function lbm_check_sums

        ! the same sums over the caller's device arrays (c_ptr device addresses),
        ! allocating nothing; the sums land in device memory at sums_dev
        integer(c_int) function lbm_check_sums_noalloc(u, v, w, up, vp, n, sums_dev, hip_stream) &
                bind(C, name='lbm_check_sums_noalloc')
            import :: c_int, c_ptr, c_long_long
            type(c_ptr), value :: u, v, w, up, vp, sums_dev, hip_stream
            integer(c_long_long), value :: n
        end function lbm_check_sums_noalloc

        ! the subdomain exchange (ex_sendrecv.f90:12-123): device buffers
        integer(c_int) function lbm_halo_layout(ctx, region, offset, count) bind(C, name='lbm_halo_layout')
            import :: c_int, c_ptr, c_long_long
            type(c_ptr), value :: ctx
            integer(c_int), value :: region
            integer(c_long_long), intent(out) :: offset, count
        end function lbm_halo_layout

        integer(c_int) function lbm_halo_buffer_bytes(ctx, bytes) bind(C, name='lbm_halo_buffer_bytes')
            import :: c_int, c_ptr, c_size_t
            type(c_ptr), value :: ctx
            integer(c_size_t), intent(out) :: bytes
        end function lbm_halo_buffer_bytes

        integer(c_int) function lbm_halo_pack(ctx, dev_send) bind(C, name='lbm_halo_pack')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, dev_send
        end function lbm_halo_pack

        integer(c_int) function lbm_halo_unpack(ctx, dev_recv) bind(C, name='lbm_halo_unpack')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, dev_recv
        end function lbm_halo_unpack

        ! collision_with_message_exchange (lid3_mpi_nonblock.f90:1108-1231):
        ! boundary shell, then the interior
        integer(c_int) function lbm_step_begin(ctx) bind(C, name='lbm_step_begin')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function lbm_step_begin

        integer(c_int) function lbm_step_end(ctx) bind(C, name='lbm_step_end')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function lbm_step_end

        ! moving spheres (Micro_particles/fortran/case4/mpi: initial.F90:78-135,
        ! the particle state of commondata.F90:14-35, obst)
        integer(c_int) function lbm_particles_init(ctx, pp, centers, radii, U, omega) &
                bind(C, name='lbm_particles_init')
            import :: c_int, c_ptr, c_double, lbm_particle_params
            type(c_ptr), value :: ctx
            type(lbm_particle_params), intent(in) :: pp
            real(c_double), intent(in) :: centers(*), radii(*), U(*), omega(*)
        end function lbm_particles_init

        integer(c_int) function lbm_particles_get(ctx, centers, U, omega, F, T) bind(C, name='lbm_particles_get')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: centers(*), U(*), omega(*), F(*), T(*)
        end function lbm_particles_get

        integer(c_int) function lbm_particles_map(ctx, owner) bind(C, name='lbm_particles_map')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), intent(out) :: owner(*)
        end function lbm_particles_map

        type(c_ptr) function lbm_last_error_c() bind(C, name='lbm_last_error')
            import :: c_ptr
        end function lbm_last_error_c
    end interface

contains

    ! The message of the last failed call, as a Fortran string.
    function lbm_last_error() result(msg)
        character(len=512) :: msg
        character(kind=c_char), pointer :: s(:)
        type(c_ptr) :: p
        integer :: i
        msg = ' '
        p = lbm_last_error_c()
        if (.not. c_associated(p)) return
        call c_f_pointer(p, s, [512])
        do i = 1, 512
            if (s(i) == c_null_char) exit
            msg(i:i) = s(i)
        enddo
    end function lbm_last_error

    ! backupData (seq/bouyancy3d.F90:1011-1030): the whole device state of
    ! ctx as one stream file
    subroutine lbm_save_state_file(ctx, path)
        type(c_ptr), intent(in) :: ctx
        character(len=*), intent(in) :: path
        integer(c_size_t) :: bytes
        integer(c_int8_t), allocatable :: blob(:)
        integer :: unit
        call lbm_require(lbm_state_bytes(ctx, bytes), 'lbm_state_bytes')
        allocate(blob(bytes))
        call lbm_require(lbm_save_state(ctx, blob, bytes), 'lbm_save_state')
        open(newunit=unit, file=path, access='stream', form='unformatted', status='replace', action='write')
        write(unit) blob
        close(unit)
        deallocate(blob)
        write(*,*) 'state saved to ', path
    end subroutine lbm_save_state_file

    ! loadInitField (3d/mpi_blocked/initial.F90:102,195-206): continue from a
    ! file written by lbm_save_state_file; returns the step count it holds
    subroutine lbm_load_state_file(ctx, path, itc)
        type(c_ptr), intent(in) :: ctx
        character(len=*), intent(in) :: path
        integer, intent(out) :: itc
        integer(c_size_t) :: bytes
        integer(c_int8_t), allocatable :: blob(:)
        integer :: unit, ios
        type(lbm_info) :: info
        call lbm_require(lbm_state_bytes(ctx, bytes), 'lbm_state_bytes')
        allocate(blob(bytes))
        open(newunit=unit, file=path, access='stream', form='unformatted', status='old', action='read', iostat=ios)
        if (ios /= 0) then
            write(*,*) 'error: cannot open restart file ', path
            error stop 1
        endif
        read(unit, iostat=ios) blob
        close(unit)
        if (ios /= 0) then
            write(*,*) 'error: restart file ', path, ' is shorter than this run''s state'
            error stop 1
        endif
        call lbm_require(lbm_load_state(ctx, blob, bytes), 'lbm_load_state')
        deallocate(blob)
        call lbm_require(lbm_get_info(ctx, info), 'lbm_get_info')
        itc = int(info%steps_done)
        write(*,*) 'restarted from ', path, ' at itc =', itc
    end subroutine lbm_load_state_file

    ! Stop with the status, in the reference's style of fatal errors
    ! (stop on failure; the reference's MPI codes call MPI_Abort).
    subroutine lbm_require(rc, where)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: where
        if (rc /= LBM_OK) then
            write(*,*) 'error: ', where, ' returned ', rc, ': ', trim(lbm_last_error())
            error stop 2
        endif
    end subroutine lbm_require

end module mglc_lbm_capi
