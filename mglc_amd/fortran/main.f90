! main.f90 -- lid-driven cavity driver on one MI355X.
!
! Same program shape as the reference (MPI/Lid_driven_cavity/fortran/3d/
! mpi_3d_blocked/main.f90:7-141): parameters, initial, a timed
! do while (errorU > eps .and. itc <= itc_max) loop with check() every
! check_every steps, wall time, output.  The per-step subroutines
! collision / exchange / streaming / bounceback / macro are one fused HIP
! kernel behind lbm_step; steps between two checks go to the GPU as one call.
!
! usage: mglc_cavity [params.nml]     (namelist &cavity, see commondata.f90)
!        mglc_cavity --abi            print the sizes of lbm_params, lbm_particle_params and
!                                     lbm_info and exit
program main
    use iso_c_binding
    use commondata
    use mglc_lbm_capi
    use cavity_output
    implicit none
    type(lbm_params) :: p
    type(c_ptr) :: ctx
    integer(c_int) :: rc
    integer :: next_itc, i, j, k
    integer(8) :: t_start, t_end, t_rate
    real(8) :: wall
    character(len=16) :: arg1
    type(lbm_particle_params) :: pp_abi
    type(lbm_info) :: info_abi

    if (command_argument_count() >= 1) then
        call get_command_argument(1, arg1)
        if (trim(arg1) == '--abi') then
            write(*, '(i0)') c_sizeof(p)
            write(*, '(i0)') c_sizeof(pp_abi)
            write(*, '(i0)') c_sizeof(info_abi)
            stop
        endif
    endif

    call read_parameters()

    p%nx = nx
    p%ny = ny
    p%nz = nz
    p%collision = merge(LBM_BGK, LBM_MRT, collision_model(1:3) == 'BGK' .or. collision_model(1:3) == 'bgk')
    select case (storage(1:3))
    case ('F32', 'f32')
        p%storage = LBM_STORE_F32
    case ('F16', 'f16')
        p%storage = LBM_STORE_F16
    case default
        p%storage = LBM_STORE_F64
    end select
    p%meq12_rho = 0
    p%f16_scale_log2 = 0
    p%thermal = 0
    p%block_x = block_x
    p%block_y = block_y
    p%halo_faces = 0
    p%halo_loopback = merge(1, 0, halo_loopback)
    p%thermal_axis = -1
    p%halo_overlap = 0
    p%waves_per_cu = 0
    p%transform = 0
    p%tau = tauf
    p%u_lid = U0
    p%rho0 = rho0
    p%para_a = 0.0d0
    p%g_beta = 0.0d0
    p%omega_rot = 0.0d0
    p%t_hot = 0.0d0
    p%t_cold = 0.0d0
    p%t_ref = 0.0d0
    p%reserved_d = 0.0d0

    write(*,*) "nx=", nx, ", ny=", ny, ",  nz = ", nz
    write(*,*) "Reynolds=", real(Reynolds)
    write(*,*) "U0=", real(U0), ",    tauf=", real(tauf)
    write(*,*) "collision=", trim(collision_model), ",  storage=", trim(storage)

    call lbm_require(lbm_create(p, ctx), 'lbm_create')

    ! initial (initial.f90:1-76): node coordinates, then the device state
    allocate(xp(0:nx+1), yp(0:ny+1), zp(0:nz+1))
    xp(0) = 0.0d0
    xp(nx+1) = dble(nx)
    do i = 1, nx
        xp(i) = dble(i) - 0.5d0
    enddo
    yp(0) = 0.0d0
    yp(ny+1) = dble(ny)
    do j = 1, ny
        yp(j) = dble(j) - 0.5d0
    enddo
    zp(0) = 0.0d0
    zp(nz+1) = dble(nz)
    do k = 1, nz
        zp(k) = dble(k) - 0.5d0
    enddo
    allocate(u(nx, ny, nz), v(nx, ny, nz), w(nx, ny, nz), rho(nx, ny, nz))
    if (len_trim(restart_file) > 0) then
        ! the restart path of the reference (loadInitField): continue a saved run
        call lbm_load_state_file(ctx, trim(restart_file), itc)
    else
        call lbm_require(lbm_initial(ctx), 'lbm_initial')
        itc = 0
    endif
    errorU = 100.0d0

    call lbm_require(lbm_synchronize(ctx), 'lbm_synchronize')
    call system_clock(t_start, t_rate)

    do while ((errorU > eps) .and. (itc <= itc_max))
        ! run to the next check step, or to itc_max + 1 (the reference's last step)
        next_itc = min((itc / check_every + 1) * check_every, itc_max + 1)
        call lbm_require(lbm_step(ctx, int(next_itc - itc, c_int)), 'lbm_step')
        itc = next_itc
        if (mod(itc, check_every) == 0) then
            rc = lbm_check(ctx, errorU)
            write(*,*) itc, ' ', errorU
            if (rc == LBM_ERR_NAN) then
                write(*,*) 'error: the residual is NaN at itc =', itc, ' (the flow diverged)'
                error stop 3
            endif
            call lbm_require(rc, 'lbm_check')
        endif
    enddo

    call lbm_require(lbm_synchronize(ctx), 'lbm_synchronize')
    call system_clock(t_end)
    wall = dble(t_end - t_start) / dble(t_rate)
    write(*,*) "Wall time = ", real(wall), "s"
    if (wall > 0.0d0) write(*,*) "MLUPS = ", real(dble(nx) * dble(ny) * dble(nz) * dble(itc) / wall / 1.0d6)

    call lbm_require(lbm_download_macro(ctx, rho, u, v, w), 'lbm_download_macro')
    if (write_tecplot) call output_Tecplot(output_dir, xp, yp, zp, u, v, w, rho, nx, ny, nz, itc)
    if (write_centerlines) call getVelocity(output_dir, xp, zp, u, w, nx, ny, nz, U0, itc)
    if (write_binary) call output_binary(output_dir, u, v, rho, nx, ny, nz, itc)
    if (write_raw) call output_raw(output_dir, rho, u, v, w, nx, ny, nz, itc)

    if (len_trim(backup_file) > 0) call lbm_save_state_file(ctx, trim(backup_file))

    call lbm_require(lbm_destroy(ctx), 'lbm_destroy')
    deallocate(u, v, w, rho, xp, yp, zp)
    write(*,*) "Successfully: DNS completed!"

end program main
