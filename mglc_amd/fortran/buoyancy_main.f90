! buoyancy_main.f90 -- buoyancy-driven cavity driver on one MI355X.
!
! Same program shape as the reference (MPI/Buoyancy_driven_cavity/fortran/3d/
! mpi_blocked/main.F90:1-180): parameters, initial, a timed
! do while (((errorU > epsU) .or. (errorT > epsT)) .and. itc <= itc_max
!           .and. statisticallyConvergeState == 0) loop with check() every
! check_every steps and, with statistics = .true., calNuRe every int(timeUnit)
! steps (its RBconvection block), wall time, output.  The per-step chain
! collision / streaming / bounceback / macro / collisionT / streamingT /
! bouncebackT / macroT is one fused HIP kernel behind lbm_step; the steps
! between two events go to the GPU as one call.
!
! usage: mglc_buoyancy [params.nml]     (namelist &buoyancy, see buoyancy_data.f90)
!        mglc_buoyancy params.nml --initial-only
!                 print the derived parameters, write the initial temperature
!                 to <output_dir>/buoyancy-initial-T.bin (raw FP64, x fastest)
!                 and exit without touching the GPU
program buoyancy_main
    use iso_c_binding
    use buoyancy_data
    use mglc_lbm_capi
    use cavity_output
    implicit none
    type(lbm_params) :: p
    type(c_ptr) :: ctx
    integer(c_int) :: rc
    integer :: next_itc, i, j, k, stat_every
    integer(8) :: t_start, t_end, t_rate
    real(8) :: wall, nu_vol, nu_wall, re_vol
    character(len=32) :: arg2
    integer :: unit

    call read_parameters()
    if (command_argument_count() >= 2) then
        call get_command_argument(2, arg2)
        if (trim(arg2) == '--initial-only') then
            write(*, '(a, es26.17e3)') 'tauf ', tauf
            write(*, '(a, es26.17e3)') 'viscosity ', viscosity
            write(*, '(a, es26.17e3)') 'diffusivity ', diffusivity
            write(*, '(a, es26.17e3)') 'omegaRatating ', omegaRatating
            write(*, '(a, es26.17e3)') 'paraA ', paraA
            write(*, '(a, es26.17e3)') 'gBeta ', gBeta
            write(*, '(a, es26.17e3)') 'timeUnit ', timeUnit
            write(*, '(a, i0)') 'itc_max ', itc_max
            allocate(T(nx, ny, nz))
            call initial_temperature()
            open(newunit=unit, file=trim(output_dir)//'/buoyancy-initial-T.bin', form='unformatted', &
                 access='stream', status='replace')
            write(unit) T
            close(unit)
            deallocate(T)
            stop
        endif
    endif
    if (heated_axis < 0 .or. heated_axis > 2) then
        write(*,*) 'error: heated_axis must be 0, 1 or 2'
        error stop 1
    endif

    p%nx = nx
    p%ny = ny
    p%nz = nz
    p%collision = LBM_MRT
    ! F64: the reference's FP64 arithmetic; F32 / F16: populations stored as
    ! FP32 deviations or scaled FP16, FP32 arithmetic, FP64 reductions
    select case (storage(1:3))
    case ('F32', 'f32')
        p%storage = LBM_STORE_F32
    case ('F16', 'f16')
        p%storage = LBM_STORE_F16
    case ('F64', 'f64')
        p%storage = LBM_STORE_F64
    case default
        write(*,*) 'error: storage must be F64, F32 or F16'
        error stop 1
    end select
    p%meq12_rho = 0
    p%f16_scale_log2 = 0
    p%thermal = 1
    p%block_x = 0
    p%block_y = 0
    p%halo_faces = 0
    p%halo_loopback = 0
    p%thermal_axis = heated_axis
    p%halo_overlap = 0
    p%waves_per_cu = 0
    p%transform = 0
    p%tau = tauf
    p%u_lid = 0.0d0
    p%rho0 = 1.0d0
    p%para_a = paraA
    p%g_beta = gBeta
    p%omega_rot = omegaRatating
    p%t_hot = Thot
    p%t_cold = Tcold
    p%t_ref = Tref
    p%reserved_d = 0.0d0

    ! initial.F90:14-34
    write(*,*) 'Mesh:', nx, ny, nz
    write(*,*) 'Rayleigh=', real(Rayleigh), ', Prandtl=', real(Prandtl), ', Mach=', real(Mach)
    write(*,*) 'Ekman=', real(Ekman)
    write(*,*) 'tauf=', real(tauf)
    write(*,*) 'paraA=', real(paraA)
    write(*,*) 'viscosity=', real(viscosity), ', diffusivity=', real(diffusivity)
    write(*,*) 'omegaRatating=', real(omegaRatating)
    write(*,*) 'itc_max=', itc_max
    write(*,*) 'Output interval is', int(timeUnit)

    call lbm_require(lbm_create(p, ctx), 'lbm_create')

    ! initial (initial.F90:52-70): node coordinates, then the device state
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
    allocate(u(nx, ny, nz), v(nx, ny, nz), w(nx, ny, nz), T(nx, ny, nz), rho(nx, ny, nz))
    allocate(NuVolAvg(0:dimensionlessTimeMax), ReVolAvg(0:dimensionlessTimeMax))
    allocate(NuVolAvg_mean(0:dimensionlessTimeMax), ReVolAvg_mean(0:dimensionlessTimeMax))
    NuVolAvg = 0.0d0
    ReVolAvg = 0.0d0
    NuVolAvg_mean = 0.0d0
    ReVolAvg_mean = 0.0d0

    if (len_trim(restart_file) > 0) then
        call lbm_load_state_file(ctx, trim(restart_file), itc)
    else if (noise_amplitude > 0.0d0) then
        ! rest state, held layers, seeded noise on the other nodes
        rho = 1.0d0
        u = 0.0d0
        v = 0.0d0
        w = 0.0d0
        call initial_temperature()
        call lbm_require(lbm_load_fields(ctx, rho, u, v, w, T), 'lbm_load_fields')
        itc = 0
    else
        ! initial.F90:97-190: rest state, T = 0 with the held layers
        call lbm_require(lbm_initial(ctx), 'lbm_initial')
        itc = 0
    endif
    errorU = 100.0d0
    errorT = 100.0d0
    stat_every = max(1, int(timeUnit))

    call lbm_require(lbm_synchronize(ctx), 'lbm_synchronize')
    call system_clock(t_start, t_rate)

    do while (((errorU > epsU) .or. (errorT > epsT)) .and. (itc <= itc_max) .and. &
              (statisticallyConvergeState == 0))
        ! run to the next check / statistics step, or to itc_max + 1 (the reference's last step)
        next_itc = min((itc / check_every + 1) * check_every, itc_max + 1)
        if (statistics) next_itc = min(next_itc, (itc / stat_every + 1) * stat_every)
        call lbm_require(lbm_step(ctx, int(next_itc - itc, c_int)), 'lbm_step')
        itc = next_itc
        if (mod(itc, check_every) == 0) then
            rc = lbm_check_thermal(ctx, errorU, errorT)
            write(*,*) itc, ' ', errorU, ' ', errorT
            if (rc == LBM_ERR_NAN) then
                write(*,*) 'error: the residual is NaN at itc =', itc, ' (the flow diverged)'
                error stop 3
            endif
            call lbm_require(rc, 'lbm_check_thermal')
        endif
        if (statistics .and. mod(itc, stat_every) == 0) then
            call lbm_require(lbm_nusselt(ctx, nu_vol, nu_wall), 'lbm_nusselt')
            call lbm_require(lbm_reynolds(ctx, re_vol), 'lbm_reynolds')
            call calNuRe(nu_vol, re_vol)
        endif
    enddo

    call lbm_require(lbm_synchronize(ctx), 'lbm_synchronize')
    call system_clock(t_end)
    wall = dble(t_end - t_start) / dble(t_rate)
    write(*,*) '---------------------------------------------'
    write(*,*) 'Time (wall) = ', real(wall), 's'
    if (wall > 0.0d0) write(*,*) 'MLUPS = ', real(dble(nx) * dble(ny) * dble(nz) * dble(itc) / wall / 1.0d6)
    write(*,*) '---------------------------------------------'
    if (statisticallyConvergeState == 1) write(*,*) 'statistically converged at time', dimensionlessTime

    call lbm_require(lbm_nusselt(ctx, nu_vol, nu_wall), 'lbm_nusselt')
    write(*,*) 'Nu_vol = ', nu_vol, ', Nu_wall = ', nu_wall

    call lbm_require(lbm_download_macro(ctx, rho, u, v, w), 'lbm_download_macro')
    call lbm_require(lbm_download_temperature(ctx, T), 'lbm_download_temperature')
    if (write_tecplot) call output_Tecplot_buoyancy(output_dir, xp, yp, zp, u, v, w, T, nx, ny, nz, itc)
    if (write_binary) call output_binary_buoyancy(output_dir, u, v, w, T, nx, ny, nz, itc)
    if (write_raw) call output_raw_buoyancy(output_dir, rho, u, v, w, T, nx, ny, nz, itc)

    if (len_trim(backup_file) > 0) call lbm_save_state_file(ctx, trim(backup_file))

    call lbm_require(lbm_destroy(ctx), 'lbm_destroy')
    deallocate(u, v, w, T, rho, xp, yp, zp)
    deallocate(NuVolAvg, ReVolAvg, NuVolAvg_mean, ReVolAvg_mean)
    write(*,*) 'Successfully: DNS completed!'

end program buoyancy_main
