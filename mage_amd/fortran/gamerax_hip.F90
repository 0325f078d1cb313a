!gamera.x-style driver over the gfx950 step (pattern of the reference's
!src/drivers/gamerax.F90:3-66 and the case-test harness
!tests/cases/testCases.pf:20-47): Fortran owns the Model/Grid/State arrays,
!the device runs Corners2Grid and stepGamera.
!
!Problem: Orszag-Tang vortex (prob.F90:74-114) on a periodic N x N x 8 box.
!Usage: gamerax_hip.x [N] [nsteps] [outfile]
!Prints t, dt and the kZCs throughput of consoleOutput (output.F90:64-85);
!with outfile, writes the initial and final states (raw FP64, stream access).
program gamerax_hip
    use, intrinsic :: iso_c_binding
    use gamhip
    implicit none

    type(gamhip_ctx) :: ctx
    type(gam_model_t) :: model
    type(gam_grid_arrays_t) :: ga
    integer :: N, nsteps, Nkp, i, j, k, ts, nargs
    character(len=256) :: arg, outfile
    real(rp), allocatable, target :: xyz(:,:,:,:), Gas(:,:,:,:), magFlux(:,:,:,:), Bxyz(:,:,:,:), &
                                     Efld(:,:,:,:), face(:,:,:,:), xyzcc(:,:,:,:)
    real(rp) :: dx, dz, rho0, P0, B0, vx, vy, xc, yc, t, dt, ms, kzcs
    real(rp), parameter :: pi = 4.0_rp*atan(1.0_rp)
    integer :: isg, ieg, jsg, jeg, ksg, keg, u

    N = 64; nsteps = 20; outfile = ''
    nargs = command_argument_count()
    if (nargs >= 1) then; call get_command_argument(1, arg); read(arg, *) N; endif
    if (nargs >= 2) then; call get_command_argument(2, arg); read(arg, *) nsteps; endif
    if (nargs >= 3) call get_command_argument(3, outfile)
    Nkp = 8

    model%rmeth = GAM_RECON_7UP
    model%tFin = 1.0e+30_rp
    model%bc = GAM_BC_PERIODIC

    call gamhip_create(ctx, N, N, Nkp, model)
    isg = ctx%isg; ieg = ctx%ieg; jsg = ctx%jsg; jeg = ctx%jeg; ksg = ctx%ksg; keg = ctx%keg

    !Uniform Cartesian corners, x = xMin + (i-1)*dx (init.F90:663-709)
    allocate(xyz(isg:ieg+1, jsg:jeg+1, ksg:keg+1, 3))
    dx = 1.0_rp/N
    dz = 1.0_rp/N
    do k = ksg, keg+1
        do j = jsg, jeg+1
            do i = isg, ieg+1
                xyz(i,j,k,1) = (i-1)*dx
                xyz(i,j,k,2) = (j-1)*dx
                xyz(i,j,k,3) = -0.5_rp*Nkp*dz + (k-1)*dz
            enddo
        enddo
    enddo
    call gamhip_grid_from_corners(ctx, xyz)

    !Cell centroids and face areas back from the device metrics
    allocate(xyzcc(isg:ieg, jsg:jeg, ksg:keg, 3), face(isg:ieg+1, jsg:jeg+1, ksg:keg+1, 3))
    ga%xyzcc = c_loc(xyzcc); ga%face = c_loc(face)
    ga%xfc = c_null_ptr; ga%di = c_null_ptr; ga%dj = c_null_ptr; ga%dk = c_null_ptr
    ga%volume = c_null_ptr; ga%Tf = c_null_ptr; ga%Te = c_null_ptr; ga%Teb = c_null_ptr
    ga%edge = c_null_ptr
    call gamhip_check(gam_get_grid(ctx%p, ga), 'gam_get_grid')

    !Orszag-Tang initial state (prob.F90:74-114)
    allocate(Gas(isg:ieg, jsg:jeg, ksg:keg, 5), Bxyz(isg:ieg, jsg:jeg, ksg:keg, 3), &
             Efld(isg:ieg, jsg:jeg, ksg:keg, 3), magFlux(isg:ieg+1, jsg:jeg+1, ksg:keg+1, 3))
    Gas = 0.0_rp; Bxyz = 0.0_rp; Efld = 0.0_rp; magFlux = 0.0_rp
    rho0 = 25.0_rp/(36.0_rp*pi)
    P0 = 5.0_rp/(12.0_rp*pi)
    B0 = 1.0_rp/sqrt(4.0_rp*pi)
    do k = ksg, keg
        do j = jsg, jeg
            do i = isg, ieg
                xc = xyzcc(i,j,k,1)
                yc = xyzcc(i,j,k,2)
                vx = -sin(2*pi*yc)
                vy = sin(2*pi*xc)
                Gas(i,j,k,1) = rho0
                Gas(i,j,k,2) = rho0*vx
                Gas(i,j,k,3) = rho0*vy
                Gas(i,j,k,4) = 0.0_rp
                Gas(i,j,k,5) = 0.5_rp*rho0*(vx*vx + vy*vy) + P0/(model%gamma-1)
                magFlux(i,j,k,1) = (-B0*sin(2*pi*yc))*face(i,j,k,1)
                magFlux(i,j,k,2) = (B0*sin(4*pi*xc))*face(i,j,k,2)
            enddo
        enddo
    enddo
    if (len_trim(outfile) > 0) then
        open(newunit=u, file=trim(outfile), access='stream', form='unformatted', status='replace')
        write(u) Gas, magFlux
    endif

    call gamhip_put_state(ctx, GAM_STATE, Gas, magFlux, Bxyz, Efld, 0.0_rp)
    dt = gamhip_init_run(ctx)
    write(*,'(a,i0,a,i0,a,i0,a,es12.5)') 'gamerax_hip: ', N, 'x', N, 'x', Nkp, '  dt0 = ', dt

    !Main loop (gamerax.F90:30-62): stepGamera, console output
    call gamhip_check(gam_time_steps(ctx%p, int(nsteps, c_int), ms), 'gam_time_steps')
    call gamhip_get_time(ctx, t, dt, ts)
    kzcs = real(nsteps, rp)*N*N*Nkp/(ms*1.0e-3_rp)/1.0e3_rp
    write(*,'(a,es14.7,a,es12.5,a,i0,a,es12.5)') '  t = ', t, '  dt = ', dt, '  ts = ', ts, '  kZCs = ', kzcs

    call gamhip_get_state(ctx, GAM_STATE, Gas, magFlux, Bxyz, Efld, t)
    if (len_trim(outfile) > 0) then
        write(u) Gas, magFlux, Bxyz
        close(u)
    endif
    write(*,'(a,es23.15)') '  sum(rho) = ', sum(Gas(1:N,1:N,1:Nkp,1))
    call gamhip_destroy(ctx)
end program gamerax_hip
