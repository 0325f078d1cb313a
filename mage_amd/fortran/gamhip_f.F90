!Fortran host binding of libgamhip (include/gamhip.h) through ISO_C_BINDING.
!
!Keeps the reference's Model_T/Grid_T/State_T array conventions
!(src/base/types/gamtypes.F90:77-271): allocatable, column-major,
!i fastest, ghost lower bounds isg = 1-nG; cell arrays
!(isg:ieg,jsg:jeg,ksg:keg,...), face arrays (isg:ieg+1,...).  The
!wrappers pass c_loc() of those arrays to the C ABI, so a Fortran host
!calls the gfx950 kernels without reshaping anything.
!
!Error handling follows the reference: a failing call prints the library's
!message and stops (step.F90:145-151 prints and stops on a collapsed dt).
module gamhip
    use, intrinsic :: iso_c_binding
    implicit none

    integer, parameter :: rp = c_double

    integer(c_int), parameter :: GAM_OK = 0
    integer(c_int), parameter :: GAM_E_DT_BAD = -4, GAM_E_DT_DISASTER = -5
    integer(c_int), parameter :: GAM_BC_PERIODIC = 0, GAM_BC_ZEROGRAD = 1, &
                                 GAM_BC_SPHIN = 2, GAM_BC_SPHOUT = 3, GAM_BC_WIND = 4
    integer(c_int), parameter :: GAM_RECON_8CENT = 0, GAM_RECON_7UP = 1
    integer(c_int), parameter :: GAM_STATE = 0, GAM_OSTATE = 1
    integer(c_int), parameter :: GAM_TUNE_SLOTS = 8

    !gam_dims (Grid_T%Nip.., Model_T%nG)
    type, bind(C) :: gam_dims_t
        integer(c_int) :: Nip, Njp, Nkp, nG
    end type gam_dims_t

    !gam_model (Model_T knobs on the hot path, defaults init.F90:312-484)
    type, bind(C) :: gam_model_t
        real(c_double) :: gamma  = 5.0_rp/3.0_rp
        real(c_double) :: CFL    = 0.3_rp
        real(c_double) :: Vd0    = 0.5_rp
        real(c_double) :: Ca     = 1.0e+12_rp
        real(c_double) :: cHogH  = 0.25_rp
        real(c_double) :: cHogM  = 0.25_rp
        real(c_double) :: pdmb   = 0.75_rp
        real(c_double) :: dFloor = 1.0e-12_rp
        real(c_double) :: pFloor = 1.0e-12_rp
        real(c_double) :: tFin   = 1.0_rp
        real(c_double) :: dt0    = 0.0_rp
        real(c_double) :: limDT0 = 1.0e-3_rp
        integer(c_int) :: doMHD = 1, doBoris = 0, doBackground = 0, doLFMLim = 1
        integer(c_int) :: rmeth = 0
        integer(c_int) :: bc(6) = 0
    end type gam_model_t

    type, bind(C) :: gam_grid_arrays_t
        type(c_ptr) :: xyzcc, xfc, di, dj, dk, volume, Tf, Te, Teb, face, edge
    end type gam_grid_arrays_t

    type :: gamhip_ctx
        type(c_ptr) :: p = c_null_ptr
        integer :: Nip, Njp, Nkp, nG
        integer :: isg, ieg, jsg, jeg, ksg, keg
    end type gamhip_ctx

    interface
        function gam_last_error() bind(C, name="gam_last_error")
            import :: c_ptr
            type(c_ptr) :: gam_last_error
        end function
        integer(c_int) function gam_create(ctx, dims, model) bind(C, name="gam_create")
            import :: c_int, c_ptr, gam_dims_t, gam_model_t
            type(c_ptr) :: ctx
            type(gam_dims_t), intent(in) :: dims
            type(gam_model_t), intent(in) :: model
        end function
        integer(c_int) function gam_destroy(ctx) bind(C, name="gam_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function gam_grid_from_corners(ctx, xyz) bind(C, name="gam_grid_from_corners")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, xyz
        end function
        integer(c_int) function gam_set_grid(ctx, g) bind(C, name="gam_set_grid")
            import :: c_int, c_ptr, gam_grid_arrays_t
            type(c_ptr), value :: ctx
            type(gam_grid_arrays_t), intent(in) :: g
        end function
        integer(c_int) function gam_get_grid(ctx, g) bind(C, name="gam_get_grid")
            import :: c_int, c_ptr, gam_grid_arrays_t
            type(c_ptr), value :: ctx
            type(gam_grid_arrays_t), intent(in) :: g
        end function
        integer(c_int) function gam_add_dipole_b0(ctx, M0) bind(C, name="gam_add_dipole_b0")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: M0
        end function
        integer(c_int) function gam_set_wind(ctx, n, tW, Q, B, Bx0, ByC, BzC, xyzW) bind(C, name="gam_set_wind")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: n
            type(c_ptr), value :: tW, Q, B, xyzW
            real(c_double), value :: Bx0, ByC, BzC
        end function
        integer(c_int) function gam_put_state(ctx, which, Gas, magFlux, Bxyz, Efld, t) bind(C, name="gam_put_state")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: which
            type(c_ptr), value :: Gas, magFlux, Bxyz, Efld
            real(c_double), value :: t
        end function
        integer(c_int) function gam_get_state(ctx, which, Gas, magFlux, Bxyz, Efld, t) bind(C, name="gam_get_state")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: which
            type(c_ptr), value :: Gas, magFlux, Bxyz, Efld
            real(c_double) :: t
        end function
        integer(c_int) function gam_set_time(ctx, t, dt, ts) bind(C, name="gam_set_time")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: t, dt
            integer(c_int), value :: ts
        end function
        integer(c_int) function gam_get_time(ctx, t, dt, ts) bind(C, name="gam_get_time")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double) :: t, dt
            integer(c_int) :: ts
        end function
        integer(c_int) function gam_init_run(ctx, dt) bind(C, name="gam_init_run")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double) :: dt
        end function
        integer(c_int) function gam_advance_mhd(ctx, dt) bind(C, name="gam_advance_mhd")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: dt
        end function
        integer(c_int) function gam_calc_dt(ctx, dt) bind(C, name="gam_calc_dt")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double) :: dt
        end function
        integer(c_int) function gam_enforce_bcs(ctx) bind(C, name="gam_enforce_bcs")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function gam_bflux2fld(ctx, which) bind(C, name="gam_bflux2fld")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: which
        end function
        integer(c_int) function gam_step(ctx, dt_next) bind(C, name="gam_step")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, dt_next
        end function
        integer(c_int) function gam_run(ctx, nsteps) bind(C, name="gam_run")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: nsteps
        end function
        integer(c_int) function gam_advance_model(ctx, dt, nsteps) bind(C, name="gam_advance_model")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: dt
            integer(c_int) :: nsteps
        end function
        integer(c_int) function gam_divb(ctx, tot) bind(C, name="gam_divb")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double) :: tot
        end function
        integer(c_int) function gam_time_steps(ctx, nsteps, ms) bind(C, name="gam_time_steps")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: nsteps
            real(c_double) :: ms
        end function
        ! launch plans (tile shapes per grid size)
        integer(c_int) function gam_autotune(ctx, reps) bind(C, name="gam_autotune")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: reps
        end function
        integer(c_int) function gam_get_plan(ctx, plan) bind(C, name="gam_get_plan")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int) :: plan(GAM_TUNE_SLOTS)
        end function
        integer(c_int) function gam_set_plan(ctx, plan) bind(C, name="gam_set_plan")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int) :: plan(GAM_TUNE_SLOTS)
        end function
        ! ensembles: several contexts stepped as one batched launch per kernel
        integer(c_int) function gam_ens_create(ens, n, members) bind(C, name="gam_ens_create")
            import :: c_int, c_ptr
            type(c_ptr) :: ens
            integer(c_int), value :: n
            type(c_ptr) :: members(*)
        end function
        integer(c_int) function gam_ens_run(ens, nsteps) bind(C, name="gam_ens_run")
            import :: c_int, c_ptr
            type(c_ptr), value :: ens
            integer(c_int), value :: nsteps
        end function
        integer(c_int) function gam_ens_time_steps(ens, nsteps, ms) bind(C, name="gam_ens_time_steps")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ens
            integer(c_int), value :: nsteps
            real(c_double) :: ms
        end function
        integer(c_int) function gam_ens_destroy(ens) bind(C, name="gam_ens_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: ens
        end function
        ! discrete adjoint (include/gamadj.h; no counterpart in the reference)
        integer(c_int) function gam_adj_create(adj, ctx, nsteps, ckpt_every) bind(C, name="gam_adj_create")
            import :: c_int, c_ptr
            type(c_ptr) :: adj
            type(c_ptr), value :: ctx
            integer(c_int), value :: nsteps, ckpt_every
        end function
        integer(c_int) function gam_adj_create_budget(adj, ctx, nsteps, ckpt_every, max_taped, overlap) &
                bind(C, name="gam_adj_create_budget")
            import :: c_int, c_ptr
            type(c_ptr) :: adj
            type(c_ptr), value :: ctx
            integer(c_int), value :: nsteps, ckpt_every, max_taped, overlap
        end function
        integer(c_int) function gam_adj_set_ostate_bc_time(adj, at_state_time) &
                bind(C, name="gam_adj_set_ostate_bc_time")
            import :: c_int, c_ptr
            type(c_ptr), value :: adj
            integer(c_int), value :: at_state_time
        end function
        integer(c_int) function gam_adj_tape_info(adj, taped_steps, overlap) bind(C, name="gam_adj_tape_info")
            import :: c_int, c_ptr
            type(c_ptr), value :: adj
            integer(c_int) :: taped_steps, overlap
        end function
        integer(c_int) function gam_adj_destroy(adj) bind(C, name="gam_adj_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: adj
        end function
        integer(c_int) function gam_adj_set_weights(adj, w) bind(C, name="gam_adj_set_weights")
            import :: c_int, c_ptr
            type(c_ptr), value :: adj, w
        end function
        integer(c_int) function gam_adj_forward(adj, J) bind(C, name="gam_adj_forward")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: adj
            real(c_double) :: J
        end function
        integer(c_int) function gam_adj_backward(adj, gGas, gGasO, gMF, gMFO) bind(C, name="gam_adj_backward")
            import :: c_int, c_ptr
            type(c_ptr), value :: adj, gGas, gGasO, gMF, gMFO
        end function
        integer(c_int) function gam_adj_replay(adj, Gas, oGas, MF, oMF, J) bind(C, name="gam_adj_replay")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: adj, Gas, oGas, MF, oMF
            real(c_double) :: J
        end function
        integer(c_int) function gam_adj_replay_dt(adj, Gas, oGas, MF, oMF, dts, J) bind(C, name="gam_adj_replay_dt")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: adj, Gas, oGas, MF, oMF, dts
            real(c_double) :: J
        end function
        integer(c_int) function gam_adj_set_dt_mode(adj, mode) bind(C, name="gam_adj_set_dt_mode")
            import :: c_int, c_ptr
            type(c_ptr), value :: adj
            integer(c_int), value :: mode
        end function
        integer(c_int) function gam_adj_dt_grad(adj, g) bind(C, name="gam_adj_dt_grad")
            import :: c_int, c_ptr
            type(c_ptr), value :: adj, g
        end function
        integer(c_int) function gam_adj_step_dt(adj, dts) bind(C, name="gam_adj_step_dt")
            import :: c_int, c_ptr
            type(c_ptr), value :: adj, dts
        end function
        integer(c_int) function gam_adj_wind_grad(adj, gQ, gB) bind(C, name="gam_adj_wind_grad")
            import :: c_int, c_ptr
            type(c_ptr), value :: adj, gQ, gB
        end function
        integer(c_int) function gam_adj_times(adj, fwd_ms, bwd_ms) bind(C, name="gam_adj_times")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: adj
            real(c_double) :: fwd_ms, bwd_ms
        end function
        function c_strlen(s) bind(C, name="strlen")
            import :: c_ptr, c_size_t
            type(c_ptr), value :: s
            integer(c_size_t) :: c_strlen
        end function
    end interface

    contains

    !Message of the last failing call
    function gamhip_errmsg() result(msg)
        character(len=:), allocatable :: msg
        type(c_ptr) :: p
        character(kind=c_char), pointer :: chars(:)
        integer :: n, q
        p = gam_last_error()
        n = int(c_strlen(p))
        call c_f_pointer(p, chars, [n])
        allocate(character(len=n) :: msg)
        do q = 1, n
            msg(q:q) = chars(q)
        enddo
    end function gamhip_errmsg

    !Stop with the library's message on failure (the reference prints and stops)
    subroutine gamhip_check(rc, where)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: where
        if (rc /= GAM_OK) then
            write(*,*) 'gamhip: ', trim(where), ' failed (', rc, '): ', gamhip_errmsg()
            error stop 1
        endif
    end subroutine gamhip_check

    !Hatch-equivalent allocation of the device Grid/State/Solver (init.F90:38-59)
    subroutine gamhip_create(ctx, Nip, Njp, Nkp, model)
        type(gamhip_ctx), intent(out) :: ctx
        integer, intent(in) :: Nip, Njp, Nkp
        type(gam_model_t), intent(in) :: model
        type(gam_dims_t) :: d
        d = gam_dims_t(Nip, Njp, Nkp, 4)
        call gamhip_check(gam_create(ctx%p, d, model), 'gam_create')
        ctx%Nip = Nip; ctx%Njp = Njp; ctx%Nkp = Nkp; ctx%nG = 4
        ctx%isg = -3; ctx%ieg = Nip + 4
        ctx%jsg = -3; ctx%jeg = Njp + 4
        ctx%ksg = -3; ctx%keg = Nkp + 4
    end subroutine gamhip_create

    subroutine gamhip_destroy(ctx)
        type(gamhip_ctx), intent(inout) :: ctx
        integer(c_int) :: rc
        if (c_associated(ctx%p)) rc = gam_destroy(ctx%p)
        ctx%p = c_null_ptr
    end subroutine gamhip_destroy

    !Corners2Grid on the device from xyz(isg:ieg+1,jsg:jeg+1,ksg:keg+1,3)
    subroutine gamhip_grid_from_corners(ctx, xyz)
        type(gamhip_ctx), intent(in) :: ctx
        real(rp), dimension(:,:,:,:), intent(in), target, contiguous :: xyz
        call gamhip_check(gam_grid_from_corners(ctx%p, c_loc(xyz)), 'gam_grid_from_corners')
    end subroutine gamhip_grid_from_corners

    !State_T upload: Gas(isg:ieg,..,1:5), magFlux(isg:ieg+1,..,1:3), Bxyz/Efld(isg:ieg,..,1:3)
    subroutine gamhip_put_state(ctx, which, Gas, magFlux, Bxyz, Efld, t)
        type(gamhip_ctx), intent(in) :: ctx
        integer(c_int), intent(in) :: which
        real(rp), dimension(:,:,:,:), intent(in), target, contiguous :: Gas, magFlux, Bxyz, Efld
        real(rp), intent(in) :: t
        call gamhip_check(gam_put_state(ctx%p, which, c_loc(Gas), c_loc(magFlux), c_loc(Bxyz), &
                                        c_loc(Efld), t), 'gam_put_state')
    end subroutine gamhip_put_state

    subroutine gamhip_get_state(ctx, which, Gas, magFlux, Bxyz, Efld, t)
        type(gamhip_ctx), intent(in) :: ctx
        integer(c_int), intent(in) :: which
        real(rp), dimension(:,:,:,:), intent(inout), target, contiguous :: Gas, magFlux, Bxyz, Efld
        real(rp), intent(out) :: t
        call gamhip_check(gam_get_state(ctx%p, which, c_loc(Gas), c_loc(magFlux), c_loc(Bxyz), &
                                        c_loc(Efld), t), 'gam_get_state')
    end subroutine gamhip_get_state

    !Tail of CalcGridInfo (init.F90:200-244): returns the first dt
    function gamhip_init_run(ctx) result(dt)
        type(gamhip_ctx), intent(in) :: ctx
        real(rp) :: dt
        call gamhip_check(gam_init_run(ctx%p, dt), 'gam_init_run')
    end function gamhip_init_run

    !stepGamera (gamapp.F90:73-91); returns the new Model%dt
    function gamhip_step(ctx) result(dt)
        type(gamhip_ctx), intent(in) :: ctx
        real(rp), target :: dt
        dt = 0.0_rp
        call gamhip_check(gam_step(ctx%p, c_loc(dt)), 'gam_step')
    end function gamhip_step

    !gamAdvanceModel (gamtypessub.F90:74-90): returns the number of steps
    function gamhip_advance_model(ctx, dt) result(nsteps)
        type(gamhip_ctx), intent(in) :: ctx
        real(rp), intent(in) :: dt
        integer :: nsteps
        integer(c_int) :: n
        call gamhip_check(gam_advance_model(ctx%p, dt, n), 'gam_advance_model')
        nsteps = n
    end function gamhip_advance_model

    subroutine gamhip_get_time(ctx, t, dt, ts)
        type(gamhip_ctx), intent(in) :: ctx
        real(rp), intent(out) :: t, dt
        integer, intent(out) :: ts
        integer(c_int) :: its
        call gamhip_check(gam_get_time(ctx%p, t, dt, its), 'gam_get_time')
        ts = its
    end subroutine gamhip_get_time

end module gamhip
