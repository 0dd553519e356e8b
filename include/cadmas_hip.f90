! cadmas_hip.f90 -- Fortran 2003 interface (ISO_C_BINDING) to the C ABI in
! cadmas_hip.h, for the reference's own Fortran driver.
!
! The reference passes explicit-shape REAL*8 (NUMI,NUMJ,NUMK) arrays, INTEGER*4
! flags and INDB(4,NUMB) by reference (Source code/vf_a1main.f:412-697); every
! array argument below is declared assumed-size so that those arrays are passed
! as they are, without copies.  The context handle is an opaque TYPE(C_PTR).
!
! Build: flang -c cadmas_hip.f90 (module cadmas_hip), link with
! -L<package dir> -lcadmas_hip.
!
! The pressure-solve drop-ins VF_M1BCGS and VF_VPSOL need no interface: the
! library exports them under the reference's own link names (vf_m1bcgs_,
! vf_vpsol_) with the reference's argument lists, so the reference's own
! implicit-interface CALLs (vf_vpsol.f:64-69, vf_v1cal.f:278-281) bind to them
! when the library is linked in place of vf_m1bcgs.o / vf_vpsol.o.
module cadmas_hip
  use, intrinsic :: iso_c_binding
  implicit none

  integer(c_int), parameter :: CAD_OK = 0
  integer(c_int), parameter :: CAD_ERR_NOT_CONVERGED = -30
  integer(c_int), parameter :: CAD_ERR_INSTABILITY = -31
  integer(c_int), parameter :: CAD_ERR_DECOMPOSITION = -32
  integer(c_int), parameter :: CAD_ERR_ORIENTATION = -40
  integer(c_int), parameter :: CAD_ERR_WAVE = -50
  integer(c_int), parameter :: CAD_PRECISION_FP64 = 0
  integer(c_int), parameter :: CAD_PRECISION_MIXED = 1
  integer(c_int), parameter :: CAD_PRECISION_LEARNED = 2
  integer(c_int), parameter :: CAD_INIT_VELOCITY = 1
  integer(c_int), parameter :: CAD_INIT_PRESSURE = 2

  ! COMMON /VF_APHYSR/, /VF_ACOMPR/, /VF_ACOMPI/, /VF_ATIMER/ scalars;
  ! IDROP / WBUB are OPTION T-DOOR DROP / BUB (vf_iiopt.f:57-77), TNOW the
  ! time that CAD_STEP advances by DTNOW before each step (vf_a1main.f:1047)
  type, bind(c) :: cad_params
    real(c_double) :: rho0, grz0, anu0, dtnow
    real(c_double) :: scmvp, vvmax, flower, fupper, plower
    real(c_double) :: cgpara, cgepsa, cgepsr, cgdiv
    integer(c_int) :: icgtyp, icgmax, ibsuw0, loops
    integer(c_int) :: precision
    integer(c_int) :: idrop
    real(c_double) :: wbub, tnow
  end type cad_params

  ! list-file line of one step (vf_a1main.f:954-961)
  type, bind(c) :: cad_step_info
    real(c_double) :: cgbnrm, cgxnrm, fsum, fcut
    integer(c_int) :: icgitr, nums
    real(c_float) :: ms_predictor, ms_poisson, ms_correction, ms_vof
  end type cad_step_info

  interface
    function cad_last_error() bind(c, name='cad_last_error')
      import :: c_ptr
      type(c_ptr) :: cad_last_error
    end function

    function cad_create(numi, numj, numk, numb, device, ctx) bind(c, name='cad_create')
      import :: c_int, c_ptr
      integer(c_int), value :: numi, numj, numk, numb, device
      type(c_ptr), intent(out) :: ctx
      integer(c_int) :: cad_create
    end function

    subroutine cad_destroy(ctx) bind(c, name='cad_destroy')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine

    function cad_set_params(ctx, prm) bind(c, name='cad_set_params')
      import :: c_int, c_ptr, cad_params
      type(c_ptr), value :: ctx
      type(cad_params), intent(in) :: prm
      integer(c_int) :: cad_set_params
    end function

    function cad_get_params(ctx, prm) bind(c, name='cad_get_params')
      import :: c_int, c_ptr, cad_params
      type(c_ptr), value :: ctx
      type(cad_params), intent(out) :: prm
      integer(c_int) :: cad_get_params
    end function

    ! TimerDoor state: TBUB(NUMK), DROPTX..DROPWW(NUMI,NUMJ,NUMK) (vf_f1cal.f:2-3)
    function cad_upload_tdoor(ctx, tbub, droptx, dropty, droptz, dropuu, dropvv, dropww) &
        bind(c, name='cad_upload_tdoor')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: tbub(*), droptx(*), dropty(*), droptz(*), dropuu(*), &
                                    dropvv(*), dropww(*)
      integer(c_int) :: cad_upload_tdoor
    end function

    function cad_download_tdoor(ctx, tbub, droptx, dropty, droptz, dropuu, dropvv, dropww) &
        bind(c, name='cad_download_tdoor')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: tbub(*), droptx(*), dropty(*), droptz(*), dropuu(*), &
                                     dropvv(*), dropww(*)
      integer(c_int) :: cad_download_tdoor
    end function

    function cad_upload_static(ctx, xx, yy, zz, anu, cd0, ggv, ggx, ggy, ggz, glv, glx, gly, glz, &
                               indx, indy, indz, indb, bcvi) bind(c, name='cad_upload_static')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: xx(*), yy(*), zz(*), anu(*), cd0(*), ggv(*), ggx(*), ggy(*), &
                                    ggz(*), glv(*), glx(*), gly(*), glz(*), bcvi(*)
      integer(c_int), intent(in) :: indx(*), indy(*), indz(*), indb(*)
      integer(c_int) :: cad_upload_static
    end function

    function cad_upload_state(ctx, uu, vv, ww, pp, ff, bcu, bcv, bcw, bcp, bcf, nf, indc) &
        bind(c, name='cad_upload_state')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: uu(*), vv(*), ww(*), pp(*), ff(*), bcu(*), bcv(*), bcw(*), &
                                    bcp(*), bcf(*)
      integer(c_int), intent(in) :: nf(*), indc(*)
      integer(c_int) :: cad_upload_state
    end function

    function cad_download_state(ctx, uu, vv, ww, pp, ff, bcu, bcv, bcw, bcp, bcf, nf, indc) &
        bind(c, name='cad_download_state')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: uu(*), vv(*), ww(*), pp(*), ff(*), bcu(*), bcv(*), bcw(*), &
                                     bcp(*), bcf(*)
      integer(c_int), intent(out) :: nf(*), indc(*)
      integer(c_int) :: cad_download_state
    end function

    function cad_init_state(ctx, wvlvl, uini, vini, wini, flags, info) bind(c, name='cad_init_state')
      import :: c_int, c_ptr, c_double, cad_step_info
      type(c_ptr), value :: ctx
      real(c_double), value :: wvlvl, uini, vini, wini
      integer(c_int), value :: flags
      type(cad_step_info), intent(out) :: info
      integer(c_int) :: cad_init_state
    end function

    function cad_v1cal(ctx, info) bind(c, name='cad_v1cal')
      import :: c_int, c_ptr, cad_step_info
      type(c_ptr), value :: ctx
      type(cad_step_info), intent(out) :: info
      integer(c_int) :: cad_v1cal
    end function

    function cad_f1cal(ctx, info) bind(c, name='cad_f1cal')
      import :: c_int, c_ptr, cad_step_info
      type(c_ptr), value :: ctx
      type(cad_step_info), intent(out) :: info
      integer(c_int) :: cad_f1cal
    end function

    function cad_step(ctx, info) bind(c, name='cad_step')
      import :: c_int, c_ptr, cad_step_info
      type(c_ptr), value :: ctx
      type(cad_step_info), intent(out) :: info
      integer(c_int) :: cad_step
    end function

    ! VF_CDTCAL: the next step size (TIME AUTO DTINIT DTSAFE, TIME LIMIT DTMIN DTMAX)
    function cad_cdtcal(ctx, nnow, dtinit, dtsafe, dtmin, dtmax, dt) bind(c, name='cad_cdtcal')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nnow
      real(c_double), value :: dtinit, dtsafe, dtmin, dtmax
      real(c_double), intent(out) :: dt
      integer(c_int) :: cad_cdtcal
    end function

    ! the LEVEL-ON table of <name>.mtb (VF_IM1INP): ZZ(NZ), TT(NT), HH(NT),
    ! UN(NZ,NT), UT(NZ,NT) as the reference's DMTB* arrays, and DMTBT0
    function cad_set_wave_table(ctx, nz, nt, zz, tt, hh, un, ut, t0) &
        bind(c, name='cad_set_wave_table')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nz, nt
      real(c_double), intent(in) :: zz(*), tt(*), hh(*), un(*), ut(*)
      real(c_double), value :: t0
      integer(c_int) :: cad_set_wave_table
    end function

    ! MODEL WAVE-BC side (1..4 = X-, X+, Y-, Y+) FUNC MATRIX: columns l1..l2
    ! (IBCTYP(3..4,side)), DEPTH, PERIOD, AMPL, and W-LEVEL
    function cad_set_wave(ctx, side, l1, l2, depth, period, ampl, wvlvl) &
        bind(c, name='cad_set_wave')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: side, l1, l2
      real(c_double), value :: depth, period, ampl, wvlvl
      integer(c_int) :: cad_set_wave
    end function

    ! the second table, <name>.mtb2 (VF_IM2INP), and a side of FUNC MATRIX2
    function cad_set_wave_table2(ctx, nz, nt, zz, tt, hh, un, ut, t0) &
        bind(c, name='cad_set_wave_table2')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nz, nt
      real(c_double), intent(in) :: zz(*), tt(*), hh(*), un(*), ut(*)
      real(c_double), value :: t0
      integer(c_int) :: cad_set_wave_table2
    end function

    function cad_set_wave2(ctx, side, l1, l2, depth, period, ampl, wvlvl) &
        bind(c, name='cad_set_wave2')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: side, l1, l2
      real(c_double), value :: depth, period, ampl, wvlvl
      integer(c_int) :: cad_set_wave2
    end function

    ! MODEL WAVE-BC side (1..4) FUNC STOKES (-2), CNOIDAL (-1) or STK-CND (0)
    function cad_set_wave_theory(ctx, side, func, l1, l2, depth, height, period, ampl, wvlvl, &
        angle, ax, ay) bind(c, name='cad_set_wave_theory')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: side, func, l1, l2
      real(c_double), value :: depth, height, period, ampl, wvlvl, angle, ax, ay
      integer(c_int) :: cad_set_wave_theory
    end function

    ! the theory wave on the host: level, velocities at nz heights, phase origin
    function cad_wave_theory_eval(func, depth, height, period, stk_state, wvt, nz, zc, un, ut, &
        level, t0, length) bind(c, name='cad_wave_theory_eval')
      import :: c_int, c_double, c_float
      integer(c_int), value :: func, nz
      real(c_double), value :: depth, height, period, wvt
      real(c_float), intent(inout) :: stk_state
      real(c_double), intent(in) :: zc(*)
      real(c_double), intent(out) :: un(*), ut(*), level, t0, length
      integer(c_int) :: cad_wave_theory_eval
    end function

    ! MODEL OPEN-BC side (1..4) FUNC TYPE1 with DEPTH and PERIOD
    function cad_set_open(ctx, side, depth, period) bind(c, name='cad_set_open')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: side
      real(c_double), value :: depth, period
      integer(c_int) :: cad_set_open
    end function

    function cad_get_wave_level(ctx, side, level) bind(c, name='cad_get_wave_level')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: side
      real(c_double), intent(out) :: level
      integer(c_int) :: cad_get_wave_level
    end function

    ! OPTION DRAG-DF yu n with its n rows dr, ap, bt (n = 0: the CD drag form)
    function cad_set_drag_df(ctx, yu, n, dr, ap, bt) bind(c, name='cad_set_drag_df')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: yu
      integer(c_int), value :: n
      real(c_double), intent(in) :: dr(*), ap(*), bt(*)
      integer(c_int) :: cad_set_drag_df
    end function

    ! MODEL DAMP side (1..4) with DEGREE (< 0: off), PARAM-XY, PARAM-Z, WIDTH, DEPTH
    function cad_set_damp(ctx, side, degree, param_xy, param_z, width, depth) &
        bind(c, name='cad_set_damp')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: side, degree
      real(c_double), value :: param_xy, param_z, width, depth
      integer(c_int) :: cad_set_damp
    end function

    ! the preconditioner application of the context's form, timed after a solve
    function cad_time_precond(ctx, reps, ms_avg, bytes) bind(c, name='cad_time_precond')
      import :: c_int, c_ptr, c_float, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: reps
      real(c_float), intent(out) :: ms_avg
      real(c_double), intent(out) :: bytes
      integer(c_int) :: cad_time_precond
    end function

    ! ||BB - A*PT|| / ||BB|| of the last pressure solve (FP64)
    function cad_poisson_residual(ctx, rel) bind(c, name='cad_poisson_residual')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: rel
      integer(c_int) :: cad_poisson_residual
    end function

    ! every form held to ||BB - A PT|| <= rel ||BB|| (rel > 0; 0 restores
    ! each form's own test)
    function cad_set_true_residual_stop(ctx, rel) bind(c, name='cad_set_true_residual_stop')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: rel
      integer(c_int) :: cad_set_true_residual_stop
    end function

    ! the learned form's network parameters (n FP32 values) and mode
    ! (0: network alone, 1: V-cycle then network) with the restart length
    function cad_set_learned_weights(ctx, params, n) bind(c, name='cad_set_learned_weights')
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: ctx
      real(c_float), intent(in) :: params(*)
      integer(c_int), value :: n
      integer(c_int) :: cad_set_learned_weights
    end function

    function cad_set_learned_mode(ctx, mode, restart) bind(c, name='cad_set_learned_mode')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: mode, restart
      integer(c_int) :: cad_set_learned_mode
    end function

    function cad_cdiv00(ctx, dv) bind(c, name='cad_cdiv00')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: dv
      integer(c_int) :: cad_cdiv00
    end function

  end interface
end module cadmas_hip
