/*
 * cadmas_hip.h -- C ABI of the MI355X (gfx950) CADMAS-SURF/3D single-phase
 * VOF-SMAC time step.
 *
 * The reference's replaceable unit is the Fortran subroutine, linked
 * statically (Simulators/CADMAS-SURF-3D/Source code/Makefile:14-83); its
 * arrays are explicit-shape column-major REAL*8 (NUMI,NUMJ,NUMK) fields,
 * INTEGER*4 flags, INDB(4,NUMB) and metric tables XX(6,NUMI), YY(6,NUMJ),
 * ZZ(6,NUMK).  Every pointer below takes exactly those host arrays.  Fields stay
 * resident in HBM inside a cad_ctx between steps; host copies happen only when
 * the caller uploads or downloads them (the reference's .rsl output steps).
 *
 * All functions return CAD_OK (0) or a negative CAD_ERR_* code; the message of
 * the last failure is available from cad_last_error().  Solver outcomes keep
 * the reference's ITER convention (vf_m1bcgs.f:60-64): >=0 iterations,
 * -1 MILU pivot <= 0, -2 divergence, >ICGMAX not converged.
 */
#ifndef CADMAS_HIP_H
#define CADMAS_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define CAD_OK 0
#define CAD_ERR_ARG (-1)
#define CAD_ERR_HIP (-2)
#define CAD_ERR_STATE (-3)
#define CAD_ERR_UNSUPPORTED (-4)
#define CAD_ERR_NOT_CONVERGED (-30)   /* VF_VPSOL 'NOT CONVERGED.'  vf_vpsol.f:71 */
#define CAD_ERR_INSTABILITY (-31)     /* VF_VPSOL 'INSTABILITY.'    vf_vpsol.f:73 */
#define CAD_ERR_DECOMPOSITION (-32)   /* VF_VPSOL 'CAN NOT DECOMPOSITION.' :75 */
#define CAD_ERR_ORIENTATION (-40)     /* VF_FNFPRV 'P.G ERROR.'     vf_fnfprv.f:223 */
#define CAD_ERR_WAVE (-50)            /* VF_CWMTB1/2 'P.G ERROR.'   vf_cwmtb1.f, vf_cwmtb2.f */

#define CAD_PRECISION_FP64 0          /* plain form: all FP64 */
#define CAD_PRECISION_MIXED 1         /* FP64 BiCGSTAB with an FP32 multigrid preconditioner,
                                         stopped on the true residual (cad_vpsol) */
#define CAD_PRECISION_LEARNED 2       /* FP64 flexible GMRES preconditioned by an FP32 V-cycle and
                                         the BF16 U-Net, MILU-BiCGSTAB as the fallback, stopped
                                         on the true residual (cad_vpsol) */

typedef struct cad_ctx cad_ctx;

/* Scalars the reference keeps in COMMON /VF_APHYSR/ (RHO0, GRZ0, ANU0, VVMAX),
 * /VF_ACOMPR/ (SCMVP, FLOWER, FUPPER, PLOWER, CGPARA, CGEPSA, CGEPSR, CGDIV,
 * WBUB), /VF_ACOMPI/ (ICGTYP, ICGMAX, IBSUW0, IDROP), /VF_ATIMER/ (DTNOW,
 * TNOW), /VF_ATIMEI/ (LOOPS).  cad_create fills the defaults of
 * vf_a2dflt.f, which include both TimerDoor models on (IDROP=1 at :151,
 * WBUB=0.2 at :205). */
typedef struct cad_params {
    double rho0, grz0, anu0, dtnow;
    double scmvp, vvmax, flower, fupper, plower;
    double cgpara, cgepsa, cgepsr, cgdiv;
    int icgtyp, icgmax, ibsuw0, loops;
    int precision;            /* CAD_PRECISION_* */
    int idrop;                /* OPTION T-DOOR DROP: 1 FREE-RUNDOWN, 0 OFF (vf_iiopt.f:68-75) */
    double wbub;              /* OPTION T-DOOR BUB rise speed [m/s]; < 1e-20 is OFF (vf_f1cal.f:208) */
    double tnow;              /* TNOW; cad_step advances it by DTNOW first (vf_a1main.f:1047) */
} cad_params;

/* Per-step outputs of the list-file line (vf_a1main.f:954-961). */
typedef struct cad_step_info {
    double cgbnrm, cgxnrm;    /* !B!, !R! */
    double fsum, fcut;        /* FSUM, FCUT */
    int icgitr;               /* ITR */
    int nums;                 /* surface cells after flagging */
    float ms_predictor, ms_poisson, ms_correction, ms_vof;  /* HIP-event phase timers */
} cad_step_info;

/* Library / context management ------------------------------------------ */
const char *cad_last_error(void);
int cad_create(int numi, int numj, int numk, int numb, int device, cad_ctx **out);
void cad_destroy(cad_ctx *ctx);
int cad_set_params(cad_ctx *ctx, const cad_params *prm);
/* The context's current scalars (TNOW included). */
int cad_get_params(cad_ctx *ctx, cad_params *prm);

/* Set-up arrays built by the reference's VF_CSETUP/VF_CGLV/VF_CGLXYZ/VF_CINDX/
 * VF_CINDB (vf_a1main.f:784) -- uploaded once. */
int cad_upload_static(cad_ctx *ctx, const double *xx, const double *yy, const double *zz,
                      const double *anu, const double *cd0,
                      const double *ggv, const double *ggx, const double *ggy, const double *ggz,
                      const double *glv, const double *glx, const double *gly, const double *glz,
                      const int *indx, const int *indy, const int *indz, const int *indb,
                      const double *bcvi);

/* Time-level state: the arrays VF_OR1TRN writes to .rsl (vf_or1trn.f:105-200)
 * plus the flags NF/INDC. */
int cad_upload_state(cad_ctx *ctx, const double *uu, const double *vv, const double *ww,
                     const double *pp, const double *ff, const double *bcu, const double *bcv,
                     const double *bcw, const double *bcp, const double *bcf, const int *nf,
                     const int *indc);
int cad_download_state(cad_ctx *ctx, double *uu, double *vv, double *ww, double *pp, double *ff,
                       double *bcu, double *bcv, double *bcw, double *bcp, double *bcf, int *nf,
                       int *indc);

/* Device part of VF_CINIT (vf_cinit.f:190-409), run once after the state is
 * uploaded with F at its initial value (W-LEVEL / DEBUG F-BOX or a restart):
 * F clip and volume sums (VF_FCUT01), boundary F (VF_BWFF), flags and INDC
 * (VF_FNFINI, VF_FNFPRV, VF_CINDC), then boundary values (VF_BWUWN, VF_BSUWT,
 * VF_BSUWT2, VF_BSUWN3, VF_BSPP, VF_BWUWT, VF_BWPP).  flags bit 0: set the
 * initial velocity (MATE I.C. V, vf_cinit.f:216-258); bit 1: set the pressure
 * RHO0*GRZ0*(W-LEVEL - z) on active cells (vf_cinit.f:283-296).  With flags=0
 * the uploaded velocity and pressure are kept, as after a restart. */
#define CAD_INIT_VELOCITY 1
#define CAD_INIT_PRESSURE 2
int cad_init_state(cad_ctx *ctx, double wvlvl, double uini, double vini, double wini, int flags,
                   cad_step_info *info);

/* TimerDoor state carried between steps, the arguments TBUB(NUMK) and
 * DROPTX, DROPTY, DROPTZ, DROPUU, DROPVV, DROPWW (NUMI,NUMJ,NUMK) of VF_F1CAL
 * (vf_f1cal.f:2-3, 59-65): the bubble-level timers and the droplet times and
 * velocities of VF_FBUBUP / VF_FDROPF.  cad_init_state sets their VF_CINIT
 * values (vf_cinit.f:412-421); upload them with the rest of a restart. */
int cad_upload_tdoor(cad_ctx *ctx, const double *tbub, const double *droptx, const double *dropty,
                     const double *droptz, const double *dropuu, const double *dropvv,
                     const double *dropww);
int cad_download_tdoor(cad_ctx *ctx, double *tbub, double *droptx, double *dropty, double *droptz,
                       double *dropuu, double *dropvv, double *dropww);

/* Matrix-table wave-maker (MODEL WAVE-BC <side> FUNC MATRIX, vf_iimdl.f:33-115,
 * applied in VF_BWUWN, vf_bwuwn.f:257-480) -------------------------------- */
/* The LEVEL-ON table of <name>.mtb as VF_IM1INP reads it (vf_im1inp.f): NZ
 * increasing heights ZZ (relative to the still-water level), NT increasing
 * times TT (fractions of a period), the level HH(NT), the normal and
 * tangential velocities UN(NZ,NT), UT(NZ,NT) in Fortran order, and the phase
 * origin T0.  A LEVEL-OFF table (no level column) is passed with hh = NULL:
 * its level is 0 and its profiles are not scaled by the measured level
 * (vf_bwuwn.f:326-329).  A LEVEL-ONLY table (a fixed water level) is passed
 * with nz = 0 and zz, un, ut NULL: its sides' faces are then taken as free faces
 * (the list pass of VF_BWUWN, VF_VPCOEF, VF_VMODIF, VF_BWUWT) and the cells next
 * to them take F from the level (VF_BWFFSF, in VF_F1CAL and the initial state).
 * Resets the time search start (MTBNOW) and clears the table sides set before
 * (cad_set_wave comes after the table). */
int cad_set_wave_table(cad_ctx *ctx, int nz, int nt, const double *zz, const double *tt,
                       const double *hh, const double *un, const double *ut, double t0);
/* One wave-maker side: 1..4 = X-, X+, Y-, Y+; columns l1..l2 along it
 * (1-based with ghosts, within 2..NUMJ-1 or 2..NUMI-1), water depth, period,
 * the ramp length in periods (AMPL; below 1e-20 no ramp) and the still-water
 * level (W-LEVEL).  The side's faces carry boundary type 5 in INDB (velocity
 * and F); every VF_BWUWN from then on (the predictor, the initial state, the
 * sub-loop passes) sets their velocities at the context's TNOW. */
int cad_set_wave(cad_ctx *ctx, int side, int l1, int l2, double depth, double period,
                 double ampl, double wvlvl);
/* The second table, FUNC MATRIX2 with <name>.mtb2 (VF_IM2INP, vf_im2inp.f;
 * VF_CWMTB12/22): the same as cad_set_wave_table / cad_set_wave for a table of
 * its own, with its own kind, search start (MTBNOW2) and phase origin. */
int cad_set_wave_table2(cad_ctx *ctx, int nz, int nt, const double *zz, const double *tt,
                        const double *hh, const double *un, const double *ut, double t0);
int cad_set_wave2(cad_ctx *ctx, int side, int l1, int l2, double depth, double period,
                  double ampl, double wvlvl);
/* A theory wave-maker side, MODEL WAVE-BC <side> FUNC STOKES / CNOIDAL / STK-CND
 * / STREAM n (vf_iimdl.f:51-70; VF_CWMAK0/1/2, vf_wstk0.f, vf_wcnd0.f,
 * vf_wsfmb2.f): func -2 fifth-order Stokes, -1 third-order cnoidal, 0 one of
 * the two by the Ursell number g H T^2 / d^2 (Stokes up to 25,
 * vf_csetup.f:101-111), 1..21 the stream function of order func (Dean's
 * method B, second definition of the wave speed; CAD_ERR_ARG if it does not
 * converge, as VF_CSETUP stops; the reference's order 22 writes past its own
 * arrays and is refused).  Columns l1..l2 as in
 * cad_set_wave; DEPTH, HEIGHT, PERIOD, AMPL and W-LEVEL.  The phase origin is
 * where the wave's level crosses zero (VF_CWZERO).  Every VF_BWUWN from then on
 * sets the side's face velocities from the wave, scaled by the columns'
 * measured levels, as for the table.  angle (degrees, ANGLE; 0 = normal
 * incidence) turns the wave about the side's normal with the phase shifted
 * along the ray through (ax, ay) by the set-up wave length (vf_bwuwn.f:297-325,
 * BCTYP(9..11)). */
int cad_set_wave_theory(cad_ctx *ctx, int side, int func, int l1, int l2, double depth,
                        double height, double period, double ampl, double wvlvl, double angle,
                        double ax, double ay);
/* The theory wave itself, on the host and without a context: one VF_CWMAK0 of
 * the wave (func -2 Stokes, -1 cnoidal, 1..21 stream function; depth, height
 * in m, period in s), then
 * optionally its wave length (length), its phase origin (VF_CWZERO, t0), its
 * level at phase wvt (VF_CWMAK1, level) and the velocities at nz heights zc
 * from the still water level (VF_CWMAK2, un, ut).  stk_state is VF_STK0's secant step carried from
 * call to call (0 before the first); cad_set_wave_theory and every boundary
 * pass advance the context's own copy the same way. */
int cad_wave_theory_eval(int func, double depth, double height, double period, float *stk_state,
                         double wvt, int nz, const double *zc, double *un, double *ut,
                         double *level, double *t0, double *length);
/* A radiation (open) side, MODEL OPEN-BC <side> FUNC TYPE1 (vf_iimdl.f:174-210,
 * applied in VF_BWUWN, vf_bwuwn.f:483-620): side 1..4 = X-, X+, Y-, Y+, the
 * DEPTH and PERIOD whose small-amplitude wave (VF_COMAK0) gives the
 * celerity.  The side's faces carry boundary type 7 in INDB (velocity and
 * F); every VF_BWUWN relaxes their values towards the inside velocities of
 * the start of the step. */
int cad_set_open(cad_ctx *ctx, int side, double depth, double period);
/* The wave-maker level of a side at the last VF_BWUWN (table level times the
 * ramp; the W-LEVEL ANS probe of FILE TRN, vf_ot1trn.f). */
int cad_get_wave_level(cad_ctx *ctx, int side, double *level);

/* Dupuit-Forchheimer porous drag, OPTION DRAG-DF yu n with n lines 'dr ap bt'
 * (vf_iiopt.f:86-108, applied in VF_VGENE, vf_vgene.f:111-146, 184-222,
 * 260-298): YU > 0 and n = 1..10 rows of a diameter DR > 0 and coefficients
 * AP, BT >= 0.  The predictor then takes the drag U(alpha + beta|U|)GV from
 * the face porosity GV and the face value of POROUS CD read as a diameter,
 * in place of the CD form.  n = 0 restores the CD form (the default, IDRGN=0
 * in vf_a2dflt.f:155). */
int cad_set_drag_df(cad_ctx *ctx, double yu, int n, const double *dr, const double *ap,
                    const double *bt);

/* A damping zone, MODEL DAMP <side> DEGREE/PARAM-XY/PARAM-Z/WIDTH/DEPTH
 * (vf_iimdl.f:216-258; IDAMP(side) and DAMP(1..4,side)), applied implicitly
 * after the predictor's Euler update (VF_VEULER, vf_veuler.f:75-166, 187-221,
 * 243-277): side 1..4 = X-, X+, Y-, Y+.  Within WIDTH of that side of the grid
 * a face's velocity is divided by 1 + DTNOW G DMAX/GV0 with DMAX the largest
 * (DEGREE+1) PARAM sqrt(GRZ0/DEPTH) (1 - dist/WIDTH)**DEGREE over the zones
 * (PARAM-XY for u and v, PARAM-Z for w).  degree < 0 removes the side's zone
 * (IDAMP = -1, the default of vf_a2dflt.f:150). */
int cad_set_damp(cad_ctx *ctx, int side, int degree, double param_xy, double param_z,
                 double width, double depth);

/* Whole phases ------------------------------------------------------------ */
/* VF_V1CAL (vf_v1cal.f:1-297): predictor, Poisson, correction, boundary values. */
int cad_v1cal(cad_ctx *ctx, cad_step_info *info);
/* VF_F1CAL (vf_f1cal.f:1-239): VOF advection, redistribution, flags with the
 * TimerDoor bubble / droplet corrections (at the context's TNOW), BCs. */
int cad_f1cal(cad_ctx *ctx, cad_step_info *info);
/* One pass of the time loop body (vf_a1main.f:1029-1104) with DT = DTNOW:
 * LOOPS passes of TNOW += DT/LOOPS and V1CAL at DTNOW = DT/LOOPS (OPTION
 * SUB-LOOP; passes 2.. first refresh the boundary values as
 * vf_v1cal.f:115-118 does), then DTNOW = DT and F1CAL. */
int cad_step(cad_ctx *ctx, cad_step_info *info);
/* VF_CDTCAL (vf_cdtcal.f, called at vf_a1main.f:940): the next step size
 * from the resident fields for TIME AUTO dtinit dtsafe / TIME LIMIT dtmin
 * dtmax, with W = DTNOW (dtinit when nnow = 0).  The cell limits are
 * combined by an exact minimum; the reference folds them in serially, which
 * can differ by the rounding of V/(V/DT) per cell (a few ulps). */
int cad_cdtcal(cad_ctx *ctx, int nnow, double dtinit, double dtsafe, double dtmin, double dtmax,
               double *dt);
/* VF_CDIV00 (vf_cdiv00.f:1-56): || sum of face fluxes ||_2 over active cells. */
int cad_cdiv00(cad_ctx *ctx, double *dv);

/* Finer-grained stages of VF_V1CAL, for stage-wise checking ---------------- */
/* VF_VFLXDU/V/W + VF_VGENE + VF_VEULER + VF_BWUWN (vf_v1cal.f:126-251) */
int cad_predictor(cad_ctx *ctx);
/* VF_VPCOEF (vf_v1cal.f:261) into the context's matrix */
int cad_vpcoef(cad_ctx *ctx);
/* VF_VPSOL (vf_vpsol.f:55-76) on the context's matrix; PT in the context.
 * With precision CAD_PRECISION_MIXED the solve is the mixed form's: FP64
 * BiCGSTAB right-preconditioned by an FP32 aggregation multigrid V-cycle,
 * stopped when ||BB - A PT|| <= CGEPSR ||BB|| + CGEPSA (the true residual,
 * rechecked explicitly); info gets ||BB||, the final true relative residual
 * and the iteration count. */
int cad_vpsol(cad_ctx *ctx, cad_step_info *info);
/* VF_VMODIF + VF_BSUWT/BSUWT2/BSUWN/BWUWT/BWPP (vf_v1cal.f:287-297) */
int cad_correction(cad_ctx *ctx);
int cad_download_matrix(cad_ctx *ctx, double *ad, double *ali, double *alj, double *alk,
                        double *aui, double *auj, double *auk, double *bb);
int cad_upload_matrix(cad_ctx *ctx, const double *ad, const double *ali, const double *alj,
                      const double *alk, const double *aui, const double *auj, const double *auk,
                      const double *bb, const int *indc);
int cad_download_pt(cad_ctx *ctx, double *pt);

/* The true relative residual ||BB - A*PT||_2 / ||BB||_2 over the active cells,
 * in FP64, of the last pressure solve (its matrix and potential).  No
 * counterpart in the reference (VF_M1BCGS stops on the preconditioned
 * residual); it is what the two forms' solves are compared on. */
int cad_poisson_residual(cad_ctx *ctx, double *rel);
/* The preconditioner application of the context's form, timed with HIP
 * events on its stream (average of reps) after a pressure solve: the MILU
 * M^-1 y (as cad_time_minv; 96 B per active cell) in the plain form, one
 * multigrid V-cycle in the mixed form (its algorithmic bytes from the level
 * sizes).  No counterpart in the reference; it is the benchmark's dominant
 * kernel. */
int cad_time_precond(cad_ctx *ctx, int reps, float *ms_avg, double *bytes);
/* Measurement helpers (no counterpart in the reference): average duration of
 * one MILU preconditioner application M^-1 y (vf_mzminv.f) on the current
 * factorisation, timed with HIP events on the context's stream; and the number
 * of active cells (INDC != -1) that the solver's vector operations touch. */
int cad_time_minv(cad_ctx *ctx, int reps, float *ms_avg);
int cad_count_active(cad_ctx *ctx, long long *active);
/* In-solve timing of the forms' dominant kernels (no counterpart in the
 * reference; measurement support).  With on = n >= 1, every n-th launch
 * inside the solves of the plain form's M^-1 (CAD_KERNEL_MINV), the mixed
 * form's FP32 V-cycle (CAD_KERNEL_VCYCLE) and the learned form's network pass
 * (CAD_KERNEL_UNET) is bracketed by HIP events on the context's stream
 * (without a system-scope fence), with its algorithmic work: the V-cycle's
 * bytes from the level sizes and the solve's K range, the pass's FLOP (2 per
 * multiply-add) on the solve's box, 0 for M^-1 (96 B per active cell, which
 * the caller counts); on = 0 stops.  Turning it on clears the totals;
 * cad_kernel_timing gives the average duration and work per timed launch and
 * the number of timed launches since then. */
#define CAD_KERNEL_MINV 0
#define CAD_KERNEL_VCYCLE 1
#define CAD_KERNEL_UNET 2
int cad_set_kernel_timing(cad_ctx *ctx, int on);
int cad_kernel_timing(cad_ctx *ctx, int which, double *ms_avg, double *work_avg, long long *count);

/* Cap on the flag-fix passes of VF_FNFINI (vf_fnfini.f:108-180, repeated
 * until no surface cell is left without fluid on one side and gas on the
 * other).  Each pass that fixes a cell turns at least one cell into gas and
 * none back, so the loop ends within (interior cells + 1) passes, the
 * default (passes = 0).  A state still unfixed after the cap makes the
 * flagging pass fail with CAD_ERR_STATE instead of going on with
 * inconsistent flags.  No counterpart in the reference, whose loop has no
 * cap. */
int cad_set_flag_pass_limit(cad_ctx *ctx, long long passes);

/* The plain form's MILU-BiCGSTAB (cad_vpsol with CAD_PRECISION_FP64) stops
 * on the reference's test ||M^-1 r||^2 <= ||M^-1 b||^2 CGEPSR^2 + CGEPSA^2
 * (vf_m1bcgs.f:144,219).  With rel > 0 it also requires the true relative
 * residual ||BB - A PT|| / ||BB|| <= rel, rechecked in FP64 whenever the
 * reference's test passes, and iterates on otherwise; 0 (the default)
 * restores the reference's test alone.  The mixed and learned forms stop on
 * the true residual ||BB - A PT|| <= CGEPSR ||BB|| + CGEPSA; with rel > 0
 * also on ||BB - A PT|| <= rel ||BB||, so that CGEPSA cannot loosen it.  No
 * counterpart in the reference: it holds the three forms to the same true
 * relative residual, so that their solves are compared at it. */
int cad_set_true_residual_stop(cad_ctx *ctx, double rel);

/* Schedule of the MILU triangular sweeps (vf_mzdcmp.f, vf_mzfrwd.f,
 * vf_mzbkwd.f).  All give the serial sweep's results bit for bit:
 *   CAD_SWEEP_PIPELINED       (default) one launch per sweep, wavefront tiles
 *                             of 64 lines walking I and handing values on
 *                             point to point; the tile shape (8x8 or 16x4
 *                             lines in J x K) is chosen per solve from the
 *                             active region;
 *   CAD_SWEEP_PIPELINED_8X8,
 *   CAD_SWEEP_PIPELINED_16X4  the same with a fixed tile shape;
 *   CAD_SWEEP_HYPERPLANE      one launch per hyperplane I+J+K = const.
 * No reference counterpart: the reference sweeps serially. */
#define CAD_SWEEP_PIPELINED 0
#define CAD_SWEEP_HYPERPLANE 1
#define CAD_SWEEP_PIPELINED_8X8 2
#define CAD_SWEEP_PIPELINED_16X4 3
int cad_set_sweep_schedule(cad_ctx *ctx, int schedule);

/* The MILU preconditioner on its own, for checking the sweeps: factor the
 * uploaded matrix (vf_mzdcmp.f:49-67; *npos = 0 when a pivot is not positive)
 * and apply x = M^-1 y (vf_mzminv.f:38-41) to a host vector (x is 0 on
 * inactive cells).  The reference calls these only inside VF_M1BCGS, so they
 * have no entry-point counterpart there. */
int cad_milu_factor(cad_ctx *ctx, int *npos);
int cad_milu_apply(cad_ctx *ctx, const double *y, double *x);

/* Link-level drop-in of the reference's pressure solve -------------------
 * The reference's own routines under their own link names (flang/gfortran:
 * lowercase with a trailing underscore) and by-reference argument lists, so
 * that the reference links this library in place of vf_m1bcgs.o, or of
 * vf_vpsol.o and vf_m1bcgs.o, with no other change.  Arrays are the
 * reference's explicit-shape column-major host arrays.  Single MPI rank
 * (loops over 2..NI-1, 2..NJ-1 as MYIS..MYIE, MYJS..MYJE are at one rank);
 * DBUF and the work arrays DD..AE are accepted and left untouched (the
 * reference's callers do not read them afterwards, vf_v1cal.f:278-297).
 * The device context is created on the first call and kept while the grid
 * size is unchanged. */

/* COMMON blocks with the reference's member order (VF_ACOMPR.h, VF_ACOMPI.h,
 * VF_ANUMBI.h; MAXNC = 10 in VF_A0PRM.h).  The library defines them weakly:
 * linked into the reference, the reference's blocks are the ones used. */
typedef struct vf_acompr_block {
    double cgpara, cgepsa, cgepsr, cgdiv, cgbnrm, cgxnrm;
    double feps, flower, fupper, fsum, fcut, plower;
    double scmvp, scmk, scmt, scmc[10];
} vf_acompr_block;
typedef struct vf_acompi_block {
    int icgtyp, icgmax, icgitr, iscmvp, iscmff, iscmk, iscmt, iscmc[10], ibsuw0;
} vf_acompi_block;
typedef struct vf_anumbi_block {
    int numi, numj, numk, numb, nums, leqk, leqt, leqc, numij, numbx;
} vf_anumbi_block;
extern vf_acompr_block vf_acompr_;
extern vf_acompi_block vf_acompi_;
extern vf_anumbi_block vf_anumbi_;

/* VF_M1BCGS (vf_m1bcgs.f:1-5): MILU-BiCGSTAB on A x = b, arrays (MI,MJ,MK)
 * with the loop range (NI,NJ,NK), MI >= NI etc.  X0 in: the initial guess;
 * out: the solution where ITER >= 0 (active cells), unchanged otherwise.
 * BNORM = ||M^-1 b||, XNORM = the last ||M^-1 r||, ITER as the reference's:
 * >= 0 iterations, -1 MILU pivot <= 0, -2 divergence, > ICGMAX not
 * converged.  A failure of the library itself (sizes, device, memory) also
 * returns ITER = ICGMAX + 1, which VF_VPSOL stops on (vf_vpsol.f:70-71),
 * with the cause on stderr and in cad_last_error(). */
void vf_m1bcgs_(double *dbuf, const double *ad, const double *ali, const double *alj,
                const double *alk, const double *aui, const double *auj, const double *auk,
                const double *bb, double *x0, double *dd, double *xw, double *yy, double *r0,
                double *rr, double *pp, double *ap, double *ae, const double *cgpara,
                const double *cgepsa, const double *cgepsr, const double *cgdiv, double *bnorm,
                double *xnorm, const int *indc, const int *mi, const int *mj, const int *mk,
                const int *ni, const int *nj, const int *nk, const int *icgtyp,
                const int *icgmax, int *iter);

/* VF_VPSOL (vf_vpsol.f:1-3): PT = 0, then VF_M1BCGS on (NUMI,NUMJ,NUMK) of
 * /VF_ANUMBI/ with CGPARA, CGEPSA, CGEPSR, CGDIV of /VF_ACOMPR/ and ICGTYP,
 * ICGMAX of /VF_ACOMPI/; CGBNRM, CGXNRM and ICGITR are written back there.
 * Its error checks (vf_vpsol.f:70-75) call the reference's VF_A2ERR when the
 * program has one (linked into the reference); other callers read ICGITR and
 * cad_last_error(). */
void vf_vpsol_(double *dbuf, const double *ad, const double *ali, const double *alj,
               const double *alk, const double *aui, const double *auj, const double *auk,
               const double *bb, double *pt, double *dd, double *xw, double *yy, double *r0,
               double *rr, double *pp, double *ap, double *ae, const int *indc);

/* No counterpart in the reference: the VF_M1BCGS solve of vf_m1bcgs_ on
 * arrays already resident on the current GPU, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  It allocates nothing.  Arguments keep
 * VF_M1BCGS's names and meaning (vf_m1bcgs.f:1-5) with MI,MJ,MK = NI,NJ,NK,
 * all arrays as device pointers: AD..BB, X0 (in: initial guess; out: the
 * solution where ITER >= 0), INDC, and the reference's work arrays DD, XW,
 * YY, R0, RR, PP, AP, AE (NI*NJ*NK doubles each, contents on return
 * unspecified).  BNORM, XNORM, ITER are host pointers with vf_m1bcgs_'s
 * meaning; the result equals vf_m1bcgs_'s on the same system bit for bit.
 * stream: a hipStream_t (NULL: the null stream); the call returns after the
 * stream has finished the solve.  Everything else the solve needs lives in
 * one library-owned device array of CAD_NOALLOC_SCRATCH_BYTES per GPU, held
 * until the process ends; grids beyond CAD_NOALLOC_MAX_CELLS cells, or whose
 * sweep mailboxes do not fit the array, are refused with CAD_ERR_ARG.  Calls
 * that overlap on one GPU share that array and race.  Returns CAD_OK, or a
 * negative code with the cause in cad_last_error(). */
#define CAD_NOALLOC_MAX_CELLS 2097152ull                                   /* 2^21 */
#define CAD_NOALLOC_SCRATCH_BYTES ((72ull << 20) + (64ull << 10))          /* 72.06 MiB */
int cad_m1bcgs_noalloc(void *stream, const double *ad, const double *ali, const double *alj,
                       const double *alk, const double *aui, const double *auj,
                       const double *auk, const double *bb, double *x0, double *dd, double *xw,
                       double *yy, double *r0, double *rr, double *pp, double *ap, double *ae,
                       const int *indc, int ni, int nj, int nk, double cgpara, double cgepsa,
                       double cgepsr, double cgdiv, int icgtyp, int icgmax, double *bnorm,
                       double *xnorm, int *iter);

/* The mixed and learned forms --------------------------------------------
 * No counterparts in the reference (its only solver is VF_M1BCGS).  Both
 * stop on the true relative residual ||BB - A PT|| / ||BB|| <= CGEPSR and,
 * when their own iteration stalls (no convergence within the cap, 200 by
 * default and at most ICGMAX, or for the learned form a restart cycle that
 * does not halve the residual), hand the solve to VF_M1BCGS from their
 * current iterate with the same true-residual stop.  cad_set_krylov_cap sets
 * the cap (0: the default); cad_mixed_stats gives the mixed form's fallback
 * count and the last fallback's BiCGSTAB iterations. */
int cad_set_krylov_cap(cad_ctx *ctx, int cap);
/* One multigrid V-cycle on the hierarchy and K range of the last mixed or
 * learned solve, timed with HIP events on the context's stream, and its
 * algorithmic bytes; fp64 != 0 runs the same V-cycle with every level in
 * FP64 (a hierarchy of its own from the same matrix), to measure what FP32
 * buys. */
int cad_time_vcycle(cad_ctx *ctx, int reps, int fp64, float *ms_avg, double *bytes);
int cad_mixed_stats(cad_ctx *ctx, long long *out2);

/* The learned form's network: a 3-level 3-D U-Net, 3x3x3 convolutions with
 * 16/32/64 channels, 8 input channels (the residual scaled by its RMS, GGV,
 * F, the cell-type one-hot, the activity), one output (the correction on the
 * same scale), cad_unet_param_count() = 390,913 FP32 parameters.  Weights and
 * activations in BF16 on the matrix cores, FP32 accumulation.  cad_set_learned_
 * weights gives a context its parameters (host FP32 vector); cad_set_learned_
 * mode selects mode 1 (the default: one V-cycle, then the network on its
 * residual) or 0 (the network alone) and the restart length (1..32, default
 * 30).  cad_learned_stats: the last solve's FGMRES and fallback iterations,
 * solves, fallbacks, all FGMRES iterations, the network box's depth.
 * cad_learned_box: the network's box for the current flags (nx, ny, nz, and
 * the first K plane).  cad_learned_sample: training pairs from the plain
 * solver on the context's current system, for the iterates listed in iters
 * (see learned.py); feat, target, mask are device buffers. */
int cad_set_learned_weights(cad_ctx *ctx, const float *params, int n);
int cad_set_learned_mode(cad_ctx *ctx, int mode, int restart);
int cad_learned_stats(cad_ctx *ctx, long long *out6);
int cad_learned_box(cad_ctx *ctx, int *box4);
/* The context's current Poisson matrix on the network's box, FP32 on the
 * device, coefficient-major (AD, ALI, AUI, ALJ, AUJ, ALK, AUK; 7*nx*ny*nz
 * values, x fastest; AD = 1 and no couplings beyond the grid): for training
 * losses that apply the matrix to the network's output.  No reference
 * counterpart. */
int cad_learned_box_matrix(cad_ctx *ctx, float *out);
/* Average duration of the network's forward pass on the current box (HIP
 * events on the context's stream) and its work, 2 FLOP per multiply-add. */
int cad_time_learned(cad_ctx *ctx, int reps, float *ms_avg, double *flops);
int cad_learned_sample(cad_ctx *ctx, int mode, const int *iters, int n, void *feat, float *target,
                       float *mask);

/* The network on its own (training and tests): a batch of nb boxes of
 * nx x ny x nz voxels (nx a multiple of 64, ny and nz of 4); inputs are BF16
 * (nb, nz+2, ny+2, nx+2, 8) with zero halos, outputs FP32 nb*nz*ny*nx (x
 * fastest).  train != 0 also holds the gradients and Adam's moments.
 * cad_unet_train_step: one Adam step on the loss mean over boxes of
 * sum m (y-e)^2 / sum m e^2; cad_unet_gradients: that loss and its gradient;
 * cad_unet_train_step_dy: one Adam step from the caller's gradient dy
 * (device FP32, the output's order) of any loss with respect to the output
 * of the last cad_unet_forward; cad_unet_gradients_dy: the parameters'
 * gradient from such a dy (host FP32), with no step.
 * stream: a hipStream_t or NULL (the network's own). */
typedef struct cad_unet cad_unet;
int cad_unet_param_count(void);
double cad_unet_mac_per_voxel(void);
int cad_unet_init_params(int seed, float *params);
int cad_unet_create(int nx, int ny, int nz, int nb, int train, void *stream, cad_unet **out);
void cad_unet_destroy(cad_unet *net);
int cad_unet_set_params(cad_unet *net, const float *params);
int cad_unet_get_params(cad_unet *net, float *params);
int cad_unet_set_input(cad_unet *net, const void *dev);
int cad_unet_forward(cad_unet *net);
int cad_unet_get_output(cad_unet *net, float *dev);
int cad_unet_set_depth(cad_unet *net, int nz);
int cad_unet_train_step(cad_unet *net, const float *target, const float *mask, float lr, float *loss);
int cad_unet_train_step_dy(cad_unet *net, const float *dy, float lr);
int cad_unet_gradients_dy(cad_unet *net, const float *dy, float *grad);
int cad_unet_gradients(cad_unet *net, const float *target, const float *mask, float *grad,
                       float *loss);

#ifdef __cplusplus
}
#endif
#endif
