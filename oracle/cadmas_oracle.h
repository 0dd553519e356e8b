/*
 * cadmas_oracle.h -- CPU restatement of the CADMAS-SURF/3D single-phase
 * VOF-SMAC time step (plain C, serial, the reference's loop and expression
 * order).
 *
 * TEST INFRASTRUCTURE: this library is the checker for the HIP product path
 * and the CPU baseline timed beside it.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it.  It is never linked into the
 * product library.
 *
 * Conventions follow the reference's Fortran (Source code/ of
 * Simulators/CADMAS-SURF-3D):
 *   - arrays are column major, (NUMI,NUMJ,NUMK), I fastest, 1-based indices
 *     in the restatement (Source code/vf_a1main.f:150-151);
 *   - metric tables XX(6,NUMI), YY(6,NUMJ), ZZ(6,NUMK) (vf_cgrid.f:44-67);
 *   - boundary list INDB(4,NUMB) (vf_cindb.f);
 *   - one rank: MYIS=MYJS=2, MYIE=NUMI-1, MYJE=NUMJ-1, MYMIS..MYMJE=1
 *     (vf_cpara.f:81-98).
 */
#ifndef CADMAS_ORACLE_H
#define CADMAS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_state {
    int numi, numj, numk, numb;
    const double *xx, *yy, *zz;
    double *uu, *vv, *ww, *pp, *ff;
    const double *anu, *cd0;
    const double *ggv, *ggx, *ggy, *ggz, *glv, *glx, *gly, *glz;
    double *bcu, *bcv, *bcw, *bcp, *bcf;
    const double *bcvi;
    int *nf;
    const int *indx, *indy, *indz;
    int *indc;
    const int *indb;
    int *inds;       /* surface-cell list, NUMI*NUMJ*NUMK capacity */
    int *nlim;       /* VF_F1CAL work (NUMI,NUMJ,NUMK); holds old NF on return */
    int nums;
    /* physics and numerics (COMMON /VF_APHYSR/, /VF_ACOMPR/, /VF_ACOMPI/) */
    double rho0, grz0, anu0, dtnow;
    double scmvp, vvmax, flower, fupper, plower;
    double cgpara, cgepsa, cgepsr, cgdiv;
    int icgtyp, icgmax, ibsuw0, loops;
    /* per-step outputs */
    double cgbnrm, cgxnrm, fsum, fcut;
    int icgitr;
    /* TimerDoor models (vf_f1cal.f:208-217): bubbles rise when
     * WBUB >= 1e-20, droplets fall when IDROP >= 1.  TNOW is the time of
     * COMMON /VF_ATIMER/ (orc_step advances it by DTNOW first, as the time
     * loop does at vf_a1main.f:1047); tbub holds TBUB(NUMK), the drop*
     * fields DROPTX..DROPWW (NUMI,NUMJ,NUMK).  Unused when both are off. */
    double tnow, wbub;
    int idrop;
    double *tbub;
    double *droptx, *dropty, *droptz, *dropuu, *dropvv, *dropww;
    /* Table wave-makers, MODEL WAVE-BC FUNC MATRIX with a LEVEL-ON table
     * (vf_bwuwn.f:257-480): per side JD = 1..4 (x-, x+, y-, y+) whether it
     * is on, its range L1..L2 (IBCTYP(3..4)), DEPTH, PERIOD, AMPL and the
     * initial phase T0; the table DMTBZZ(MTBZZ), DMTBTT/DMTBHH(MTBTT),
     * DMTBUN/DMTBUT(MTBZZ,MTBTT); MTBNOW, the search position the table
     * lookup keeps between calls; wv_level the side's level (BCTYP(7)).
     * wvlvl is MATE W-LEVEL. */
    double wvlvl;
    int wv_on[4], wv_l1[4], wv_l2[4];
    double wv_d[4], wv_t[4], wv_a[4], wv_t0[4], wv_level[4];
    int mtbzz, mtbtt, mtbnow;
    const double *mtb_zz, *mtb_tt, *mtb_hh, *mtb_un, *mtb_ut;
    /* Radiation boundaries, MODEL OPEN-BC side FUNC TYPE1 (vf_bwuwn.f:483-620):
     * per side whether it is on, with the DEPTH and PERIOD that give its
     * celerity (VF_COMAK0). */
    int op_on[4];
    double op_d[4], op_t[4];
    /* Dupuit-Forchheimer drag, OPTION DRAG-DF (vf_iiopt.f:86-108): IDRGN
     * table rows (0 = off, the CD form), DRGYU and per row the diameter
     * DRGDR and the coefficients DRGAP, DRGBT (MAXDRG = 10 rows). */
    int idrgn;
    double drgyu, drgdr[10], drgap[10], drgbt[10];
    /* Damping zones, MODEL DAMP side (vf_iimdl.f:216-258): per side JD =
     * 1..4 (x-, x+, y-, y+) whether it is on (IDAMP(JD) >= 0), its DEGREE
     * IDAMP and DAMP(1..4,JD) = PARAM-XY, PARAM-Z, WIDTH, DEPTH; the grid's
     * extent GLXMIN, GLXMAX, GLYMIN, GLYMAX (vf_iigrid.f:124-131). */
    int dm_on[4], dm_deg[4];
    double dm_par[4][4];
    double glxmin, glxmax, glymin, glymax;
    /* the table is LEVEL-OFF (MTBTYP=2): no level column, profiles unscaled;
     * LEVEL-ONLY (MTBTYP=3): a fixed level, no profiles */
    int mtb_off, mtb_only;
    /* the second table (FUNC MATRIX2, <name>.mtb2; VF_IM2INP, VF_CWMTB12/22)
     * and per side the table it reads (0 the first, 1 the second) */
    int mtbzz2, mtbtt2, mtbnow2;
    const double *mtb_zz2, *mtb_tt2, *mtb_hh2, *mtb_un2, *mtb_ut2;
    int mtb_off2, mtb_only2;
    int wv_tab[4];
} orc_state;

/* Poisson matrix and solver (vf_vpcoef.f, vf_m1bcgs.f and vf_mz*.f) */
typedef struct orc_matrix {
    double *ad, *ali, *alj, *alk, *aui, *auj, *auk, *bb;
} orc_matrix;

/* phase (a): fluxes, sources, Euler update, normal wall velocities */
int orc_predictor(orc_state *s);
/* phase (b): matrix assembly; solve with MILU-BiCGSTAB into pt */
int orc_vpcoef(orc_state *s, orc_matrix *m);
int orc_m1bcgs(int numi, int numj, int numk, const orc_matrix *m, double *x0,
               const int *indc, double cgpara, double cgepsa, double cgepsr, double cgdiv,
               int icgtyp, int icgmax, double *bnorm, double *xnorm, int *iter);
int orc_mzdcmp(int numi, int numj, int numk, const orc_matrix *m, double *dd,
               double cgpara, const int *indc, int icgtyp);
void orc_mzminv(int numi, int numj, int numk, const double *dd, const orc_matrix *m,
                double *x, const double *y, const int *indc);
void orc_mzax(int numi, int numj, int numk, const orc_matrix *m, const double *x,
              double *y, const int *indc);
double orc_mzip(int numi, int numj, int numk, const double *x, const double *y,
                const int *indc);
/* Summation order of orc_mzip, for studying how the solve depends on the
 * rounding of its dot products: ORC_DOT_SERIAL (the default) adds the terms
 * one at a time in the reference's loop order; ORC_DOT_EXTENDED accumulates
 * them in long double and rounds once; ORC_DOT_BLOCKED adds blocks of 1024
 * terms serially and then the block sums.  Returns the previous order, or -1
 * for an unknown one. */
enum { ORC_DOT_SERIAL = 0, ORC_DOT_EXTENDED = 1, ORC_DOT_BLOCKED = 2 };
int orc_set_dot_order(int order);
/* phase (d): velocity / pressure correction and the boundary values after it */
int orc_vmodif(orc_state *s, const double *pt);
int orc_post_bc(orc_state *s);
/* VF_V1CAL: (a) + (b) + (d); pt receives the potential (WK09) */
int orc_v1cal(orc_state *s, double *pt);
/* VF_F1CAL: phase (c) */
int orc_f1cal(orc_state *s);
/* TimerDoor corrections on their own (vf_fbubup.f, vf_fdropf.f) */
void orc_fbubup(orc_state *s);
int orc_fdropf(orc_state *s);
/* VF_CDIV00 */
double orc_cdiv00(const orc_state *s);
/* the part of VF_CINIT that follows the initial F (clip, flags, p, BCs) */
int orc_init_state(orc_state *s, double wvlvl, double uini, double vini, double wini, int flags);
/* one SMAC step at DT = dtnow: LOOPS passes of V1CAL at DT/LOOPS, F1CAL at DT */
int orc_step(orc_state *s, double *pt);
/* VF_V1CAL pass iloop of OPTION SUB-LOOP (iloop >= 2 refreshes boundary values) */
int orc_v1cal_loop(orc_state *s, double *pt, int iloop);
/* VF_CDTCAL: the next step size (TIME AUTO / LIMIT values as arguments) */
double orc_cdtcal(const orc_state *s, int nnow, double dtinit, double dtsafe, double dtmin,
                  double dtmax);

#ifdef __cplusplus
}
#endif
#endif
