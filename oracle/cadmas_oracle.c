/*
 * cadmas_oracle.c -- serial CPU restatement of the CADMAS-SURF/3D single-phase
 * VOF-SMAC time step.
 *
 * TEST INFRASTRUCTURE ONLY (see cadmas_oracle.h).  Every routine follows one
 * reference subroutine of Simulators/CADMAS-SURF-3D/Source code/ with the same
 * loop order, loop ranges and floating-point expression order, so that on the
 * same inputs it reproduces the reference bit for bit (build with
 * -ffp-contract=off; x86-64 baseline has no FMA).  Reference citations are
 * given per routine as file:line.
 *
 * Supported feature set (anything else returns a negative code):
 *   boundary types VP SLIP / NON-S / FIX-V / FREE (INDB(3,*) = 1..4),
 *   F FIX / FREE (INDB(4,*) = 1,2), static porosity (IPRNT<=1), donor scheme
 *   (ISCMVP=0, ISCMFF=0), drag by CD (IDRGN=0), no sponge (IDAMP<0), no wave
 *   maker / wave source / open boundary, no air-pressure model (PVCP0<ZERO),
 *   LOOPS=1.  The TimerDoor bubble and droplet models (OPTION T-DOOR) are
 *   restated (orc_fbubup, orc_fdropf).
 */
#include "cadmas_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ZERO 1.0e-20

/* ---- indexing (1-based, column major) ---------------------------------- */
#define DECL_DIMS(s)                                   \
    const int ni = (s)->numi, nj = (s)->numj, nk = (s)->numk; \
    const size_t nij = (size_t)ni * (size_t)nj;        \
    (void)nk; (void)nij
#define ID(i, j, k) ((size_t)((i) - 1) + (size_t)ni * (size_t)((j) - 1) + nij * (size_t)((k) - 1))
#define XX(m, i) (s->xx[(size_t)((i) - 1) * 6 + (m) - 1])
#define YY(m, j) (s->yy[(size_t)((j) - 1) * 6 + (m) - 1])
#define ZZ(m, k) (s->zz[(size_t)((k) - 1) * 6 + (m) - 1])
#define INDB(m, l) (s->indb[(size_t)((l) - 1) * 4 + (m) - 1])

static double dmax(double a, double b) { return a > b ? a : b; }
static double dmin(double a, double b) { return a < b ? a : b; }

/* metric of axis a (0=x,1=y,2=z), entry m (1..6), index i (1-based) */
static inline double MET(const orc_state *s, int a, int m, int i)
{
    const double *t = a == 0 ? s->xx : (a == 1 ? s->yy : s->zz);
    return t[(size_t)(i - 1) * 6 + m - 1];
}

static void decode(const orc_state *s, int ijk, int *i, int *j, int *k)
{
    DECL_DIMS(s);
    int r = ijk;
    *k = (r - 1) / (int)nij + 1;
    r = r - (int)nij * (*k - 1);
    *j = (r - 1) / ni + 1;
    *i = r - ni * (*j - 1);
}

/* The boundary type a face takes in VF_BWUWN's list pass, VF_VPCOEF,
 * VF_VMODIF and VF_BWUWT: a wave-maker face (5) of a side driven by a
 * LEVEL-ONLY table (a fixed level) is treated there as a free face (4)
 * (vf_bwuwn.f:161-211, vf_vpcoef.f:162-175, vf_vmodif.f:87-110,
 * vf_bwuwt.f:80-116). */
static int ib_lv(const orc_state *s, int l)
{
    const int ib = INDB(3, l), ns = INDB(2, l);
    if (ib == 5 && ns >= 1 && ns <= 4 && s->wv_on[ns - 1] &&
        (s->wv_tab[ns - 1] == 1 ? s->mtb_only2 : s->mtb_only))
        return 4;
    return ib;
}

/* ======================================================================
 * Phase (a): momentum predictor
 * ====================================================================== */

/* Normal flux of component a at cell centres: the (G u) u term blended by
 * SCMVP and G nu 2 du/dx.  vf_vflxdu.f:71-90 (u), vf_vflxdv.f:71-90 (v),
 * vf_vflxdw.f:66-85 (w). */
static void flux_normal(orc_state *s, int a, const double *q, const double *gl,
                        const double *gg, double *fl)
{
    DECL_DIMS(s);
    const size_t st[3] = {1, (size_t)ni, nij};
    const size_t sa = st[a];
    const double su = s->scmvp, sc = 1.0 - s->scmvp;
    for (int k = 2; k <= nk - 1; k++)
        for (int j = 2; j <= nj - 1; j++)
            for (int i = 2; i <= ni - 1; i++) {
                size_t n = ID(i, j, k);
                if (s->indc[n] == -1) continue;
                int ia = a == 0 ? i : (a == 1 ? j : k);
                double win = (gl[n] * q[n] + gl[n + sa] * q[n + sa]) * 0.5;
                double fg = win * (q[n] + q[n + sa]) * 0.5;
                fl[n] = fl[n] - su * (dmax(win, 0.0) * q[n] - dmax(-win, 0.0) * q[n + sa]) - sc * fg;
                fl[n] = fl[n] + (gg[n] + gg[n + sa]) * s->anu[n] * MET(s, a, 4, ia) * (q[n + sa] - q[n]);
            }
}

/* Cross flux of component a (q) through faces normal to axis b: (G_b v_b) q
 * and G_b nu (dq/db + dv_b/da) on the cell edge (a-face, b-face), including
 * the wall/free-surface special branch.  The six instances in the reference
 * are axis permutations of one formula:
 *   q=u,b=y vf_vflxdu.f:94-219   q=u,b=z vf_vflxdu.f:222-347
 *   q=v,b=x vf_vflxdv.f:94-219   q=v,b=z vf_vflxdv.f:222-347
 *   q=w,b=x vf_vflxdw.f:88-213   q=w,b=y vf_vflxdw.f:216-341
 * Loop ranges per instance are passed in (lo/hi for i, j, k). */
static void flux_cross(orc_state *s, int a, int b, const double *q, const double *vb,
                       const double *glb, const double *ggb, const int *ind_a,
                       const int *ind_b, const double *bcq, double *fl,
                       int i0, int i1, int j0, int j1, int k0, int k1)
{
    DECL_DIMS(s);
    const size_t st[3] = {1, (size_t)ni, nij};
    const size_t sa = st[a], sb = st[b];
    const double su = s->scmvp, sc = 1.0 - s->scmvp;
    const int nsb = 2 * b + 1; /* INDB(2,*) of a boundary on the minus-b side */
    for (int k = k0; k <= k1; k++)
        for (int j = j0; j <= j1; j++)
            for (int i = i0; i <= i1; i++) {
                size_t n = ID(i, j, k);
                const int *nf = s->nf;
                int ic = nf[n - sa - sb] * nf[n - sb] * nf[n - sa] * nf[n];
                int it = ind_a[n];
                int ib = ind_a[n - sb];
                int ia = a == 0 ? i : (a == 1 ? j : k);
                int jb = b == 0 ? i : (b == 1 ? j : k);
                if (ic != 0 || (it != 0 && ib != 0)) {
                    continue;
                } else if (it == 0 && ib == 0) {
                    double win = MET(s, a, 6, ia) * (MET(s, a, 2, ia - 1) * glb[n] * vb[n]
                                                     + MET(s, a, 2, ia) * glb[n - sa] * vb[n - sa]);
                    double fg = win * MET(s, b, 6, jb) * (MET(s, b, 2, jb - 1) * q[n] + MET(s, b, 2, jb) * q[n - sb]);
                    fl[n] = fl[n] - su * (dmax(win, 0.0) * q[n - sb] - dmax(-win, 0.0) * q[n]) - sc * fg;
                    double gg = MET(s, a, 6, ia) * (MET(s, a, 2, ia - 1) * ggb[n] + MET(s, a, 2, ia) * ggb[n - sa]);
                    double anr = MET(s, b, 6, jb) * (MET(s, b, 2, jb - 1) * s->anu[n] + MET(s, b, 2, jb) * s->anu[n - sb]);
                    double anl = MET(s, b, 6, jb) * (MET(s, b, 2, jb - 1) * s->anu[n - sa] + MET(s, b, 2, jb) * s->anu[n - sa - sb]);
                    double an = MET(s, a, 6, ia) * (MET(s, a, 2, ia - 1) * anr + MET(s, a, 2, ia) * anl);
                    double d1 = MET(s, b, 5, jb) * (q[n] - q[n - sb]);
                    double d2 = MET(s, a, 5, ia) * (vb[n] - vb[n - sa]);
                    fl[n] = fl[n] + gg * an * (d1 + d2);
                } else {
                    double br, bl;
                    int is = ind_b[n];
                    if (is >= 1) {
                        if (INDB(3, is) == 1 || INDB(3, is) == 4)
                            br = INDB(2, is) == nsb ? q[n] : q[n - sb];
                        else
                            br = bcq[is - 1];
                    } else {
                        br = MET(s, b, 6, jb) * (MET(s, b, 2, jb - 1) * (q[n] + q[n + sa])
                                                 + MET(s, b, 2, jb) * (q[n - sb] + q[n - sb + sa])) * 0.5;
                    }
                    is = ind_b[n - sa];
                    if (is >= 1) {
                        if (INDB(3, is) == 1 || INDB(3, is) == 4)
                            bl = INDB(2, is) == nsb ? q[n] : q[n - sb];
                        else
                            bl = bcq[is - 1];
                    } else {
                        bl = MET(s, b, 6, jb) * (MET(s, b, 2, jb - 1) * (q[n] + q[n - sa])
                                                 + MET(s, b, 2, jb) * (q[n - sb] + q[n - sb - sa])) * 0.5;
                    }
                    double win = MET(s, a, 6, ia) * (MET(s, a, 2, ia - 1) * glb[n] * vb[n]
                                                     + MET(s, a, 2, ia) * glb[n - sa] * vb[n - sa]);
                    double fg = MET(s, a, 6, ia) * (MET(s, a, 2, ia) * glb[n] * vb[n] * br
                                                    + MET(s, a, 2, ia - 1) * glb[n - sa] * vb[n - sa] * bl);
                    if (it == 0 && win < 0.0)
                        fl[n] = fl[n] - su * win * q[n] - sc * fg;
                    else if (ib == 0 && win > 0.0)
                        fl[n] = fl[n] - su * win * q[n - sb] - sc * fg;
                    else
                        fl[n] = fl[n] - fg;
                    double gg = MET(s, a, 6, ia) * (MET(s, a, 2, ia - 1) * ggb[n] + MET(s, a, 2, ia) * ggb[n - sa]);
                    fg = MET(s, a, 6, ia) * (MET(s, a, 2, ia) * br + MET(s, a, 2, ia - 1) * bl);
                    double anr, anl;
                    if (nf[n] == -1)
                        anr = s->anu[n - sb];
                    else if (nf[n - sb] == -1)
                        anr = s->anu[n];
                    else
                        anr = MET(s, b, 6, jb) * (MET(s, b, 2, jb - 1) * s->anu[n] + MET(s, b, 2, jb) * s->anu[n - sb]);
                    if (nf[n - sa] == -1)
                        anl = s->anu[n - sa - sb];
                    else if (nf[n - sa - sb] == -1)
                        anl = s->anu[n - sa];
                    else
                        anl = MET(s, b, 6, jb) * (MET(s, b, 2, jb - 1) * s->anu[n - sa] + MET(s, b, 2, jb) * s->anu[n - sa - sb]);
                    double an = MET(s, a, 6, ia) * (MET(s, a, 2, ia - 1) * anr + MET(s, a, 2, ia) * anl);
                    double d1;
                    if (it == 0)
                        d1 = MET(s, b, 4, jb) * (q[n] - fg) * 2.0;
                    else
                        d1 = MET(s, b, 4, jb - 1) * (fg - q[n - sb]) * 2.0;
                    double d2 = MET(s, a, 5, ia) * (vb[n] - vb[n - sa]);
                    fl[n] = fl[n] + gg * an * (d1 + d2);
                }
            }
}

/* Source terms of component a: pressure gradient, porous drag, gravity.
 * vf_vgene.f:84-163 (u), :166-241 (v), :244-330 (w).  The drag is the
 * CD form (IDRGN<=0, :102-110) or the Dupuit-Forchheimer form of OPTION
 * DRAG-DF (:111-146): alpha and beta from the porosity GV and the face
 * diameter CD through the DRGDR table, applied to the porosity-weighted
 * velocities. */
static void vgene(orc_state *s, int a, double *qsrc)
{
    DECL_DIMS(s);
    const size_t st[3] = {1, (size_t)ni, nij};
    const size_t sa = st[a];
    const int *ind = a == 0 ? s->indx : (a == 1 ? s->indy : s->indz);
    const double *gga = a == 0 ? s->ggx : (a == 1 ? s->ggy : s->ggz);
    const double *vel[3] = {s->uu, s->vv, s->ww};
    const double ri = 1.0 / s->rho0;
    int i0 = 2, j0 = 2, k0 = 2;
    if (a == 0) i0 = 3;
    if (a == 1) j0 = 3;
    if (a == 2) k0 = 3;
    for (int k = k0; k <= nk - 1; k++)
        for (int j = j0; j <= nj - 1; j++)
            for (int i = i0; i <= ni - 1; i++) {
                size_t n = ID(i, j, k);
                if (!(ind[n] == 0 && s->nf[n - sa] * s->nf[n] == 0)) continue;
                int ia = a == 0 ? i : (a == 1 ? j : k);
                double gv = MET(s, a, 6, ia) * (MET(s, a, 2, ia) * s->ggv[n] + MET(s, a, 2, ia - 1) * s->ggv[n - sa]);
                double cd = MET(s, a, 6, ia) * (MET(s, a, 2, ia) * s->cd0[n] + MET(s, a, 2, ia - 1) * s->cd0[n - sa]);
                qsrc[n] = qsrc[n] + gv * ri * MET(s, a, 5, ia) * (s->pp[n - sa] - s->pp[n]);
                double c[3];
                if (s->idrgn <= 0) {
                    for (int d = 0; d < 3; d++) {
                        if (d == a) {
                            c[d] = vel[a][n];
                        } else {
                            const double *vd = vel[d];
                            size_t sd = st[d];
                            c[d] = MET(s, a, 6, ia) * (MET(s, a, 2, ia - 1) * (vd[n] + vd[n + sd])
                                                       + MET(s, a, 2, ia) * (vd[n - sa] + vd[n - sa + sd])) * 0.5;
                        }
                    }
                    double uvw = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
                    qsrc[n] = qsrc[n] - 0.5 * cd * MET(s, a, 5, ia) * (1.0 - gga[n]) * c[a] * uvw;
                } else {
                    /* vf_vgene.f:112-130: below PLOWER or the first diameter
                     * the drag is off; otherwise the row below the first
                     * diameter DRGDR(L) >= CD, the last row past the end */
                    double rn, dr, ap, bt;
                    if (gv <= s->plower || cd <= s->drgdr[0]) {
                        rn = s->plower;
                        dr = s->drgdr[0];
                        ap = 0.0;
                        bt = 0.0;
                    } else {
                        rn = gv;
                        dr = cd;
                        ap = s->drgap[s->idrgn - 1];
                        bt = s->drgbt[s->idrgn - 1];
                        for (int l = 1; l < s->idrgn; l++)
                            if (dr <= s->drgdr[l]) {
                                ap = s->drgap[l - 1];
                                bt = s->drgbt[l - 1];
                                break;
                            }
                    }
                    double rn1 = 1.0 - rn;
                    double alpha = ap * rn1 * rn1 * rn1 * s->drgyu / (rn * rn * dr * dr);
                    double beta = bt * rn1 / (rn * rn * rn * dr);
                    const double *ggd[3] = {s->ggx, s->ggy, s->ggz};
                    for (int d = 0; d < 3; d++) {
                        if (d == a) {
                            c[d] = vel[a][n] * gga[n];
                        } else {
                            const double *vd = vel[d], *gd = ggd[d];
                            size_t sd = st[d];
                            c[d] = MET(s, a, 6, ia) * (MET(s, a, 2, ia - 1) * (vd[n] * gd[n] + vd[n + sd] * gd[n + sd])
                                                       + MET(s, a, 2, ia) * (vd[n - sa] * gd[n - sa]
                                                                             + vd[n - sa + sd] * gd[n - sa + sd])) * 0.5;
                        }
                    }
                    double uvw = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
                    qsrc[n] = qsrc[n] - c[a] * (alpha + beta * uvw) * gv;
                }
                if (a == 2) {
                    qsrc[n] = qsrc[n] - gv * s->grz0;
                    double dr = 0.0;
                    qsrc[n] = qsrc[n] - gv * s->grz0 * dr / s->rho0;
                }
            }
}

/* x**n for a non-negative integer n by squaring, the lowest bit first
 * (Fortran's REAL**INTEGER as the reference's compiler evaluates it). */
static double ipow(double x, int n)
{
    double r = 1.0;
    while (n > 0) {
        if (n & 1) r *= x;
        n >>= 1;
        if (n) x *= x;
    }
    return r;
}

/* Explicit update of component a and the VVMAX clip.
 * vf_veuler.f:116-158 (u), :161-203 (v), :206-249 (w); GV0 = GV for static
 * porosity (vf_v1cal.f:218-227 passes GLV twice). */
static void veuler(orc_state *s, int a, double *const fl[3], const double *qsrc)
{
    DECL_DIMS(s);
    const size_t st[3] = {1, (size_t)ni, nij};
    const size_t sa = st[a];
    const int *ind = a == 0 ? s->indx : (a == 1 ? s->indy : s->indz);
    double *q = a == 0 ? s->uu : (a == 1 ? s->vv : s->ww);
    int i0 = 2, j0 = 2, k0 = 2;
    if (a == 0) i0 = 3;
    if (a == 1) j0 = 3;
    if (a == 2) k0 = 3;
    /* damping parameters, vf_veuler.f:75-113: P = (n+1) PARAM sqrt(g/DEPTH),
     * PARAM-XY for u and v, PARAM-Z for w */
    int ipa = 0;
    double pw[4], pp[4];
    for (int jd = 0; jd < 4; jd++) {
        if (!s->dm_on[jd]) continue;
        ipa++;
        double d = sqrt(s->grz0 / s->dm_par[jd][3]);
        pw[jd] = s->dm_par[jd][2];
        pp[jd] = (double)(s->dm_deg[jd] + 1) * s->dm_par[jd][a == 2 ? 1 : 0] * d;
    }
    for (int k = k0; k <= nk - 1; k++)
        for (int j = j0; j <= nj - 1; j++)
            for (int i = i0; i <= ni - 1; i++) {
                size_t n = ID(i, j, k);
                if (!(ind[n] == 0 && s->nf[n - sa] * s->nf[n] == 0)) continue;
                int idx[3] = {i, j, k};
                int ia = idx[a];
                double gv = MET(s, a, 6, ia) * (MET(s, a, 2, ia) * s->glv[n] + MET(s, a, 2, ia - 1) * s->glv[n - sa]);
                double gv0 = MET(s, a, 6, ia) * (MET(s, a, 2, ia) * s->glv[n] + MET(s, a, 2, ia - 1) * s->glv[n - sa]);
                double t[3];
                for (int d = 0; d < 3; d++) {
                    const double *f = fl[d];
                    if (d == a)
                        t[d] = MET(s, d, 5, idx[d]) * (f[n] - f[n - st[d]]);
                    else
                        t[d] = MET(s, d, 4, idx[d]) * (f[n + st[d]] - f[n]);
                }
                double du = (t[0] + t[1] + t[2] + qsrc[n]);
                q[n] = (gv * q[n] + s->dtnow * du) / gv0;
                if (q[n] > s->vvmax) q[n] = s->vvmax;
                if (q[n] < -s->vvmax) q[n] = -s->vvmax;
                /* implicit damping, vf_veuler.f:131-166 (u), :187-221 (v),
                 * :243-277 (w): the largest P(1 - dist/WIDTH)**n over the
                 * zones within WIDTH of the face's (x, y) */
                if (ipa != 0) {
                    double x = a == 0 ? MET(s, 0, 1, i) : (MET(s, 0, 1, i) + MET(s, 0, 1, i + 1)) * 0.5;
                    double y = a == 1 ? MET(s, 1, 1, j) : (MET(s, 1, 1, j) + MET(s, 1, 1, j + 1)) * 0.5;
                    const double dist[4] = {x - s->glxmin, s->glxmax - x, y - s->glymin, s->glymax - y};
                    double dm = 0.0;
                    for (int jd = 0; jd < 4; jd++) {
                        if (!s->dm_on[jd]) continue;
                        double d = dist[jd];
                        if (d <= pw[jd]) {
                            d = pp[jd] * ipow(1.0 - d / pw[jd], s->dm_deg[jd]);
                            if (d > dm) dm = d;
                        }
                    }
                    double g = MET(s, a, 6, ia) * (MET(s, a, 2, ia) * s->ggv[n] + MET(s, a, 2, ia - 1) * s->ggv[n - sa]);
                    double d = gv0 / (gv0 + s->dtnow * g * dm);
                    q[n] = q[n] * d;
                }
            }
}

/* Normal velocity on boundary faces.  vf_bwuwn.f:84-254 (types 1-4, 6-8). */
/* the reference's PI (Source code/VF_A0PRM.h:23) */
static const double PI_REF = 3.141592653589794;

/* VF_CWMTB1 (vf_cwmtb1.f): time weights of the table at phase wvt, searching
 * down from MTBNOW first, then down from the top; the level there. */
static int mtb_time(orc_state *s, double wvt, double *w1o, double *w2o, double *wvz)
{
    const int nt = s->mtbtt;
    const double *tt = s->mtb_tt - 1;   /* 1-based */
    double w1 = 0.0, w2 = 0.0;
    if (s->mtbnow <= 0) s->mtbnow = nt - 1;
    if (wvt <= tt[1]) {
        s->mtbnow = 1;
        w1 = 1.0;
        w2 = 0.0;
    } else if (wvt >= tt[nt]) {
        s->mtbnow = nt - 1;
        w1 = 0.0;
        w2 = 1.0;
    } else {
        int ip = 0;
        for (int i = s->mtbnow; i >= 1 && !ip; i--)
            if (tt[i] <= wvt && wvt <= tt[i + 1]) {
                ip = i;
                w2 = 1.0 / (tt[i + 1] - tt[i]);
                w1 = (tt[i + 1] - wvt) * w2;
                w2 = (wvt - tt[i]) * w2;
            }
        for (int i = nt - 1; i >= s->mtbnow + 1 && !ip; i--)
            if (tt[i] <= wvt && wvt <= tt[i + 1]) {
                ip = i;
                w2 = 1.0 / (tt[i + 1] - tt[i]);
                w1 = (tt[i + 1] - wvt) * w2;
                w2 = (wvt - tt[i]) * w2;
            }
        if (!ip) return -13;
        s->mtbnow = ip;
    }
    *wvz = w1 * s->mtb_hh[s->mtbnow - 1] + w2 * s->mtb_hh[s->mtbnow];
    *w1o = w1;
    *w2o = w2;
    return 0;
}

/* VF_CWMTB2 (vf_cwmtb2.f): normal and tangential velocity at height zc,
 * searching up from km, between the two phases of the time weights. */
static int mtb_depth(const orc_state *s, int *km, double wt1, double wt2, double zc, double *un,
                     double *ut)
{
    const int nz = s->mtbzz;
    const double *zz = s->mtb_zz - 1;
    double w1 = 0.0, w2 = 0.0;
    if (zc <= zz[1]) {
        *km = 1;
        w1 = 1.0;
        w2 = 0.0;
    } else if (zc >= zz[nz]) {
        *km = nz - 1;
        w1 = 0.0;
        w2 = 1.0;
    } else {
        int kp = 0;
        for (int k = *km; k <= nz - 1 && !kp; k++)
            if (zz[k] <= zc && zc <= zz[k + 1]) {
                kp = k;
                w2 = 1.0 / (zz[k + 1] - zz[k]);
                w1 = (zz[k + 1] - zc) * w2;
                w2 = (zc - zz[k]) * w2;
            }
        if (!kp) return -14;
        *km = kp;
    }
    const int km1 = *km + 1, im = s->mtbnow, im1 = s->mtbnow + 1;
#define TAB(a, k, i) (a)[(size_t)((k) - 1) + (size_t)nz * (size_t)((i) - 1)]
    double p1 = w1 * TAB(s->mtb_un, *km, im) + w2 * TAB(s->mtb_un, km1, im);
    double p2 = w1 * TAB(s->mtb_un, *km, im1) + w2 * TAB(s->mtb_un, km1, im1);
    *un = wt1 * p1 + wt2 * p2;
    p1 = w1 * TAB(s->mtb_ut, *km, im) + w2 * TAB(s->mtb_ut, km1, im);
    p2 = w1 * TAB(s->mtb_ut, *km, im1) + w2 * TAB(s->mtb_ut, km1, im1);
    *ut = wt1 * p1 + wt2 * p2;
#undef TAB
    return 0;
}

/* The table wave-maker sides of VF_BWUWN (vf_bwuwn.f:257-480, FUNC MATRIX
 * with a LEVEL-ON table, normal incidence): for every boundary column the
 * table profile, scaled by (level + depth) / (measured level + depth), gives
 * the normal and tangential velocity of each face. */
static int wave_bc_side(orc_state *s, int jd)
{
    DECL_DIMS(s);
    {
        const double d = s->wv_d[jd - 1], t = s->wv_t[jd - 1], t0 = s->wv_t0[jd - 1];
        double a = s->wv_a[jd - 1];
        /* the table phase in single precision, as the reference keeps it */
        float r4tn = (float)(t0 - s->tnow / t);
        r4tn = r4tn - (float)(int)r4tn;
        if (r4tn < 0.0f) r4tn = r4tn + 1.0f;
        double aw = 1.0;
        if (a >= ZERO) {
            a = s->tnow / t / a;
            if (a < 1.0) aw = 0.5 * sin(PI_REF * (a - 0.5)) + 0.5;
        }
        const double wvt = (double)r4tn;
        double wmt1, wmt2, wvz;
        int rc = mtb_time(s, wvt, &wmt1, &wmt2, &wvz);
        if (rc) return rc;
        s->wv_level[jd - 1] = wvz * aw;
        /* a LEVEL-ONLY table sets no velocity here (vf_bwuwn.f:326-327) */
        if (s->mtb_only) return 0;
        const double sa = jd == 1 ? 1.0 : (jd == 2 ? -1.0 : 0.0);
        const double sb = jd == 3 ? 1.0 : (jd == 4 ? -1.0 : 0.0);
        const int xside = jd <= 2;
        const int pos = xside ? (jd == 1 ? 2 : ni) : (jd == 3 ? 2 : nj);
        const int cpos = xside ? (jd == 1 ? 2 : ni - 1) : (jd == 3 ? 2 : nj - 1);
        const int lim = xside ? nj : ni;
        for (int q = s->wv_l1[jd - 1]; q <= s->wv_l2[jd - 1]; q++) {
            if (q < 2 || q > lim - 1) continue;
            double val = 0.0;
            for (int k = 2; k <= nk - 1; k++) {
                const size_t c = xside ? ID(cpos, q, k) : ID(q, cpos, k);
                const int nw = s->nf[c];
                if (nw == -1) {
                    val = val + ZZ(2, k) * 1.0;
                } else if (nw == 0) {
                    val = val + ZZ(2, k) * s->ff[c];
                } else if (nw != 8) {
                    val = val + ZZ(2, k) * s->ff[c];
                    break;
                }
            }
            val = val - (s->wvlvl - ZZ(1, 2));
            /* a LEVEL-OFF table (MTBTYP=2) takes its profiles unscaled
             * (vf_bwuwn.f:326-329) */
            const double vws = s->mtb_off ? 1.0 : (wvz + d) / (val + d);
            int kmt = 1;
            for (int k = 2; k <= nk - 1; k++) {
                const size_t f = xside ? ID(pos, q, k) : ID(q, pos, k);
                const int l = xside ? s->indx[f] : s->indy[f];
                if (l < 1) continue;
                double zc = (ZZ(1, k) + ZZ(1, k + 1)) * 0.5 - s->wvlvl;
                zc = vws * (zc + d) - d;
                double un, ut;
                rc = mtb_depth(s, &kmt, wmt1, wmt2, zc, &un, &ut);
                if (rc) return rc;
                s->bcu[l - 1] = un * vws * aw * sa;
                s->bcv[l - 1] = un * vws * aw * sb;
                s->bcw[l - 1] = ut * vws * aw;
                if (xside)
                    s->uu[f] = s->bcu[l - 1];
                else
                    s->vv[f] = s->bcv[l - 1];
            }
        }
    }
    return 0;
}

/* a side of the second table (FUNC MATRIX2) runs with the two tables'
 * fields exchanged, so one restatement serves both */
static void swap_tables(orc_state *s)
{
#define SWP(T, a, b) do { T t_ = s->a; s->a = s->b; s->b = t_; } while (0)
    SWP(int, mtbzz, mtbzz2);
    SWP(int, mtbtt, mtbtt2);
    SWP(int, mtbnow, mtbnow2);
    SWP(const double *, mtb_zz, mtb_zz2);
    SWP(const double *, mtb_tt, mtb_tt2);
    SWP(const double *, mtb_hh, mtb_hh2);
    SWP(const double *, mtb_un, mtb_un2);
    SWP(const double *, mtb_ut, mtb_ut2);
    SWP(int, mtb_off, mtb_off2);
    SWP(int, mtb_only, mtb_only2);
#undef SWP
}

static int wave_bc(orc_state *s)
{
    for (int jd = 1; jd <= 4; jd++) {
        if (!s->wv_on[jd - 1]) continue;
        const int t2 = s->wv_tab[jd - 1] == 1;
        if (t2) swap_tables(s);
        const int rc = wave_bc_side(s, jd);
        if (t2) swap_tables(s);
        if (rc) return rc;
    }
    return 0;
}

/* VF_COMAK0 (vf_comak0.f): the wave length of a small-amplitude wave of
 * period t in depth d, by Newton's method on the dispersion relation
 * (at most 10 iterations, stop at |dX/X| <= 1e-10). */
static double comak0(double d, double t, double gg)
{
    const double dd = 4.0 * PI_REF * PI_REF * d / gg / t / t;
    double x = dd >= 1.0 ? dd : sqrt(dd);
    for (int itr = 1; itr <= 10; itr++) {
        const double ex = exp(x), exi = 1.0 / ex;
        const double coth = (ex + exi) / (ex - exi);
        const double dx = (x - dd * coth) / (1.0 + dd * (coth * coth - 1.0));
        x = x - dx;
        if (fabs(dx / x) <= 1.0e-10) break;
    }
    return 2.0 * PI_REF * d / x;
}

/* The radiation sides of VF_BWUWN (vf_bwuwn.f:483-620): on each face of the
 * side next to a fluid cell the boundary values relax towards the inside
 * velocities of the start of the step (UB, VB, WB) with the weights
 * ALN = min(C*DTNOW/dx, 1) (normal) and ALT = min(2*ALN, 1) (tangential),
 * C = wave length / period; both are 1 in the initial pass (NNOW = 0).
 * Other faces get 0; then faces next to a surface cell copy the face below
 * (NF = 5) or above (NF = 6) it. */
static void open_bc(orc_state *s, const double *ub, const double *vb, const double *wb, int first)
{
    DECL_DIMS(s);
    for (int jd = 1; jd <= 4; jd++) {
        if (!s->op_on[jd - 1]) continue;
        const double c = comak0(s->op_d[jd - 1], s->op_t[jd - 1], s->grz0) / s->op_t[jd - 1];
        const int xside = jd <= 2;
        const int pos = xside ? (jd == 1 ? 2 : ni) : (jd == 3 ? 2 : nj);
        const int ic = xside ? (jd == 1 ? 2 : ni - 1) : (jd == 3 ? 2 : nj - 1);
        const int i2 = jd == 1 || jd == 3 ? 3 : ic;
        double aln = 1.0, alt = 1.0;
        if (!first) {
            aln = c * s->dtnow / (xside ? XX(2, ic) : YY(2, ic));
            if (aln > 1.0) aln = 1.0;
            alt = aln * 2.0;
            if (alt > 1.0) alt = 1.0;
        }
        const int lim = xside ? nj : ni;
        const int *ind = xside ? s->indx : s->indy;
        double *vel = xside ? s->uu : s->vv;
        /* faces next to fluid cells, and the rest (vf_bwuwn.f:505-527) */
        for (int k = 2; k <= nk - 1; k++)
            for (int q = 2; q <= lim - 1; q++) {
                const size_t f = xside ? ID(pos, q, k) : ID(q, pos, k);
                const int l = ind[f];
                if (l < 1) continue;
                const size_t cc = xside ? ID(ic, q, k) : ID(q, ic, k);
                if (s->nf[cc] == 0) {
                    if (xside) {
                        const double vc = (vb[cc] + vb[cc + ni]) * 0.5;
                        const double wc = (wb[cc] + wb[cc + nij]) * 0.5;
                        s->bcu[l - 1] = (1.0 - aln) * s->bcu[l - 1] + aln * ub[ID(i2, q, k)];
                        s->bcv[l - 1] = (1.0 - alt) * s->bcv[l - 1] + alt * vc;
                        s->bcw[l - 1] = (1.0 - alt) * s->bcw[l - 1] + alt * wc;
                        vel[f] = s->bcu[l - 1];
                    } else {
                        const double uc = (ub[cc] + ub[cc + 1]) * 0.5;
                        const double wc = (wb[cc] + wb[cc + nij]) * 0.5;
                        s->bcu[l - 1] = (1.0 - alt) * s->bcu[l - 1] + alt * uc;
                        s->bcv[l - 1] = (1.0 - aln) * s->bcv[l - 1] + aln * vb[ID(q, i2, k)];
                        s->bcw[l - 1] = (1.0 - alt) * s->bcw[l - 1] + alt * wc;
                        vel[f] = s->bcv[l - 1];
                    }
                } else {
                    s->bcu[l - 1] = 0.0;
                    s->bcv[l - 1] = 0.0;
                    s->bcw[l - 1] = 0.0;
                    vel[f] = 0.0;
                }
            }
        /* faces next to surface cells (vf_bwuwn.f:528-552); on an x side a
         * face under a surface cell with water above gets U = 0, on a y side
         * V = BCV (vf_bwuwn.f:546, :614) */
        for (int k = 2; k <= nk - 1; k++)
            for (int q = 2; q <= lim - 1; q++) {
                const size_t f = xside ? ID(pos, q, k) : ID(q, pos, k);
                const int l = ind[f];
                if (l < 1) continue;
                const int nw = s->nf[xside ? ID(ic, q, k) : ID(q, ic, k)];
                if (nw != 5 && nw != 6) continue;
                const int l2 = ind[nw == 5 ? f - nij : f + nij];
                if (l2 < 1) continue;   /* no face there: out of the reference's reach */
                s->bcu[l - 1] = s->bcu[l2 - 1];
                s->bcv[l - 1] = s->bcv[l2 - 1];
                s->bcw[l - 1] = s->bcw[l2 - 1];
                if (nw == 5)
                    vel[f] = xside ? s->bcu[l - 1] : s->bcv[l - 1];
                else
                    vel[f] = xside ? 0.0 : s->bcv[l - 1];
            }
    }
}

static int bwuwn(orc_state *s, const double *ub, const double *vb, const double *wb, int first)
{
    DECL_DIMS(s);
    for (int l = 1; l <= s->numb; l++) {
        int i, j, k;
        decode(s, INDB(1, l), &i, &j, &k);
        int ns = INDB(2, l), ib = ib_lv(s, l);
        size_t n = ID(i, j, k);
        if (ib == 1 || ib == 6 || ib == 8) {
            if (ns == 1 || ns == 2) { s->bcu[l - 1] = 0.0; s->uu[n] = 0.0; }
            else if (ns == 3 || ns == 4) { s->bcv[l - 1] = 0.0; s->vv[n] = 0.0; }
            else { s->bcw[l - 1] = 0.0; s->ww[n] = 0.0; }
        } else if (ib == 2) {
            s->bcu[l - 1] = 0.0; s->bcv[l - 1] = 0.0; s->bcw[l - 1] = 0.0;
            if (ns == 1 || ns == 2) s->uu[n] = s->bcu[l - 1];
            else if (ns == 3 || ns == 4) s->vv[n] = s->bcv[l - 1];
            else s->ww[n] = s->bcw[l - 1];
        } else if (ib == 3) {
            if (ns == 1 || ns == 2) s->uu[n] = s->bcu[l - 1];
            else if (ns == 3 || ns == 4) s->vv[n] = s->bcv[l - 1];
            else s->ww[n] = s->bcw[l - 1];
        } else if (ib == 4) {
            double val;
            switch (ns) {
            case 1: val = s->uu[n + 1]; s->bcu[l - 1] = val; s->uu[n] = val; break;
            case 2: val = s->uu[n - 1]; s->bcu[l - 1] = val; s->uu[n] = val; break;
            case 3: val = s->vv[n + ni]; s->bcv[l - 1] = val; s->vv[n] = val; break;
            case 4: val = s->vv[n - ni]; s->bcv[l - 1] = val; s->vv[n] = val; break;
            case 5: val = s->ww[n + nij]; s->bcw[l - 1] = val; s->ww[n] = val; break;
            case 6: val = s->ww[n - nij]; s->bcw[l - 1] = val; s->ww[n] = val; break;
            }
        } else if (ib == 5) {
            /* a LEVEL-ON table sets these faces below (vf_bwuwn.f:171-223) */
        } else if (ib == 7) {
            /* a radiation side, set below (vf_bwuwn.f:483-620) */
        } else {
            return -10;
        }
    }
    int rc = wave_bc(s);
    if (rc) return rc;
    open_bc(s, ub, vb, wb, first);
    return 0;
}

int orc_predictor(orc_state *s)
{
    DECL_DIMS(s);
    size_t nc = nij * (size_t)nk;
    double *w = (double *)calloc(15 * nc, sizeof(double));
    if (!w) return -1;
    /* WK01..WK12 of vf_v1cal.f:126-146, and WK13..WK15, the velocities of
     * the start of the step that the radiation sides read */
    double *ub = w + 12 * nc, *vb = w + 13 * nc, *wb = w + 14 * nc;
    memcpy(ub, s->uu, nc * sizeof(double));
    memcpy(vb, s->vv, nc * sizeof(double));
    memcpy(wb, s->ww, nc * sizeof(double));
    double *flu[3] = {w + 0 * nc, w + 1 * nc, w + 2 * nc};
    double *qu = w + 3 * nc;
    double *flv[3] = {w + 4 * nc, w + 5 * nc, w + 6 * nc};
    double *qv = w + 7 * nc;
    double *flw[3] = {w + 8 * nc, w + 9 * nc, w + 10 * nc};
    double *qw = w + 11 * nc;

    /* VF_VFLXDU (vf_v1cal.f:155) */
    flux_normal(s, 0, s->uu, s->glx, s->ggx, flu[0]);
    flux_cross(s, 0, 1, s->uu, s->vv, s->gly, s->ggy, s->indx, s->indy, s->bcu, flu[1],
               3, ni - 1, 2, nj, 2, nk);
    flux_cross(s, 0, 2, s->uu, s->ww, s->glz, s->ggz, s->indx, s->indz, s->bcu, flu[2],
               3, ni - 1, 2, nj - 1, 2, nk);
    /* VF_VFLXDV (vf_v1cal.f:169) */
    flux_normal(s, 1, s->vv, s->gly, s->ggy, flv[1]);
    flux_cross(s, 1, 0, s->vv, s->uu, s->glx, s->ggx, s->indy, s->indx, s->bcv, flv[0],
               2, ni, 3, nj - 1, 2, nk);
    flux_cross(s, 1, 2, s->vv, s->ww, s->glz, s->ggz, s->indy, s->indz, s->bcv, flv[2],
               2, ni - 1, 3, nj - 1, 2, nk);
    /* VF_VFLXDW (vf_v1cal.f:183) */
    flux_normal(s, 2, s->ww, s->glz, s->ggz, flw[2]);
    flux_cross(s, 2, 0, s->ww, s->uu, s->glx, s->ggx, s->indz, s->indx, s->bcw, flw[0],
               2, ni, 2, nj - 1, 3, nk - 1);
    flux_cross(s, 2, 1, s->ww, s->vv, s->gly, s->ggy, s->indz, s->indy, s->bcw, flw[1],
               2, ni - 1, 2, nj, 3, nk - 1);
    /* VF_VGENE (vf_v1cal.f:199) */
    vgene(s, 0, qu);
    vgene(s, 1, qv);
    vgene(s, 2, qw);
    /* VF_VEULER (vf_v1cal.f:218) */
    veuler(s, 0, flu, qu);
    veuler(s, 1, flv, qv);
    veuler(s, 2, flw, qw);
    /* VF_BWUWN (vf_v1cal.f:251) */
    int rc = bwuwn(s, ub, vb, wb, 0);
    free(w);
    return rc;
}

/* ======================================================================
 * Phase (b): pressure Poisson
 * ====================================================================== */

/* Matrix and right-hand side.  vf_vpcoef.f:76-437 (wave-source and
 * air-pressure branches excluded). */
int orc_vpcoef(orc_state *s, orc_matrix *m)
{
    DECL_DIMS(s);
    size_t nc = nij * (size_t)nk;
    memset(m->ad, 0, nc * sizeof(double));
    memset(m->ali, 0, nc * sizeof(double));
    memset(m->alj, 0, nc * sizeof(double));
    memset(m->alk, 0, nc * sizeof(double));
    memset(m->aui, 0, nc * sizeof(double));
    memset(m->auj, 0, nc * sizeof(double));
    memset(m->auk, 0, nc * sizeof(double));
    memset(m->bb, 0, nc * sizeof(double));
    const double dti = 1.0 / s->dtnow;
    const double tr = s->dtnow / s->rho0;
    const double pps = 0.0;
    for (int k = 1; k <= nk; k++)
        for (int j = 1; j <= nj; j++)
            for (int i = 1; i <= ni; i++) {
                size_t n = ID(i, j, k);
                int nn = s->nf[n];
                if (s->indc[n] == -1) {
                    m->ad[n] = 1.0;
                    m->ali[n] = 0.0; m->alj[n] = 0.0; m->alk[n] = 0.0;
                    m->aui[n] = 0.0; m->auj[n] = 0.0; m->auk[n] = 0.0;
                    m->bb[n] = 0.0;
                } else if (nn == 0) {
                    double axy = XX(2, i) * YY(2, j);
                    double axz = XX(2, i) * ZZ(2, k);
                    double ayz = YY(2, j) * ZZ(2, k);
                    double avl = XX(2, i) * YY(2, j) * ZZ(2, k);
                    double aim, aip, ajm, ajp, akm, akp, gv;
                    const double *ggv = s->ggv, *glv = s->glv;
                    int is;
                    /* ALI */
                    is = s->indx[n];
                    if (is == 0) {
                        gv = XX(6, i) * (XX(2, i) * ggv[n] / glv[n] + XX(2, i - 1) * ggv[n - 1] / glv[n - 1]);
                        aim = ayz * XX(5, i) * s->ggx[n] * gv;
                        m->ali[n] = -aim;
                    } else if (is >= 1) {
                        if (ib_lv(s, is) == 4) aim = 2.0 * ayz * XX(4, i) * s->ggx[n] * ggv[n] / glv[n];
                        else aim = 0.0;
                        m->ali[n] = 0.0;
                    } else return -20;
                    /* AUI */
                    is = s->indx[n + 1];
                    if (is == 0) {
                        gv = XX(6, i + 1) * (XX(2, i + 1) * ggv[n + 1] / glv[n + 1] + XX(2, i) * ggv[n] / glv[n]);
                        aip = ayz * XX(5, i + 1) * s->ggx[n + 1] * gv;
                        m->aui[n] = -aip;
                    } else if (is >= 1) {
                        if (ib_lv(s, is) == 4) aip = 2.0 * ayz * XX(4, i) * s->ggx[n + 1] * ggv[n] / glv[n];
                        else aip = 0.0;
                        m->aui[n] = 0.0;
                    } else return -20;
                    /* ALJ */
                    is = s->indy[n];
                    if (is == 0) {
                        gv = YY(6, j) * (YY(2, j) * ggv[n] / glv[n] + YY(2, j - 1) * ggv[n - ni] / glv[n - ni]);
                        ajm = axz * YY(5, j) * s->ggy[n] * gv;
                        m->alj[n] = -ajm;
                    } else if (is >= 1) {
                        if (ib_lv(s, is) == 4) ajm = 2.0 * axz * YY(4, j) * s->ggy[n] * ggv[n] / glv[n];
                        else ajm = 0.0;
                        m->alj[n] = 0.0;
                    } else return -20;
                    /* AUJ */
                    is = s->indy[n + ni];
                    if (is == 0) {
                        gv = YY(6, j + 1) * (YY(2, j + 1) * ggv[n + ni] / glv[n + ni] + YY(2, j) * ggv[n] / glv[n]);
                        ajp = axz * YY(5, j + 1) * s->ggy[n + ni] * gv;
                        m->auj[n] = -ajp;
                    } else if (is >= 1) {
                        if (ib_lv(s, is) == 4) ajp = 2.0 * axz * YY(4, j) * s->ggy[n + ni] * ggv[n] / glv[n];
                        else ajp = 0.0;
                        m->auj[n] = 0.0;
                    } else return -20;
                    /* ALK */
                    is = s->indz[n];
                    if (is == 0) {
                        gv = ZZ(6, k) * (ZZ(2, k) * ggv[n] / glv[n] + ZZ(2, k - 1) * ggv[n - nij] / glv[n - nij]);
                        akm = axy * ZZ(5, k) * s->ggz[n] * gv;
                        m->alk[n] = -akm;
                    } else if (is >= 1) {
                        if (INDB(3, is) != 4) akm = 0.0;
                        else akm = 2.0 * axy * ZZ(4, k) * s->ggz[n] * ggv[n] / glv[n];
                        m->alk[n] = 0.0;
                    } else return -20;
                    /* AUK */
                    is = s->indz[n + nij];
                    if (is == 0) {
                        gv = ZZ(6, k + 1) * (ZZ(2, k + 1) * ggv[n + nij] / glv[n + nij] + ZZ(2, k) * ggv[n] / glv[n]);
                        akp = axy * ZZ(5, k + 1) * s->ggz[n + nij] * gv;
                        m->auk[n] = -akp;
                    } else if (is >= 1) {
                        if (INDB(3, is) != 4) akp = 0.0;
                        else akp = 2.0 * axy * ZZ(4, k) * s->ggz[n + nij] * ggv[n] / glv[n];
                        m->auk[n] = 0.0;
                    } else return -20;
                    m->ad[n] = akm + ajm + aim + aip + ajp + akp;
                    double vd = ayz * (s->ggx[n + 1] * s->uu[n + 1] - s->ggx[n] * s->uu[n])
                              + axz * (s->ggy[n + ni] * s->vv[n + ni] - s->ggy[n] * s->vv[n])
                              + axy * (s->ggz[n + nij] * s->ww[n + nij] - s->ggz[n] * s->ww[n]);
                    double vq = -(ggv[n] - ggv[n]) * dti * avl;
                    m->bb[n] = vd - vq;
                } else if (nn >= 1 && nn <= 6) {
                    /* surface cell: Dirichlet row toward the fluid neighbour */
                    double psi, pf;
                    m->ad[n] = 1.0;
                    m->ali[n] = 0.0; m->alj[n] = 0.0; m->alk[n] = 0.0;
                    m->aui[n] = 0.0; m->auj[n] = 0.0; m->auk[n] = 0.0;
                    switch (nn) {
                    case 5: psi = ZZ(3, k) / (0.5 * ZZ(2, k - 1) + s->ff[n] * ZZ(2, k)); m->alk[n] = psi - 1.0; pf = s->pp[n - nij]; break;
                    case 6: psi = ZZ(3, k + 1) / (0.5 * ZZ(2, k + 1) + s->ff[n] * ZZ(2, k)); m->auk[n] = psi - 1.0; pf = s->pp[n + nij]; break;
                    case 3: psi = YY(3, j) / (0.5 * YY(2, j - 1) + s->ff[n] * YY(2, j)); m->alj[n] = psi - 1.0; pf = s->pp[n - ni]; break;
                    case 4: psi = YY(3, j + 1) / (0.5 * YY(2, j + 1) + s->ff[n] * YY(2, j)); m->auj[n] = psi - 1.0; pf = s->pp[n + ni]; break;
                    case 1: psi = XX(3, i) / (0.5 * XX(2, i - 1) + s->ff[n] * XX(2, i)); m->ali[n] = psi - 1.0; pf = s->pp[n - 1]; break;
                    default: psi = XX(3, i + 1) / (0.5 * XX(2, i + 1) + s->ff[n] * XX(2, i)); m->aui[n] = psi - 1.0; pf = s->pp[n + 1]; break;
                    }
                    m->bb[n] = -tr * ((1.0 - psi) * pf + psi * pps - s->pp[n]);
                } else {
                    return -21;
                }
            }
    return 0;
}

/* MILU factorisation.  vf_mzdcmp.f:49-67. */
int orc_mzdcmp(int ni, int nj, int nk, const orc_matrix *m, double *dd, double cgpara,
               const int *indc, int icgtyp)
{
    const size_t nij = (size_t)ni * nj;
    double u = 0.0;
    if (icgtyp != 0) u = cgpara;
    int npos = 1;
    for (int k = 2; k <= nk - 1; k++)
        for (int j = 2; j <= nj - 1; j++)
            for (int i = 2; i <= ni - 1; i++) {
                size_t n = ID(i, j, k);
                if (indc[n] == -1) continue;
                double au1 = m->aui[n - 1] + u * (m->auj[n - 1] + m->auk[n - 1]);
                double au2 = m->auj[n - ni] + u * (m->aui[n - ni] + m->auk[n - ni]);
                double au3 = m->auk[n - nij] + u * (m->aui[n - nij] + m->auj[n - nij]);
                dd[n] = 1.0 / (m->ad[n] - m->ali[n] * dd[n - 1] * au1
                                         - m->alj[n] * dd[n - ni] * au2
                                         - m->alk[n] * dd[n - nij] * au3);
                if (dd[n] <= 0.0) npos = 0;
            }
    return npos;
}

/* (L D^-1 + I)(D + U) x = y.  vf_mzminv.f:38-41, vf_mzfrwd.f:37-48,
 * vf_mzbkwd.f:37-47 (called with X aliased to Y). */
void orc_mzminv(int ni, int nj, int nk, const double *dd, const orc_matrix *m, double *x,
                const double *y, const int *indc)
{
    const size_t nij = (size_t)ni * nj;
    for (int k = 2; k <= nk - 1; k++)
        for (int j = 2; j <= nj - 1; j++)
            for (int i = 2; i <= ni - 1; i++) {
                size_t n = ID(i, j, k);
                if (indc[n] == -1) continue;
                x[n] = y[n] - m->ali[n] * dd[n - 1] * x[n - 1]
                            - m->alj[n] * dd[n - ni] * x[n - ni]
                            - m->alk[n] * dd[n - nij] * x[n - nij];
            }
    for (int k = nk - 1; k >= 2; k--)
        for (int j = nj - 1; j >= 2; j--)
            for (int i = ni - 1; i >= 2; i--) {
                size_t n = ID(i, j, k);
                if (indc[n] == -1) continue;
                x[n] = (x[n] - m->aui[n] * x[n + 1]
                             - m->auj[n] * x[n + ni]
                             - m->auk[n] * x[n + nij]) * dd[n];
            }
}

/* y = A x.  vf_mzax.f:42-56. */
void orc_mzax(int ni, int nj, int nk, const orc_matrix *m, const double *x, double *y,
              const int *indc)
{
    const size_t nij = (size_t)ni * nj;
    for (int k = 2; k <= nk - 1; k++)
        for (int j = 2; j <= nj - 1; j++)
            for (int i = 2; i <= ni - 1; i++) {
                size_t n = ID(i, j, k);
                if (indc[n] == -1) continue;
                y[n] = m->ad[n] * x[n] + m->ali[n] * x[n - 1] + m->alj[n] * x[n - ni]
                     + m->alk[n] * x[n - nij] + m->aui[n] * x[n + 1] + m->auj[n] * x[n + ni]
                     + m->auk[n] * x[n + nij];
            }
}

static int dot_order = ORC_DOT_SERIAL;

int orc_set_dot_order(int order)
{
    if (order < ORC_DOT_SERIAL || order > ORC_DOT_BLOCKED) return -1;
    const int prev = dot_order;
    dot_order = order;
    return prev;
}

/* Masked dot product, lexicographic order.  vf_mzip.f:32-43.  The other
 * summation orders (orc_set_dot_order) keep the same terms x*y, each rounded
 * to double as in the reference. */
double orc_mzip(int ni, int nj, int nk, const double *x, const double *y, const int *indc)
{
    const size_t nij = (size_t)ni * nj;
    double sum = 0.0, blk = 0.0;
    long double ext = 0.0L;
    size_t cnt = 0;
    for (int k = 2; k <= nk - 1; k++)
        for (int j = 2; j <= nj - 1; j++)
            for (int i = 2; i <= ni - 1; i++) {
                size_t n = ID(i, j, k);
                if (indc[n] == -1) continue;
                const double t = x[n] * y[n];
                if (dot_order == ORC_DOT_SERIAL) {
                    sum = sum + t;
                } else if (dot_order == ORC_DOT_EXTENDED) {
                    ext += (long double)t;
                } else {
                    blk = blk + t;
                    if (++cnt % 1024 == 0) { sum = sum + blk; blk = 0.0; }
                }
            }
    if (dot_order == ORC_DOT_EXTENDED) return (double)ext;
    if (dot_order == ORC_DOT_BLOCKED) return sum + blk;
    return sum;
}

/* MILU-preconditioned BiCGSTAB on M^-1 A x = M^-1 b.  vf_m1bcgs.f:78-262. */
int orc_m1bcgs(int ni, int nj, int nk, const orc_matrix *m, double *x0, const int *indc,
               double cgpara, double cgepsa, double cgepsr, double cgdiv, int icgtyp, int icgmax,
               double *bnorm_out, double *xnorm_out, int *iter_out)
{
    const size_t nij = (size_t)ni * nj;
    const size_t nc = nij * (size_t)nk;
    double *w = (double *)calloc(8 * nc, sizeof(double));
    if (!w) return -1;
    double *dd = w, *xw = w + nc, *yy = w + 2 * nc, *r0 = w + 3 * nc, *rr = w + 4 * nc;
    double *pp = w + 5 * nc, *ap = w + 6 * nc, *ae = w + 7 * nc;
    memcpy(xw, x0, nc * sizeof(double));
    int iter = -1;
    double bnorm = 0.0, xnorm = 0.0, s = 0.0;
    int isw = orc_mzdcmp(ni, nj, nk, m, dd, cgpara, indc, icgtyp);
    if (isw == 0) goto done_noconv;
    {
        double epsa2 = cgepsa * cgepsa, epsr2 = cgepsr * cgepsr, div2 = cgdiv * cgdiv;
        orc_mzminv(ni, nj, nk, dd, m, yy, m->bb, indc);
        bnorm = orc_mzip(ni, nj, nk, yy, yy, indc);
        iter = 0;
        orc_mzax(ni, nj, nk, m, xw, yy, indc);
        for (int k = 2; k <= nk - 1; k++)
            for (int j = 2; j <= nj - 1; j++)
                for (int i = 2; i <= ni - 1; i++) {
                    size_t n = ID(i, j, k);
                    if (indc[n] != -1) yy[n] = m->bb[n] - yy[n];
                }
        orc_mzminv(ni, nj, nk, dd, m, r0, yy, indc);
        s = orc_mzip(ni, nj, nk, r0, r0, indc);
        xnorm = s;
        if (xnorm <= bnorm * epsr2 + epsa2) goto converged;
        for (size_t n = 0; n < nc; n++)
            if (indc[n] != -1) { rr[n] = r0[n]; pp[n] = r0[n]; }
        for (;;) {
            iter = iter + 1;
            if (iter > icgmax) break;
            orc_mzax(ni, nj, nk, m, pp, yy, indc);
            orc_mzminv(ni, nj, nk, dd, m, ap, yy, indc);
            double w1 = orc_mzip(ni, nj, nk, r0, ap, indc);
            double alpha = s / w1;
            for (size_t n = 0; n < nc; n++)
                if (indc[n] != -1) rr[n] = rr[n] - alpha * ap[n];
            orc_mzax(ni, nj, nk, m, rr, yy, indc);
            orc_mzminv(ni, nj, nk, dd, m, ae, yy, indc);
            w1 = orc_mzip(ni, nj, nk, rr, ae, indc);
            double w2 = orc_mzip(ni, nj, nk, ae, ae, indc);
            double omega = w1 / w2;
            for (size_t n = 0; n < nc; n++)
                if (indc[n] != -1) {
                    xw[n] = xw[n] + (alpha * pp[n] + omega * rr[n]);
                    rr[n] = rr[n] - omega * ae[n];
                }
            xnorm = orc_mzip(ni, nj, nk, rr, rr, indc);
            if (xnorm <= bnorm * epsr2 + epsa2) break;
            if (xnorm > div2) { iter = -2; break; }
            w1 = orc_mzip(ni, nj, nk, r0, rr, indc);
            double beta = w1 * alpha / (s * omega);
            for (size_t n = 0; n < nc; n++)
                if (indc[n] != -1) pp[n] = rr[n] + beta * (pp[n] - omega * ap[n]);
            s = w1;
        }
    }
converged:
    bnorm = sqrt(bnorm);
    xnorm = sqrt(xnorm);
    if (iter >= 0) {
        for (int k = 2; k <= nk - 1; k++)
            for (int j = 2; j <= nj - 1; j++)
                for (int i = 2; i <= ni - 1; i++) {
                    size_t n = ID(i, j, k);
                    if (indc[n] != -1) x0[n] = xw[n];
                }
    }
done_noconv:
    free(w);
    *bnorm_out = bnorm;
    *xnorm_out = xnorm;
    *iter_out = iter;
    return 0;
}

/* ======================================================================
 * Phase (d): correction and boundary values
 * ====================================================================== */

/* vf_vmodif.f:65-210 (wave-boundary branches excluded). */
int orc_vmodif(orc_state *s, const double *pt)
{
    DECL_DIMS(s);
    const double *ggv = s->ggv, *glv = s->glv;
    for (int k = 2; k <= nk - 1; k++)
        for (int j = 2; j <= nj - 1; j++)
            for (int i = 2; i <= ni; i++) {
                size_t n = ID(i, j, k);
                if (s->nf[n - 1] * s->nf[n] != 0) continue;
                int is = s->indx[n];
                if (is == 0) {
                    double gv = XX(6, i) * (XX(2, i) * ggv[n] / glv[n] + XX(2, i - 1) * ggv[n - 1] / glv[n - 1]);
                    s->uu[n] = s->uu[n] + XX(5, i) * gv * (pt[n] - pt[n - 1]);
                } else if (is >= 1 && ib_lv(s, is) == 4) {
                    if (INDB(2, is) == 1)
                        s->uu[n] = s->uu[n] + 2.0 * XX(4, i) * ggv[n] / glv[n] * pt[n];
                    else
                        s->uu[n] = s->uu[n] - 2.0 * XX(4, i - 1) * ggv[n - 1] / glv[n - 1] * pt[n - 1];
                    s->bcu[is - 1] = s->uu[n];
                }
            }
    for (int k = 2; k <= nk - 1; k++)
        for (int j = 2; j <= nj; j++)
            for (int i = 2; i <= ni - 1; i++) {
                size_t n = ID(i, j, k);
                if (s->nf[n - ni] * s->nf[n] != 0) continue;
                int is = s->indy[n];
                if (is == 0) {
                    double gv = YY(6, j) * (YY(2, j) * ggv[n] / glv[n] + YY(2, j - 1) * ggv[n - ni] / glv[n - ni]);
                    s->vv[n] = s->vv[n] + YY(5, j) * gv * (pt[n] - pt[n - ni]);
                } else if (is >= 1 && ib_lv(s, is) == 4) {
                    if (INDB(2, is) == 3)
                        s->vv[n] = s->vv[n] + 2.0 * YY(4, j) * ggv[n] / glv[n] * pt[n];
                    else
                        s->vv[n] = s->vv[n] - 2.0 * YY(4, j - 1) * ggv[n - ni] / glv[n - ni] * pt[n - ni];
                    s->bcv[is - 1] = s->vv[n];
                }
            }
    for (int k = 2; k <= nk; k++)
        for (int j = 2; j <= nj - 1; j++)
            for (int i = 2; i <= ni - 1; i++) {
                size_t n = ID(i, j, k);
                if (s->nf[n - nij] * s->nf[n] != 0) continue;
                int is = s->indz[n];
                if (is == 0) {
                    double gv = ZZ(6, k) * (ZZ(2, k) * ggv[n] / glv[n] + ZZ(2, k - 1) * ggv[n - nij] / glv[n - nij]);
                    s->ww[n] = s->ww[n] + ZZ(5, k) * gv * (pt[n] - pt[n - nij]);
                } else if (is >= 1 && INDB(3, is) == 4) {
                    if (INDB(2, is) == 5)
                        s->ww[n] = s->ww[n] + 2.0 * ZZ(4, k) * ggv[n] / glv[n] * pt[n];
                    else
                        s->ww[n] = s->ww[n] - 2.0 * ZZ(4, k - 1) * ggv[n - nij] / glv[n - nij] * pt[n - nij];
                    s->bcw[is - 1] = s->ww[n];
                }
            }
    const double rt = s->rho0 / s->dtnow;
    for (int k = 2; k <= nk - 1; k++)
        for (int j = 2; j <= nj - 1; j++)
            for (int i = 2; i <= ni - 1; i++) {
                size_t n = ID(i, j, k);
                if (s->indc[n] != -1) s->pp[n] = s->pp[n] - rt * pt[n];
            }
    return 0;
}

/* Tangential velocity on faces that do not touch fluid.  vf_bsuwt.f:47-463.
 * Face of component a between cells n-sa and n; the surface orientation of
 * each side selects a neighbouring face along axis d (1..6 -> -x,+x,-y,+y,
 * -z,+z) whose value is extrapolated (IBSUW0=0) and blended by distance. */
static void bsuwt_axis(orc_state *s, int a)
{
    DECL_DIMS(s);
    const size_t st[3] = {1, (size_t)ni, nij};
    const size_t sa = st[a];
    const int *ind = a == 0 ? s->indx : (a == 1 ? s->indy : s->indz);
    double *q = a == 0 ? s->uu : (a == 1 ? s->vv : s->ww);
    int i0 = 2, j0 = 2, k0 = 2;
    if (a == 0) i0 = 3;
    if (a == 1) j0 = 3;
    if (a == 2) k0 = 3;
    for (int k = k0; k <= nk - 1; k++)
        for (int j = j0; j <= nj - 1; j++)
            for (int i = i0; i <= ni - 1; i++) {
                size_t n = ID(i, j, k);
                if (ind[n] != 0) continue;
                int nl = s->nf[n - sa], nr = s->nf[n];
                if (nl == 8 && nr == 8) {
                    q[n] = 0.0;
                } else if (nl != 0 && nr != 0) {
                    double dl = 0.0, vl = 0.0, dr = 0.0, vr = 0.0;
                    int idx[3] = {i, j, k};
                    /* orientations in the reference's order; skip those normal to a */
                    for (int t = 1; t <= 6; t++) {
                        int d = (t - 1) / 2;
                        if (d == a) continue;
                        if (!(nl == t || nr == t)) continue;
                        int plus = (t % 2 == 0);
                        size_t sd = st[d];
                        int id = idx[d];
                        double v0 = plus ? q[n + sd] : q[n - sd];
                        if (s->ibsuw0 == 0) {
                            size_t n1 = plus ? n + sd : n - sd;
                            size_t n2 = plus ? n + 2 * sd : n - 2 * sd;
                            if (s->nf[n1 - sa] != -1 && s->nf[n1] != -1) {
                                if (s->nf[n2 - sa] != -1 && s->nf[n2] != -1) {
                                    if (s->nf[n2 - sa] == 0 || s->nf[n2] == 0) {
                                        if (plus)
                                            v0 = v0 + (v0 - q[n2]) * MET(s, d, 3, id + 1) / MET(s, d, 3, id + 2);
                                        else
                                            v0 = v0 + (v0 - q[n2]) * MET(s, d, 3, id) / MET(s, d, 3, id - 1);
                                    }
                                }
                            }
                        }
                        double dist = plus ? MET(s, d, 3, id + 1) : MET(s, d, 3, id);
                        if (nl == t) { dl = dist; vl = v0; }
                        if (nr == t) { dr = dist; vr = v0; }
                    }
                    if (nl == 8) q[n] = vr;
                    else if (nr == 8) q[n] = vl;
                    else q[n] = (dr * vl + dl * vr) / (dr + dl);
                }
            }
}

/* vf_bsuwt2.f:39-117: tangential velocities on the gas side of a surface
 * cell copied across a wall corner. */
static void bsuwt2(orc_state *s)
{
    DECL_DIMS(s);
    const size_t st[3] = {1, (size_t)ni, nij};
    double *vel[3] = {s->uu, s->vv, s->ww};
    const int *ind[3] = {s->indx, s->indy, s->indz};
    for (int l = 1; l <= s->nums; l++) {
        int i, j, k;
        decode(s, s->inds[l - 1], &i, &j, &k);
        size_t n = ID(i, j, k);
        int nf0 = s->nf[n];
        if (nf0 < 1 || nf0 > 6) continue;
        int d = (nf0 - 1) / 2;                 /* normal axis */
        size_t lb = (nf0 % 2 == 1) ? n - st[d] : n + st[d];
        /* the two tangential axes in the reference's order (x before y before z) */
        for (int t = 0; t < 3; t++) {
            if (t == d) continue;
            size_t s_t = st[t];
            size_t l1 = n, l2 = n;
            if (s->nf[n - s_t] == 8 && ind[t][lb] != 0) l2 = l2 + s_t;
            if (s->nf[n + s_t] == 8 && ind[t][lb + s_t] != 0) l1 = l1 + s_t;
            vel[t][l1] = vel[t][l2];
        }
    }
}

/* vf_bsuwn.f:43-76: normal velocity on the gas-side face of a surface cell
 * from discrete continuity. */
static void bsuwn(orc_state *s)
{
    DECL_DIMS(s);
    for (int l = 1; l <= s->nums; l++) {
        int i, j, k;
        decode(s, s->inds[l - 1], &i, &j, &k);
        size_t n = ID(i, j, k);
        int nf0 = s->nf[n];
        double fxp = s->ggx[n + 1] * s->uu[n + 1];
        double fxm = s->ggx[n] * s->uu[n];
        double fyp = s->ggy[n + ni] * s->vv[n + ni];
        double fym = s->ggy[n] * s->vv[n];
        double fzp = s->ggz[n + nij] * s->ww[n + nij];
        double fzm = s->ggz[n] * s->ww[n];
        if (nf0 == 5)
            s->ww[n + nij] = (fzm - ZZ(2, k) * (XX(4, i) * (fxp - fxm) + YY(4, j) * (fyp - fym))) / s->ggz[n + nij];
        else if (nf0 == 6)
            s->ww[n] = (fzp + ZZ(2, k) * (XX(4, i) * (fxp - fxm) + YY(4, j) * (fyp - fym))) / s->ggz[n];
        else if (nf0 == 3)
            s->vv[n + ni] = (fym - YY(2, j) * (XX(4, i) * (fxp - fxm) + ZZ(4, k) * (fzp - fzm))) / s->ggy[n + ni];
        else if (nf0 == 4)
            s->vv[n] = (fyp + YY(2, j) * (XX(4, i) * (fxp - fxm) + ZZ(4, k) * (fzp - fzm))) / s->ggy[n];
        else if (nf0 == 1)
            s->uu[n + 1] = (fxm - XX(2, i) * (YY(4, j) * (fyp - fym) + ZZ(4, k) * (fzp - fzm))) / s->ggx[n + 1];
        else if (nf0 == 2)
            s->uu[n] = (fxp + XX(2, i) * (YY(4, j) * (fyp - fym) + ZZ(4, k) * (fzp - fzm))) / s->ggx[n];
    }
}

/* Tangential wall velocities on slip/free boundaries.  vf_bwuwt.f:42-176
 * (types 1,4; 2,3 leave them; 7 radiation does nothing; 5, a wave-maker
 * face, sets them only for LEVEL-ONLY tables, vf_bwuwt.f:80-116, as 4). */
static int bwuwt(orc_state *s)
{
    DECL_DIMS(s);
    for (int l = 1; l <= s->numb; l++) {
        int i, j, k;
        decode(s, INDB(1, l), &i, &j, &k);
        int ns = INDB(2, l), ib = ib_lv(s, l);
        size_t n = ID(i, j, k);
        const double *u = s->uu, *v = s->vv, *w = s->ww;
        if (ib == 1 || ib == 4) {
            switch (ns) {
            case 1: s->bcv[l - 1] = (v[n] + v[n + ni]) * 0.5; s->bcw[l - 1] = (w[n] + w[n + nij]) * 0.5; break;
            case 2: s->bcv[l - 1] = (v[n - 1] + v[n - 1 + ni]) * 0.5; s->bcw[l - 1] = (w[n - 1] + w[n - 1 + nij]) * 0.5; break;
            case 3: s->bcu[l - 1] = (u[n] + u[n + 1]) * 0.5; s->bcw[l - 1] = (w[n] + w[n + nij]) * 0.5; break;
            case 4: s->bcu[l - 1] = (u[n - ni] + u[n - ni + 1]) * 0.5; s->bcw[l - 1] = (w[n - ni] + w[n - ni + nij]) * 0.5; break;
            case 5: s->bcu[l - 1] = (u[n] + u[n + 1]) * 0.5; s->bcv[l - 1] = (v[n] + v[n + ni]) * 0.5; break;
            case 6: s->bcu[l - 1] = (u[n - nij] + u[n - nij + 1]) * 0.5; s->bcv[l - 1] = (v[n - nij] + v[n - nij + ni]) * 0.5; break;
            }
        } else if (ib == 2 || ib == 3 || ib == 5 || ib == 7) {
        } else {
            return -11; /* log-law and wave boundaries are outside the restated path */
        }
    }
    return 0;
}

/* vf_bwpp.f:28-47 */
static void bwpp(orc_state *s)
{
    DECL_DIMS(s);
    for (int l = 1; l <= s->numb; l++) {
        int i, j, k;
        decode(s, INDB(1, l), &i, &j, &k);
        size_t n = ID(i, j, k);
        switch (INDB(2, l)) {
        case 1: s->bcp[l - 1] = s->pp[n]; break;
        case 2: s->bcp[l - 1] = s->pp[n - 1]; break;
        case 3: s->bcp[l - 1] = s->pp[n]; break;
        case 4: s->bcp[l - 1] = s->pp[n - ni]; break;
        case 5: s->bcp[l - 1] = s->pp[n]; break;
        case 6: s->bcp[l - 1] = s->pp[n - nij]; break;
        }
    }
}

/* vf_v1cal.f:293-297 */
int orc_post_bc(orc_state *s)
{
    bsuwt_axis(s, 0);
    bsuwt_axis(s, 1);
    bsuwt_axis(s, 2);
    bsuwt2(s);
    bsuwn(s);
    int rc = bwuwt(s);
    if (rc) return rc;
    bwpp(s);
    return 0;
}

/* VF_V1CAL (vf_v1cal.f:111-297) for ILOOP=1, static porosity. */
static void bsuwn3(orc_state *s);

/* VF_V1CAL pass iloop of a step; passes 2.. of OPTION SUB-LOOP first refresh
 * the boundary values (vf_v1cal.f:115-118). */
int orc_v1cal_loop(orc_state *s, double *pt, int iloop)
{
    DECL_DIMS(s);
    size_t nc = nij * (size_t)nk;
    if (iloop >= 2) {
        bsuwn3(s);
        int rc0 = bwuwt(s);
        if (rc0) return rc0;
    }
    int rc = orc_predictor(s);
    if (rc) return rc;
    double *w = (double *)calloc(8 * nc, sizeof(double));
    if (!w) return -1;
    orc_matrix m = {w, w + nc, w + 2 * nc, w + 3 * nc, w + 4 * nc, w + 5 * nc, w + 6 * nc, w + 7 * nc};
    rc = orc_vpcoef(s, &m);
    if (rc) { free(w); return rc; }
    /* VF_VPSOL: PT=0 then MILU-BiCGSTAB (vf_vpsol.f:55-76) */
    memset(pt, 0, nc * sizeof(double));
    orc_m1bcgs(ni, nj, nk, &m, pt, s->indc, s->cgpara, s->cgepsa, s->cgepsr, s->cgdiv,
               s->icgtyp, s->icgmax, &s->cgbnrm, &s->cgxnrm, &s->icgitr);
    free(w);
    if (s->icgitr > s->icgmax) return -30;
    if (s->icgitr == -2) return -31;
    if (s->icgitr == -1) return -32;
    rc = orc_vmodif(s, pt);
    if (rc) return rc;
    return orc_post_bc(s);
}

int orc_v1cal(orc_state *s, double *pt) { return orc_v1cal_loop(s, pt, 1); }

/* ======================================================================
 * Phase (c): VOF
 * ====================================================================== */

/* Donor-acceptor F flux through faces normal to axis a.  vf_fconv.f:56-200. */
static void fconv_axis(orc_state *s, int a, double *flf)
{
    DECL_DIMS(s);
    const size_t st[3] = {1, (size_t)ni, nij};
    const size_t sa = st[a];
    const int *ind = a == 0 ? s->indx : (a == 1 ? s->indy : s->indz);
    const double *gga = a == 0 ? s->ggx : (a == 1 ? s->ggy : s->ggz);
    const double *q = a == 0 ? s->uu : (a == 1 ? s->vv : s->ww);
    int i1 = ni - 1, j1 = nj - 1, k1 = nk - 1;
    if (a == 0) i1 = ni;
    if (a == 1) j1 = nj;
    if (a == 2) k1 = nk;
    /* tangential orientations that let FAD be replaced by FD */
    for (int k = 2; k <= k1; k++)
        for (int j = 2; j <= j1; j++)
            for (int i = 2; i <= i1; i++) {
                size_t n = ID(i, j, k);
                if (ind[n] == -1) continue;
                int idx[3] = {i, j, k};
                double c = s->dtnow * gga[n] * q[n];
                double ac = fabs(c);
                int lsn;
                size_t la, ld, ldm;
                int ild;
                if (c >= 0.0) { lsn = 1; la = n; ld = n - sa; ldm = n - 2 * sa; ild = idx[a] - 1; }
                else { lsn = -1; la = n - sa; ld = n; ldm = n + sa; ild = idx[a]; }
                int nn = s->nf[ld];
                double fd, fdmw, v, fa;
                if (nn != -1) {
                    fd = s->ff[ld];
                    fdmw = 1.0;
                    if (s->nf[ldm] != -1) fdmw = s->ff[ldm];
                    v = MET(s, a, 2, ild) * s->ggv[ld];
                } else {
                    fd = s->bcf[ind[n] - 1];
                    fdmw = 1.0;
                    v = MET(s, a, 2, ild);
                }
                if (s->nf[la] != -1) fa = s->ff[la];
                else fa = fd;
                double fad = fa;
                int tang = (nn == -1) || ((nn >= 1 && nn <= 6) && ((nn - 1) / 2 != a));
                if (tang) {
                    if (fdmw != 0.0 && fa != 0.0) fad = fd;
                }
                double fdm = dmax(fdmw, fd);
                if (fdmw < s->flower && fa < s->flower) fdm = dmax(fdm, 0.1);
                double cfx = dmax((fdm - fad) * ac - (fdm - fd) * v, 0.0);
                double fvx = fad * ac + cfx;
                double fvm = fd * v;
                flf[n] = flf[n] - (double)lsn * dmin(fvx, fvm);
                if (fvx > fvm) s->nlim[ld] = 1;
            }
}

/* vf_feuler.f:43-58 (static porosity: G0 = GGV, LOOPS = 1, QF = 0) */
static void feuler(orc_state *s, const double *flu, const double *flv, const double *flw)
{
    DECL_DIMS(s);
    for (int k = 2; k <= nk - 1; k++)
        for (int j = 2; j <= nj - 1; j++)
            for (int i = 2; i <= ni - 1; i++) {
                size_t n = ID(i, j, k);
                if (s->nf[n] == -1) continue;
                double g0 = s->ggv[n] - (s->ggv[n] - s->ggv[n]) * (double)s->loops;
                double df = (XX(4, i) * (flu[n + 1] - flu[n])
                           + YY(4, j) * (flv[n + ni] - flv[n])
                           + ZZ(4, k) * (flw[n + nij] - flw[n])
                           + s->dtnow * 0.0);
                s->ff[n] = (g0 * s->ff[n] + df) / s->ggv[n];
            }
}

static void surf_pair(const orc_state *s, int i, int j, int k, int nn, int *e, int *f)
{
    /* (IEC,JEC,KEC) gas side and (IFC,JFC,KFC) fluid side; vf_fmod1.f:64-86 */
    e[0] = f[0] = i; e[1] = f[1] = j; e[2] = f[2] = k;
    (void)s;
    switch (nn) {
    case 5: e[2] = k + 1; f[2] = k - 1; break;
    case 6: e[2] = k - 1; f[2] = k + 1; break;
    case 3: e[1] = j + 1; f[1] = j - 1; break;
    case 4: e[1] = j - 1; f[1] = j + 1; break;
    case 1: e[0] = i + 1; f[0] = i - 1; break;
    case 2: e[0] = i - 1; f[0] = i + 1; break;
    }
}

/* Surface-cell redistribution.  vf_fmod1.f:46-191. */
static void fmod1(orc_state *s)
{
    DECL_DIMS(s);
    size_t nc = nij * (size_t)nk;
    memset(s->nlim, 0, nc * sizeof(int));
    for (int l = 1; l <= s->nums; l++) {
        int i, j, k, e[3], f[3];
        decode(s, s->inds[l - 1], &i, &j, &k);
        surf_pair(s, i, j, k, s->nf[ID(i, j, k)], e, f);
        s->nlim[ID(e[0], e[1], e[2])] += 1;
        s->nlim[ID(f[0], f[1], f[2])] += 1;
    }
    for (int l = 1; l <= s->nums; l++) {
        int i, j, k, e[3], f[3];
        decode(s, s->inds[l - 1], &i, &j, &k);
        surf_pair(s, i, j, k, s->nf[ID(i, j, k)], e, f);
        size_t ne = ID(e[0], e[1], e[2]), n = ID(i, j, k), nf_ = ID(f[0], f[1], f[2]);
        double ve = XX(2, e[0]) * YY(2, e[1]) * ZZ(2, e[2]) * s->ggv[ne];
        double vs = XX(2, i) * YY(2, j) * ZZ(2, k) * s->ggv[n];
        double vf = XX(2, f[0]) * YY(2, f[1]) * ZZ(2, f[2]) * s->ggv[nf_];
        double ss = vs * s->ff[n];
        if (s->nlim[ne] == 1) ss = ss + ve * s->ff[ne];
        if (s->nlim[nf_] == 1) ss = ss + vf * s->ff[nf_];
        double ds;
        if (s->nlim[nf_] == 1) {
            ds = dmin(vf, ss);
            ss = ss - ds;
            s->ff[nf_] = ds / vf;
        }
        ds = dmin(vs, ss);
        ss = ss - ds;
        s->ff[n] = ds / vs;
        if (s->nlim[ne] == 1) {
            ds = dmin(ve, ss);
            ss = ss - ds;
            s->ff[ne] = ds / ve;
        }
        if (ss > 0.0) {
            if (s->nlim[ne] == 1) s->ff[ne] = s->ff[ne] + ss / ve;
            else s->ff[n] = s->ff[n] + ss / vs;
        }
    }
    for (int l = 1; l <= s->nums; l++) {
        int i, j, k, e[3], f[3];
        decode(s, s->inds[l - 1], &i, &j, &k);
        surf_pair(s, i, j, k, s->nf[ID(i, j, k)], e, f);
        size_t ne = ID(e[0], e[1], e[2]), n = ID(i, j, k), nf_ = ID(f[0], f[1], f[2]);
        double ve = XX(2, e[0]) * YY(2, e[1]) * ZZ(2, e[2]) * s->ggv[ne];
        double vs = XX(2, i) * YY(2, j) * ZZ(2, k) * s->ggv[n];
        double vf = XX(2, f[0]) * YY(2, f[1]) * ZZ(2, f[2]) * s->ggv[nf_];
        if (s->ff[n] > 1.0) {
            s->ff[ne] = s->ff[ne] + vs * (s->ff[n] - 1.0) / ve;
            s->ff[n] = 1.0;
        }
        if (s->ff[n] < 0.0) {
            s->ff[nf_] = s->ff[nf_] + vs * s->ff[n] / vf;
            s->ff[n] = 0.0;
        }
    }
}

/* Clip F and volume sums.  vf_fcut01.f:34-55. */
static void fcut01(orc_state *s)
{
    DECL_DIMS(s);
    double fsum = 0.0, fcut = 0.0;
    for (int k = 2; k <= nk - 1; k++)
        for (int j = 2; j <= nj - 1; j++)
            for (int i = 2; i <= ni - 1; i++) {
                size_t n = ID(i, j, k);
                if (s->nf[n] == -1) continue;
                double v = XX(2, i) * YY(2, j) * ZZ(2, k) * s->ggv[n];
                if (s->ff[n] > s->fupper) {
                    fcut = fcut + (s->ff[n] - 1.0) * v;
                    s->ff[n] = 1.0;
                } else if (s->ff[n] < s->flower) {
                    fcut = fcut + (s->ff[n] - 0.0) * v;
                    s->ff[n] = 0.0;
                }
                fsum = fsum + s->ff[n] * v;
            }
    s->fsum = fsum;
    s->fcut = fcut;
}

/* vf_bwff.f:33-87 (fix / free) */
static int bwff(orc_state *s)
{
    DECL_DIMS(s);
    for (int l = 1; l <= s->numb; l++) {
        int i, j, k;
        decode(s, INDB(1, l), &i, &j, &k);
        int ns = INDB(2, l), ib = INDB(4, l);
        size_t n = ID(i, j, k);
        if (ib == 1) {
        } else if (ib == 2 || ((ib == 5 || ib == 7) && ns <= 4)) {
            /* free, and the wave-maker and radiation faces (vf_bwff.f:41-87) */
            switch (ns) {
            case 1: s->bcf[l - 1] = s->ff[n]; break;
            case 2: s->bcf[l - 1] = s->ff[n - 1]; break;
            case 3: s->bcf[l - 1] = s->ff[n]; break;
            case 4: s->bcf[l - 1] = s->ff[n - ni]; break;
            case 5: s->bcf[l - 1] = s->ff[n]; break;
            case 6: s->bcf[l - 1] = s->ff[n - nij]; break;
            }
        } else {
            return -12;
        }
    }
    return 0;
}

/* Normal velocity on the gas side of a surface cell from its tangential
 * neighbours.  vf_bsuwn3.f:37-335. */
static void bsuwn3(orc_state *s)
{
    DECL_DIMS(s);
    const size_t st[3] = {1, (size_t)ni, nij};
    double *vel[3] = {s->uu, s->vv, s->ww};
    for (int l = 1; l <= s->nums; l++) {
        int i, j, k;
        decode(s, s->inds[l - 1], &i, &j, &k);
        size_t n = ID(i, j, k);
        int nf0 = s->nf[n];
        if (nf0 < 1 || nf0 > 6) continue;
        int idx[3] = {i, j, k};
        int d = (nf0 - 1) / 2;
        int plus = (nf0 % 2 == 1); /* fluid on the minus side -> gas-side face is n+sd */
        size_t sd = st[d];
        double *q = vel[d];
        double d0 = MET(s, d, 2, idx[d]);
        double v0 = plus ? q[n] : q[n + sd];
        size_t l1 = plus ? n + sd : n - sd;  /* neighbour cell across the gas-side face */
        size_t wface = plus ? n + sd : n;    /* the face that is written */
        int is_[4] = {0, 0, 0, 0};
        double dd[4] = {0, 0, 0, 0}, vv_[4] = {0, 0, 0, 0};
        int c = 0;
        for (int t = 0; t < 3; t++) {
            if (t == d) continue;
            for (int pm = 0; pm < 2; pm++) {
                size_t ls = pm ? n + st[t] : n - st[t];
                size_t ls1 = pm ? l1 + st[t] : l1 - st[t];
                if (s->nf[ls] == 0 || s->nf[ls1] == 0) {
                    if (s->nf[ls] != -1 && s->nf[ls1] != -1) {
                        is_[c] = 1;
                        dd[c] = pm ? MET(s, t, 3, idx[t] + 1) : MET(s, t, 3, idx[t]);
                        vv_[c] = pm ? q[wface + st[t]] : q[wface - st[t]];
                    }
                }
                c++;
            }
        }
        int iss = is_[0] + is_[1] + is_[2] + is_[3];
        if (iss <= 0) {
            if (plus) {
                if (s->ibsuw0 == 0 && s->nf[n - 2 * sd] == 0)
                    v0 = v0 + (v0 - q[n - sd]) * MET(s, d, 2, idx[d]) / MET(s, d, 2, idx[d] - 1);
            } else {
                if (s->ibsuw0 == 0 && s->nf[n + 2 * sd] == 0)
                    v0 = v0 + (v0 - q[n + 2 * sd]) * MET(s, d, 2, idx[d]) / MET(s, d, 2, idx[d] + 1);
            }
            q[wface] = v0;
        } else {
            double ds = d0 + dd[0] + dd[1] + dd[2] + dd[3];
            q[wface] = ((ds - d0) * v0 + (ds - dd[0]) * vv_[0] + (ds - dd[1]) * vv_[1]
                        + (ds - dd[2]) * vv_[2] + (ds - dd[3]) * vv_[3]) / ds / (double)iss;
        }
    }
}

/* Velocities of gas cells boxed in by surface cells.  vf_bsuwem.f:36-94. */
static void bsuwem(orc_state *s)
{
    DECL_DIMS(s);
    size_t nc = nij * (size_t)nk;
    int *nl = s->nlim;
    memset(nl, 0, nc * sizeof(int));
    for (int l = 1; l <= s->nums; l++) {
        int i, j, k;
        decode(s, s->inds[l - 1], &i, &j, &k);
        size_t n = ID(i, j, k);
        int nn = s->nf[n];
        if (nn == 5 || nn == 6) {
            nl[n + 1] += 1; nl[n - 1] += 1; nl[n + ni] += 1; nl[n - ni] += 1;
        } else if (nn == 3 || nn == 4) {
            nl[n + 1] += 1; nl[n - 1] += 1; nl[n + nij] += 1; nl[n - nij] += 1;
        } else {
            nl[n + ni] += 1; nl[n - ni] += 1; nl[n + nij] += 1; nl[n - nij] += 1;
        }
    }
    const int *nf = s->nf;
    for (int k = 2; k <= nk - 1; k++)
        for (int j = 2; j <= nj - 1; j++)
            for (int i = 2; i <= ni - 1; i++) {
                size_t n = ID(i, j, k);
                if (!(nl[n] >= 2 && nf[n] == 8)) continue;
                if (nf[n - 1] == 8 && nl[n - 1] < 2) { if (nf[n + 1] != 8) s->uu[n] = s->uu[n + 1]; }
                if (nf[n + 1] == 8 && nl[n + 1] < 2) { if (nf[n - 1] != 8) s->uu[n + 1] = s->uu[n]; }
                if (nf[n - ni] == 8 && nl[n - ni] < 2) { if (nf[n + ni] != 8) s->vv[n] = s->vv[n + ni]; }
                if (nf[n + ni] == 8 && nl[n + ni] < 2) { if (nf[n - ni] != 8) s->vv[n + ni] = s->vv[n]; }
                if (nf[n - nij] == 8 && nl[n - nij] < 2) { if (nf[n + nij] != 8) s->ww[n] = s->ww[n + nij]; }
                if (nf[n + nij] == 8 && nl[n + nij] < 2) { if (nf[n - nij] != 8) s->ww[n + nij] = s->ww[n]; }
            }
}

/* F value of the face neighbour across the face with index ind (BCF on a
 * boundary face); used by vf_fnfini.f and vf_fnfprv.f. */
static double fnb(const orc_state *s, int ind, size_t nb)
{
    return ind >= 1 ? s->bcf[ind - 1] : s->ff[nb];
}

/* Cell flags fluid(0) / surface(7) / gas(8) and the surface list.
 * vf_fnfini.f:36-189. */
static void fnfini(orc_state *s)
{
    DECL_DIMS(s);
    int *nf = s->nf;
    for (int k = 2; k <= nk - 1; k++)
        for (int j = 1; j <= nj; j++)
            for (int i = 1; i <= ni; i++) {
                size_t n = ID(i, j, k);
                if (nf[n] != -1) nf[n] = 0;
            }
    for (int k = 2; k <= nk - 1; k++)
        for (int j = 2; j <= nj - 1; j++)
            for (int i = 2; i <= ni - 1; i++) {
                size_t n = ID(i, j, k);
                if (nf[n] == -1 || !(s->ff[n] < s->flower)) continue;
                int ix1 = s->indx[n], ix2 = s->indx[n + 1];
                int iy1 = s->indy[n], iy2 = s->indy[n + ni];
                int iz1 = s->indz[n], iz2 = s->indz[n + nij];
                double fx1 = fnb(s, ix1, n - 1), fx2 = fnb(s, ix2, n + 1);
                double fy1 = fnb(s, iy1, n - ni), fy2 = fnb(s, iy2, n + ni);
                double fz1 = fnb(s, iz1, n - nij), fz2 = fnb(s, iz2, n + nij);
                if (!(fz1 > s->fupper && iz1 == 0) && !(fy1 > s->fupper && iy1 == 0)
                    && !(fy2 > s->fupper && iy2 == 0) && !(fx1 > s->fupper && ix1 == 0)
                    && !(fx2 > s->fupper && ix2 == 0) && !(fz2 > s->fupper && iz2 == 0))
                    nf[n] = 8;
            }
    for (;;) {
        int ns = 0;
        for (int k = 2; k <= nk - 1; k++)
            for (int j = 2; j <= nj - 1; j++)
                for (int i = 2; i <= ni - 1; i++) {
                    size_t n = ID(i, j, k);
                    if (nf[n] == -1 || nf[n] == 8) continue;
                    if (nf[n + nij] == 8 || nf[n + ni] == 8 || nf[n - ni] == 8 || nf[n + 1] == 8
                        || nf[n - 1] == 8 || nf[n - nij] == 8) {
                        s->inds[ns++] = (int)n + 1;
                        nf[n] = 7;
                    }
                }
        int nums = 0;
        for (int l = 0; l < ns; l++) {
            int i00 = s->inds[l];
            size_t n = (size_t)i00 - 1;
            if (nf[n + nij] != 0 && nf[n + ni] != 0 && nf[n - ni] != 0 && nf[n + 1] != 0
                && nf[n - 1] != 0 && nf[n - nij] != 0)
                nf[n] = 8;
            else
                s->inds[nums++] = i00;
        }
        s->nums = nums;
        int nch = 0;
        for (int l = 0; l < nums; l++) {
            size_t n = (size_t)s->inds[l] - 1;
            if ((nf[n - nij] != 0 || nf[n + nij] != 8) && (nf[n - 1] != 0 || nf[n + 1] != 8)
                && (nf[n + 1] != 0 || nf[n - 1] != 8) && (nf[n - ni] != 0 || nf[n + ni] != 8)
                && (nf[n + ni] != 0 || nf[n - ni] != 8) && (nf[n + nij] != 0 || nf[n - nij] != 8)) {
                nf[n] = -7;
                nch++;
            }
        }
        if (nch <= 0) break;
        for (int l = 0; l < nums; l++) {
            size_t n = (size_t)s->inds[l] - 1;
            if (nf[n] == -7) nf[n] = 8;
        }
    }
}

/* Surface orientation.  vf_fnfprv.f:34-223. */
static int fnfprv(orc_state *s)
{
    DECL_DIMS(s);
    const size_t st[3] = {1, (size_t)ni, nij};
    const int *ind[3] = {s->indx, s->indy, s->indz};
    int *nf = s->nf;
    /* candidate order: z-, z+, y-, y+, x-, x+ -> NF 5, 6, 3, 4, 1, 2 */
    static const int order[6] = {5, 6, 3, 4, 1, 2};
    for (int l = 0; l < s->nums; l++) {
        size_t n = (size_t)s->inds[l] - 1;
        double fm = -1.0;
        for (int o = 0; o < 6; o++) {
            int code = order[o];
            int d = (code - 1) / 2;
            int minus = (code % 2 == 1);
            size_t nb = minus ? n - st[d] : n + st[d];     /* fluid neighbour */
            size_t op = minus ? n + st[d] : n - st[d];     /* must be gas */
            if (!(nf[nb] == 0 && nf[op] == 8)) continue;
            double fs = 2.0 * s->ff[nb];
            for (int t = 0; t < 3; t++) {
                if (t == d) continue;
                fs = fs + fnb(s, ind[t][nb], nb - st[t]);
                fs = fs + fnb(s, ind[t][nb + st[t]], nb + st[t]);
            }
            if (fs > fm) { nf[n] = code; fm = fs; }
        }
        if (nf[n] == 7) return -40;
    }
    return 0;
}

/* vf_cindc.f:35-50 */
static void cindc(orc_state *s)
{
    DECL_DIMS(s);
    for (int k = 2; k <= nk - 1; k++)
        for (int j = 2; j <= nj - 1; j++)
            for (int i = 2; i <= ni - 1; i++) {
                size_t n = ID(i, j, k);
                int v = s->nf[n];
                s->indc[n] = (v == 8 || v == -1) ? -1 : 0;
            }
}

/* Pressure of cells that turned from gas to fluid.  vf_bsppfl.f:40-62. */
static void bsppfl(orc_state *s)
{
    DECL_DIMS(s);
    size_t nc = nij * (size_t)nk;
    double *pp0 = (double *)malloc(nc * sizeof(double));
    memcpy(pp0, s->pp, nc * sizeof(double));
    for (int k = 2; k <= nk - 1; k++)
        for (int j = 2; j <= nj - 1; j++)
            for (int i = 2; i <= ni - 1; i++) {
                size_t n = ID(i, j, k);
                if (!(s->nf[n] == 0 && s->nlim[n] == 8)) continue;
                double sumf = 0.0, sump = 0.0;
                for (int kp = -1; kp <= 1; kp++)
                    for (int jp = -1; jp <= 1; jp++)
                        for (int ip = -1; ip <= 1; ip++) {
                            size_t m = ID(i + ip, j + jp, k + kp);
                            if (s->nf[m] == 0) {
                                double f = s->ff[m];
                                sumf = sumf + f;
                                sump = sump + f * pp0[m];
                            }
                        }
                if (sumf < ZERO) s->pp[n] = 0.0;
                else s->pp[n] = sump / sumf;
            }
    free(pp0);
}

/* Surface and gas pressures.  vf_bspp.f:36-62 (no air-pressure model). */
static void bspp(orc_state *s)
{
    DECL_DIMS(s);
    const double pps = 0.0;
    for (int k = 2; k <= nk - 1; k++)
        for (int j = 2; j <= nj - 1; j++)
            for (int i = 2; i <= ni - 1; i++) {
                size_t n = ID(i, j, k);
                int v = s->nf[n];
                double psi;
                switch (v) {
                case 8: s->pp[n] = 0.0; break;
                case 5: psi = ZZ(3, k) / (0.5 * ZZ(2, k - 1) + s->ff[n] * ZZ(2, k)); s->pp[n] = (1.0 - psi) * s->pp[n - nij] + psi * pps; break;
                case 6: psi = ZZ(3, k + 1) / (0.5 * ZZ(2, k + 1) + s->ff[n] * ZZ(2, k)); s->pp[n] = (1.0 - psi) * s->pp[n + nij] + psi * pps; break;
                case 3: psi = YY(3, j) / (0.5 * YY(2, j - 1) + s->ff[n] * YY(2, j)); s->pp[n] = (1.0 - psi) * s->pp[n - ni] + psi * pps; break;
                case 4: psi = YY(3, j + 1) / (0.5 * YY(2, j + 1) + s->ff[n] * YY(2, j)); s->pp[n] = (1.0 - psi) * s->pp[n + ni] + psi * pps; break;
                case 1: psi = XX(3, i) / (0.5 * XX(2, i - 1) + s->ff[n] * XX(2, i)); s->pp[n] = (1.0 - psi) * s->pp[n - 1] + psi * pps; break;
                case 2: psi = XX(3, i + 1) / (0.5 * XX(2, i + 1) + s->ff[n] * XX(2, i)); s->pp[n] = (1.0 - psi) * s->pp[n + 1] + psi * pps; break;
                default: break;
                }
            }
}

/* VF_F1CAL (vf_f1cal.f:124-239) for the restated feature set. */
/* ======================================================================
 * TimerDoor corrections
 * ====================================================================== */

/* Bubble rise: in every level K whose timer has run out, water moves down
 * from K+1 into fluid cells of K that are not full.  vf_fbubup.f:35-51. */
void orc_fbubup(orc_state *s)
{
    DECL_DIMS(s);
    for (int k = nk - 2; k >= 2; k--) {
        const double dtbub = ZZ(2, k) / s->wbub;
        if (s->tnow - s->tbub[k - 1] > dtbub) {
            const int k1 = k + 1;
            for (int j = 2; j <= nj - 1; j++)
                for (int i = 2; i <= ni - 1; i++) {
                    const size_t n = ID(i, j, k), n1 = ID(i, j, k1);
                    if (s->ff[n] < s->fupper && s->nf[n] == 0 && s->nf[n1] != -1) {
                        const double v0 = ZZ(2, k) * s->ggv[n];
                        const double v1 = ZZ(2, k1) * s->ggv[n1];
                        const double fmove = dmin((1.0 - s->ff[n]) * v0, s->ff[n1] * v1);
                        s->ff[n1] = s->ff[n1] - fmove / v1;
                        s->ff[n] = s->ff[n] + fmove / v0;
                    }
                }
            s->tbub[k - 1] = s->tbub[k - 1] + dtbub;
        }
    }
}

/* droplet state of a cell that holds no droplet (vf_fdropf.f:299-306) */
static void drop_clear(orc_state *s, size_t n)
{
    s->droptx[n] = -1.0e10;
    s->dropty[n] = 0.0;
    s->droptz[n] = 0.0;
    s->dropuu[n] = 0.0;
    s->dropvv[n] = 0.0;
    s->dropww[n] = 0.0;
}

/* cell-centre velocity of cell c added to (up,vp,wp) (vf_fdropf.f:82-84) */
static void drop_addvel(const orc_state *s, size_t c, double *up, double *vp, double *wp)
{
    const size_t ni = (size_t)s->numi, nij = ni * (size_t)s->numj;
    *up = *up + 0.5 * (s->uu[c + 1] + s->uu[c]);
    *vp = *vp + 0.5 * (s->vv[c + ni] + s->vv[c]);
    *wp = *wp + 0.5 * (s->ww[c + nij] + s->ww[c]);
}

/* weighted droplet values moved to (or kept in) cell m, from droplet n, with
 * weight q; the moving axis a has its time advanced by dt and, along z, its
 * velocity reduced by dt*g (vf_fdropf.f:359-367, 371-380, 599-609) */
typedef struct { double *tx, *ty, *tz, *uu, *vv, *ww, *ff; } drop_wk;

static void drop_acc(const orc_state *s, const drop_wk *w, size_t m, size_t n, double q, int a,
                     double dt)
{
    w->tx[m] = w->tx[m] + q * (a == 0 ? s->droptx[n] + dt : s->droptx[n]);
    w->ty[m] = w->ty[m] + q * (a == 1 ? s->dropty[n] + dt : s->dropty[n]);
    w->tz[m] = w->tz[m] + q * (a == 2 ? s->droptz[n] + dt : s->droptz[n]);
    w->uu[m] = w->uu[m] + q * s->dropuu[n];
    w->vv[m] = w->vv[m] + q * s->dropvv[n];
    w->ww[m] = w->ww[m] + q * (a == 2 ? s->dropww[n] - dt * s->grz0 : s->dropww[n]);
    w->ff[m] = w->ff[m] + q;
}

static int is_drop(const orc_state *s, size_t n) { return s->ff[n] > s->flower && s->nf[n] == 8; }

/* New droplets take the mean velocity of their fluid (else surface)
 * neighbours; old droplets resting on an obstacle slide down the local
 * obstacle/porosity slope.  vf_fdropf.f:66-307. */
static void fdropf_start(orc_state *s)
{
    DECL_DIMS(s);
    /* edge neighbours in the reference's order: offsets (di,dj,dk) */
    static const int edge[12][3] = {
        {-1, 0, -1}, {1, 0, -1}, {-1, 0, 1}, {1, 0, 1},
        {-1, -1, 0}, {1, -1, 0}, {-1, 1, 0}, {1, 1, 0},
        {0, -1, -1}, {0, 1, -1}, {0, -1, 1}, {0, 1, 1}};
    static const int face[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, 0, -1}, {0, 0, 1}, {0, -1, 0}, {0, 1, 0}};
    for (int k = 2; k <= nk - 1; k++)
        for (int j = 2; j <= nj - 1; j++)
            for (int i = 2; i <= ni - 1; i++) {
                const size_t n = ID(i, j, k);
                if (!is_drop(s, n)) {
                    drop_clear(s, n);
                    continue;
                }
                if (s->droptx[n] < 0.0) {
                    s->droptx[n] = s->tnow;
                    s->dropty[n] = s->tnow;
                    s->droptz[n] = s->tnow;
                    int ip = 0;
                    double up = 0.0, vp = 0.0, wp = 0.0;
                    for (int e = 0; e < 12; e++) {
                        const int *o = edge[e];
                        const size_t c = ID(i + o[0], j + o[1], k + o[2]);
                        /* the two face cells between n and c: first along the
                         * lower axis of the pair, then along the higher one */
                        size_t fa, fb;
                        if (o[0] != 0) {
                            fa = ID(i + o[0], j, k);
                            fb = o[1] != 0 ? ID(i, j + o[1], k) : ID(i, j, k + o[2]);
                        } else {
                            fa = ID(i, j + o[1], k);
                            fb = ID(i, j, k + o[2]);
                        }
                        if (s->nf[c] == 0 && (s->nf[fa] != -1 || s->nf[fb] != -1)) {
                            ip++;
                            drop_addvel(s, c, &up, &vp, &wp);
                        }
                    }
                    if (ip == 0)
                        for (int e = 0; e < 6; e++) {
                            const size_t c = ID(i + face[e][0], j + face[e][1], k + face[e][2]);
                            if (s->nf[c] != -1 && s->nf[c] != 8) {
                                ip++;
                                drop_addvel(s, c, &up, &vp, &wp);
                            }
                        }
                    if (ip >= 1) {
                        const double cp = 1.0 / (double)ip;
                        s->dropuu[n] = up * cp;
                        s->dropvv[n] = vp * cp;
                        s->dropww[n] = wp * cp;
                    } else {
                        s->dropuu[n] = 0.5 * (s->uu[n] + s->uu[n + 1]);
                        s->dropvv[n] = 0.5 * (s->vv[n] + s->vv[n + ni]);
                        s->dropww[n] = 0.5 * (s->ww[n] + s->ww[n + nij]);
                    }
                    if (s->nf[n - nij] == -1) {
                        s->dropuu[n] = 0.0;
                        s->dropvv[n] = 0.0;
                        s->dropww[n] = 0.0;
                    }
                } else if (s->nf[n - nij] == -1) {
                    double h1 = s->nf[n - 1] == -1 ? 1.0 : 1.0 - s->ggv[n - 1];
                    double h2 = s->nf[n + 1] == -1 ? 1.0 : 1.0 - s->ggv[n + 1];
                    double hx = 0.5 * (h2 - h1) * ZZ(2, k);
                    h1 = s->nf[n - ni] == -1 ? 1.0 : 1.0 - s->ggv[n - ni];
                    h2 = s->nf[n + ni] == -1 ? 1.0 : 1.0 - s->ggv[n + ni];
                    double hy = 0.5 * (h2 - h1) * ZZ(2, k);
                    hx = hx / XX(2, i);
                    hy = hy / YY(2, j);
                    const double dn = 1.0 / (hx * hx + hy * hy + 1.0);
                    hx = -s->grz0 * hx * dn;
                    hy = -s->grz0 * hy * dn;
                    s->dropuu[n] = s->dropuu[n] + s->dtnow * hx;
                    s->dropvv[n] = s->dropvv[n] + s->dtnow * hy;
                }
            }
}

/* One pass of droplet moves along axis a (0 x, 1 y, 2 z) towards +a (l = 1)
 * or -a (l = 2), then the weighted means.  vf_fdropf.f:317-398 (x),
 * :409-490 (y), :501-601 (z). */
static void fdropf_pass(orc_state *s, const drop_wk *w, int a, int l)
{
    DECL_DIMS(s);
    const int nn[3] = {ni, nj, nk};
    const size_t st[3] = {1, (size_t)ni, nij};
    for (int k = 2; k <= nk - 1; k++)
        for (int j = 2; j <= nj - 1; j++)
            for (int i = 2; i <= ni - 1; i++) {
                const size_t n = ID(i, j, k);
                w->tx[n] = 0.0; w->ty[n] = 0.0; w->tz[n] = 0.0;
                w->uu[n] = 0.0; w->vv[n] = 0.0; w->ww[n] = 0.0; w->ff[n] = 0.0;
            }
    /* sweep range of the moving index: interior, against the move */
    const int dl = l == 1 ? -1 : 1;
    const int p0 = l == 1 ? nn[a] - 1 : 2, p1 = l == 1 ? 2 : nn[a] - 1;
    const double sg = l == 1 ? 1.0 : -1.0;
    for (int p = p0; dl > 0 ? p <= p1 : p >= p1; p += dl) {
        /* the reference's nests: x (I; K; J), y (J; K; I), z (K; J; I) */
        const double xz = a == 2 ? sqrt(2.0 * s->grz0 * ZZ(2, p)) : 0.0;
        for (int o = 2; o <= (a == 2 ? nj : nk) - 1; o++)
            for (int q = 2; q <= (a == 0 ? nj : ni) - 1; q++) {
                int i, j, k;
                if (a == 0) { i = p; k = o; j = q; }
                else if (a == 1) { j = p; k = o; i = q; }
                else { k = p; j = o; i = q; }
                const size_t n = ID(i, j, k);
                const size_t n1 = dl > 0 ? n - st[a] : n + st[a];
                double dt = 0.0;
                if (is_drop(s, n) && s->nf[n1] != -1) {
                    int go = 0;
                    double tlast;
                    if (a == 2) {
                        const double wv = s->dropww[n];
                        if (l == 1) {
                            if (wv > xz) dt = (wv - sqrt(wv * wv - xz * xz)) / s->grz0;
                        } else if (wv <= xz) {
                            if (wv <= 0.0)
                                dt = (wv + sqrt(wv * wv + xz * xz)) / s->grz0;
                            else
                                dt = 2.0 * wv / s->grz0;
                        }
                        tlast = s->droptz[n];
                        go = dt > 0.0 && s->tnow - tlast > dt;
                    } else {
                        const double vel = a == 0 ? s->dropuu[n] : s->dropvv[n];
                        if (sg * vel > 0.0) {
                            dt = (a == 0 ? XX(2, i) : YY(2, j)) / fabs(vel);
                            tlast = a == 0 ? s->droptx[n] : s->dropty[n];
                            go = s->tnow - tlast > dt;
                            if (!go) dt = 0.0;
                        }
                    }
                    if (a == 2 && !go) dt = 0.0;
                    if (go) {
                        int i1 = i, j1 = j, k1 = k;
                        if (a == 0) i1 = i - dl; else if (a == 1) j1 = j - dl; else k1 = k - dl;
                        const double v0 = XX(2, i) * YY(2, j) * ZZ(2, k) * s->ggv[n];
                        const double v1 = XX(2, i1) * YY(2, j1) * ZZ(2, k1) * s->ggv[n1];
                        const double fmove = dmin((1.0 - s->ff[n1]) * v1, s->ff[n] * v0);
                        s->ff[n] = s->ff[n] - fmove / v0;
                        s->ff[n1] = s->ff[n1] + fmove / v1;
                        if (s->nf[n1] == 8) drop_acc(s, w, n1, n, fmove, a, dt);
                    }
                }
                if (is_drop(s, n)) {
                    const double fold = s->ff[n] * XX(2, i) * YY(2, j) * ZZ(2, k) * s->ggv[n];
                    drop_acc(s, w, n, n, fold, a, dt);
                } else {
                    drop_clear(s, n);
                }
            }
    }
    for (int k = 2; k <= nk - 1; k++)
        for (int j = 2; j <= nj - 1; j++)
            for (int i = 2; i <= ni - 1; i++) {
                const size_t n = ID(i, j, k);
                if (!is_drop(s, n)) continue;
                const double wq = 1.0 / w->ff[n];
                s->droptx[n] = w->tx[n] * wq;
                s->dropty[n] = w->ty[n] * wq;
                s->droptz[n] = w->tz[n] * wq;
                s->dropuu[n] = w->uu[n] * wq;
                s->dropvv[n] = w->vv[n] * wq;
                s->dropww[n] = w->ww[n] * wq;
                if (a == 2 && l == 1) {
                    const double xz = sqrt(2.0 * s->grz0 * ZZ(2, k));
                    if (s->nf[n + nij] == -1 && s->dropww[n] > xz) {
                        s->dropuu[n] = 0.0;
                        s->dropvv[n] = 0.0;
                        s->dropww[n] = 0.0;
                    }
                }
            }
}

/* Droplet free fall and run-down.  vf_fdropf.f:1-678. */
int orc_fdropf(orc_state *s)
{
    DECL_DIMS(s);
    const size_t nc = nij * (size_t)nk;
    double *buf = (double *)malloc(7 * nc * sizeof(double));
    if (!buf) return -1;
    drop_wk w = {buf, buf + nc, buf + 2 * nc, buf + 3 * nc, buf + 4 * nc, buf + 5 * nc, buf + 6 * nc};
    fdropf_start(s);
    for (int a = 0; a < 3; a++)
        for (int l = 1; l <= 2; l++) fdropf_pass(s, &w, a, l);
    free(buf);
    return 0;
}

/* VF_BWFFSF (vf_bwffsf.f:44-125): the cells next to a side driven by a
 * LEVEL-ONLY table take F from the table's level at TNOW (ramped):
 * 0 above it, 1 below it, the filled fraction in the cell it cuts. */
static int bwffsf_side(orc_state *s, int jd)
{
    DECL_DIMS(s);
    {
        const double t = s->wv_t[jd - 1], t0 = s->wv_t0[jd - 1];
        double a = s->wv_a[jd - 1];
        float r4tn = (float)(t0 - s->tnow / t);
        r4tn = r4tn - (float)(int)r4tn;
        if (r4tn < 0.0f) r4tn = r4tn + 1.0f;
        double aw = 1.0;
        if (a >= ZERO) {
            a = s->tnow / t / a;
            if (a < 1.0) aw = 0.5 * sin(PI_REF * (a - 0.5)) + 0.5;
        }
        double wmt1, wmt2, wvz;
        int rc = mtb_time(s, (double)r4tn, &wmt1, &wmt2, &wvz);
        if (rc) return rc;
        s->wv_level[jd - 1] = wvz * aw;
        const double wl = s->wvlvl + s->wv_level[jd - 1];
        const int xside = jd <= 2;
        const int cpos = xside ? (jd == 1 ? 2 : ni - 1) : (jd == 3 ? 2 : nj - 1);
        for (int k = 2; k <= nk - 1; k++) {
            double val;
            if (ZZ(1, k) > wl)
                val = 0.0;
            else if (ZZ(1, k + 1) < wl)
                val = 1.0;
            else
                val = (wl - ZZ(1, k)) / (ZZ(1, k + 1) - ZZ(1, k));
            for (int q = s->wv_l1[jd - 1]; q <= s->wv_l2[jd - 1]; q++) {
                const size_t c = xside ? ID(cpos, q, k) : ID(q, cpos, k);
                if (s->nf[c] != -1) s->ff[c] = val;
            }
        }
    }
    return 0;
}

static int bwffsf(orc_state *s)
{
    for (int jd = 1; jd <= 4; jd++) {
        if (!s->wv_on[jd - 1]) continue;
        const int t2 = s->wv_tab[jd - 1] == 1;
        if (!(t2 ? s->mtb_only2 : s->mtb_only)) continue;
        if (t2) swap_tables(s);
        const int rc = bwffsf_side(s, jd);
        if (t2) swap_tables(s);
        if (rc) return rc;
    }
    return 0;
}

int orc_f1cal(orc_state *s)
{
    DECL_DIMS(s);
    size_t nc = nij * (size_t)nk;
    double *w = (double *)calloc(3 * nc, sizeof(double));
    if (!w) return -1;
    double *flfu = w, *flfv = w + nc, *flfw = w + 2 * nc;
    memset(s->nlim, 0, nc * sizeof(int));
    fconv_axis(s, 0, flfu);
    fconv_axis(s, 1, flfv);
    fconv_axis(s, 2, flfw);
    feuler(s, flfu, flfv, flfw);
    free(w);
    int rc0 = bwffsf(s);
    if (rc0) return rc0;
    fmod1(s);
    fcut01(s);
    int rc = bwff(s);
    if (rc) return rc;
    bsuwn3(s);
    bsuwem(s);
    memcpy(s->nlim, s->nf, nc * sizeof(int));
    fnfini(s);
    if (s->wbub >= ZERO) {
        orc_fbubup(s);
        fnfini(s);
    }
    if (s->idrop >= 1) {
        rc = orc_fdropf(s);
        if (rc) return rc;
        fnfini(s);
    }
    rc = fnfprv(s);
    if (rc) return rc;
    cindc(s);
    bsppfl(s);
    bsuwt_axis(s, 0);
    bsuwt_axis(s, 1);
    bsuwt_axis(s, 2);
    bsuwt2(s);
    bsuwn3(s);
    bspp(s);
    rc = bwuwt(s);
    if (rc) return rc;
    bwpp(s);
    return 0;
}

/* vf_cdiv00.f:38-56 */
double orc_cdiv00(const orc_state *s)
{
    DECL_DIMS(s);
    double val = 0.0;
    for (int k = 2; k <= nk - 1; k++)
        for (int j = 2; j <= nj - 1; j++)
            for (int i = 2; i <= ni - 1; i++) {
                size_t n = ID(i, j, k);
                if (s->indc[n] == -1) continue;
                double vd = YY(2, j) * ZZ(2, k) * (s->ggx[n + 1] * s->uu[n + 1] - s->ggx[n] * s->uu[n])
                          + XX(2, i) * ZZ(2, k) * (s->ggy[n + ni] * s->vv[n + ni] - s->ggy[n] * s->vv[n])
                          + XX(2, i) * YY(2, j) * (s->ggz[n + nij] * s->ww[n + nij] - s->ggz[n] * s->ww[n]);
                val = val + vd * vd;
            }
    return sqrt(val);
}

/* Device-side part of VF_CINIT restated: vf_cinit.f:190-409 after F holds its
 * initial value.  flags bit 0: initial velocity uvw on non-boundary faces with
 * a fluid side (vf_cinit.f:216-258); bit 1: pressure RHO0*GRZ0*(WVLVL - z) on
 * active cells (vf_cinit.f:283-296). */
int orc_init_state(orc_state *s, double wvlvl, double uini, double vini, double wini, int flags)
{
    DECL_DIMS(s);
    s->wvlvl = wvlvl;
    size_t nc = nij * (size_t)nk;
    /* VF_BWFFSF before the clip (vf_cinit.f:177) */
    int rc = bwffsf(s);
    if (rc) return rc;
    fcut01(s);
    rc = bwff(s);
    if (rc) return rc;
    fnfini(s);
    rc = fnfprv(s);
    if (rc) return rc;
    for (size_t n = 0; n < nc; n++) s->indc[n] = -1;
    cindc(s);
    if (flags & 1) {
        const double ini[3] = {uini, vini, wini};
        double *vel[3] = {s->uu, s->vv, s->ww};
        const int *ind[3] = {s->indx, s->indy, s->indz};
        const size_t st[3] = {1, (size_t)ni, nij};
        for (int a = 0; a < 3; a++) {
            memset(vel[a], 0, nc * sizeof(double));
            for (int k = (a == 2 ? 3 : 2); k <= nk - 1; k++)
                for (int j = (a == 1 ? 3 : 2); j <= nj - 1; j++)
                    for (int i = (a == 0 ? 3 : 2); i <= ni - 1; i++) {
                        size_t n = ID(i, j, k);
                        if (ind[a][n] == 0 && s->nf[n - st[a]] * s->nf[n] == 0) vel[a][n] = ini[a];
                    }
        }
    }
    if (flags & 2) {
        memset(s->pp, 0, nc * sizeof(double));
        for (int k = 2; k <= nk - 1; k++) {
            double val = s->rho0 * s->grz0 * (wvlvl - (ZZ(1, k) + ZZ(1, k + 1)) * 0.5);
            for (int j = 2; j <= nj - 1; j++)
                for (int i = 2; i <= ni - 1; i++) {
                    size_t n = ID(i, j, k);
                    if (s->indc[n] != -1) s->pp[n] = val;
                }
        }
    }
    /* VF_BWUWN with UB = a copy of the velocities, NNOW = 0 (vf_cinit.f:376-390) */
    double *bk = (double *)malloc(3 * nc * sizeof(double));
    if (!bk) return -1;
    memcpy(bk, s->uu, nc * sizeof(double));
    memcpy(bk + nc, s->vv, nc * sizeof(double));
    memcpy(bk + 2 * nc, s->ww, nc * sizeof(double));
    rc = bwuwn(s, bk, bk + nc, bk + 2 * nc, 1);
    free(bk);
    if (rc) return rc;
    bsuwt_axis(s, 0);
    bsuwt_axis(s, 1);
    bsuwt_axis(s, 2);
    bsuwt2(s);
    bsuwn3(s);
    bspp(s);
    rc = bwuwt(s);
    if (rc) return rc;
    bwpp(s);
    /* TimerDoor timers (vf_cinit.f:412-421) */
    if (s->tbub)
        for (int k = 0; k < nk; k++) s->tbub[k] = 0.0;
    if (s->droptx)
        for (size_t n = 0; n < nc; n++) drop_clear(s, n);
    return 0;
}

/* One pass of the time loop body with DT = DTNOW (vf_a1main.f:1029-1104):
 * LOOPS passes of TNOW += DT/LOOPS and VF_V1CAL at DTNOW = DT/LOOPS, then
 * DTNOW = DT and VF_F1CAL. */
int orc_step(orc_state *s, double *pt)
{
    const double dt = s->dtnow;
    s->dtnow = dt / (double)s->loops;
    for (int l = 1; l <= s->loops; l++) {
        s->tnow = s->tnow + s->dtnow;
        int rc = orc_v1cal_loop(s, pt, l);
        if (rc) return rc;
    }
    s->dtnow = dt;
    return orc_f1cal(s);
}

/* Step size of the next step (VF_CDTCAL, vf_cdtcal.f, one rank so that
 * VF_P0MIND is the identity).  The limit is folded cell by cell into DT in
 * the reference's loop order; each cell's flux limit V/max|face flux| and
 * diffusion limit (0.5/sum 1/dx^2)/nu enter through DT = V/MAX(..., V/DT).
 * nnow = 0 takes DTINIT as the previous step. */
double orc_cdtcal(const orc_state *s, int nnow, double dtinit, double dtsafe, double dtmin,
                  double dtmax)
{
    DECL_DIMS(s);
    double dt = dtmax / dtsafe;
    for (int k = 2; k <= nk - 1; k++)
        for (int j = 2; j <= nj - 1; j++)
            for (int i = 2; i <= ni - 1; i++) {
                const size_t n = ID(i, j, k);
                if (s->nf[n] == -1) continue;
                const double dx = XX(2, i), dy = YY(2, j), dz = ZZ(2, k);
                const double dxy = dx * dy, dxz = dx * dz, dyz = dy * dz;
                const double fc = s->ff[n];
                double v = s->ggv[n] * dx * dy * dz;
                double fl[6];
                const size_t face[6] = {n, n + 1, n, n + ni, n, n + nij};
                const double *g[3] = {s->ggx, s->ggy, s->ggz};
                const double *q[3] = {s->uu, s->vv, s->ww};
                const int *ind[3] = {s->indx, s->indy, s->indz};
                const double ar[3] = {dyz, dxz, dxy};
                for (int e = 0; e < 6; e++) {
                    const int a = e / 2;
                    const size_t f = face[e];
                    fl[e] = fabs(ar[a] * g[a][f] * q[a][f]);
                    const int b = ind[a][f];
                    if (b >= 1 && dmax(fc, s->bcf[b - 1]) <= 0.0) fl[e] = 0.0;
                }
                double m = fl[0];
                for (int e = 1; e < 6; e++) m = dmax(m, fl[e]);
                dt = v / dmax(m, v / dt);
                v = 0.5 / (1.0 / (dx * dx) + 1.0 / (dy * dy) + 1.0 / (dz * dz));
                dt = v / dmax(s->anu[n], v / dt);
            }
    const double w = nnow == 0 ? dtinit : s->dtnow;
    return dmax(dmin(dt * dtsafe, 1.20 * w), dtmin);
}
