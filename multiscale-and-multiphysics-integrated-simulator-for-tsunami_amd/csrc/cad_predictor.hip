// cad_predictor.hip -- phase (a): explicit momentum predictor on gfx950.
//
// The reference computes nine flux arrays and three source arrays in separate
// passes (VF_VFLXDU/V/W, VF_VGENE), stores them in WK01..WK12 and re-reads them
// in VF_VEULER (Source code/vf_v1cal.f:126-227): about fifteen full-field
// streams.  Here one kernel evaluates, for each cell's three faces, the six
// fluxes and the source each face's update needs and writes the predicted
// velocities directly: the flux arrays never reach HBM.  A flux value
// depends only on its own location's inputs, so re-evaluating it for each of
// the two faces that difference it gives bit-identical results to the stored
// array.  The kernel is out of place (old u, v, w read, new component written to
// a spare buffer) because every update reads the old values of all three
// components.
#include "cad_common.h"

#include <cmath>

namespace {

// Normal flux of component A at cell n (FLUU/FLVV/FLWW):
// vf_vflxdu.f:71-90, vf_vflxdv.f:71-90, vf_vflxdw.f:66-85.
template <int A>
__device__ __forceinline__ double flux_normal(const Static &s, const Dims &d, const double *q,
                                              const double *gl, const double *gg, const int *indc,
                                              double su, double sc, long long n, int ia)
{
    if (indc[n] == -1) return 0.0;
    const long long sa = stride(d, A);
    double fl = 0.0;
    double win = (gl[n] * q[n] + gl[n + sa] * q[n + sa]) * 0.5;
    double fg = win * (q[n] + q[n + sa]) * 0.5;
    fl = fl - su * (dmaxd(win, 0.0) * q[n] - dmaxd(-win, 0.0) * q[n + sa]) - sc * fg;
    fl = fl + (gg[n] + gg[n + sa]) * s.anu[n] * met<A>(s, 4, ia) * (q[n + sa] - q[n]);
    return fl;
}

// Cross flux of component A through faces normal to B on the (A-face, B-face)
// edge at n (FLUV, FLUW, FLVU, FLVW, FLWU, FLWV), including the wall /
// free-surface branch: vf_vflxdu.f:94-347, vf_vflxdv.f:94-347,
// vf_vflxdw.f:88-341 (the six are axis permutations of one formula).
template <int A, int B>
__device__ __forceinline__ double flux_cross(const Static &s, const Dims &d, const double *q,
                                             const double *vb, const double *glb,
                                             const double *ggb, const int *ind_a,
                                             const int *ind_b, const double *bcq, const int *nf,
                                             double su, double sc, long long n, int ia, int jb)
{
    const long long sa = stride(d, A), sb = stride(d, B);
    const int ic = nf[n - sa - sb] * nf[n - sb] * nf[n - sa] * nf[n];
    const int it = ind_a[n];
    const int ib = ind_a[n - sb];
    if (ic != 0 || (it != 0 && ib != 0)) return 0.0;
    const double *anu = s.anu;
    double fl = 0.0;
    if (it == 0 && ib == 0) {
        double win = met<A>(s, 6, ia) * (met<A>(s, 2, ia - 1) * glb[n] * vb[n]
                                         + met<A>(s, 2, ia) * glb[n - sa] * vb[n - sa]);
        double fg = win * met<B>(s, 6, jb) * (met<B>(s, 2, jb - 1) * q[n] + met<B>(s, 2, jb) * q[n - sb]);
        fl = fl - su * (dmaxd(win, 0.0) * q[n - sb] - dmaxd(-win, 0.0) * q[n]) - sc * fg;
        double gg = met<A>(s, 6, ia) * (met<A>(s, 2, ia - 1) * ggb[n] + met<A>(s, 2, ia) * ggb[n - sa]);
        double anr = met<B>(s, 6, jb) * (met<B>(s, 2, jb - 1) * anu[n] + met<B>(s, 2, jb) * anu[n - sb]);
        double anl = met<B>(s, 6, jb) * (met<B>(s, 2, jb - 1) * anu[n - sa] + met<B>(s, 2, jb) * anu[n - sa - sb]);
        double an = met<A>(s, 6, ia) * (met<A>(s, 2, ia - 1) * anr + met<A>(s, 2, ia) * anl);
        double d1 = met<B>(s, 5, jb) * (q[n] - q[n - sb]);
        double d2 = met<A>(s, 5, ia) * (vb[n] - vb[n - sa]);
        fl = fl + gg * an * (d1 + d2);
        return fl;
    }
    const int nsb = 2 * B + 1;
    double br, bl;
    int is = ind_b[n];
    if (is >= 1) {
        const int *e = s.indb + (long long)(is - 1) * 4;
        if (e[2] == 1 || e[2] == 4) br = e[1] == nsb ? q[n] : q[n - sb];
        else br = bcq[is - 1];
    } else {
        br = met<B>(s, 6, jb) * (met<B>(s, 2, jb - 1) * (q[n] + q[n + sa])
                                 + met<B>(s, 2, jb) * (q[n - sb] + q[n - sb + sa])) * 0.5;
    }
    is = ind_b[n - sa];
    if (is >= 1) {
        const int *e = s.indb + (long long)(is - 1) * 4;
        if (e[2] == 1 || e[2] == 4) bl = e[1] == nsb ? q[n] : q[n - sb];
        else bl = bcq[is - 1];
    } else {
        bl = met<B>(s, 6, jb) * (met<B>(s, 2, jb - 1) * (q[n] + q[n - sa])
                                 + met<B>(s, 2, jb) * (q[n - sb] + q[n - sb - sa])) * 0.5;
    }
    double win = met<A>(s, 6, ia) * (met<A>(s, 2, ia - 1) * glb[n] * vb[n]
                                     + met<A>(s, 2, ia) * glb[n - sa] * vb[n - sa]);
    double fg = met<A>(s, 6, ia) * (met<A>(s, 2, ia) * glb[n] * vb[n] * br
                                    + met<A>(s, 2, ia - 1) * glb[n - sa] * vb[n - sa] * bl);
    if (it == 0 && win < 0.0)
        fl = fl - su * win * q[n] - sc * fg;
    else if (ib == 0 && win > 0.0)
        fl = fl - su * win * q[n - sb] - sc * fg;
    else
        fl = fl - fg;
    double gg = met<A>(s, 6, ia) * (met<A>(s, 2, ia - 1) * ggb[n] + met<A>(s, 2, ia) * ggb[n - sa]);
    fg = met<A>(s, 6, ia) * (met<A>(s, 2, ia) * br + met<A>(s, 2, ia - 1) * bl);
    double anr, anl;
    if (nf[n] == -1) anr = anu[n - sb];
    else if (nf[n - sb] == -1) anr = anu[n];
    else anr = met<B>(s, 6, jb) * (met<B>(s, 2, jb - 1) * anu[n] + met<B>(s, 2, jb) * anu[n - sb]);
    if (nf[n - sa] == -1) anl = anu[n - sa - sb];
    else if (nf[n - sa - sb] == -1) anl = anu[n - sa];
    else anl = met<B>(s, 6, jb) * (met<B>(s, 2, jb - 1) * anu[n - sa] + met<B>(s, 2, jb) * anu[n - sa - sb]);
    double an = met<A>(s, 6, ia) * (met<A>(s, 2, ia - 1) * anr + met<A>(s, 2, ia) * anl);
    double d1;
    if (it == 0) d1 = met<B>(s, 4, jb) * (q[n] - fg) * 2.0;
    else d1 = met<B>(s, 4, jb - 1) * (fg - q[n - sb]) * 2.0;
    double d2 = met<A>(s, 5, ia) * (vb[n] - vb[n - sa]);
    fl = fl + gg * an * (d1 + d2);
    return fl;
}

template <int A> __device__ __forceinline__ const double *vel(const double *u, const double *v, const double *w)
{
    return A == 0 ? u : (A == 1 ? v : w);
}
template <int A> __device__ __forceinline__ const double *gl_of(const Static &s)
{
    return A == 0 ? s.glx : (A == 1 ? s.gly : s.glz);
}
template <int A> __device__ __forceinline__ const double *gg_of(const Static &s)
{
    return A == 0 ? s.ggx : (A == 1 ? s.ggy : s.ggz);
}
template <int A> __device__ __forceinline__ const int *ind_of(const Static &s)
{
    return A == 0 ? s.indx : (A == 1 ? s.indy : s.indz);
}

// Cross flux of component A through B-faces evaluated at the edge location n.
template <int A, int B>
__device__ __forceinline__ double cross_at(const Static &s, const Dims &d, const double *u,
                                           const double *v, const double *w, const double *bcq,
                                           const int *nf, double su, double sc, long long n,
                                           const int idx[3])
{
    return flux_cross<A, B>(s, d, vel<A>(u, v, w), vel<B>(u, v, w), gl_of<B>(s), gg_of<B>(s),
                            ind_of<A>(s), ind_of<B>(s), bcq, nf, su, sc, n, idx[A], idx[B]);
}

// x**n for n >= 0 by squaring, the lowest bit first (Fortran's REAL**INTEGER
// as the reference's compiler evaluates it)
__device__ __forceinline__ double ipow(double x, int n)
{
    double r = 1.0;
    while (n > 0) {
        if (n & 1) r *= x;
        n >>= 1;
        if (n) x *= x;
    }
    return r;
}

// One face of component A (its value when the face is not updated): VF_VGENE
// source (vf_vgene.f:84-330) and
// VF_VEULER update with the VVMAX clip (vf_veuler.f:116-249).  DF selects
// the Dupuit-Forchheimer drag of OPTION DRAG-DF (vf_vgene.f:111-146) in
// place of the CD form (:102-110); SP the implicit damping of MODEL DAMP
// after the clip (vf_veuler.f:131-166, 187-221, 243-277).
template <int A, bool DF, bool SP>
__device__ __forceinline__ double predict_face(const Dims &d, const Static &s, const DragDF &df,
                                               const Damp &dz, const double *__restrict__ u,
                                               const double *__restrict__ v,
                                               const double *__restrict__ w,
                                               const double *__restrict__ p,
                                               const double *__restrict__ bcq,
                                               const int *__restrict__ nf,
                                               const int *__restrict__ indc, const cad_params &prm,
                                               int i, int j, int k, long long n)
{
    const double *q = vel<A>(u, v, w);
    const int idx[3] = {i, j, k};
    const int lo[3] = {A == 0 ? 3 : 2, A == 1 ? 3 : 2, A == 2 ? 3 : 2};
    const int hi[3] = {d.ni - 1, d.nj - 1, d.nk - 1};
    const long long sa = stride(d, A);
    const int *ind = ind_of<A>(s);
    bool upd = idx[0] >= lo[0] && idx[0] <= hi[0] && idx[1] >= lo[1] && idx[1] <= hi[1] &&
               idx[2] >= lo[2] && idx[2] <= hi[2];
    if (upd) upd = ind[n] == 0 && nf[n - sa] * nf[n] == 0;
    if (!upd) return q[n];
    const double su = prm.scmvp, sc = 1.0 - prm.scmvp;
    const int ia = idx[A];
    // fluxes differenced by this face, in the x, y, z order of VF_VEULER
    double t[3];
#pragma unroll
    for (int dd = 0; dd < 3; dd++) {
        if (dd == A) {
            double fn = flux_normal<A>(s, d, q, gl_of<A>(s), gg_of<A>(s), indc, su, sc, n, ia);
            double fm = flux_normal<A>(s, d, q, gl_of<A>(s), gg_of<A>(s), indc, su, sc, n - sa, ia - 1);
            t[dd] = met<A>(s, 5, ia) * (fn - fm);
        } else {
            const long long sd = stride(d, dd);
            int idp[3] = {i, j, k};
            idp[dd] += 1;
            double fn, fp;
            if (dd == 0) {
                fn = cross_at<A, (A == 0 ? 1 : 0)>(s, d, u, v, w, bcq, nf, su, sc, n, idx);
                fp = cross_at<A, (A == 0 ? 1 : 0)>(s, d, u, v, w, bcq, nf, su, sc, n + sd, idp);
            } else if (dd == 1) {
                fn = cross_at<A, (A == 1 ? 0 : 1)>(s, d, u, v, w, bcq, nf, su, sc, n, idx);
                fp = cross_at<A, (A == 1 ? 0 : 1)>(s, d, u, v, w, bcq, nf, su, sc, n + sd, idp);
            } else {
                fn = cross_at<A, (A == 2 ? 0 : 2)>(s, d, u, v, w, bcq, nf, su, sc, n, idx);
                fp = cross_at<A, (A == 2 ? 0 : 2)>(s, d, u, v, w, bcq, nf, su, sc, n + sd, idp);
            }
            t[dd] = metr(s, dd, 4, idx[dd]) * (fp - fn);
        }
    }
    // source term (QU/QV/QW start from zero in VF_V1CAL)
    const double ri = 1.0 / prm.rho0;
    double qs = 0.0;
    double gv = met<A>(s, 6, ia) * (met<A>(s, 2, ia) * s.ggv[n] + met<A>(s, 2, ia - 1) * s.ggv[n - sa]);
    double cd = met<A>(s, 6, ia) * (met<A>(s, 2, ia) * s.cd0[n] + met<A>(s, 2, ia - 1) * s.cd0[n - sa]);
    qs = qs + gv * ri * met<A>(s, 5, ia) * (p[n - sa] - p[n]);
    double c[3];
    const double *vv3[3] = {u, v, w};
    if (!DF) {
#pragma unroll
        for (int dd = 0; dd < 3; dd++) {
            if (dd == A) {
                c[dd] = q[n];
            } else {
                const double *vd = vv3[dd];
                const long long sd = stride(d, dd);
                c[dd] = met<A>(s, 6, ia) * (met<A>(s, 2, ia - 1) * (vd[n] + vd[n + sd])
                                            + met<A>(s, 2, ia) * (vd[n - sa] + vd[n - sa + sd])) * 0.5;
            }
        }
        double uvw = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
        qs = qs - 0.5 * cd * met<A>(s, 5, ia) * (1.0 - gg_of<A>(s)[n]) * c[A] * uvw;
    } else {
        // table row: off below PLOWER or the first diameter, else the row
        // before the first DR(L) >= CD, the last row past the end
        // (vf_vgene.f:112-130)
        double rn, dr, ap, bt;
        if (gv <= prm.plower || cd <= df.dr[0]) {
            rn = prm.plower;
            dr = df.dr[0];
            ap = 0.0;
            bt = 0.0;
        } else {
            rn = gv;
            dr = cd;
            ap = df.ap[df.n - 1];
            bt = df.bt[df.n - 1];
            for (int l = 1; l < df.n; l++)
                if (dr <= df.dr[l]) {
                    ap = df.ap[l - 1];
                    bt = df.bt[l - 1];
                    break;
                }
        }
        const double rn1 = 1.0 - rn;
        const double alpha = ap * rn1 * rn1 * rn1 * df.yu / (rn * rn * dr * dr);
        const double beta = bt * rn1 / (rn * rn * rn * dr);
        // velocities weighted by their face porosities (vf_vgene.f:133-141)
        const double *gd3[3] = {s.ggx, s.ggy, s.ggz};
#pragma unroll
        for (int dd = 0; dd < 3; dd++) {
            const double *gd = gd3[dd];
            if (dd == A) {
                c[dd] = q[n] * gd[n];
            } else {
                const double *vd = vv3[dd];
                const long long sd = stride(d, dd);
                c[dd] = met<A>(s, 6, ia) * (met<A>(s, 2, ia - 1) * (vd[n] * gd[n] + vd[n + sd] * gd[n + sd])
                                            + met<A>(s, 2, ia) * (vd[n - sa] * gd[n - sa]
                                                                  + vd[n - sa + sd] * gd[n - sa + sd])) * 0.5;
            }
        }
        double uvw = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
        qs = qs - c[A] * (alpha + beta * uvw) * gv;
    }
    if (A == 2) {
        qs = qs - gv * prm.grz0;
        double dr = 0.0;
        qs = qs - gv * prm.grz0 * dr / prm.rho0;
    }
    // Euler update; GV0 = GV for static porosity (vf_v1cal.f:218-227)
    double g = met<A>(s, 6, ia) * (met<A>(s, 2, ia) * s.glv[n] + met<A>(s, 2, ia - 1) * s.glv[n - sa]);
    double g0 = met<A>(s, 6, ia) * (met<A>(s, 2, ia) * s.glv[n] + met<A>(s, 2, ia - 1) * s.glv[n - sa]);
    double du = (t[0] + t[1] + t[2] + qs);
    double r = (g * q[n] + prm.dtnow * du) / g0;
    if (r > prm.vvmax) r = prm.vvmax;
    if (r < -prm.vvmax) r = -prm.vvmax;
    if (SP) {
        // the largest P (1 - dist/WIDTH)**n over the zones within WIDTH of
        // the face's (x, y): x at the face for u, at the cell centre else
        const double x = A == 0 ? metr(s, 0, 1, i) : (metr(s, 0, 1, i) + metr(s, 0, 1, i + 1)) * 0.5;
        const double y = A == 1 ? metr(s, 1, 1, j) : (metr(s, 1, 1, j) + metr(s, 1, 1, j + 1)) * 0.5;
        const double dist[4] = {x - dz.xmin, dz.xmax - x, y - dz.ymin, dz.ymax - y};
        double dmx = 0.0;
#pragma unroll
        for (int jd = 0; jd < 4; jd++) {
            if (dz.deg[jd] < 0) continue;
            double e = dist[jd];
            if (e <= dz.w[jd]) {
                e = (A == 2 ? dz.pz[jd] : dz.px[jd]) * ipow(1.0 - e / dz.w[jd], dz.deg[jd]);
                if (e > dmx) dmx = e;
            }
        }
        const double gg = met<A>(s, 6, ia) * (met<A>(s, 2, ia) * s.ggv[n] + met<A>(s, 2, ia - 1) * s.ggv[n - sa]);
        r = r * (g0 / (g0 + prm.dtnow * gg * dmx));
    }
    return r;
}

// The three components of one cell's faces in one pass: the neighbourhood
// each component reads (velocities, porosities, flags, viscosity) is shared,
// so it comes from HBM once for all three instead of once per component.
// A thread takes PRED_KU cells of its (I, J) column, so the K neighbours the
// cells share are loaded once and all their loads are in flight together.
constexpr int PRED_KU = 2;
template <bool DF, bool SP>
__global__ void __launch_bounds__(256) k_predict(Dims d, Static s, DragDF df, Damp dz,
                                                 const double *__restrict__ u,
                                                 const double *__restrict__ v,
                                                 const double *__restrict__ w,
                                                 const double *__restrict__ p,
                                                 const double *__restrict__ bcu,
                                                 const double *__restrict__ bcv,
                                                 const double *__restrict__ bcw,
                                                 const int *__restrict__ nf,
                                                 const int *__restrict__ indc, double *__restrict__ u2,
                                                 double *__restrict__ v2, double *__restrict__ w2,
                                                 cad_params prm)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    if (i > d.ni || j > d.nj) return;
#pragma unroll
    for (int kk = 0; kk < PRED_KU; kk++) {
        const int k = blockIdx.z * PRED_KU + kk + 1;
        if (k > d.nk) break;
        const long long n = cid(d, i, j, k);
        u2[n] = predict_face<0, DF, SP>(d, s, df, dz, u, v, w, p, bcu, nf, indc, prm, i, j, k, n);
        v2[n] = predict_face<1, DF, SP>(d, s, df, dz, u, v, w, p, bcv, nf, indc, prm, i, j, k, n);
        w2[n] = predict_face<2, DF, SP>(d, s, df, dz, u, v, w, p, bcw, nf, indc, prm, i, j, k, n);
    }
}

// VF_BWUWN boundary-list pass (vf_bwuwn.f:84-254): normal velocity on wall and
// free boundary faces.  Type 5 (wave-maker) and type 7 (radiation) faces are
// set afterwards by the special sides (cad_wave.hip).
__global__ void k_bwuwn(Dims d, const int *__restrict__ indb, int lvo, double *u, double *v, double *w,
                        double *bcu, double *bcv, double *bcw)
{
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= d.numb) return;
    const int *e = indb + (long long)l * 4;
    const long long n = (long long)e[0] - 1;
    const int ns = e[1], ib = face_type(indb, lvo, l + 1);
    if (ib == 1 || ib == 6 || ib == 8) {
        if (ns == 1 || ns == 2) { bcu[l] = 0.0; u[n] = 0.0; }
        else if (ns == 3 || ns == 4) { bcv[l] = 0.0; v[n] = 0.0; }
        else { bcw[l] = 0.0; w[n] = 0.0; }
    } else if (ib == 2) {
        bcu[l] = 0.0; bcv[l] = 0.0; bcw[l] = 0.0;
        if (ns == 1 || ns == 2) u[n] = 0.0;
        else if (ns == 3 || ns == 4) v[n] = 0.0;
        else w[n] = 0.0;
    } else if (ib == 3) {
        if (ns == 1 || ns == 2) u[n] = bcu[l];
        else if (ns == 3 || ns == 4) v[n] = bcv[l];
        else w[n] = bcw[l];
    } else if (ib == 4) {
        double val;
        switch (ns) {
        case 1: val = u[n + 1]; bcu[l] = val; u[n] = val; break;
        case 2: val = u[n - 1]; bcu[l] = val; u[n] = val; break;
        case 3: val = v[n + d.ni]; bcv[l] = val; v[n] = val; break;
        case 4: val = v[n - d.ni]; bcv[l] = val; v[n] = val; break;
        case 5: val = w[n + d.nij]; bcw[l] = val; w[n] = val; break;
        case 6: val = w[n - d.nij]; bcw[l] = val; w[n] = val; break;
        default: break;
        }
    }
}

// the three components of the predictor, out of place into u2, v2, w2
template <bool DF, bool SP>
void launch_predict(cad_ctx *c, const Static &s, const DragDF &df, const Damp &dz)
{
    const Dims &d = c->d;
    dim3 b = cell_block(), g = cell_grid(d);
    g.z = (unsigned)((d.nk + PRED_KU - 1) / PRED_KU);
    Fields &f = c->f;
    hipLaunchKernelGGL((k_predict<DF, SP>), g, b, 0, c->stream, d, s, df, dz, f.uu, f.vv, f.ww, f.pp,
                       f.bcu, f.bcv, f.bcw, f.nf, f.indc, c->u2, c->v2, c->w2, c->prm);
}

// The damping zones at the context's GRZ0, as VF_VEULER derives them on each
// call (vf_veuler.f:75-113); false when no side has one.
bool damp_zones(const cad_ctx *c, Damp *dz)
{
    int ipa = 0;
    for (int jd = 0; jd < 4; jd++) {
        dz->deg[jd] = -1;
        dz->w[jd] = dz->px[jd] = dz->pz[jd] = 0.0;
        if (!c->dm_on[jd]) continue;
        ipa++;
        const double e = std::sqrt(c->prm.grz0 / c->dm_par[jd][3]);
        dz->deg[jd] = c->dm_deg[jd];
        dz->w[jd] = c->dm_par[jd][2];
        dz->px[jd] = (double)(c->dm_deg[jd] + 1) * c->dm_par[jd][0] * e;
        dz->pz[jd] = (double)(c->dm_deg[jd] + 1) * c->dm_par[jd][1] * e;
    }
    // GLXMIN..GLYMAX: the first and last grid coordinates, XX(1,2), XX(1,NUMI)
    dz->xmin = c->xx_host[6 * 1];
    dz->xmax = c->xx_host[6 * (size_t)(c->d.ni - 1)];
    dz->ymin = c->yy_host[6 * 1];
    dz->ymax = c->yy_host[6 * (size_t)(c->d.nj - 1)];
    return ipa > 0;
}

}  // namespace

int cad_launch_predictor(cad_ctx *c)
{
    Static s = cad_static(c);
    Fields &f = c->f;
    const DragDF &df = c->drag;
    Damp dz;
    const bool sp = damp_zones(c, &dz);
    if (df.n > 0 && sp) launch_predict<true, true>(c, s, df, dz);
    else if (df.n > 0) launch_predict<true, false>(c, s, df, dz);
    else if (sp) launch_predict<false, true>(c, s, df, dz);
    else launch_predict<false, false>(c, s, df, dz);
    CAD_CHECK(hipGetLastError());
    double *t;
    t = f.uu; f.uu = c->u2; c->u2 = t;
    t = f.vv; f.vv = c->v2; c->v2 = t;
    t = f.ww; f.ww = c->w2; c->w2 = t;
    return CAD_OK;
}

int cad_launch_bwuwn(cad_ctx *c, bool initial)
{
    const Dims &d = c->d;
    if (d.numb <= 0) return CAD_OK;
    hipLaunchKernelGGL(k_bwuwn, dim3((d.numb + 255) / 256), dim3(256), 0, c->stream, d, c->indb,
                       cad_level_sides(c),
                       c->f.uu, c->f.vv, c->f.ww, c->f.bcu, c->f.bcv, c->f.bcw);
    CAD_CHECK(hipGetLastError());
    // the wave-maker and radiation sides (vf_bwuwn.f:257-620)
    return cad_wave_apply(c, initial);
}
