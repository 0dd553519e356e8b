// cad_correct.hip -- phase (d): velocity / pressure correction and the
// boundary values that follow it (Source code/vf_v1cal.f:287-297), plus the
// surface-cell boundary routines shared with phase (c).
//
// The reference walks the surface list INDS in VF_BSUWT2/VF_BSUWN/VF_BSUWN3.
// Each surface cell writes only faces between itself and a gas cell, and reads
// faces no other surface cell writes (a surface cell's orientation always has
// fluid on one side and gas on the other, vf_fnfprv.f:42-43), so the list loop
// is order independent and runs here as one thread per cell with the cell's
// flag selecting the work; no list needs to be built on the device.
#include "cad_common.h"

namespace {

// Velocity correction of component A, vf_vmodif.f:65-197 (types 0 and 4).
template <int A>
__global__ void __launch_bounds__(256) k_vmodif(Dims d, Static s, const double *__restrict__ pt,
                                                const int *__restrict__ nf, double *q, double *bcq)
{
    int i, j, k;
    if (!cell_of_thread(d, i, j, k)) return;
    const int idx[3] = {i, j, k};
    const int hi[3] = {A == 0 ? d.ni : d.ni - 1, A == 1 ? d.nj : d.nj - 1, A == 2 ? d.nk : d.nk - 1};
    if (i < 2 || j < 2 || k < 2 || i > hi[0] || j > hi[1] || k > hi[2]) return;
    const long long n = cid(d, i, j, k);
    const long long sa = stride(d, A);
    if (nf[n - sa] * nf[n] != 0) return;
    const int *ind = A == 0 ? s.indx : (A == 1 ? s.indy : s.indz);
    const int is = ind[n];
    const int ia = idx[A];
    const double *ggv = s.ggv, *glv = s.glv;
    if (is == 0) {
        double gv = met<A>(s, 6, ia) * (met<A>(s, 2, ia) * ggv[n] / glv[n] + met<A>(s, 2, ia - 1) * ggv[n - sa] / glv[n - sa]);
        q[n] = q[n] + met<A>(s, 5, ia) * gv * (pt[n] - pt[n - sa]);
    } else if (is >= 1 && face_type(s.indb, s.lvo, is) == 4) {
        if (s.indb[(long long)(is - 1) * 4 + 1] == 2 * A + 1)
            q[n] = q[n] + 2.0 * met<A>(s, 4, ia) * ggv[n] / glv[n] * pt[n];
        else
            q[n] = q[n] - 2.0 * met<A>(s, 4, ia - 1) * ggv[n - sa] / glv[n - sa] * pt[n - sa];
        bcq[is - 1] = q[n];
    }
}

// Pressure correction p -= rho/dt * phi, vf_vmodif.f:199-206
__global__ void __launch_bounds__(256) k_pmodif(Dims d, const double *__restrict__ pt,
                                                const int *__restrict__ indc, double *p, double rt)
{
    int i, j, k;
    if (!cell_of_thread(d, i, j, k)) return;
    if (i < 2 || i > d.ni - 1 || j < 2 || j > d.nj - 1 || k < 2 || k > d.nk - 1) return;
    const long long n = cid(d, i, j, k);
    if (indc[n] != -1) p[n] = p[n] - rt * pt[n];
}

// Tangential velocity on faces with no fluid side, vf_bsuwt.f:47-463: the
// face normal to A at cell (i, j, k) of component q.
template <int A>
__device__ __forceinline__ void bsuwt_face(const Dims &d, const Static &s, const int *__restrict__ nf, double *q,
                                           int ibsuw0, int i, int j, int k)
{
    const int idx[3] = {i, j, k};
    if (i < (A == 0 ? 3 : 2) || j < (A == 1 ? 3 : 2) || k < (A == 2 ? 3 : 2)) return;
    if (i > d.ni - 1 || j > d.nj - 1 || k > d.nk - 1) return;
    const long long n = cid(d, i, j, k);
    const int *ind = A == 0 ? s.indx : (A == 1 ? s.indy : s.indz);
    if (ind[n] != 0) return;
    const long long sa = stride(d, A);
    const int nl = nf[n - sa], nr = nf[n];
    if (nl == 8 && nr == 8) {
        q[n] = 0.0;
        return;
    }
    if (nl == 0 || nr == 0) return;
    double dl = 0.0, vl = 0.0, dr = 0.0, vr = 0.0;
#pragma unroll
    for (int t = 1; t <= 6; t++) {
        const int dd = (t - 1) / 2;
        if (dd == A) continue;
        if (!(nl == t || nr == t)) continue;
        const bool plus = (t % 2 == 0);
        const long long sd = stride(d, dd);
        const int id = idx[dd];
        double v0 = plus ? q[n + sd] : q[n - sd];
        if (ibsuw0 == 0) {
            const long long n1 = plus ? n + sd : n - sd;
            const long long n2 = plus ? n + 2 * sd : n - 2 * sd;
            if (nf[n1 - sa] != -1 && nf[n1] != -1) {
                if (nf[n2 - sa] != -1 && nf[n2] != -1) {
                    if (nf[n2 - sa] == 0 || nf[n2] == 0) {
                        if (plus)
                            v0 = v0 + (v0 - q[n2]) * metr(s, dd, 3, id + 1) / metr(s, dd, 3, id + 2);
                        else
                            v0 = v0 + (v0 - q[n2]) * metr(s, dd, 3, id) / metr(s, dd, 3, id - 1);
                    }
                }
            }
        }
        const double dist = plus ? metr(s, dd, 3, id + 1) : metr(s, dd, 3, id);
        if (nl == t) { dl = dist; vl = v0; }
        if (nr == t) { dr = dist; vr = v0; }
    }
    if (nl == 8) q[n] = vr;
    else if (nr == 8) q[n] = vl;
    else q[n] = (dr * vl + dl * vr) / (dr + dl);
}

// The three components in one pass over the cells: each reads and writes
// only its own velocity array, so the pass gives each component's result,
// and the flags and face indices they share are read once.
__global__ void __launch_bounds__(256) k_bsuwt(Dims d, Static s, const int *__restrict__ nf, double *u, double *v,
                                               double *w, int ibsuw0)
{
    int i, j, k;
    if (!cell_of_thread(d, i, j, k)) return;
    bsuwt_face<0>(d, s, nf, u, ibsuw0, i, j, k);
    bsuwt_face<1>(d, s, nf, v, ibsuw0, i, j, k);
    bsuwt_face<2>(d, s, nf, w, ibsuw0, i, j, k);
}

// Tangential velocities across wall corners of surface cells, vf_bsuwt2.f:39-117.
__global__ void __launch_bounds__(256) k_bsuwt2(Dims d, Static s, const int *__restrict__ nf,
                                                double *u, double *v, double *w)
{
    int i, j, k;
    if (!cell_of_thread(d, i, j, k)) return;
    if (i < 2 || i > d.ni - 1 || j < 2 || j > d.nj - 1 || k < 2 || k > d.nk - 1) return;
    const long long n = cid(d, i, j, k);
    const int nf0 = nf[n];
    if (nf0 < 1 || nf0 > 6) return;
    const int dn = (nf0 - 1) / 2;
    const long long lb = (nf0 % 2 == 1) ? n - stride(d, dn) : n + stride(d, dn);
    double *vel[3] = {u, v, w};
    const int *ind[3] = {s.indx, s.indy, s.indz};
#pragma unroll
    for (int t = 0; t < 3; t++) {
        if (t == dn) continue;
        const long long st = stride(d, t);
        long long l1 = n, l2 = n;
        if (nf[n - st] == 8 && ind[t][lb] != 0) l2 = l2 + st;
        if (nf[n + st] == 8 && ind[t][lb + st] != 0) l1 = l1 + st;
        if (l1 != l2) vel[t][l1] = vel[t][l2];
    }
}

// Normal velocity of the gas-side face from continuity, vf_bsuwn.f:43-76.
__global__ void __launch_bounds__(256) k_bsuwn(Dims d, Static s, const int *__restrict__ nf,
                                               double *u, double *v, double *w)
{
    int i, j, k;
    if (!cell_of_thread(d, i, j, k)) return;
    if (i < 2 || i > d.ni - 1 || j < 2 || j > d.nj - 1 || k < 2 || k > d.nk - 1) return;
    const long long n = cid(d, i, j, k);
    const int nf0 = nf[n];
    if (nf0 < 1 || nf0 > 6) return;
    const long long ni = d.ni, nij = d.nij;
    double fxp = s.ggx[n + 1] * u[n + 1];
    double fxm = s.ggx[n] * u[n];
    double fyp = s.ggy[n + ni] * v[n + ni];
    double fym = s.ggy[n] * v[n];
    double fzp = s.ggz[n + nij] * w[n + nij];
    double fzm = s.ggz[n] * w[n];
    if (nf0 == 5)
        w[n + nij] = (fzm - met<2>(s, 2, k) * (met<0>(s, 4, i) * (fxp - fxm) + met<1>(s, 4, j) * (fyp - fym))) / s.ggz[n + nij];
    else if (nf0 == 6)
        w[n] = (fzp + met<2>(s, 2, k) * (met<0>(s, 4, i) * (fxp - fxm) + met<1>(s, 4, j) * (fyp - fym))) / s.ggz[n];
    else if (nf0 == 3)
        v[n + ni] = (fym - met<1>(s, 2, j) * (met<0>(s, 4, i) * (fxp - fxm) + met<2>(s, 4, k) * (fzp - fzm))) / s.ggy[n + ni];
    else if (nf0 == 4)
        v[n] = (fyp + met<1>(s, 2, j) * (met<0>(s, 4, i) * (fxp - fxm) + met<2>(s, 4, k) * (fzp - fzm))) / s.ggy[n];
    else if (nf0 == 1)
        u[n + 1] = (fxm - met<0>(s, 2, i) * (met<1>(s, 4, j) * (fyp - fym) + met<2>(s, 4, k) * (fzp - fzm))) / s.ggx[n + 1];
    else
        u[n] = (fxp + met<0>(s, 2, i) * (met<1>(s, 4, j) * (fyp - fym) + met<2>(s, 4, k) * (fzp - fzm))) / s.ggx[n];
}

// Tangential wall velocities, vf_bwuwt.f:42-176 (types 1 and 4).
__global__ void k_bwuwt(Dims d, const int *__restrict__ indb, int lvo, const double *__restrict__ u,
                        const double *__restrict__ v, const double *__restrict__ w, double *bcu,
                        double *bcv, double *bcw)
{
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= d.numb) return;
    const int *e = indb + (long long)l * 4;
    const long long n = (long long)e[0] - 1;
    const int ns = e[1], ib = face_type(indb, lvo, l + 1);
    if (!(ib == 1 || ib == 4)) return;
    const long long ni = d.ni, nij = d.nij;
    switch (ns) {
    case 1: bcv[l] = (v[n] + v[n + ni]) * 0.5; bcw[l] = (w[n] + w[n + nij]) * 0.5; break;
    case 2: bcv[l] = (v[n - 1] + v[n - 1 + ni]) * 0.5; bcw[l] = (w[n - 1] + w[n - 1 + nij]) * 0.5; break;
    case 3: bcu[l] = (u[n] + u[n + 1]) * 0.5; bcw[l] = (w[n] + w[n + nij]) * 0.5; break;
    case 4: bcu[l] = (u[n - ni] + u[n - ni + 1]) * 0.5; bcw[l] = (w[n - ni] + w[n - ni + nij]) * 0.5; break;
    case 5: bcu[l] = (u[n] + u[n + 1]) * 0.5; bcv[l] = (v[n] + v[n + ni]) * 0.5; break;
    case 6: bcu[l] = (u[n - nij] + u[n - nij + 1]) * 0.5; bcv[l] = (v[n - nij] + v[n - nij + ni]) * 0.5; break;
    default: break;
    }
}

// Boundary pressure, vf_bwpp.f:28-47.
__global__ void k_bwpp(Dims d, const int *__restrict__ indb, const double *__restrict__ p, double *bcp)
{
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= d.numb) return;
    const int *e = indb + (long long)l * 4;
    const long long n = (long long)e[0] - 1;
    const long long ni = d.ni, nij = d.nij;
    switch (e[1]) {
    case 1: bcp[l] = p[n]; break;
    case 2: bcp[l] = p[n - 1]; break;
    case 3: bcp[l] = p[n]; break;
    case 4: bcp[l] = p[n - ni]; break;
    case 5: bcp[l] = p[n]; break;
    case 6: bcp[l] = p[n - nij]; break;
    default: break;
    }
}

}  // namespace

int cad_launch_vmodif(cad_ctx *c)
{
    Static s = cad_static(c);
    Fields &f = c->f;
    dim3 g = cell_grid(c->d), b = cell_block();
    hipLaunchKernelGGL(k_vmodif<0>, g, b, 0, c->stream, c->d, s, c->pt, f.nf, f.uu, f.bcu);
    hipLaunchKernelGGL(k_vmodif<1>, g, b, 0, c->stream, c->d, s, c->pt, f.nf, f.vv, f.bcv);
    hipLaunchKernelGGL(k_vmodif<2>, g, b, 0, c->stream, c->d, s, c->pt, f.nf, f.ww, f.bcw);
    hipLaunchKernelGGL(k_pmodif, g, b, 0, c->stream, c->d, c->pt, f.indc, f.pp, c->prm.rho0 / c->prm.dtnow);
    CAD_CHECK(hipGetLastError());
    return CAD_OK;
}

int cad_launch_bsuwt(cad_ctx *c)
{
    Static s = cad_static(c);
    Fields &f = c->f;
    dim3 g = cell_grid(c->d), b = cell_block();
    hipLaunchKernelGGL(k_bsuwt, g, b, 0, c->stream, c->d, s, f.nf, f.uu, f.vv, f.ww, c->prm.ibsuw0);
    CAD_CHECK(hipGetLastError());
    return CAD_OK;
}

int cad_launch_bsuwt2(cad_ctx *c)
{
    Static s = cad_static(c);
    hipLaunchKernelGGL(k_bsuwt2, cell_grid(c->d), cell_block(), 0, c->stream, c->d, s, c->f.nf,
                       c->f.uu, c->f.vv, c->f.ww);
    CAD_CHECK(hipGetLastError());
    return CAD_OK;
}

int cad_launch_bsuwn(cad_ctx *c)
{
    Static s = cad_static(c);
    hipLaunchKernelGGL(k_bsuwn, cell_grid(c->d), cell_block(), 0, c->stream, c->d, s, c->f.nf,
                       c->f.uu, c->f.vv, c->f.ww);
    CAD_CHECK(hipGetLastError());
    return CAD_OK;
}

int cad_launch_bwuwt(cad_ctx *c)
{
    if (c->d.numb <= 0) return CAD_OK;
    hipLaunchKernelGGL(k_bwuwt, dim3((c->d.numb + 255) / 256), dim3(256), 0, c->stream, c->d, c->indb,
                       cad_level_sides(c),
                       c->f.uu, c->f.vv, c->f.ww, c->f.bcu, c->f.bcv, c->f.bcw);
    CAD_CHECK(hipGetLastError());
    return CAD_OK;
}

int cad_launch_bwpp(cad_ctx *c)
{
    if (c->d.numb <= 0) return CAD_OK;
    hipLaunchKernelGGL(k_bwpp, dim3((c->d.numb + 255) / 256), dim3(256), 0, c->stream, c->d, c->indb,
                       c->f.pp, c->f.bcp);
    CAD_CHECK(hipGetLastError());
    return CAD_OK;
}
