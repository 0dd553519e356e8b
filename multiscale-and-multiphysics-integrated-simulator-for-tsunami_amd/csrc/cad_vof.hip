// cad_vof.hip -- phase (c): VOF F-function advection, surface-cell
// redistribution, clipping with the volume sums, and cell-type flagging
// (Source code/vf_f1cal.f:124-239).
//
// Order-dependent list loops of the reference are restated as gathers:
//  * VF_FMOD1 counts, per fluid / gas cell, the surface cells pointing at it
//    (vf_fmod1.f:46-86); here each cell counts its pointing neighbours.
//  * VF_FMOD1's overflow / underflow spill (vf_fmod1.f:152-191) adds into a
//    fluid or gas cell from every surface cell pointing at it, in the order of
//    the surface list, which is lexicographic (vf_fnfini.f:112-126); the gather
//    visits the six neighbours in increasing linear index, reproducing the
//    reference's summation order exactly.
//  * VF_FNFINI's surface list is not materialised: NF = 7 marks exactly the
//    listed cells at every stage of its iteration.
#include "cad_common.h"

#include <algorithm>
#include <utility>

namespace {

// fluid-side offset axis/sign of surface orientation c (1..6): c = 2*axis+1
// has fluid on the minus side, c = 2*axis+2 on the plus side
__device__ __forceinline__ long long fluid_off(const Dims &d, int c)
{
    const int a = (c - 1) / 2;
    return (c % 2 == 1) ? -stride(d, a) : stride(d, a);
}

// ------------------------------------------------------------------ FCONV
// Donor-acceptor flux through a face normal to A, vf_fconv.f:56-200: the
// flux, and NLIM(donor) = 1 where the flux is limited.  The face is at offset
// FO (0: the cell's lower face, 1: its upper one) of a cell whose F, NF and
// GGV along A at offset o (-2..+2) are fat(o), nat5(o), gat(o); q, gg, ind:
// the face's velocity, porosity and index; idxf: the face's 1-based index
// along A; inside: the face lies within the faces the reference updates
// (2..NUMI on X, 2..NUMI-1 across); nc: the cell's linear index.
template <int A, int FO, class FG, class NG, class GG>
__device__ __forceinline__ double fconv_face_v(const Dims &d, const Static &s, const double *__restrict__ bcf,
                                               int *nlim, double dtnow, double flower, double q, double gg, int ind,
                                               FG fat, NG nat5, GG gat, int idxf, bool inside, long long nc)
{
    if (!inside || ind == -1) return 0.0;
    double c = dtnow * gg * q;
    double ac = fabs(c);
    const bool pos = c >= 0.0;
    // c >= 0: acceptor FO, donor FO-1, donor's upwind FO-2; else FO-1, FO, FO+1
    const int lsn = pos ? 1 : -1;
    const int ild = pos ? idxf - 1 : idxf;
    const int ldo = pos ? FO - 1 : FO;
    const int nn = pos ? nat5(FO - 1) : nat5(FO);
    double fd, fdmw, v, fa;
    if (nn != -1) {
        fd = pos ? fat(FO - 1) : fat(FO);
        fdmw = 1.0;
        if ((pos ? nat5(FO - 2) : nat5(FO + 1)) != -1) fdmw = pos ? fat(FO - 2) : fat(FO + 1);
        v = met<A>(s, 2, ild) * (pos ? gat(FO - 1) : gat(FO));
    } else {
        fd = ind >= 1 ? bcf[ind - 1] : 0.0;  // a donor obstacle face is a boundary face
        fdmw = 1.0;
        v = met<A>(s, 2, ild);
    }
    if ((pos ? nat5(FO) : nat5(FO - 1)) != -1) fa = pos ? fat(FO) : fat(FO - 1);
    else fa = fd;
    double fad = fa;
    const bool tang = (nn == -1) || ((nn >= 1 && nn <= 6) && ((nn - 1) / 2 != A));
    if (tang) {
        if (fdmw != 0.0 && fa != 0.0) fad = fd;
    }
    double fdm = dmaxd(fdmw, fd);
    if (fdmw < flower && fa < flower) fdm = dmaxd(fdm, 0.1);
    double cfx = dmaxd((fdm - fad) * ac - (fdm - fd) * v, 0.0);
    double fvx = fad * ac + cfx;
    double fvm = fd * v;
    if (fvx > fvm) nlim[nc + ldo * stride(d, A)] = 1;
    return 0.0 - (double)lsn * dmind(fvx, fvm);
}

// One plane's tile of the staged VOF kernel: a block of VT_X x VT_Y cells
// reads F and NF to +-2 in I and J, GGV to +-1, the X faces' U, GGX, INDX at
// I and I+1 and the Y faces' V, GGY, INDY at J and J+1.  Tile positions
// outside the arrays (ghost rows' neighbours) hold NF = -1 and zeros; inside
// them, wrapped rows included, they hold what a direct read would.
constexpr int VT_X = 64, VT_Y = 4;
template <typename T, int W, int H>
struct VRegion {
    static constexpr int N = W * H, PER = (N + 255) / 256;
    T r[PER];
    // loads the region's cells of plane k at (ilo.., jlo..) into registers
    __device__ __forceinline__ void fetch(const Dims &d, const T *__restrict__ a, T fill, int ilo, int jlo, int k)
    {
        const long long base = (long long)(k - 1) * d.nij;   // the plane
#pragma unroll
        for (int p = 0; p < PER; p++) {
            const int e = (int)threadIdx.x + p * 256;
            const int ii = e % W, jj = e / W;
            const long long n = base + ((ilo + ii - 1) + d.ni * (jlo + jj - 1));
            r[p] = (e < N && n >= 0 && n < d.nc) ? a[n] : fill;
        }
    }
    __device__ __forceinline__ void store(T *lds) const
    {
#pragma unroll
        for (int p = 0; p < PER; p++) {
            const int e = (int)threadIdx.x + p * 256;
            if (e < N) lds[e] = r[p];
        }
    }
};

// VF_FCONV and VF_FEULER in one pass (vf_f1cal.f:124-160, vf_feuler.f:43-58;
// static porosity, LOOPS=1, no source): each cell evaluates the fluxes of its
// six faces -- the same expression on the same operands as the face's other
// cell, so both get its bits -- and writes its new F out of place (fo), so
// that every flux reads the old F.  The flux arrays never reach HBM.  A face
// with a limited flux sets NLIM of its donor from both sides (the same 1).
// A block of 64 x 4 columns walks a run of kl K planes: each plane's I / J
// neighbourhood is staged in LDS, the K neighbours of a thread's own column
// roll through its registers, so every input is read from memory about once.
__global__ void __launch_bounds__(256) k_fconv_feuler(Dims d, Static s, const double *__restrict__ u,
                                                      const double *__restrict__ v,
                                                      const double *__restrict__ w,
                                                      const double *__restrict__ f,
                                                      const double *__restrict__ bcf,
                                                      const int *__restrict__ nf, int *nlim,
                                                      double *__restrict__ fo, double dtnow,
                                                      double flower, double loops, int kl)
{
    constexpr int FW = VT_X + 4, GW = VT_X + 2, UW = VT_X + 1;
    __shared__ double sf[FW * (VT_Y + 4)], sg[GW * (VT_Y + 2)], su[UW * VT_Y], sgx[UW * VT_Y],
        sv[VT_X * (VT_Y + 1)], sgy[VT_X * (VT_Y + 1)];
    __shared__ int snf[FW * (VT_Y + 4)], six[UW * VT_Y], siy[VT_X * (VT_Y + 1)];
    const int i0 = (int)blockIdx.x * VT_X + 1, j0 = (int)blockIdx.y * VT_Y + 1;
    const int k0 = (int)blockIdx.z * kl + 1, k1 = min(d.nk, k0 + kl - 1);
    const int tx = (int)threadIdx.x % VT_X, ty = (int)threadIdx.x / VT_X;
    const int i = i0 + tx, j = j0 + ty;
    const bool cell = i <= d.ni && j <= d.nj;
    VRegion<double, FW, VT_Y + 4> rf;
    VRegion<int, FW, VT_Y + 4> rn;
    VRegion<double, GW, VT_Y + 2> rg;
    VRegion<double, UW, VT_Y> ru, rgx;
    VRegion<int, UW, VT_Y> rix;
    VRegion<double, VT_X, VT_Y + 1> rv, rgy;
    VRegion<int, VT_X, VT_Y + 1> riy;
    auto fetch = [&](int k) {
        rf.fetch(d, f, 0.0, i0 - 2, j0 - 2, k);
        rn.fetch(d, nf, -1, i0 - 2, j0 - 2, k);
        rg.fetch(d, s.ggv, 0.0, i0 - 1, j0 - 1, k);
        ru.fetch(d, u, 0.0, i0, j0, k);
        rgx.fetch(d, s.ggx, 0.0, i0, j0, k);
        rix.fetch(d, s.indx, -1, i0, j0, k);
        rv.fetch(d, v, 0.0, i0, j0, k);
        rgy.fetch(d, s.ggy, 0.0, i0, j0, k);
        riy.fetch(d, s.indy, -1, i0, j0, k);
    };
    // the own column along K: F, NF at k-2..k+2, GGV at k-1..k+1, W, GGZ, INDZ at k, k+1
    double cf[5], cg[3], cw[2], cgz[2];
    int cn[5], ciz[2];
    auto col = [&](int kk, auto &dst, auto *a, auto fill) {
        const long long n = cid(d, i, j, kk);
        dst = (cell && kk >= 1 && kk <= d.nk && n >= 0 && n < d.nc) ? a[n] : fill;
    };
#pragma unroll
    for (int o = 0; o < 5; o++) {
        col(k0 - 2 + o, cf[o], f, 0.0);
        col(k0 - 2 + o, cn[o], nf, -1);
    }
#pragma unroll
    for (int o = 0; o < 3; o++) col(k0 - 1 + o, cg[o], s.ggv, 0.0);
#pragma unroll
    for (int o = 0; o < 2; o++) {
        col(k0 + o, cw[o], w, 0.0);
        col(k0 + o, cgz[o], s.ggz, 0.0);
        col(k0 + o, ciz[o], s.indz, -1);
    }
    for (int k = k0; k <= k1; k++) {
        fetch(k);
        __syncthreads();   // the previous plane's tile is read
        rf.store(sf);
        rn.store(snf);
        rg.store(sg);
        ru.store(su);
        rgx.store(sgx);
        rix.store(six);
        rv.store(sv);
        rgy.store(sgy);
        riy.store(siy);
        __syncthreads();
        // the column's next values, in flight during this plane
        double nfv, ngv, nwv, ngz;
        int nnv, nizv;
        col(k + 3, nfv, f, 0.0);
        col(k + 3, nnv, nf, -1);
        col(k + 2, ngv, s.ggv, 0.0);
        col(k + 2, nwv, w, 0.0);
        col(k + 2, ngz, s.ggz, 0.0);
        col(k + 2, nizv, s.indz, -1);
        if (cell) {
            const long long n = cid(d, i, j, k);
            if (i < 2 || i > d.ni - 1 || j < 2 || j > d.nj - 1 || k < 2 || k > d.nk - 1 || cn[2] == -1) {
                fo[n] = cf[2];
            } else {
                // the cell in the F tile and in the GGV tile
                const int fc = (ty + 2) * FW + tx + 2, gc = (ty + 1) * GW + tx + 1;
                auto xf = [&](int o) { return sf[fc + o]; };
                auto xn = [&](int o) { return snf[fc + o]; };
                auto xg = [&](int o) { return sg[gc + o]; };
                auto yf = [&](int o) { return sf[fc + o * FW]; };
                auto yn = [&](int o) { return snf[fc + o * FW]; };
                auto yg = [&](int o) { return sg[gc + o * GW]; };
                auto zf = [&](int o) { return cf[o + 2]; };
                auto zn = [&](int o) { return cn[o + 2]; };
                auto zg = [&](int o) { return cg[o + 1]; };
                const int ux = ty * UW + tx, vy = ty * VT_X + tx;
                const int hi0 = d.ni, hi1 = d.nj - 1, hi2 = d.nk - 1;   // X faces run to ni
                const int hj0 = d.ni - 1, hj1 = d.nj, hj2 = d.nk - 1;   // Y faces to nj
                const int hk0 = d.ni - 1, hk1 = d.nj - 1, hk2 = d.nk;   // Z faces to nk
                auto in_box = [](int a, int b, int c, int ha, int hb, int hc) {
                    return !(a < 2 || b < 2 || c < 2 || a > ha || b > hb || c > hc);
                };
                const double fu0 = fconv_face_v<0, 0>(d, s, bcf, nlim, dtnow, flower, su[ux], sgx[ux], six[ux], xf,
                                                      xn, xg, i, in_box(i, j, k, hi0, hi1, hi2), n);
                const double fu1 = fconv_face_v<0, 1>(d, s, bcf, nlim, dtnow, flower, su[ux + 1], sgx[ux + 1],
                                                      six[ux + 1], xf, xn, xg, i + 1,
                                                      in_box(i + 1, j, k, hi0, hi1, hi2), n);
                const double fv0 = fconv_face_v<1, 0>(d, s, bcf, nlim, dtnow, flower, sv[vy], sgy[vy], siy[vy], yf,
                                                      yn, yg, j, in_box(i, j, k, hj0, hj1, hj2), n);
                const double fv1 = fconv_face_v<1, 1>(d, s, bcf, nlim, dtnow, flower, sv[vy + VT_X],
                                                      sgy[vy + VT_X], siy[vy + VT_X], yf, yn, yg, j + 1,
                                                      in_box(i, j + 1, k, hj0, hj1, hj2), n);
                const double fw0 = fconv_face_v<2, 0>(d, s, bcf, nlim, dtnow, flower, cw[0], cgz[0], ciz[0], zf, zn,
                                                      zg, k, in_box(i, j, k, hk0, hk1, hk2), n);
                const double fw1 = fconv_face_v<2, 1>(d, s, bcf, nlim, dtnow, flower, cw[1], cgz[1], ciz[1], zf, zn,
                                                      zg, k + 1, in_box(i, j, k + 1, hk0, hk1, hk2), n);
                double g0 = cg[1] - (cg[1] - cg[1]) * loops;
                double df = (met<0>(s, 4, i) * (fu1 - fu0) + met<1>(s, 4, j) * (fv1 - fv0)
                             + met<2>(s, 4, k) * (fw1 - fw0) + dtnow * 0.0);
                fo[n] = (g0 * cf[2] + df) / cg[1];
            }
        }
        // roll the column
#pragma unroll
        for (int o = 0; o < 4; o++) {
            cf[o] = cf[o + 1];
            cn[o] = cn[o + 1];
        }
        cf[4] = nfv;
        cn[4] = nnv;
        cg[0] = cg[1];
        cg[1] = cg[2];
        cg[2] = ngv;
        cw[0] = cw[1];
        cw[1] = nwv;
        cgz[0] = cgz[1];
        cgz[1] = ngz;
        ciz[0] = ciz[1];
        ciz[1] = nizv;
    }
}

__device__ __forceinline__ double cell_vol(const Static &s, int i, int j, int k, double g)
{
    return met<0>(s, 2, i) * met<1>(s, 2, j) * met<2>(s, 2, k) * g;
}

// decode a linear index into 1-based (i,j,k)
__device__ __forceinline__ void ijk_of(const Dims &d, long long n, int &i, int &j, int &k)
{
    k = (int)(n / d.nij) + 1;
    long long r = n - (long long)(k - 1) * d.nij;
    j = (int)(r / d.ni) + 1;
    i = (int)(r - (long long)(j - 1) * d.ni) + 1;
}

// FMOD1 step 1: surface cells pointing at each cell, vf_fmod1.f:46-86.
__global__ void __launch_bounds__(256) k_fmod_count(Dims d, const int *__restrict__ nf, int *nlim)
{
    int i, j, k;
    if (!cell_of_thread(d, i, j, k)) return;
    const long long n = cid(d, i, j, k);
    int cnt = 0;
    if (i >= 2 && i <= d.ni - 1 && j >= 2 && j <= d.nj - 1 && k >= 2 && k <= d.nk - 1) {
        const long long nb[6] = {n - d.nij, n - d.ni, n - 1, n + 1, n + d.ni, n + d.nij};
#pragma unroll
        for (int q = 0; q < 6; q++) {
            const int c = nf[nb[q]];
            if (c < 1 || c > 6) continue;
            const long long fo = fluid_off(d, c);
            // the neighbour's fluid cell (IFC) or gas cell (IEC) is this cell
            if (nb[q] + fo == n) cnt++;
            if (nb[q] - fo == n) cnt++;
        }
    }
    nlim[n] = cnt;
}

// FMOD1 step 2: redistribution among the fluid-surface-gas triple,
// vf_fmod1.f:85-142.  The targets with NLIM==1 belong to this cell alone.
__global__ void __launch_bounds__(256) k_fmod_redist(Dims d, Static s, const int *__restrict__ nf,
                                                     const int *__restrict__ nlim, double *f)
{
    int i, j, k;
    if (!cell_of_thread(d, i, j, k)) return;
    if (i < 2 || i > d.ni - 1 || j < 2 || j > d.nj - 1 || k < 2 || k > d.nk - 1) return;
    const long long n = cid(d, i, j, k);
    const int c = nf[n];
    if (c < 1 || c > 6) return;
    const long long fo = fluid_off(d, c);
    const long long ne = n - fo, nfc = n + fo;
    int ie, je, ke, i2, j2, k2;
    ijk_of(d, ne, ie, je, ke);
    ijk_of(d, nfc, i2, j2, k2);
    double ve = cell_vol(s, ie, je, ke, s.ggv[ne]);
    double vs = cell_vol(s, i, j, k, s.ggv[n]);
    double vf = cell_vol(s, i2, j2, k2, s.ggv[nfc]);
    double ss = vs * f[n];
    const bool le = nlim[ne] == 1, lf = nlim[nfc] == 1;
    if (le) ss = ss + ve * f[ne];
    if (lf) ss = ss + vf * f[nfc];
    double ds;
    if (lf) {
        ds = dmind(vf, ss);
        ss = ss - ds;
        f[nfc] = ds / vf;
    }
    ds = dmind(vs, ss);
    ss = ss - ds;
    f[n] = ds / vs;
    if (le) {
        ds = dmind(ve, ss);
        ss = ss - ds;
        f[ne] = ds / ve;
    }
    if (ss > 0.0) {
        if (le) f[ne] = f[ne] + ss / ve;
        else f[n] = f[n] + ss / vs;
    }
}

// FMOD1 step 3a: spill of surface over/underflow into the gas / fluid cell,
// gathered per receiving cell in surface-list order, vf_fmod1.f:152-188.
__global__ void __launch_bounds__(256) k_fmod_spill(Dims d, Static s, const int *__restrict__ nf,
                                                    double *f)
{
    int i, j, k;
    if (!cell_of_thread(d, i, j, k)) return;
    if (i < 2 || i > d.ni - 1 || j < 2 || j > d.nj - 1 || k < 2 || k > d.nk - 1) return;
    const long long n = cid(d, i, j, k);
    const int cx = nf[n];
    if (cx >= 1 && cx <= 6) return;
    const long long nb[6] = {n - d.nij, n - d.ni, n - 1, n + 1, n + d.ni, n + d.nij};
    double fx = f[n];
    bool touched = false;
    double vx = 0.0;
#pragma unroll
    for (int q = 0; q < 6; q++) {
        const int c = nf[nb[q]];
        if (c < 1 || c > 6) continue;
        const double fs = f[nb[q]];
        const long long fo = fluid_off(d, c);
        const bool gas_side = (nb[q] - fo == n), fluid_side = (nb[q] + fo == n);
        if (!(gas_side && fs > 1.0) && !(fluid_side && !(fs > 1.0) && fs < 0.0)) continue;
        if (!touched) {
            vx = cell_vol(s, i, j, k, s.ggv[n]);
            touched = true;
        }
        int is, js, ks;
        ijk_of(d, nb[q], is, js, ks);
        const double vs = cell_vol(s, is, js, ks, s.ggv[nb[q]]);
        if (gas_side && fs > 1.0) fx = fx + vs * (fs - 1.0) / vx;
        else fx = fx + vs * fs / vx;
    }
    if (touched) f[n] = fx;
}

// FMOD1 step 3b: clamp the surface cells themselves.
__global__ void __launch_bounds__(256) k_fmod_clamp(Dims d, const int *__restrict__ nf, double *f)
{
    int i, j, k;
    if (!cell_of_thread(d, i, j, k)) return;
    if (i < 2 || i > d.ni - 1 || j < 2 || j > d.nj - 1 || k < 2 || k > d.nk - 1) return;
    const long long n = cid(d, i, j, k);
    const int c = nf[n];
    if (c < 1 || c > 6) return;
    if (f[n] > 1.0) f[n] = 1.0;
    if (f[n] < 0.0) f[n] = 0.0;
}

// Compensated summation (Neumaier): s + c carries the running sum with the
// rounding error of every addition kept in c.
__device__ __forceinline__ void comp_add(double &s, double &c, double x)
{
    const double t = s + x;
    c += fabs(s) >= fabs(x) ? (s - t) + x : (x - t) + s;
    s = t;
}

// Clip F to {0,1} beyond FLOWER/FUPPER and partial sums of FSUM and FCUT,
// vf_fcut01.f:34-55.  The volume sums are compensated sums in FP64: each
// thread, each block's tree and the final pass over the blocks carry the
// rounding errors along, so FSUM and FCUT are accurate to a few ulps of the
// exact sums over any number of cells.  Partials: FSUM, FCUT, then their
// compensations, REDUCE_BLOCKS each.
__global__ void __launch_bounds__(REDUCE_THREADS) k_fcut(Dims d, Static s, const int *__restrict__ nf,
                                                         double *f, double flower, double fupper,
                                                         double *partial)
{
    __shared__ double sh[4][REDUCE_THREADS];
    double fsum = 0.0, fcut = 0.0, csum = 0.0, ccut = 0.0;
    for (long long n = (long long)blockIdx.x * REDUCE_THREADS + threadIdx.x; n < d.nc;
         n += (long long)REDUCE_BLOCKS * REDUCE_THREADS) {
        int i, j, k;
        ijk_of(d, n, i, j, k);
        if (i < 2 || i > d.ni - 1 || j < 2 || j > d.nj - 1 || k < 2 || k > d.nk - 1) continue;
        if (nf[n] == -1) continue;
        double v = cell_vol(s, i, j, k, s.ggv[n]);
        double fv = f[n];
        if (fv > fupper) {
            comp_add(fcut, ccut, (fv - 1.0) * v);
            fv = 1.0;
            f[n] = fv;
        } else if (fv < flower) {
            comp_add(fcut, ccut, (fv - 0.0) * v);
            fv = 0.0;
            f[n] = fv;
        }
        comp_add(fsum, csum, fv * v);
    }
    sh[0][threadIdx.x] = fsum;
    sh[1][threadIdx.x] = fcut;
    sh[2][threadIdx.x] = csum;
    sh[3][threadIdx.x] = ccut;
    __syncthreads();
    for (int off = REDUCE_THREADS / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
#pragma unroll
            for (int q = 0; q < 2; q++) {
                double a = sh[q][threadIdx.x], ca = sh[q + 2][threadIdx.x] + sh[q + 2][threadIdx.x + off];
                comp_add(a, ca, sh[q][threadIdx.x + off]);
                sh[q][threadIdx.x] = a;
                sh[q + 2][threadIdx.x] = ca;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < 4; q++) partial[q * REDUCE_BLOCKS + blockIdx.x] = sh[q][0];
    }
}

// FSUM and FCUT from k_fcut's partials, compensated, in block order
__global__ void k_fcut_final(const double *__restrict__ partial, double *out)
{
    if (threadIdx.x >= 2) return;
    const int q = threadIdx.x;
    double a = 0.0, c = 0.0;
    for (int b = 0; b < REDUCE_BLOCKS; b++) {
        comp_add(a, c, partial[q * REDUCE_BLOCKS + b]);
        c += partial[(q + 2) * REDUCE_BLOCKS + b];
    }
    out[q] = a + c;
}

// Boundary F, vf_bwff.f:33-87 (types FIX and FREE; wave-maker and radiation
// faces, types 5 and 7 on the sides 1-4, take F like FREE).
__global__ void k_bwff(Dims d, const int *__restrict__ indb, const double *__restrict__ f, double *bcf)
{
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= d.numb) return;
    const int *e = indb + (long long)l * 4;
    if (!(e[3] == 2 || ((e[3] == 5 || e[3] == 7) && e[1] <= 4))) return;
    const long long n = (long long)e[0] - 1;
    switch (e[1]) {
    case 1: bcf[l] = f[n]; break;
    case 2: bcf[l] = f[n - 1]; break;
    case 3: bcf[l] = f[n]; break;
    case 4: bcf[l] = f[n - d.ni]; break;
    case 5: bcf[l] = f[n]; break;
    case 6: bcf[l] = f[n - d.nij]; break;
    default: break;
    }
}

// Gas-side normal velocity of a surface cell from its tangential neighbours,
// vf_bsuwn3.f:37-335.
__global__ void __launch_bounds__(256) k_bsuwn3(Dims d, Static s, const int *__restrict__ nf, double *u,
                                                double *v, double *w, int ibsuw0)
{
    int i, j, k;
    if (!cell_of_thread(d, i, j, k)) return;
    if (i < 2 || i > d.ni - 1 || j < 2 || j > d.nj - 1 || k < 2 || k > d.nk - 1) return;
    const long long n = cid(d, i, j, k);
    const int nf0 = nf[n];
    if (nf0 < 1 || nf0 > 6) return;
    const int idx[3] = {i, j, k};
    const int dn = (nf0 - 1) / 2;
    const bool plus = (nf0 % 2 == 1);
    const long long sd = stride(d, dn);
    double *vel[3] = {u, v, w};
    double *q = vel[dn];
    const double d0 = metr(s, dn, 2, idx[dn]);
    double v0 = plus ? q[n] : q[n + sd];
    const long long l1 = plus ? n + sd : n - sd;
    const long long wface = plus ? n + sd : n;
    int is_[4] = {0, 0, 0, 0};
    double dd[4] = {0.0, 0.0, 0.0, 0.0}, vv[4] = {0.0, 0.0, 0.0, 0.0};
    int c = 0;
#pragma unroll
    for (int t = 0; t < 3; t++) {
        if (t == dn) continue;
        const long long st = stride(d, t);
#pragma unroll
        for (int pm = 0; pm < 2; pm++) {
            const long long ls = pm ? n + st : n - st;
            const long long ls1 = pm ? l1 + st : l1 - st;
            if (nf[ls] == 0 || nf[ls1] == 0) {
                if (nf[ls] != -1 && nf[ls1] != -1) {
                    is_[c] = 1;
                    dd[c] = pm ? metr(s, t, 3, idx[t] + 1) : metr(s, t, 3, idx[t]);
                    vv[c] = pm ? q[wface + st] : q[wface - st];
                }
            }
            c++;
        }
    }
    const int iss = is_[0] + is_[1] + is_[2] + is_[3];
    if (iss <= 0) {
        // the fluid neighbour is interior, so the cell two steps away exists
        if (plus) {
            if (ibsuw0 == 0 && idx[dn] >= 3 && nf[n - 2 * sd] == 0)
                v0 = v0 + (v0 - q[n - sd]) * metr(s, dn, 2, idx[dn]) / metr(s, dn, 2, idx[dn] - 1);
        } else {
            if (ibsuw0 == 0 && idx[dn] + 2 <= (dn == 0 ? d.ni : (dn == 1 ? d.nj : d.nk)) &&
                nf[n + 2 * sd] == 0)
                v0 = v0 + (v0 - q[n + 2 * sd]) * metr(s, dn, 2, idx[dn]) / metr(s, dn, 2, idx[dn] + 1);
        }
        q[wface] = v0;
    } else {
        const double ds = d0 + dd[0] + dd[1] + dd[2] + dd[3];
        q[wface] = ((ds - d0) * v0 + (ds - dd[0]) * vv[0] + (ds - dd[1]) * vv[1] + (ds - dd[2]) * vv[2]
                    + (ds - dd[3]) * vv[3]) / ds / (double)iss;
    }
}

// VF_BSUWEM count: surface cells for which this cell is a tangential
// neighbour, vf_bsuwem.f:36-61.
__global__ void __launch_bounds__(256) k_bsuwem_count(Dims d, const int *__restrict__ nf, int *nlim)
{
    int i, j, k;
    if (!cell_of_thread(d, i, j, k)) return;
    const long long n = cid(d, i, j, k);
    int cnt = 0;
    if (i >= 2 && i <= d.ni - 1 && j >= 2 && j <= d.nj - 1 && k >= 2 && k <= d.nk - 1) {
        const long long nb[6] = {n - d.nij, n - d.ni, n - 1, n + 1, n + d.ni, n + d.nij};
        const int ax[6] = {2, 1, 0, 0, 1, 2};
#pragma unroll
        for (int q = 0; q < 6; q++) {
            const int c = nf[nb[q]];
            if (c < 1 || c > 6) continue;
            // normal axis of the neighbour: z for 5/6, y for 3/4, x otherwise
            const int na = (c == 5 || c == 6) ? 2 : ((c == 3 || c == 4) ? 1 : 0);
            if (ax[q] != na) cnt++;
        }
    }
    nlim[n] = cnt;
}

// Velocities of gas cells enclosed by surface cells, vf_bsuwem.f:66-88.
__global__ void __launch_bounds__(256) k_bsuwem(Dims d, const int *__restrict__ nf,
                                                const int *__restrict__ nl, double *u, double *v,
                                                double *w)
{
    int i, j, k;
    if (!cell_of_thread(d, i, j, k)) return;
    if (i < 2 || i > d.ni - 1 || j < 2 || j > d.nj - 1 || k < 2 || k > d.nk - 1) return;
    const long long n = cid(d, i, j, k);
    if (!(nl[n] >= 2 && nf[n] == 8)) return;
    const long long ni = d.ni, nij = d.nij;
    if (nf[n - 1] == 8 && nl[n - 1] < 2) { if (nf[n + 1] != 8) u[n] = u[n + 1]; }
    if (nf[n + 1] == 8 && nl[n + 1] < 2) { if (nf[n - 1] != 8) u[n + 1] = u[n]; }
    if (nf[n - ni] == 8 && nl[n - ni] < 2) { if (nf[n + ni] != 8) v[n] = v[n + ni]; }
    if (nf[n + ni] == 8 && nl[n + ni] < 2) { if (nf[n - ni] != 8) v[n + ni] = v[n]; }
    if (nf[n - nij] == 8 && nl[n - nij] < 2) { if (nf[n + nij] != 8) w[n] = w[n + nij]; }
    if (nf[n + nij] == 8 && nl[n + nij] < 2) { if (nf[n - nij] != 8) w[n + nij] = w[n]; }
}

__global__ void k_icopy(long long nc, const int *__restrict__ a, int *__restrict__ b)
{
    for (long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x; n < nc;
         n += (long long)gridDim.x * blockDim.x)
        b[n] = a[n];
}

__global__ void k_dcopy(long long nc, const double *__restrict__ a, double *__restrict__ b)
{
    for (long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x; n < nc;
         n += (long long)gridDim.x * blockDim.x)
        b[n] = a[n];
}

__global__ void k_iset(int *p, int v) { *p = v; }

// --------------------------------------------------------------- FNFINI
// vf_fnfini.f:36-42: every non-obstacle cell of layers 2..NK-1 becomes fluid.
__global__ void __launch_bounds__(256) k_nf_reset(Dims d, int *nf)
{
    int i, j, k;
    if (!cell_of_thread(d, i, j, k)) return;
    if (k < 2 || k > d.nk - 1) return;
    const long long n = cid(d, i, j, k);
    if (nf[n] != -1) nf[n] = 0;
}

__device__ __forceinline__ double f_across(const double *f, const double *bcf, int ind, long long nb)
{
    return ind >= 1 ? bcf[ind - 1] : f[nb];
}

// vf_fnfini.f:45-102: empty cells with no full face neighbour become gas.
__global__ void __launch_bounds__(256) k_nf_gas(Dims d, Static s, const double *__restrict__ f,
                                                const double *__restrict__ bcf, int *nf,
                                                double flower, double fupper)
{
    int i, j, k;
    if (!cell_of_thread(d, i, j, k)) return;
    if (i < 2 || i > d.ni - 1 || j < 2 || j > d.nj - 1 || k < 2 || k > d.nk - 1) return;
    const long long n = cid(d, i, j, k);
    if (nf[n] == -1 || !(f[n] < flower)) return;
    const long long ni = d.ni, nij = d.nij;
    const int ix1 = s.indx[n], ix2 = s.indx[n + 1];
    const int iy1 = s.indy[n], iy2 = s.indy[n + ni];
    const int iz1 = s.indz[n], iz2 = s.indz[n + nij];
    const double fx1 = f_across(f, bcf, ix1, n - 1), fx2 = f_across(f, bcf, ix2, n + 1);
    const double fy1 = f_across(f, bcf, iy1, n - ni), fy2 = f_across(f, bcf, iy2, n + ni);
    const double fz1 = f_across(f, bcf, iz1, n - nij), fz2 = f_across(f, bcf, iz2, n + nij);
    if (!(fz1 > fupper && iz1 == 0) && !(fy1 > fupper && iy1 == 0) && !(fy2 > fupper && iy2 == 0) &&
        !(fx1 > fupper && ix1 == 0) && !(fx2 > fupper && ix2 == 0) && !(fz2 > fupper && iz2 == 0))
        nf[n] = 8;
}

// vf_fnfini.f:112-126: non-gas cells next to gas become surface (7).
__global__ void __launch_bounds__(256) k_nf_surface(Dims d, int *nf)
{
    int i, j, k;
    if (!cell_of_thread(d, i, j, k)) return;
    if (i < 2 || i > d.ni - 1 || j < 2 || j > d.nj - 1 || k < 2 || k > d.nk - 1) return;
    const long long n = cid(d, i, j, k);
    const int c = nf[n];
    if (c == -1 || c == 8) return;
    const long long ni = d.ni, nij = d.nij;
    if (nf[n + nij] == 8 || nf[n + ni] == 8 || nf[n - ni] == 8 || nf[n + 1] == 8 || nf[n - 1] == 8 ||
        nf[n - nij] == 8)
        nf[n] = 7;
}

// vf_fnfini.f:130-148: surface cells with no fluid neighbour become gas.
__global__ void __launch_bounds__(256) k_nf_nofluid(Dims d, int *nf)
{
    int i, j, k;
    if (!cell_of_thread(d, i, j, k)) return;
    if (i < 2 || i > d.ni - 1 || j < 2 || j > d.nj - 1 || k < 2 || k > d.nk - 1) return;
    const long long n = cid(d, i, j, k);
    if (nf[n] != 7) return;
    const long long ni = d.ni, nij = d.nij;
    if (nf[n + nij] != 0 && nf[n + ni] != 0 && nf[n - ni] != 0 && nf[n + 1] != 0 && nf[n - 1] != 0 &&
        nf[n - nij] != 0)
        nf[n] = 8;
}

// vf_fnfini.f:152-172: surface cells not sandwiched between fluid and gas
// along any axis are marked -7 and counted.
__global__ void __launch_bounds__(256) k_nf_sandwich(Dims d, int *nf, int *nch)
{
    int i, j, k;
    if (!cell_of_thread(d, i, j, k)) return;
    if (i < 2 || i > d.ni - 1 || j < 2 || j > d.nj - 1 || k < 2 || k > d.nk - 1) return;
    const long long n = cid(d, i, j, k);
    if (nf[n] != 7) return;
    const long long ni = d.ni, nij = d.nij;
    if ((nf[n - nij] != 0 || nf[n + nij] != 8) && (nf[n - 1] != 0 || nf[n + 1] != 8) &&
        (nf[n + 1] != 0 || nf[n - 1] != 8) && (nf[n - ni] != 0 || nf[n + ni] != 8) &&
        (nf[n + ni] != 0 || nf[n - ni] != 8) && (nf[n + nij] != 0 || nf[n - nij] != 8)) {
        nf[n] = -7;
        atomicAdd(nch, 1);
    }
}

// vf_fnfini.f:178-186
__global__ void __launch_bounds__(256) k_nf_fix(Dims d, int *nf)
{
    int i, j, k;
    if (!cell_of_thread(d, i, j, k)) return;
    const long long n = cid(d, i, j, k);
    if (nf[n] == -7) nf[n] = 8;
}

// Surface orientation, vf_fnfprv.f:34-223; counts surface cells.
// the oriented / unorientable surface counts go to ORIENT_SLOTS pairs of
// counters picked by the block, then one block sums them: atomics from every
// wavefront with a surface cell on one address serialise in L2
#define ORIENT_SLOTS 1024

__global__ void __launch_bounds__(256) k_nf_orient(Dims d, Static s, const double *__restrict__ f,
                                                   const double *__restrict__ bcf, int *nf, int *slots)
{
    int i, j, k;
    if (!cell_of_thread(d, i, j, k)) return;
    if (i < 2 || i > d.ni - 1 || j < 2 || j > d.nj - 1 || k < 2 || k > d.nk - 1) return;
    const long long n = cid(d, i, j, k);
    if (nf[n] != 7) return;
    const int *ind[3] = {s.indx, s.indy, s.indz};
    const int order[6] = {5, 6, 3, 4, 1, 2};
    double fm = -1.0;
    int code = 7;
#pragma unroll
    for (int o = 0; o < 6; o++) {
        const int cc = order[o];
        const int a = (cc - 1) / 2;
        const bool minus = (cc % 2 == 1);
        const long long st = stride(d, a);
        const long long nb = minus ? n - st : n + st;
        const long long op = minus ? n + st : n - st;
        if (!(nf[nb] == 0 && nf[op] == 8)) continue;
        double fs = 2.0 * f[nb];
#pragma unroll
        for (int t = 0; t < 3; t++) {
            if (t == a) continue;
            const long long tt = stride(d, t);
            fs = fs + f_across(f, bcf, ind[t][nb], nb - tt);
            fs = fs + f_across(f, bcf, ind[t][nb + tt], nb + tt);
        }
        if (fs > fm) {
            code = cc;
            fm = fs;
        }
    }
    // orientation written after all candidates: the tests read only ==0 / ==8
    nf[n] = code;
    // counts aggregated per wavefront, one atomic per wave into its block's slot
    const unsigned long long all = __ballot(1), bad = __ballot(code == 7);
    if ((int)__lane_id() == __builtin_ctzll(all)) {
        const int nb = __popcll(bad), ng = __popcll(all) - nb;
        const unsigned blk = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
        int *sl = slots + 2 * (int)(blk % ORIENT_SLOTS);
        if (ng) atomicAdd(sl, ng);
        if (nb) atomicAdd(sl + 1, nb);
    }
}

// the two counts of k_nf_orient's slots into out[0], out[1]
__global__ void __launch_bounds__(256) k_orient_sum(const int *__restrict__ slots, int *out)
{
    __shared__ int a[256], b[256];
    const int t = threadIdx.x;
    int x = 0, y = 0;
    for (int q = t; q < ORIENT_SLOTS; q += 256) {
        x += slots[2 * q];
        y += slots[2 * q + 1];
    }
    a[t] = x;
    b[t] = y;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (t < h) {
            a[t] += a[t + h];
            b[t] += b[t + h];
        }
        __syncthreads();
    }
    if (t == 0) {
        out[0] = a[0];
        out[1] = b[0];
    }
}

// vf_cindc.f:35-50
__global__ void __launch_bounds__(256) k_cindc(Dims d, const int *__restrict__ nf, int *indc)
{
    int i, j, k;
    if (!cell_of_thread(d, i, j, k)) return;
    if (i < 2 || i > d.ni - 1 || j < 2 || j > d.nj - 1 || k < 2 || k > d.nk - 1) return;
    const long long n = cid(d, i, j, k);
    const int v = nf[n];
    indc[n] = (v == 8 || v == -1) ? -1 : 0;
}

// Pressure of cells that turned gas -> fluid: F-weighted mean of the fluid
// pressures of the 27-neighbourhood, vf_bsppfl.f:40-62 (p0 is the copy).
__global__ void __launch_bounds__(256) k_bsppfl(Dims d, const int *__restrict__ nf,
                                                const int *__restrict__ nf0,
                                                const double *__restrict__ f,
                                                const double *__restrict__ p0, double *p)
{
    int i, j, k;
    if (!cell_of_thread(d, i, j, k)) return;
    if (i < 2 || i > d.ni - 1 || j < 2 || j > d.nj - 1 || k < 2 || k > d.nk - 1) return;
    const long long n = cid(d, i, j, k);
    if (!(nf[n] == 0 && nf0[n] == 8)) return;
    double sumf = 0.0, sump = 0.0;
    for (int kp = -1; kp <= 1; kp++)
        for (int jp = -1; jp <= 1; jp++)
            for (int ip = -1; ip <= 1; ip++) {
                const long long m = n + ip + (long long)jp * d.ni + (long long)kp * d.nij;
                if (nf[m] == 0) {
                    const double fv = f[m];
                    sumf = sumf + fv;
                    sump = sump + fv * p0[m];
                }
            }
    if (sumf < 1.0e-20) p[n] = 0.0;
    else p[n] = sump / sumf;
}

// Surface and gas pressures, vf_bspp.f:36-62 (no air-pressure model).
__global__ void __launch_bounds__(256) k_bspp(Dims d, Static s, const int *__restrict__ nf,
                                              const double *__restrict__ f, double *p)
{
    int i, j, k;
    if (!cell_of_thread(d, i, j, k)) return;
    if (i < 2 || i > d.ni - 1 || j < 2 || j > d.nj - 1 || k < 2 || k > d.nk - 1) return;
    const long long n = cid(d, i, j, k);
    const int v = nf[n];
    const double pps = 0.0;
    const long long ni = d.ni, nij = d.nij;
    double psi;
    switch (v) {
    case 8: p[n] = 0.0; break;
    case 5: psi = met<2>(s, 3, k) / (0.5 * met<2>(s, 2, k - 1) + f[n] * met<2>(s, 2, k)); p[n] = (1.0 - psi) * p[n - nij] + psi * pps; break;
    case 6: psi = met<2>(s, 3, k + 1) / (0.5 * met<2>(s, 2, k + 1) + f[n] * met<2>(s, 2, k)); p[n] = (1.0 - psi) * p[n + nij] + psi * pps; break;
    case 3: psi = met<1>(s, 3, j) / (0.5 * met<1>(s, 2, j - 1) + f[n] * met<1>(s, 2, j)); p[n] = (1.0 - psi) * p[n - ni] + psi * pps; break;
    case 4: psi = met<1>(s, 3, j + 1) / (0.5 * met<1>(s, 2, j + 1) + f[n] * met<1>(s, 2, j)); p[n] = (1.0 - psi) * p[n + ni] + psi * pps; break;
    case 1: psi = met<0>(s, 3, i) / (0.5 * met<0>(s, 2, i - 1) + f[n] * met<0>(s, 2, i)); p[n] = (1.0 - psi) * p[n - 1] + psi * pps; break;
    case 2: psi = met<0>(s, 3, i + 1) / (0.5 * met<0>(s, 2, i + 1) + f[n] * met<0>(s, 2, i)); p[n] = (1.0 - psi) * p[n + 1] + psi * pps; break;
    default: break;
    }
}

// Initial velocity of VF_CINIT (vf_cinit.f:216-258): UINI on faces with a
// fluid side that are not boundary faces.
template <int A>
__global__ void __launch_bounds__(256) k_init_vel(Dims d, Static s, const int *__restrict__ nf,
                                                  double *q, double val)
{
    int i, j, k;
    if (!cell_of_thread(d, i, j, k)) return;
    const long long n = cid(d, i, j, k);
    const int idx[3] = {i, j, k};
    double out = 0.0;
    if (idx[A] >= 3 && i >= 2 && j >= 2 && k >= 2 && i <= d.ni - 1 && j <= d.nj - 1 && k <= d.nk - 1) {
        const int *ind = A == 0 ? s.indx : (A == 1 ? s.indy : s.indz);
        if (ind[n] == 0 && nf[n - stride(d, A)] * nf[n] == 0) out = val;
    }
    q[n] = out;
}

// Initial pressure of VF_CINIT (vf_cinit.f:283-296): RHO0*GRZ0*(W-LEVEL - z_c)
// on active cells, zero elsewhere.
__global__ void __launch_bounds__(256) k_init_p(Dims d, Static s, const int *__restrict__ indc,
                                                double *p, double rho0, double grz0, double wvlvl)
{
    int i, j, k;
    if (!cell_of_thread(d, i, j, k)) return;
    const long long n = cid(d, i, j, k);
    double out = 0.0;
    if (i >= 2 && i <= d.ni - 1 && j >= 2 && j <= d.nj - 1 && k >= 2 && k <= d.nk - 1 && indc[n] != -1)
        out = rho0 * grz0 * (wvlvl - (met<2>(s, 1, k) + met<2>(s, 1, k + 1)) * 0.5);
    p[n] = out;
}

__global__ void k_ifill(long long nc, int *a, int v)
{
    for (long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x; n < nc;
         n += (long long)gridDim.x * blockDim.x)
        a[n] = v;
}

inline dim3 sgrid(long long nc)
{
    long long b = (nc + 255) / 256;
    if (b > 8192) b = 8192;
    return dim3((unsigned)b);
}

}  // namespace

#define RET_IF(x)              \
    do {                       \
        int rc_ = (x);         \
        if (rc_) return rc_;   \
    } while (0)

// VF_FNFINI (vf_fnfini.f:36-197) on the device: fluid / gas / surface flags.
int cad_fnfini(cad_ctx *c)
{
    Static s = cad_static(c);
    const Dims &d = c->d;
    Fields &f = c->f;
    dim3 g = cell_grid(d), b = cell_block();
    hipStream_t st = c->stream;
    hipLaunchKernelGGL(k_nf_reset, g, b, 0, st, d, f.nf);
    hipLaunchKernelGGL(k_nf_gas, g, b, 0, st, d, s, f.ff, f.bcf, f.nf, c->prm.flower, c->prm.fupper);
    // Every pass that finds a cell to fix turns at least one non-gas cell into
    // gas, and no pass turns gas back, so the loop ends within (interior
    // cells + 1) passes.  c->flag_pass_limit (0: that bound) caps it; a count
    // still positive after the cap is an error, not a silent exit.
    const long long bound = (long long)(d.ni - 2) * (d.nj - 2) * (d.nk - 2) + 1;
    const long long cap = c->flag_pass_limit > 0 ? c->flag_pass_limit : bound;
    int left = 0;
    for (long long it = 0; it < cap; it++) {
        hipLaunchKernelGGL(k_nf_surface, g, b, 0, st, d, f.nf);
        hipLaunchKernelGGL(k_nf_nofluid, g, b, 0, st, d, f.nf);
        hipLaunchKernelGGL(k_iset, dim3(1), dim3(1), 0, st, c->icount + 2, 0);
        hipLaunchKernelGGL(k_nf_sandwich, g, b, 0, st, d, f.nf, c->icount + 2);
        CAD_CHECK(hipGetLastError());
        CAD_CHECK(hipMemcpyAsync(c->host_cnt + 2, c->icount + 2, sizeof(int), hipMemcpyDeviceToHost, st));
        CAD_CHECK(hipStreamSynchronize(st));
        left = c->host_cnt[2];
        if (left <= 0) break;
        hipLaunchKernelGGL(k_nf_fix, g, b, 0, st, d, f.nf);
    }
    if (left > 0) {
        cad_set_error("VF_FNFINI: %d surface cells still not between fluid and gas after %lld "
                      "flag passes", left, cap);
        return CAD_ERR_STATE;
    }
    return CAD_OK;
}

// VF_FNFPRV + VF_CINDC (vf_fnfprv.f:34-223, vf_cindc.f:35-50); returns the
// surface count.
static int orient_cells(cad_ctx *c, int *nums)
{
    Static s = cad_static(c);
    const Dims &d = c->d;
    Fields &f = c->f;
    dim3 g = cell_grid(d), b = cell_block();
    hipStream_t st = c->stream;
    int *slots = c->icount + 16;   // 2 * ORIENT_SLOTS counters after the scalar counts
    hipLaunchKernelGGL(k_ifill, dim3(8), dim3(256), 0, st, (long long)(2 * ORIENT_SLOTS), slots, 0);
    hipLaunchKernelGGL(k_nf_orient, g, b, 0, st, d, s, f.ff, f.bcf, f.nf, slots);
    hipLaunchKernelGGL(k_orient_sum, dim3(1), dim3(256), 0, st, slots, c->icount + 3);
    hipLaunchKernelGGL(k_cindc, g, b, 0, st, d, f.nf, f.indc);
    CAD_CHECK(hipGetLastError());
    CAD_CHECK(hipMemcpyAsync(c->host_cnt + 3, c->icount + 3, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
    CAD_CHECK(hipStreamSynchronize(st));
    if (c->host_cnt[4] != 0) {
        cad_set_error("VF_FNFPRV: %d surface cells without an orientation", c->host_cnt[4]);
        return CAD_ERR_ORIENTATION;
    }
    *nums = c->host_cnt[3];
    return CAD_OK;
}

// VF_FNFINI + VF_FNFPRV + VF_CINDC
static int flag_cells(cad_ctx *c, int *nums)
{
    RET_IF(cad_fnfini(c));
    return orient_cells(c, nums);
}

// VF_F1CAL, vf_f1cal.f:124-239
int cad_vof_phase(cad_ctx *c, double *fsum, double *fcut)
{
    Static s = cad_static(c);
    const Dims &d = c->d;
    Fields &f = c->f;
    dim3 g = cell_grid(d), b = cell_block();
    hipStream_t st = c->stream;
    const double dt = c->prm.dtnow;
    // NLIM = 0 then FCONV (vf_f1cal.f:124-137)
    CAD_CHECK(hipMemsetAsync(f.nlim, 0, (size_t)d.nc * sizeof(int), st));
    // VF_FCONV and VF_FEULER (vf_f1cal.f:124-160), F out of place into the
    // spare array, which then becomes F
    // runs of K planes: about two resident blocks' worth of columns
    const long long cols = (long long)((d.ni + VT_X - 1) / VT_X) * ((d.nj + VT_Y - 1) / VT_Y);
    const int runs = (int)std::max(1LL, std::min<long long>(d.nk, (8LL * c->ncu + cols - 1) / cols));
    const int kl = (d.nk + runs - 1) / runs;
    const dim3 gv((unsigned)((d.ni + VT_X - 1) / VT_X), (unsigned)((d.nj + VT_Y - 1) / VT_Y),
                  (unsigned)((d.nk + kl - 1) / kl));
    hipLaunchKernelGGL(k_fconv_feuler, gv, dim3(256), 0, st, d, s, f.uu, f.vv, f.ww, f.ff, f.bcf, f.nf, f.nlim,
                       c->flf[0], dt, c->prm.flower, (double)c->prm.loops, kl);
    CAD_CHECK(hipGetLastError());
    std::swap(f.ff, c->flf[0]);
    // VF_BWFFSF: fixed-level sides (vf_f1cal.f:180)
    RET_IF(cad_launch_bwffsf(c));
    // VF_FMOD1 (vf_f1cal.f:185)
    hipLaunchKernelGGL(k_fmod_count, g, b, 0, st, d, f.nf, f.nlim);
    hipLaunchKernelGGL(k_fmod_redist, g, b, 0, st, d, s, f.nf, f.nlim, f.ff);
    hipLaunchKernelGGL(k_fmod_spill, g, b, 0, st, d, s, f.nf, f.ff);
    hipLaunchKernelGGL(k_fmod_clamp, g, b, 0, st, d, f.nf, f.ff);
    // VF_FCUT01 (vf_f1cal.f:188)
    hipLaunchKernelGGL(k_fcut, dim3(REDUCE_BLOCKS), dim3(REDUCE_THREADS), 0, st, d, s, f.nf, f.ff,
                       c->prm.flower, c->prm.fupper, c->red_partial);
    hipLaunchKernelGGL(k_fcut_final, dim3(1), dim3(64), 0, st, c->red_partial, c->red_out);
    CAD_CHECK(hipGetLastError());
    CAD_CHECK(hipMemcpyAsync(c->host_red, c->red_out, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    CAD_CHECK(hipStreamSynchronize(st));
    *fsum = c->host_red[0];
    *fcut = c->host_red[1];
    // VF_BWFF, VF_BSUWN3, VF_BSUWEM (vf_f1cal.f:192-194)
    if (d.numb > 0)
        hipLaunchKernelGGL(k_bwff, dim3((d.numb + 255) / 256), dim3(256), 0, st, d, c->indb, f.ff, f.bcf);
    hipLaunchKernelGGL(k_bsuwn3, g, b, 0, st, d, s, f.nf, f.uu, f.vv, f.ww, c->prm.ibsuw0);
    hipLaunchKernelGGL(k_bsuwem_count, g, b, 0, st, d, f.nf, f.nlim);
    hipLaunchKernelGGL(k_bsuwem, g, b, 0, st, d, f.nf, f.nlim, f.uu, f.vv, f.ww);
    // NLIM = old NF (vf_f1cal.f:197-203)
    hipLaunchKernelGGL(k_icopy, sgrid(d.nc), dim3(256), 0, st, d.nc, f.nf, f.nlim);
    CAD_CHECK(hipGetLastError());
    // VF_FNFINI, the TimerDoor corrections each followed by VF_FNFINI when
    // they moved water, then VF_FNFPRV and VF_CINDC (vf_f1cal.f:206-219)
    int nums = 0;
    RET_IF(cad_fnfini(c));
    RET_IF(cad_tdoor_phase(c));
    RET_IF(orient_cells(c, &nums));
    c->nums = nums;
    // VF_BSPPFL with the old flags in NLIM (vf_f1cal.f:230)
    hipLaunchKernelGGL(k_dcopy, sgrid(d.nc), dim3(256), 0, st, d.nc, f.pp, c->yv);
    hipLaunchKernelGGL(k_bsppfl, g, b, 0, st, d, f.nf, f.nlim, f.ff, c->yv, f.pp);
    CAD_CHECK(hipGetLastError());
    // boundary values (vf_f1cal.f:234-239)
    RET_IF(cad_launch_bsuwt(c));
    RET_IF(cad_launch_bsuwt2(c));
    hipLaunchKernelGGL(k_bsuwn3, g, b, 0, st, d, s, f.nf, f.uu, f.vv, f.ww, c->prm.ibsuw0);
    hipLaunchKernelGGL(k_bspp, g, b, 0, st, d, s, f.nf, f.ff, f.pp);
    CAD_CHECK(hipGetLastError());
    RET_IF(cad_launch_bwuwt(c));
    RET_IF(cad_launch_bwpp(c));
    return CAD_OK;
}

int cad_flag_cells(cad_ctx *c, int *nums) { return flag_cells(c, nums); }

// VF_BSUWN3 on its own (the boundary refresh of sub-loop passes 2..)
int cad_launch_bsuwn3(cad_ctx *c)
{
    hipLaunchKernelGGL(k_bsuwn3, cell_grid(c->d), cell_block(), 0, c->stream, c->d, cad_static(c), c->f.nf,
                       c->f.uu, c->f.vv, c->f.ww, c->prm.ibsuw0);
    CAD_CHECK(hipGetLastError());
    return CAD_OK;
}

// Device part of VF_CINIT (vf_cinit.f:190-409) after F holds its initial
// value: clip and volume sums, boundary F, flags, INDC, optional initial
// velocity / pressure, and the boundary values.
int cad_init_phase(cad_ctx *c, double wvlvl, const double *uvw, int flags, double *fsum, double *fcut)
{
    Static s = cad_static(c);
    const Dims &d = c->d;
    Fields &f = c->f;
    dim3 g = cell_grid(d), b = cell_block();
    hipStream_t st = c->stream;
    // VF_BWFFSF: fixed-level sides (vf_cinit.f:177)
    RET_IF(cad_launch_bwffsf(c));
    hipLaunchKernelGGL(k_fcut, dim3(REDUCE_BLOCKS), dim3(REDUCE_THREADS), 0, st, d, s, f.nf, f.ff,
                       c->prm.flower, c->prm.fupper, c->red_partial);
    hipLaunchKernelGGL(k_fcut_final, dim3(1), dim3(64), 0, st, c->red_partial, c->red_out);
    CAD_CHECK(hipGetLastError());
    CAD_CHECK(hipMemcpyAsync(c->host_red, c->red_out, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    CAD_CHECK(hipStreamSynchronize(st));
    *fsum = c->host_red[0];
    *fcut = c->host_red[1];
    if (d.numb > 0)
        hipLaunchKernelGGL(k_bwff, dim3((d.numb + 255) / 256), dim3(256), 0, st, d, c->indb, f.ff, f.bcf);
    hipLaunchKernelGGL(k_ifill, sgrid(d.nc), dim3(256), 0, st, d.nc, f.indc, -1);
    CAD_CHECK(hipGetLastError());
    int nums = 0;
    RET_IF(flag_cells(c, &nums));
    c->nums = nums;
    if (flags & 1) {
        hipLaunchKernelGGL(k_init_vel<0>, g, b, 0, st, d, s, f.nf, f.uu, uvw[0]);
        hipLaunchKernelGGL(k_init_vel<1>, g, b, 0, st, d, s, f.nf, f.vv, uvw[1]);
        hipLaunchKernelGGL(k_init_vel<2>, g, b, 0, st, d, s, f.nf, f.ww, uvw[2]);
    }
    if (flags & 2)
        hipLaunchKernelGGL(k_init_p, g, b, 0, st, d, s, f.indc, f.pp, c->prm.rho0, c->prm.grz0, wvlvl);
    CAD_CHECK(hipGetLastError());
    RET_IF(cad_launch_bwuwn(c, true));
    RET_IF(cad_launch_bsuwt(c));
    RET_IF(cad_launch_bsuwt2(c));
    hipLaunchKernelGGL(k_bsuwn3, g, b, 0, st, d, s, f.nf, f.uu, f.vv, f.ww, c->prm.ibsuw0);
    hipLaunchKernelGGL(k_bspp, g, b, 0, st, d, s, f.nf, f.ff, f.pp);
    CAD_CHECK(hipGetLastError());
    RET_IF(cad_launch_bwuwt(c));
    RET_IF(cad_launch_bwpp(c));
    CAD_CHECK(hipStreamSynchronize(st));
    return CAD_OK;
}
