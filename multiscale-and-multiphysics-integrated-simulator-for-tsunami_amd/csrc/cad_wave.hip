// cad_wave.hip -- the special sides of VF_BWUWN: the matrix-table
// wave-maker and the radiation (open) boundary (the theory wave-makers are
// in cad_wave_theory.hip).
//
// MODEL WAVE-BC <side> FUNC MATRIX with a LEVEL-ON table (<name>.mtb, read by
// VF_IM1INP, Source code/vf_im1inp.f) sets, at every call of VF_BWUWN
// (vf_bwuwn.f:257-480), the velocity of each boundary face of the side from
// the table.  The table's phase and the ramp are scalars of the call and are
// worked out on the host in the reference's precision (the phase in single
// precision, the ramp with the C library's sin, the time weights of VF_CWMTB1
// with its persistent search start MTBNOW).  The device then takes one thread
// per boundary column, which measures the column's water level (the
// F-weighted cell heights up to the first surface cell), scales the table's
// depth axis by (table level + depth) / (measured level + depth), and walks up
// the column's faces interpolating the normal and tangential velocity
// (VF_CWMTB2), in the reference's order of operations.  Faces of type 5
// outside a configured side are left as they are, as in the reference's
// boundary-list pass.
//
// MODEL OPEN-BC <side> FUNC TYPE1 (vf_bwuwn.f:483-620) relaxes the boundary
// values of the side's faces towards the inside velocities of the start of
// the step with the weights ALN = min(C*DTNOW/dx, 1) and ALT = min(2*ALN, 1),
// C the celerity of VF_COMAK0's wave length over the period (both weights 1
// in VF_CINIT's pass); faces next to a surface cell then copy the face below
// or above.  One thread per column does both passes in the reference's order.
#include "cad_common.h"

#include <cmath>
#include <cstdlib>
#include <cstring>

namespace {

// the reference's PI (VF_A0PRM.h): one ulp above M_PI
constexpr double PI_REF = 3.141592653589794;
constexpr double ZERO_REF = 1.0e-20;

struct WaveLaunch {
    int xside;           // faces normal to x (sides 1, 2) or to y (3, 4)
    int pos;             // face index of the side (2, or NUMI / NUMJ)
    int cpos;            // the cell column next to it (2, or NUMI-1 / NUMJ-1)
    int q0, nq;          // first column along the side and column count
    int nz;              // table depths
    int im;              // MTBNOW: the table's time interval for this call
    int off;             // LEVEL-OFF: profiles unscaled
    double d, wvz, aw, wmt1, wmt2, sa, sb, wvlvl;
};

__global__ void __launch_bounds__(64) k_wave_maker(Dims d, const double *__restrict__ zz,
                                                   const int *__restrict__ nf,
                                                   const double *__restrict__ ff,
                                                   const int *__restrict__ indf,
                                                   const double *__restrict__ tz,
                                                   const double *__restrict__ tun,
                                                   const double *__restrict__ tut, double *vel,
                                                   double *bcu, double *bcv, double *bcw,
                                                   WaveLaunch w, int *err)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= w.nq) return;
    const int q = w.q0 + t;
    // water level of the column next to the side (vf_bwuwn.f:300-330)
    double val = 0.0;
    for (int k = 2; k <= d.nk - 1; k++) {
        const long long c = w.xside ? cid(d, w.cpos, q, k) : cid(d, q, w.cpos, k);
        const int nw = nf[c];
        const double dz = zz[(long long)(k - 1) * 6 + 1];
        if (nw == -1) {
            val = val + dz * 1.0;
        } else if (nw == 0) {
            val = val + dz * ff[c];
        } else if (nw != 8) {
            val = val + dz * ff[c];
            break;
        }
    }
    val = val - (w.wvlvl - zz[6]);
    const double vws = w.off ? 1.0 : (w.wvz + w.d) / (val + w.d);
    // the column's faces bottom up (vf_bwuwn.f:331-380); VF_CWMTB2's depth
    // search starts where the face below found its interval
    const int nz = w.nz;
    const long long c0 = (long long)(w.im - 1) * nz, c1 = (long long)w.im * nz;
    int km = 1;
    for (int k = 2; k <= d.nk - 1; k++) {
        const long long f = w.xside ? cid(d, w.pos, q, k) : cid(d, q, w.pos, k);
        const int l = indf[f];
        if (l < 1) continue;
        double zc = (zz[(long long)(k - 1) * 6] + zz[(long long)k * 6]) * 0.5 - w.wvlvl;
        zc = vws * (zc + w.d) - w.d;
        double w1 = 0.0, w2 = 0.0;
        if (zc <= tz[0]) {
            km = 1;
            w1 = 1.0;
            w2 = 0.0;
        } else if (zc >= tz[nz - 1]) {
            km = nz - 1;
            w1 = 0.0;
            w2 = 1.0;
        } else {
            int kp = 0;
            for (int m = km; m <= nz - 1 && !kp; m++) {
                if (tz[m - 1] <= zc && zc <= tz[m]) {
                    kp = m;
                    w2 = 1.0 / (tz[m] - tz[m - 1]);
                    w1 = (tz[m] - zc) * w2;
                    w2 = (zc - tz[m - 1]) * w2;
                }
            }
            if (!kp) {
                // VF_CWMTB2: P.G ERROR. (no interval above the face below)
                err[0] = 1;
                return;
            }
            km = kp;
        }
        const long long a0 = km - 1, a1 = km;
        double p1 = w1 * tun[a0 + c0] + w2 * tun[a1 + c0];
        double p2 = w1 * tun[a0 + c1] + w2 * tun[a1 + c1];
        const double un = w.wmt1 * p1 + w.wmt2 * p2;
        p1 = w1 * tut[a0 + c0] + w2 * tut[a1 + c0];
        p2 = w1 * tut[a0 + c1] + w2 * tut[a1 + c1];
        const double ut = w.wmt1 * p1 + w.wmt2 * p2;
        const double u = un * vws * w.aw * w.sa;
        const double v = un * vws * w.aw * w.sb;
        bcu[l - 1] = u;
        bcv[l - 1] = v;
        bcw[l - 1] = ut * vws * w.aw;
        vel[f] = w.xside ? u : v;
    }
}

// VF_BWFFSF (vf_bwffsf.f:84-122): the cells next to a fixed-level side take F
// from the level wl: 0 above it, 1 below it, the filled fraction in the cell
// it cuts; obstacle cells are left.  One thread per (column, K).
__global__ void __launch_bounds__(64) k_bwffsf(Dims d, const double *__restrict__ zz,
                                               const int *__restrict__ nf, double *ff, int xside,
                                               int cpos, int q0, int nq, double wl)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nk = d.nk - 2;
    if (t >= nq * nk) return;
    const int q = q0 + t / nk, k = 2 + t % nk;
    const double z0 = zz[(long long)(k - 1) * 6], z1 = zz[(long long)k * 6];
    double val;
    if (z0 > wl)
        val = 0.0;
    else if (z1 < wl)
        val = 1.0;
    else
        val = (wl - z0) / (z1 - z0);
    const long long c = xside ? cid(d, cpos, q, k) : cid(d, q, cpos, k);
    if (nf[c] != -1) ff[c] = val;
}

struct OpenLaunch {
    int xside, pos, ic, i2, nq;
    double aln, alt;
};

__global__ void __launch_bounds__(64) k_open_side(Dims d, const int *__restrict__ nf,
                                                  const int *__restrict__ indf,
                                                  const double *__restrict__ ub,
                                                  const double *__restrict__ vb,
                                                  const double *__restrict__ wb, double *vel,
                                                  double *bcu, double *bcv, double *bcw,
                                                  OpenLaunch o)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= o.nq) return;
    const int q = 2 + t;
    const double aln = o.aln, alt = o.alt;
    // faces next to fluid cells, and the rest (vf_bwuwn.f:505-527)
    for (int k = 2; k <= d.nk - 1; k++) {
        const long long f = o.xside ? cid(d, o.pos, q, k) : cid(d, q, o.pos, k);
        const int l = indf[f];
        if (l < 1) continue;
        const long long cc = o.xside ? cid(d, o.ic, q, k) : cid(d, q, o.ic, k);
        if (nf[cc] == 0) {
            if (o.xside) {
                const double vc = (vb[cc] + vb[cc + d.ni]) * 0.5;
                const double wc = (wb[cc] + wb[cc + d.nij]) * 0.5;
                const double u = (1.0 - aln) * bcu[l - 1] + aln * ub[cid(d, o.i2, q, k)];
                bcu[l - 1] = u;
                bcv[l - 1] = (1.0 - alt) * bcv[l - 1] + alt * vc;
                bcw[l - 1] = (1.0 - alt) * bcw[l - 1] + alt * wc;
                vel[f] = u;
            } else {
                const double uc = (ub[cc] + ub[cc + 1]) * 0.5;
                const double wc = (wb[cc] + wb[cc + d.nij]) * 0.5;
                bcu[l - 1] = (1.0 - alt) * bcu[l - 1] + alt * uc;
                const double v = (1.0 - aln) * bcv[l - 1] + aln * vb[cid(d, q, o.i2, k)];
                bcv[l - 1] = v;
                bcw[l - 1] = (1.0 - alt) * bcw[l - 1] + alt * wc;
                vel[f] = v;
            }
        } else {
            bcu[l - 1] = 0.0;
            bcv[l - 1] = 0.0;
            bcw[l - 1] = 0.0;
            vel[f] = 0.0;
        }
    }
    // faces next to surface cells copy the face below (NF = 5) or above
    // (NF = 6); an x side then has U = 0 under water above, a y side V = BCV
    // (vf_bwuwn.f:528-552, :596-617)
    for (int k = 2; k <= d.nk - 1; k++) {
        const long long f = o.xside ? cid(d, o.pos, q, k) : cid(d, q, o.pos, k);
        const int l = indf[f];
        if (l < 1) continue;
        const int nw = nf[o.xside ? cid(d, o.ic, q, k) : cid(d, q, o.ic, k)];
        if (nw != 5 && nw != 6) continue;
        const int l2 = indf[nw == 5 ? f - d.nij : f + d.nij];
        if (l2 < 1) continue;
        const double a = bcu[l2 - 1], b = bcv[l2 - 1], cw = bcw[l2 - 1];
        bcu[l - 1] = a;
        bcv[l - 1] = b;
        bcw[l - 1] = cw;
        if (nw == 5)
            vel[f] = o.xside ? a : b;
        else
            vel[f] = o.xside ? 0.0 : b;
    }
}

// VF_COMAK0 (vf_comak0.f): the wave length of a small-amplitude wave of
// period t in depth dep, Newton's method on the dispersion relation (at most
// 10 iterations, stop at |dX/X| <= 1e-10).
double comak0(double dep, double t, double gg)
{
    const double dd = 4.0 * PI_REF * PI_REF * dep / gg / t / t;
    double x = dd >= 1.0 ? dd : std::sqrt(dd);
    for (int itr = 1; itr <= 10; itr++) {
        const double ex = std::exp(x), exi = 1.0 / ex;
        const double coth = (ex + exi) / (ex - exi);
        const double dx = (x - dd * coth) / (1.0 + dd * (coth * coth - 1.0));
        x = x - dx;
        if (std::fabs(dx / x) <= 1.0e-10) break;
    }
    return 2.0 * PI_REF * dep / x;
}

// VF_CWMTB1 (vf_cwmtb1.f; VF_CWMTB12 for the second table): the time weights
// of table T at phase wvt and the table level there; searches down from its
// MTBNOW, then down from the top.
int mtb_time(cad_ctx::WaveTable &T, double wvt, double *w1o, double *w2o, double *wvz)
{
    const int nt = T.nt;
    const double *tt = T.tt - 1;  // 1-based
    double w1 = 0.0, w2 = 0.0;
    if (T.mtbnow <= 0) T.mtbnow = nt - 1;
    if (wvt <= tt[1]) {
        T.mtbnow = 1;
        w1 = 1.0;
        w2 = 0.0;
    } else if (wvt >= tt[nt]) {
        T.mtbnow = nt - 1;
        w1 = 0.0;
        w2 = 1.0;
    } else {
        int ip = 0;
        for (int i = T.mtbnow; i >= 1 && !ip; i--)
            if (tt[i] <= wvt && wvt <= tt[i + 1]) {
                ip = i;
                w2 = 1.0 / (tt[i + 1] - tt[i]);
                w1 = (tt[i + 1] - wvt) * w2;
                w2 = (wvt - tt[i]) * w2;
            }
        for (int i = nt - 1; i >= T.mtbnow + 1 && !ip; i--)
            if (tt[i] <= wvt && wvt <= tt[i + 1]) {
                ip = i;
                w2 = 1.0 / (tt[i + 1] - tt[i]);
                w1 = (tt[i + 1] - wvt) * w2;
                w2 = (wvt - tt[i]) * w2;
            }
        if (!ip) {
            cad_set_error("VF_CWMTB1: P.G ERROR. (phase %.17g outside the table's times)", wvt);
            return CAD_ERR_WAVE;
        }
        T.mtbnow = ip;
    }
    *wvz = w1 * T.hh[T.mtbnow - 1] + w2 * T.hh[T.mtbnow];
    *w1o = w1;
    *w2o = w2;
    return CAD_OK;
}

// the table a wave-maker side reads (-3: the first, -4: the second)
inline int table_of(const cad_ctx *c, int q) { return c->wv_kind[q] == -4 ? 1 : 0; }

void table_free(cad_ctx *c, int which)
{
    cad_ctx::WaveTable &T = c->wt[which];
    if (T.tab) hipFree(T.tab);
    std::free(T.tt);
    std::free(T.hh);
    T = cad_ctx::WaveTable{};
    // the table's sides refer to it: a new table needs them set again
    for (int q = 0; q < 4; q++)
        if (c->wv_kind[q] == -3 - which) c->wv_on[q] = 0;
}

}  // namespace

void cad_wave_free(cad_ctx *c)
{
    table_free(c, 0);
    table_free(c, 1);
    if (c->wv_err) hipFree(c->wv_err);
    if (c->wv_err_host) hipHostFree(c->wv_err_host);
    c->wv_err = nullptr;
    c->wv_err_host = nullptr;
}

int cad_wave_table(cad_ctx *c, int which, int nz, int nt, const double *zz, const double *tt,
                   const double *hh, const double *un, const double *ut, double t0)
{
    if (which != 0 && which != 1) {
        cad_set_error("cad_set_wave_table: table 0 (MATRIX) or 1 (MATRIX2)");
        return CAD_ERR_ARG;
    }
    // nz = 0: a LEVEL-ONLY table (times and levels only, vf_im1inp.f:71-80)
    const bool only = nz == 0;
    if ((!only && nz < 2) || nt < 2 || !tt || (!only && (!zz || !un || !ut)) || (only && !hh)) {
        cad_set_error("cad_set_wave_table: need 2 or more depths (0 for LEVEL-ONLY, which needs the "
                      "levels) and 2 or more times, non-null arrays");
        return CAD_ERR_ARG;
    }
    for (int k = 1; k < nz; k++)
        if (!(zz[k] > zz[k - 1])) {
            cad_set_error("cad_set_wave_table: depths must increase (ZZ(%d) <= ZZ(%d))", k + 1, k);
            return CAD_ERR_ARG;
        }
    for (int i = 1; i < nt; i++)
        if (!(tt[i] > tt[i - 1])) {
            cad_set_error("cad_set_wave_table: times must increase (TT(%d) <= TT(%d))", i + 1, i);
            return CAD_ERR_ARG;
        }
    CAD_CHECK(hipSetDevice(c->device));
    CAD_CHECK(hipStreamSynchronize(c->stream));
    table_free(c, which);
    cad_ctx::WaveTable &T = c->wt[which];
    const size_t tab = (size_t)nz * (size_t)nt;
    if (!only) CAD_CHECK(hipMalloc((void **)&T.tab, (nz + 2 * tab) * sizeof(double)));
    if (!c->wv_err) {
        CAD_CHECK(hipMalloc((void **)&c->wv_err, 64));
        CAD_CHECK(hipHostMalloc((void **)&c->wv_err_host, 64, hipHostMallocDefault));
    }
    T.tt = (double *)std::malloc(nt * sizeof(double));
    T.hh = (double *)std::malloc(nt * sizeof(double));
    if (!T.tt || !T.hh) {
        cad_set_error("cad_set_wave_table: out of host memory");
        return CAD_ERR_HIP;
    }
    std::memcpy(T.tt, tt, nt * sizeof(double));
    // LEVEL-OFF (no level column): the level is 0 and the profiles are not
    // scaled (vf_im1inp.f:148-155, vf_bwuwn.f:326-329)
    T.off = hh == nullptr;
    if (hh)
        std::memcpy(T.hh, hh, nt * sizeof(double));
    else
        std::memset(T.hh, 0, nt * sizeof(double));
    T.only = only;
    if (!only) {
        CAD_CHECK(hipMemcpy(T.tab, zz, nz * sizeof(double), hipMemcpyHostToDevice));
        CAD_CHECK(hipMemcpy(T.tab + nz, un, tab * sizeof(double), hipMemcpyHostToDevice));
        CAD_CHECK(hipMemcpy(T.tab + nz + tab, ut, tab * sizeof(double), hipMemcpyHostToDevice));
    }
    CAD_CHECK(hipMemset(c->wv_err, 0, 64));
    c->wv_err_host[0] = 0;
    T.nz = nz;
    T.nt = nt;
    T.t0 = t0;
    T.mtbnow = 0;
    return CAD_OK;
}

int cad_open_side(cad_ctx *c, int side, double depth, double period)
{
    if (side < 1 || side > 4) {
        cad_set_error("cad_set_open: side %d (1..4 = X-, X+, Y-, Y+)", side);
        return CAD_ERR_ARG;
    }
    if (!(depth > 0.0) || !(period > 0.0)) {
        cad_set_error("cad_set_open: (OPEN-BC DEPTH / PERIOD) NOT FOUND. (vf_csetup.f:128-137)");
        return CAD_ERR_ARG;
    }
    if (c->wv_on[side - 1]) {
        cad_set_error("cad_set_open: side %d is a wave-maker (DEFINED DOUBLY.)", side);
        return CAD_ERR_ARG;
    }
    c->op_on[side - 1] = 1;
    c->op_d[side - 1] = depth;
    c->op_t[side - 1] = period;
    return CAD_OK;
}

int cad_wave_side(cad_ctx *c, int which, int side, int l1, int l2, double depth, double period,
                  double ampl, double wvlvl)
{
    if (which != 0 && which != 1) {
        cad_set_error("cad_set_wave: table 0 (MATRIX) or 1 (MATRIX2)");
        return CAD_ERR_ARG;
    }
    if (side < 1 || side > 4) {
        cad_set_error("cad_set_wave: side %d (1..4 = X-, X+, Y-, Y+)", side);
        return CAD_ERR_ARG;
    }
    if (c->op_on[side - 1]) {
        cad_set_error("cad_set_wave: side %d is a radiation side (DEFINED DOUBLY.)", side);
        return CAD_ERR_ARG;
    }
    if (!c->wt[which].tt) {
        cad_set_error("cad_set_wave: set the table first (cad_set_wave_table)");
        return CAD_ERR_STATE;
    }
    if (!(depth > 0.0) || !(period > 0.0)) {
        cad_set_error("cad_set_wave: depth and period must be positive (vf_iimdl.f)");
        return CAD_ERR_ARG;
    }
    const int lim = side <= 2 ? c->d.nj : c->d.ni;
    if (l1 < 2 || l2 > lim - 1 || l1 > l2) {
        cad_set_error("cad_set_wave: columns %d..%d outside 2..%d", l1, l2, lim - 1);
        return CAD_ERR_ARG;
    }
    const int q = side - 1;
    c->wv_on[q] = 1;
    c->wv_kind[q] = -3 - which;
    c->wv_l1[q] = l1;
    c->wv_l2[q] = l2;
    c->wv_d[q] = depth;
    c->wv_t[q] = period;
    c->wv_a[q] = ampl;
    c->wv_wvlvl = wvlvl;
    c->wv_level[q] = 0.0;
    return CAD_OK;
}

int cad_wave_apply(cad_ctx *c, bool initial)
{
    const Dims &d = c->d;
    bool any = false;
    for (int jd = 1; jd <= 4; jd++) {
        const int q = jd - 1;
        if (!c->wv_on[q]) continue;
        any = true;
        if (c->wv_kind[q] == -2 || c->wv_kind[q] == -1 || c->wv_kind[q] >= 1) {
            // a theory side: STOKES, CNOIDAL or STREAM n (cad_wave_theory.hip)
            const int rc = cad_wave_theory_apply(c, jd);
            if (rc) return rc;
            continue;
        }
        cad_ctx::WaveTable &T = c->wt[table_of(c, q)];
        const double dep = c->wv_d[q], t = c->wv_t[q], t0 = T.t0;
        double a = c->wv_a[q];
        // the table's phase in single precision (vf_bwuwn.f:270-276)
        float r4tn = (float)(t0 - c->prm.tnow / t);
        r4tn = r4tn - (float)(int)r4tn;
        if (r4tn < 0.0f) r4tn = r4tn + 1.0f;
        // ramp over the first AMPL periods (vf_bwuwn.f:277-283)
        double aw = 1.0;
        if (a >= ZERO_REF) {
            a = c->prm.tnow / t / a;
            if (a < 1.0) aw = 0.5 * std::sin(PI_REF * (a - 0.5)) + 0.5;
        }
        double wmt1, wmt2, wvz;
        const int rc = mtb_time(T, (double)r4tn, &wmt1, &wmt2, &wvz);
        if (rc) return rc;
        c->wv_level[q] = wvz * aw;
        // a fixed level (LEVEL-ONLY) sets no velocity here: its faces are
        // taken as free faces by the list pass (vf_bwuwn.f:326-327)
        if (T.only) continue;
        WaveLaunch w;
        w.xside = jd <= 2;
        w.pos = w.xside ? (jd == 1 ? 2 : d.ni) : (jd == 3 ? 2 : d.nj);
        w.cpos = w.xside ? (jd == 1 ? 2 : d.ni - 1) : (jd == 3 ? 2 : d.nj - 1);
        w.q0 = c->wv_l1[q];
        w.nq = c->wv_l2[q] - c->wv_l1[q] + 1;
        w.nz = T.nz;
        w.im = T.mtbnow;
        w.off = T.off;
        w.d = dep;
        w.wvz = wvz;
        w.aw = aw;
        w.wmt1 = wmt1;
        w.wmt2 = wmt2;
        w.sa = jd == 1 ? 1.0 : (jd == 2 ? -1.0 : 0.0);
        w.sb = jd == 3 ? 1.0 : (jd == 4 ? -1.0 : 0.0);
        w.wvlvl = c->wv_wvlvl;
        const size_t tab = (size_t)T.nz * (size_t)T.nt;
        const double *tz = T.tab, *tun = T.tab + T.nz, *tut = T.tab + T.nz + tab;
        hipLaunchKernelGGL(k_wave_maker, dim3((w.nq + 63) / 64), dim3(64), 0, c->stream, d, c->zz,
                           c->f.nf, c->f.ff, w.xside ? c->indx : c->indy, tz, tun, tut,
                           w.xside ? c->f.uu : c->f.vv, c->f.bcu, c->f.bcv, c->f.bcw, w, c->wv_err);
        CAD_CHECK(hipGetLastError());
    }
    if (any && c->wv_err)
        CAD_CHECK(hipMemcpyAsync(c->wv_err_host, c->wv_err, sizeof(int), hipMemcpyDeviceToHost,
                                 c->stream));
    // radiation sides (vf_bwuwn.f:483-620); the inside velocities of the
    // start of the step are the predictor's spare buffers (the pre-step
    // velocities after its swap), or in VF_CINIT's pass a copy taken here
    // (vf_cinit.f:376-384)
    if (!(c->op_on[0] || c->op_on[1] || c->op_on[2] || c->op_on[3])) return CAD_OK;
    const size_t bytes = (size_t)d.nc * sizeof(double);
    if (initial) {
        CAD_CHECK(hipMemcpyAsync(c->u2, c->f.uu, bytes, hipMemcpyDeviceToDevice, c->stream));
        CAD_CHECK(hipMemcpyAsync(c->v2, c->f.vv, bytes, hipMemcpyDeviceToDevice, c->stream));
        CAD_CHECK(hipMemcpyAsync(c->w2, c->f.ww, bytes, hipMemcpyDeviceToDevice, c->stream));
    }
    for (int jd = 1; jd <= 4; jd++) {
        const int q = jd - 1;
        if (!c->op_on[q]) continue;
        const double cel = comak0(c->op_d[q], c->op_t[q], c->prm.grz0) / c->op_t[q];
        OpenLaunch o;
        o.xside = jd <= 2;
        o.pos = o.xside ? (jd == 1 ? 2 : d.ni) : (jd == 3 ? 2 : d.nj);
        o.ic = o.xside ? (jd == 1 ? 2 : d.ni - 1) : (jd == 3 ? 2 : d.nj - 1);
        o.i2 = (jd == 1 || jd == 3) ? 3 : o.ic;
        o.nq = (o.xside ? d.nj : d.ni) - 2;
        o.aln = 1.0;
        o.alt = 1.0;
        if (!initial) {
            // XX(2,IC) or YY(2,IC): the width of the cell next to the side
            const double *m = o.xside ? c->xx_host : c->yy_host;
            o.aln = cel * c->prm.dtnow / m[(size_t)(o.ic - 1) * 6 + 1];
            if (o.aln > 1.0) o.aln = 1.0;
            o.alt = o.aln * 2.0;
            if (o.alt > 1.0) o.alt = 1.0;
        }
        if (o.nq <= 0) continue;
        hipLaunchKernelGGL(k_open_side, dim3((o.nq + 63) / 64), dim3(64), 0, c->stream, d, c->f.nf,
                           o.xside ? c->indx : c->indy, c->u2, c->v2, c->w2,
                           o.xside ? c->f.uu : c->f.vv, c->f.bcu, c->f.bcv, c->f.bcw, o);
        CAD_CHECK(hipGetLastError());
    }
    return CAD_OK;
}

int cad_level_sides(const cad_ctx *c)
{
    int m = 0;
    for (int q = 0; q < 4; q++)
        if (c->wv_on[q] && (c->wv_kind[q] == -3 || c->wv_kind[q] == -4) && c->wt[table_of(c, q)].only)
            m |= 1 << q;
    return m;
}

int cad_launch_bwffsf(cad_ctx *c)
{
    const int lvo = cad_level_sides(c);
    if (!lvo) return CAD_OK;
    const Dims &d = c->d;
    for (int jd = 1; jd <= 4; jd++) {
        const int q = jd - 1;
        if (!((lvo >> q) & 1)) continue;
        // the level at TNOW, ramped (vf_bwffsf.f:58-82)
        cad_ctx::WaveTable &T = c->wt[table_of(c, q)];
        const double t = c->wv_t[q], t0 = T.t0;
        double a = c->wv_a[q];
        float r4tn = (float)(t0 - c->prm.tnow / t);
        r4tn = r4tn - (float)(int)r4tn;
        if (r4tn < 0.0f) r4tn = r4tn + 1.0f;
        double aw = 1.0;
        if (a >= ZERO_REF) {
            a = c->prm.tnow / t / a;
            if (a < 1.0) aw = 0.5 * std::sin(PI_REF * (a - 0.5)) + 0.5;
        }
        double wmt1, wmt2, wvz;
        const int rc = mtb_time(T, (double)r4tn, &wmt1, &wmt2, &wvz);
        if (rc) return rc;
        c->wv_level[q] = wvz * aw;
        const double wl = c->wv_wvlvl + c->wv_level[q];
        const int xside = jd <= 2;
        const int cpos = xside ? (jd == 1 ? 2 : d.ni - 1) : (jd == 3 ? 2 : d.nj - 1);
        const int q0 = c->wv_l1[q], nq = c->wv_l2[q] - c->wv_l1[q] + 1;
        if (nq <= 0) continue;
        const int nk = d.nk - 2;
        hipLaunchKernelGGL(k_bwffsf, dim3((nq * nk + 63) / 64), dim3(64), 0, c->stream, d, c->zz,
                           c->f.nf, c->f.ff, xside, cpos, q0, nq, wl);
        CAD_CHECK(hipGetLastError());
    }
    return CAD_OK;
}

int cad_wave_check(cad_ctx *c)
{
    if (c->wv_err_host && c->wv_err_host[0]) {
        cad_set_error("VF_CWMTB2: P.G ERROR. (a face height has no table interval above the "
                      "face below it)");
        return CAD_ERR_WAVE;
    }
    return CAD_OK;
}
