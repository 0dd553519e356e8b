// cad_tdoor.hip -- the TimerDoor corrections of VF_F1CAL on the device:
// bubble rise (Source code/vf_fbubup.f) and droplet free fall / run-down
// (Source code/vf_fdropf.f), called between the flagging passes exactly where
// vf_f1cal.f:206-217 calls them.
//
// Both models move water along one axis at a time, in place, in a fixed
// sweep order, so a cell's outflow depends on what its downstream neighbour
// already received: the dependence runs along single grid lines and never
// across them.  The kernels therefore give each line to one thread:
//   - bubbles: one thread per (I,J) column walks K downwards (threads along I,
//     so every step is one coalesced row);
//   - droplets: one thread per line of the moving axis.  A line holds water
//     to move only if it holds a droplet (a gas-flagged cell with water in
//     it); those are rare, so a per-line flag written by the previous pass
//     lets every other thread exit at once.
// The weighted droplet values that the reference accumulates in seven work
// arrays (WKTX..WKFF) are carried in registers for the one cell still
// receiving water: a cell receives only from the cell visited after it, so
// it is final one step later and is written then.  The results are those of
// the reference's loop nest bit for bit (same expressions, same order of
// additions per cell).
//
// Droplet state invariant used to skip stores: every interior cell that holds
// no droplet has DROPTX = -1e10 and DROPTY..DROPWW = 0 after each pass (the
// reference clears them, vf_fdropf.f:299-306), and a droplet's DROPTX is a
// time >= 0, never -1e10.  So a cell without a droplet whose DROPTX reads
// -1e10 is already clear.  After an upload of droplet state the first pass
// clears unconditionally.
#include "cad_common.h"

#include <cmath>
#include <vector>

#define RET_IF(x)                     \
    do {                              \
        int r_ = (x);                 \
        if (r_ != CAD_OK) return r_;  \
    } while (0)

namespace {

constexpr double DROP_CLEAR_T = -1.0e10;

struct Drop {
    double *tx, *ty, *tz, *u, *v, *w;
};

__device__ __forceinline__ void drop_clear(const Drop &dp, long long n)
{
    dp.tx[n] = DROP_CLEAR_T;
    dp.ty[n] = 0.0;
    dp.tz[n] = 0.0;
    dp.u[n] = 0.0;
    dp.v[n] = 0.0;
    dp.w[n] = 0.0;
}

// one wave-aggregated atomic add of the lanes whose predicate holds
__device__ __forceinline__ void wave_count(bool pred, int *cnt)
{
    const unsigned long long m = __ballot(pred);
    if (m && (int)__lane_id() == __builtin_ctzll(m)) atomicAdd(cnt, __popcll(m));
}

// ---------------------------------------------------------------------------
// Bubble rise, vf_fbubup.f:35-51.  lvl[k] = 1 for the levels whose timer ran
// out this step (decided on the host from TBUB, TNOW and WBUB).
__global__ void __launch_bounds__(256) k_bubble(Dims d, Static s, const unsigned char *__restrict__ lvl,
                                                const int *__restrict__ nf, double *f, double fupper,
                                                int *moved)
{
    const int i = blockIdx.x * 64 + threadIdx.x + 2;
    const int j = blockIdx.y * 4 + threadIdx.y + 2;
    bool any = false;
    if (i <= d.ni - 1 && j <= d.nj - 1) {
        long long n1 = cid(d, i, j, d.nk - 1);
        double f1 = f[n1];          // FF(I,J,K+1) as the previous level left it
        for (int k = d.nk - 2; k >= 2; k--) {
            const long long n = n1 - d.nij;
            double f0 = f[n];
            if (lvl[k] && f0 < fupper && nf[n] == 0 && nf[n1] != -1) {
                const double v0 = met<2>(s, 2, k) * s.ggv[n];
                const double v1 = met<2>(s, 2, k + 1) * s.ggv[n1];
                const double fmove = dmind((1.0 - f0) * v0, f1 * v1);
                f1 = f1 - fmove / v1;
                f0 = f0 + fmove / v0;
                f[n1] = f1;
                f[n] = f0;
                any = true;
            }
            f1 = f0;
            n1 = n;
        }
    }
    wave_count(any, moved);
}

// ---------------------------------------------------------------------------
// Droplet start, vf_fdropf.f:66-307: new droplets take the mean velocity of
// their fluid (else surface) neighbours, old droplets resting on an obstacle
// slide down its slope, cells without a droplet are cleared.  Marks the
// x-lines (J,K) that hold a droplet and counts droplets.

// cell-centre velocity of cell c added to (up,vp,wp) (vf_fdropf.f:82-84);
// reads past the last plane (possible only for a ghost cell that is not
// flagged -1, which the set-up never produces) give 0
__device__ __forceinline__ void add_cell_vel(const Dims &d, const double *uu, const double *vv,
                                             const double *ww, long long c, double &up, double &vp,
                                             double &wp)
{
    const double u1 = c + 1 < d.nc ? uu[c + 1] : 0.0;
    const double v1 = c + d.ni < d.nc ? vv[c + d.ni] : 0.0;
    const double w1 = c + d.nij < d.nc ? ww[c + d.nij] : 0.0;
    up = up + 0.5 * (u1 + uu[c]);
    vp = vp + 0.5 * (v1 + vv[c]);
    wp = wp + 0.5 * (w1 + ww[c]);
}

template <bool FULL>
__global__ void __launch_bounds__(256) k_drop_start(Dims d, Static s, Drop dp, const double *__restrict__ f,
                                                    const int *__restrict__ nf, const double *__restrict__ uu,
                                                    const double *__restrict__ vv, const double *__restrict__ ww,
                                                    double tnow, double dtnow, double grz0, double flower,
                                                    int *xflag, int *cnt)
{
    int i, j, k;
    const bool in = cell_of_thread(d, i, j, k) && i >= 2 && i <= d.ni - 1 && j >= 2 && j <= d.nj - 1 &&
                    k >= 2 && k <= d.nk - 1;
    bool drop = false;
    if (in) {
        const long long n = cid(d, i, j, k);
        drop = f[n] > flower && nf[n] == 8;
        if (!drop) {
            if (FULL || dp.tx[n] != DROP_CLEAR_T) drop_clear(dp, n);
        } else {
            xflag[(j - 1) + (long long)d.nj * (k - 1)] = 1;
            const long long si = 1, sj = d.ni, sk = d.nij;
            if (dp.tx[n] < 0.0) {
                dp.tx[n] = tnow;
                dp.ty[n] = tnow;
                dp.tz[n] = tnow;
                int ip = 0;
                double up = 0.0, vp = 0.0, wp = 0.0;
                // edge neighbours in the reference's order (vf_fdropf.f:78-201):
                // (x,z) pairs, (x,y) pairs, (y,z) pairs; the two face cells
                // between are taken along the lower axis, then the higher
                const long long sa[3] = {si, si, sj}, sb[3] = {sk, sj, sk};
#pragma unroll
                for (int pr = 0; pr < 3; pr++)
#pragma unroll
                    for (int e = 0; e < 4; e++) {
                        const long long oa = (e & 1) ? sa[pr] : -sa[pr];
                        const long long ob = (e & 2) ? sb[pr] : -sb[pr];
                        const long long c = n + oa + ob;
                        if (nf[c] == 0 && (nf[n + oa] != -1 || nf[n + ob] != -1)) {
                            ip++;
                            add_cell_vel(d, uu, vv, ww, c, up, vp, wp);
                        }
                    }
                if (ip == 0) {
                    // face neighbours (vf_fdropf.f:203-246): x-, x+, z-, z+, y-, y+
                    const long long fo[6] = {-si, si, -sk, sk, -sj, sj};
#pragma unroll
                    for (int e = 0; e < 6; e++) {
                        const long long c = n + fo[e];
                        if (nf[c] != -1 && nf[c] != 8) {
                            ip++;
                            add_cell_vel(d, uu, vv, ww, c, up, vp, wp);
                        }
                    }
                }
                double du, dv, dw;
                if (ip >= 1) {
                    const double cp = 1.0 / (double)ip;
                    du = up * cp;
                    dv = vp * cp;
                    dw = wp * cp;
                } else {
                    du = 0.5 * (uu[n] + uu[n + 1]);
                    dv = 0.5 * (vv[n] + vv[n + sj]);
                    dw = 0.5 * (ww[n] + ww[n + sk]);
                }
                if (nf[n - sk] == -1) du = dv = dw = 0.0;
                dp.u[n] = du;
                dp.v[n] = dv;
                dp.w[n] = dw;
            } else if (nf[n - sk] == -1) {
                // vf_fdropf.f:265-296
                double h1 = nf[n - si] == -1 ? 1.0 : 1.0 - s.ggv[n - si];
                double h2 = nf[n + si] == -1 ? 1.0 : 1.0 - s.ggv[n + si];
                double hx = 0.5 * (h2 - h1) * met<2>(s, 2, k);
                h1 = nf[n - sj] == -1 ? 1.0 : 1.0 - s.ggv[n - sj];
                h2 = nf[n + sj] == -1 ? 1.0 : 1.0 - s.ggv[n + sj];
                double hy = 0.5 * (h2 - h1) * met<2>(s, 2, k);
                hx = hx / met<0>(s, 2, i);
                hy = hy / met<1>(s, 2, j);
                const double dn = 1.0 / (hx * hx + hy * hy + 1.0);
                hx = -grz0 * hx * dn;
                hy = -grz0 * hy * dn;
                dp.u[n] = dp.u[n] + dtnow * hx;
                dp.v[n] = dp.v[n] + dtnow * hy;
            }
        }
    }
    wave_count(drop, cnt);
}

// ---------------------------------------------------------------------------
// One pass of droplet moves along axis A towards +A (L = 1, lines walked
// downwards) or -A (L = 2, walked upwards), with the weighted means
// (vf_fdropf.f:317-398 x, :409-490 y, :501-601 z).  Thread = one line of
// axis A; the line's other two indices come from (blockIdx, threadIdx) with
// the lower axis fastest.  flag marks the lines to walk; next (pass L = 2
// only) receives the marks of the next axis's lines that hold a droplet.

struct Wk {
    double tx, ty, tz, u, v, w, ff;
};

// wk += q * (droplet values of cell n; axis A's time + dt, z velocity - dt*g)
template <int A>
__device__ __forceinline__ void wk_add(Wk &wk, double q, double tx, double ty, double tz, double u, double v,
                                       double w, double dt, double grz0)
{
    wk.tx = wk.tx + q * (A == 0 ? tx + dt : tx);
    wk.ty = wk.ty + q * (A == 1 ? ty + dt : ty);
    wk.tz = wk.tz + q * (A == 2 ? tz + dt : tz);
    wk.u = wk.u + q * u;
    wk.v = wk.v + q * v;
    wk.w = wk.w + q * (A == 2 ? w - dt * grz0 : w);
    wk.ff = wk.ff + q;
}

// Cell m (index pm along A) receives no more water: store its F if it
// changed and, for an interior droplet, the weighted means (vf_fdropf.f:
// 392-398; on the upward z pass a droplet under a ceiling that could not
// rise one cell stops, :589-597).  next marks the next axis's line through
// a droplet: the x pass marks y-lines (I,K), the y pass z-lines (I,J).
template <int A, int L>
__device__ __forceinline__ void settle(const Dims &d, const Static &s, const Drop &dp, double *f,
                                       const int *nf, int *next, const int *idx, long long m, int pm,
                                       double fm, double fm0, int nfm, const Wk &wm, double grz0,
                                       double flower)
{
    const int na = A == 0 ? d.ni : (A == 1 ? d.nj : d.nk);
    if (__double_as_longlong(fm) != __double_as_longlong(fm0)) f[m] = fm;
    if (pm < 2 || pm > na - 1 || !(fm > flower && nfm == 8)) return;
    const double wq = 1.0 / wm.ff;
    double mu = wm.u * wq, mv = wm.v * wq, mw = wm.w * wq;
    if (A == 2 && L == 1) {
        const double xz = sqrt(2.0 * grz0 * metr(s, 2, 2, pm));
        if (nf[m + d.nij] == -1 && mw > xz) mu = mv = mw = 0.0;
    }
    dp.tx[m] = wm.tx * wq;
    dp.ty[m] = wm.ty * wq;
    dp.tz[m] = wm.tz * wq;
    dp.u[m] = mu;
    dp.v[m] = mv;
    dp.w[m] = mw;
    if (next) {
        // x pass: y-line (I,K) with I = pm; y pass: z-line (I,J) with J = pm
        const int ii = A == 0 ? pm : idx[0];
        const int second = A == 0 ? idx[2] : pm;
        next[(ii - 1) + (long long)d.ni * (second - 1)] = 1;
    }
}

template <int A, int L>
__global__ void __launch_bounds__(256) k_drop_line(Dims d, Static s, Drop dp, double *f, const int *__restrict__ nf,
                                                   const int *__restrict__ flag, int *next, int *moved,
                                                   double tnow, double grz0, double flower)
{
    // the two other axes: B (fast, thread x) and C (slow); line index q
    constexpr int B = A == 0 ? 1 : 0;
    constexpr int C = A == 2 ? 1 : 2;
    const int nn[3] = {d.ni, d.nj, d.nk};
    const long long st[3] = {1, d.ni, d.nij};
    const int b = blockIdx.x * blockDim.x + threadIdx.x + 2;
    const int c = blockIdx.y + 2;
    bool any = false;
    if (b <= nn[B] - 1 && c <= nn[C] - 1 && flag[(b - 1) + (long long)nn[B] * (c - 1)]) {
        constexpr int dl = L == 1 ? -1 : 1;      // walk direction (IDL)
        const double sg = L == 1 ? 1.0 : -1.0;
        const int p0 = L == 1 ? nn[A] - 1 : 2;
        const int p1 = L == 1 ? 2 : nn[A] - 1;
        int idx[3];
        idx[B] = b;
        idx[C] = c;
        idx[A] = p0;
        const long long sa = st[A];
        long long n = cid(d, idx[0], idx[1], idx[2]);
        // the receiving cell m = I1 = I - IDL: first the cell beyond the walk
        long long m = n - dl * sa;
        int pm = p0 - dl;
        double fm = f[m], fm0 = fm;
        int nfm = nf[m];
        Wk wm{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        // the other two metric factors of this line
        const double hb = metr(s, B, 2, idx[B]), hc = metr(s, C, 2, idx[C]);
        for (int p = p0;; p += dl) {
            double fn = f[n];
            const double fn0 = fn;
            const int nfn = nf[n];
            const bool dr = fn > flower && nfn == 8;
            double tx = 0.0, ty = 0.0, tz = 0.0, u = 0.0, v = 0.0, w = 0.0;
            if (dr) {
                tx = dp.tx[n]; ty = dp.ty[n]; tz = dp.tz[n];
                u = dp.u[n]; v = dp.v[n]; w = dp.w[n];
            }
            const double hn = metr(s, A, 2, p);
            double dt = 0.0;
            if (dr && nfm != -1) {
                bool go = false;
                if (A == 2) {
                    const double xz = sqrt(2.0 * grz0 * hn);
                    if (L == 1) {
                        if (w > xz) dt = (w - sqrt(w * w - xz * xz)) / grz0;
                    } else if (w <= xz) {
                        if (w <= 0.0)
                            dt = (w + sqrt(w * w + xz * xz)) / grz0;
                        else
                            dt = 2.0 * w / grz0;
                    }
                    go = dt > 0.0 && tnow - tz > dt;
                } else {
                    const double vel = A == 0 ? u : v;
                    if (sg * vel > 0.0) {
                        dt = hn / fabs(vel);
                        go = tnow - (A == 0 ? tx : ty) > dt;
                    }
                }
                if (!go) dt = 0.0;
                if (go) {
                    // V0 = XX(2,I)*YY(2,J)*ZZ(2,K)*GGV, factors in x, y, z order
                    const double hm = metr(s, A, 2, pm);
                    double v0, v1;
                    if (A == 0) {
                        v0 = hn * hb * hc * s.ggv[n];
                        v1 = hm * hb * hc * s.ggv[m];
                    } else if (A == 1) {
                        v0 = hb * hn * hc * s.ggv[n];
                        v1 = hb * hm * hc * s.ggv[m];
                    } else {
                        v0 = hb * hc * hn * s.ggv[n];
                        v1 = hb * hc * hm * s.ggv[m];
                    }
                    const double fmove = dmind((1.0 - fm) * v1, fn * v0);
                    fn = fn - fmove / v0;
                    fm = fm + fmove / v1;
                    if (nfm == 8) wk_add<A>(wm, fmove, tx, ty, tz, u, v, w, dt, grz0);
                    any = true;
                }
            }
            // m receives nothing more: store its water and its means
            settle<A, L>(d, s, dp, f, nf, next, idx, m, pm, fm, fm0, nfm, wm, grz0, flower);
            // the remaining water of n (vf_fdropf.f:371-389)
            Wk wn{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
            if (fn > flower && nfn == 8) {
                double fold;
                if (A == 0) fold = fn * hn * hb * hc * s.ggv[n];
                else if (A == 1) fold = fn * hb * hn * hc * s.ggv[n];
                else fold = fn * hb * hc * hn * s.ggv[n];
                wk_add<A>(wn, fold, tx, ty, tz, u, v, w, dt, grz0);
            } else if (dr) {
                drop_clear(dp, n);
            }
            m = n;
            pm = p;
            fm = fn;
            fm0 = fn0;
            nfm = nfn;
            wm = wn;
            if (p == p1) break;
            n += dl * sa;
        }
        // the last cell of the walk
        settle<A, L>(d, s, dp, f, nf, next, idx, m, pm, fm, fm0, nfm, wm, grz0, flower);
    }
    wave_count(any, moved);
}

__global__ void k_iclear(int *p, long long n)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0;
}

}  // namespace

// Allocates the TimerDoor state on first use (a context that only solves
// the Poisson system never needs it) and clears it to its VF_CINIT value.
int cad_tdoor_alloc(cad_ctx *c)
{
    if (c->td_alloc) return CAD_OK;
    const size_t nc = (size_t)c->d.nc;
    double **dp[6] = {&c->drop[0], &c->drop[1], &c->drop[2], &c->drop[3], &c->drop[4], &c->drop[5]};
    for (double **p : dp)
        if (hipMalloc((void **)p, nc * sizeof(double)) != hipSuccess) {
            cad_set_error("hipMalloc of the droplet fields failed");
            return CAD_ERR_HIP;
        }
    const size_t nfl = (size_t)c->d.nj * c->d.nk + (size_t)c->d.ni * c->d.nk + (size_t)c->d.ni * c->d.nj;
    if (hipMalloc((void **)&c->td_flag, nfl * sizeof(int)) != hipSuccess ||
        hipMalloc((void **)&c->td_lvl, (size_t)c->d.nk + 1) != hipSuccess ||
        hipMalloc((void **)&c->td_cnt, 4 * sizeof(int)) != hipSuccess ||
        hipHostMalloc((void **)&c->td_host, 4 * sizeof(int) + (size_t)c->d.nk + 1, hipHostMallocDefault) != hipSuccess) {
        cad_set_error("allocation of the TimerDoor work arrays failed");
        return CAD_ERR_HIP;
    }
    c->tbub = new double[c->d.nk]();
    c->td_alloc = true;
    return cad_tdoor_init(c);
}

void cad_tdoor_free(cad_ctx *c)
{
    for (double *&p : c->drop)
        if (p) hipFree(p);
    if (c->td_flag) hipFree(c->td_flag);
    if (c->td_lvl) hipFree(c->td_lvl);
    if (c->td_cnt) hipFree(c->td_cnt);
    if (c->td_host) hipHostFree(c->td_host);
    delete[] c->tbub;
}

// VF_CINIT's TimerDoor state (vf_cinit.f:412-421): TBUB = 0, every droplet
// field cleared.
__global__ void k_drop_init(Drop dp, long long nc)
{
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n < nc) drop_clear(dp, n);
}

int cad_tdoor_init(cad_ctx *c)
{
    for (int k = 0; k < c->d.nk; k++) c->tbub[k] = 0.0;
    Drop dp{c->drop[0], c->drop[1], c->drop[2], c->drop[3], c->drop[4], c->drop[5]};
    hipLaunchKernelGGL(k_drop_init, dim3((unsigned)((c->d.nc + 255) / 256)), dim3(256), 0, c->stream, dp, c->d.nc);
    CAD_CHECK(hipGetLastError());
    c->td_full = false;
    return CAD_OK;
}

static int fetch_counts(cad_ctx *c, int n)
{
    CAD_CHECK(hipMemcpyAsync(c->td_host, c->td_cnt, n * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    CAD_CHECK(hipStreamSynchronize(c->stream));
    return CAD_OK;
}

// vf_f1cal.f:208-217, after the first VF_FNFINI of the step
int cad_tdoor_phase(cad_ctx *c)
{
    if (c->prm.wbub < 1.0e-20 && c->prm.idrop < 1) return CAD_OK;
    RET_IF(cad_tdoor_alloc(c));
    const Dims &d = c->d;
    const cad_params &p = c->prm;
    hipStream_t st = c->stream;
    Static s = cad_static(c);
    const dim3 cb(64, 4, 1);
    if (p.wbub >= 1.0e-20) {
        // bubble level timers on the host (vf_fbubup.f:37-38, 49); lvl[k]
        // for the 1-based level k
        unsigned char *lvl = reinterpret_cast<unsigned char *>(c->td_host + 4);
        bool fire = false;
        for (int k = 0; k <= d.nk; k++) lvl[k] = 0;
        for (int k = d.nk - 2; k >= 2; k--) {
            const double dtbub = c->zz_host[(size_t)(k - 1) * 6 + 1] / p.wbub;
            if (p.tnow - c->tbub[k - 1] > dtbub) {
                lvl[k] = 1;
                fire = true;
                c->tbub[k - 1] = c->tbub[k - 1] + dtbub;
            }
        }
        if (fire) {
            CAD_CHECK(hipMemcpyAsync(c->td_lvl, lvl, (size_t)d.nk + 1, hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(k_iclear, dim3(1), dim3(64), 0, st, c->td_cnt, 4LL);
            const dim3 g((unsigned)((d.ni - 2 + 63) / 64), (unsigned)((d.nj - 2 + 3) / 4), 1);
            hipLaunchKernelGGL(k_bubble, g, cb, 0, st, d, s, c->td_lvl, c->f.nf, c->f.ff, p.fupper, c->td_cnt);
            CAD_CHECK(hipGetLastError());
            RET_IF(fetch_counts(c, 1));
            if (c->td_host[0] > 0) RET_IF(cad_fnfini(c));
        }
    }
    if (p.idrop >= 1) {
        Drop dp{c->drop[0], c->drop[1], c->drop[2], c->drop[3], c->drop[4], c->drop[5]};
        int *xflag = c->td_flag;
        int *yflag = xflag + (size_t)d.nj * d.nk;
        int *zflag = yflag + (size_t)d.ni * d.nk;
        const long long nfl = (long long)d.nj * d.nk + (long long)d.ni * d.nk + (long long)d.ni * d.nj;
        hipLaunchKernelGGL(k_iclear, dim3((unsigned)((nfl + 255) / 256)), dim3(256), 0, st, c->td_flag, nfl);
        hipLaunchKernelGGL(k_iclear, dim3(1), dim3(64), 0, st, c->td_cnt, 4LL);
        const dim3 g = cell_grid(d);
        if (c->td_full)
            hipLaunchKernelGGL(k_drop_start<true>, g, cb, 0, st, d, s, dp, c->f.ff, c->f.nf, c->f.uu, c->f.vv,
                               c->f.ww, p.tnow, p.dtnow, p.grz0, p.flower, xflag, c->td_cnt);
        else
            hipLaunchKernelGGL(k_drop_start<false>, g, cb, 0, st, d, s, dp, c->f.ff, c->f.nf, c->f.uu, c->f.vv,
                               c->f.ww, p.tnow, p.dtnow, p.grz0, p.flower, xflag, c->td_cnt);
        CAD_CHECK(hipGetLastError());
        c->td_full = false;
        RET_IF(fetch_counts(c, 1));
        if (c->td_host[0] > 0) {
            int *mv = c->td_cnt + 1;
            const unsigned bx = 64;
            // x lines (J,K), y lines (I,K), z lines (I,J)
            const dim3 gx((unsigned)((d.nj - 2 + bx - 1) / bx), (unsigned)(d.nk - 2));
            const dim3 gy((unsigned)((d.ni - 2 + bx - 1) / bx), (unsigned)(d.nk - 2));
            const dim3 gz((unsigned)((d.ni - 2 + bx - 1) / bx), (unsigned)(d.nj - 2));
            hipLaunchKernelGGL((k_drop_line<0, 1>), gx, dim3(bx), 0, st, d, s, dp, c->f.ff, c->f.nf, xflag,
                               (int *)nullptr, mv, p.tnow, p.grz0, p.flower);
            hipLaunchKernelGGL((k_drop_line<0, 2>), gx, dim3(bx), 0, st, d, s, dp, c->f.ff, c->f.nf, xflag, yflag,
                               mv, p.tnow, p.grz0, p.flower);
            hipLaunchKernelGGL((k_drop_line<1, 1>), gy, dim3(bx), 0, st, d, s, dp, c->f.ff, c->f.nf, yflag,
                               (int *)nullptr, mv, p.tnow, p.grz0, p.flower);
            hipLaunchKernelGGL((k_drop_line<1, 2>), gy, dim3(bx), 0, st, d, s, dp, c->f.ff, c->f.nf, yflag, zflag,
                               mv, p.tnow, p.grz0, p.flower);
            hipLaunchKernelGGL((k_drop_line<2, 1>), gz, dim3(bx), 0, st, d, s, dp, c->f.ff, c->f.nf, zflag,
                               (int *)nullptr, mv, p.tnow, p.grz0, p.flower);
            hipLaunchKernelGGL((k_drop_line<2, 2>), gz, dim3(bx), 0, st, d, s, dp, c->f.ff, c->f.nf, zflag,
                               (int *)nullptr, mv, p.tnow, p.grz0, p.flower);
            CAD_CHECK(hipGetLastError());
            RET_IF(fetch_counts(c, 2));
            if (c->td_host[1] > 0) RET_IF(cad_fnfini(c));
        }
    }
    return CAD_OK;
}
