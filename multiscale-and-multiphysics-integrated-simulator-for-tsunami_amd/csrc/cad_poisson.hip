// cad_poisson.hip -- phase (b): the SMAC pressure-Poisson equation on gfx950.
//
// Assembly (VF_VPCOEF) is one per-cell kernel.  The solve is the reference's
// left-preconditioned MILU-BiCGSTAB (Source code/vf_m1bcgs.f:78-262).  The MILU
// factorisation (vf_mzdcmp.f:49-67) and the two triangular sweeps of M^-1
// (vf_mzfrwd.f:37-48, vf_mzbkwd.f:37-47) are recurrences on (i-1,j-1,k-1) and
// (i+1,j+1,k+1); all cells of one hyperplane i+j+k=h depend only on plane h-1
// (h+1 backwards), so processing planes in order performs exactly the
// reference's arithmetic on exactly the reference's operands: the results are
// bit-identical to the lexicographic loop.  Dot products are two-stage
// fixed-shape reductions, deterministic run to run (their rounding differs
// from the reference's serial sum only in summation order).
#include "cad_common.h"

#include <cmath>

namespace {

// ---------------------------------------------------------------- VPCOEF
// vf_vpcoef.f:76-437 (wave-source and air-pressure branches excluded; the
// host rejects decks that need them).
__global__ void __launch_bounds__(256) k_vpcoef(Dims d, Static s, const double *__restrict__ u,
                                                const double *__restrict__ v,
                                                const double *__restrict__ w,
                                                const double *__restrict__ p,
                                                const double *__restrict__ f,
                                                const int *__restrict__ nf,
                                                const int *__restrict__ indc, double *ad,
                                                double *ali, double *alj, double *alk, double *aui,
                                                double *auj, double *auk, double *bb,
                                                cad_params prm, int *err)
{
    int i, j, k;
    if (!cell_of_thread(d, i, j, k)) return;
    const long long n = cid(d, i, j, k);
    const long long ni = d.ni, nij = d.nij;
    double vad = 0.0, vli = 0.0, vlj = 0.0, vlk = 0.0, vui = 0.0, vuj = 0.0, vuk = 0.0, vbb = 0.0;
    const int nn = nf[n];
    if (indc[n] == -1) {
        vad = 1.0;
    } else if (nn == 0) {
        const double *ggv = s.ggv, *glv = s.glv;
        const double dti = 1.0 / prm.dtnow;
        double axy = met<0>(s, 2, i) * met<1>(s, 2, j);
        double axz = met<0>(s, 2, i) * met<2>(s, 2, k);
        double ayz = met<1>(s, 2, j) * met<2>(s, 2, k);
        double avl = met<0>(s, 2, i) * met<1>(s, 2, j) * met<2>(s, 2, k);
        double aim, aip, ajm, ajp, akm, akp, gv;
        int is;
        is = s.indx[n];
        if (is == 0) {
            gv = met<0>(s, 6, i) * (met<0>(s, 2, i) * ggv[n] / glv[n] + met<0>(s, 2, i - 1) * ggv[n - 1] / glv[n - 1]);
            aim = ayz * met<0>(s, 5, i) * s.ggx[n] * gv;
            vli = -aim;
        } else if (is >= 1) {
            aim = face_type(s.indb, s.lvo, is) == 4 ? 2.0 * ayz * met<0>(s, 4, i) * s.ggx[n] * ggv[n] / glv[n] : 0.0;
        } else { atomicExch(err, 1); aim = 0.0; }
        is = s.indx[n + 1];
        if (is == 0) {
            gv = met<0>(s, 6, i + 1) * (met<0>(s, 2, i + 1) * ggv[n + 1] / glv[n + 1] + met<0>(s, 2, i) * ggv[n] / glv[n]);
            aip = ayz * met<0>(s, 5, i + 1) * s.ggx[n + 1] * gv;
            vui = -aip;
        } else if (is >= 1) {
            aip = face_type(s.indb, s.lvo, is) == 4 ? 2.0 * ayz * met<0>(s, 4, i) * s.ggx[n + 1] * ggv[n] / glv[n] : 0.0;
        } else { atomicExch(err, 1); aip = 0.0; }
        is = s.indy[n];
        if (is == 0) {
            gv = met<1>(s, 6, j) * (met<1>(s, 2, j) * ggv[n] / glv[n] + met<1>(s, 2, j - 1) * ggv[n - ni] / glv[n - ni]);
            ajm = axz * met<1>(s, 5, j) * s.ggy[n] * gv;
            vlj = -ajm;
        } else if (is >= 1) {
            ajm = face_type(s.indb, s.lvo, is) == 4 ? 2.0 * axz * met<1>(s, 4, j) * s.ggy[n] * ggv[n] / glv[n] : 0.0;
        } else { atomicExch(err, 1); ajm = 0.0; }
        is = s.indy[n + ni];
        if (is == 0) {
            gv = met<1>(s, 6, j + 1) * (met<1>(s, 2, j + 1) * ggv[n + ni] / glv[n + ni] + met<1>(s, 2, j) * ggv[n] / glv[n]);
            ajp = axz * met<1>(s, 5, j + 1) * s.ggy[n + ni] * gv;
            vuj = -ajp;
        } else if (is >= 1) {
            ajp = face_type(s.indb, s.lvo, is) == 4 ? 2.0 * axz * met<1>(s, 4, j) * s.ggy[n + ni] * ggv[n] / glv[n] : 0.0;
        } else { atomicExch(err, 1); ajp = 0.0; }
        is = s.indz[n];
        if (is == 0) {
            gv = met<2>(s, 6, k) * (met<2>(s, 2, k) * ggv[n] / glv[n] + met<2>(s, 2, k - 1) * ggv[n - nij] / glv[n - nij]);
            akm = axy * met<2>(s, 5, k) * s.ggz[n] * gv;
            vlk = -akm;
        } else if (is >= 1) {
            akm = s.indb[(long long)(is - 1) * 4 + 2] != 4 ? 0.0 : 2.0 * axy * met<2>(s, 4, k) * s.ggz[n] * ggv[n] / glv[n];
        } else { atomicExch(err, 1); akm = 0.0; }
        is = s.indz[n + nij];
        if (is == 0) {
            gv = met<2>(s, 6, k + 1) * (met<2>(s, 2, k + 1) * ggv[n + nij] / glv[n + nij] + met<2>(s, 2, k) * ggv[n] / glv[n]);
            akp = axy * met<2>(s, 5, k + 1) * s.ggz[n + nij] * gv;
            vuk = -akp;
        } else if (is >= 1) {
            akp = s.indb[(long long)(is - 1) * 4 + 2] != 4 ? 0.0 : 2.0 * axy * met<2>(s, 4, k) * s.ggz[n + nij] * ggv[n] / glv[n];
        } else { atomicExch(err, 1); akp = 0.0; }
        vad = akm + ajm + aim + aip + ajp + akp;
        double vd = ayz * (s.ggx[n + 1] * u[n + 1] - s.ggx[n] * u[n])
                  + axz * (s.ggy[n + ni] * v[n + ni] - s.ggy[n] * v[n])
                  + axy * (s.ggz[n + nij] * w[n + nij] - s.ggz[n] * w[n]);
        double vq = -(ggv[n] - ggv[n]) * dti * avl;
        vbb = vd - vq;
    } else if (nn >= 1 && nn <= 6) {
        const double tr = prm.dtnow / prm.rho0;
        const double pps = 0.0;
        double psi, pf;
        vad = 1.0;
        switch (nn) {
        case 5: psi = met<2>(s, 3, k) / (0.5 * met<2>(s, 2, k - 1) + f[n] * met<2>(s, 2, k)); vlk = psi - 1.0; pf = p[n - nij]; break;
        case 6: psi = met<2>(s, 3, k + 1) / (0.5 * met<2>(s, 2, k + 1) + f[n] * met<2>(s, 2, k)); vuk = psi - 1.0; pf = p[n + nij]; break;
        case 3: psi = met<1>(s, 3, j) / (0.5 * met<1>(s, 2, j - 1) + f[n] * met<1>(s, 2, j)); vlj = psi - 1.0; pf = p[n - ni]; break;
        case 4: psi = met<1>(s, 3, j + 1) / (0.5 * met<1>(s, 2, j + 1) + f[n] * met<1>(s, 2, j)); vuj = psi - 1.0; pf = p[n + ni]; break;
        case 1: psi = met<0>(s, 3, i) / (0.5 * met<0>(s, 2, i - 1) + f[n] * met<0>(s, 2, i)); vli = psi - 1.0; pf = p[n - 1]; break;
        default: psi = met<0>(s, 3, i + 1) / (0.5 * met<0>(s, 2, i + 1) + f[n] * met<0>(s, 2, i)); vui = psi - 1.0; pf = p[n + 1]; break;
        }
        vbb = -tr * ((1.0 - psi) * pf + psi * pps - p[n]);
    } else {
        atomicExch(err, 2);
    }
    ad[n] = vad; ali[n] = vli; alj[n] = vlj; alk[n] = vlk;
    aui[n] = vui; auj[n] = vuj; auk[n] = vuk; bb[n] = vbb;
}

// ------------------------------------------------------- hyperplane sweeps
// Thread t of the launch for plane h handles (j,k) = (2 + t % (nj-2), k0 + t / (nj-2))
// and i = h - j - k; cells outside the interior exit.
struct Plane {
    int h, k0;
};

__device__ __forceinline__ bool plane_cell(const Dims &d, Plane pl, int &i, int &j, int &k)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nj2 = d.nj - 2;
    j = 2 + t % nj2;
    k = pl.k0 + t / nj2;
    i = pl.h - j - k;
    return k <= d.nk - 1 && i >= 2 && i <= d.ni - 1;
}

// MILU diagonal, vf_mzdcmp.f:53-64
__global__ void k_mzdcmp_plane(Dims d, Plane pl, const double *__restrict__ ad,
                               const double *__restrict__ ali, const double *__restrict__ alj,
                               const double *__restrict__ alk, const double *__restrict__ aui,
                               const double *__restrict__ auj, const double *__restrict__ auk,
                               const int *__restrict__ indc, double *dd, double u, int *npos)
{
    int i, j, k;
    if (!plane_cell(d, pl, i, j, k)) return;
    const long long n = cid(d, i, j, k);
    if (indc[n] == -1) return;
    const long long ni = d.ni, nij = d.nij;
    double au1 = aui[n - 1] + u * (auj[n - 1] + auk[n - 1]);
    double au2 = auj[n - ni] + u * (aui[n - ni] + auk[n - ni]);
    double au3 = auk[n - nij] + u * (aui[n - nij] + auj[n - nij]);
    double r = 1.0 / (ad[n] - ali[n] * dd[n - 1] * au1 - alj[n] * dd[n - ni] * au2 - alk[n] * dd[n - nij] * au3);
    dd[n] = r;
    if (r <= 0.0) atomicExch(npos, 0);
}

// (L D^-1 + I) x = y, vf_mzfrwd.f:37-48
__global__ void k_frwd_plane(Dims d, Plane pl, const double *__restrict__ dd,
                             const double *__restrict__ ali, const double *__restrict__ alj,
                             const double *__restrict__ alk, const int *__restrict__ indc,
                             double *x, const double *y)
{
    int i, j, k;
    if (!plane_cell(d, pl, i, j, k)) return;
    const long long n = cid(d, i, j, k);
    if (indc[n] == -1) return;
    const long long ni = d.ni, nij = d.nij;
    x[n] = y[n] - ali[n] * dd[n - 1] * x[n - 1] - alj[n] * dd[n - ni] * x[n - ni]
                - alk[n] * dd[n - nij] * x[n - nij];
}

// (D + U) x = x, vf_mzbkwd.f:37-47
__global__ void k_bkwd_plane(Dims d, Plane pl, const double *__restrict__ dd,
                             const double *__restrict__ aui, const double *__restrict__ auj,
                             const double *__restrict__ auk, const int *__restrict__ indc,
                             double *x)
{
    int i, j, k;
    if (!plane_cell(d, pl, i, j, k)) return;
    const long long n = cid(d, i, j, k);
    if (indc[n] == -1) return;
    const long long ni = d.ni, nij = d.nij;
    x[n] = (x[n] - aui[n] * x[n + 1] - auj[n] * x[n + ni] - auk[n] * x[n + nij]) * dd[n];
}

// ------------------------------------------------------- Krylov kernels
// The vector kernels of the solver run over the flat index range [n0, n1)
// that holds the active cells (found once per solve, k_active_range), in a
// fixed grid-stride order; cells outside it are inactive everywhere.  The
// BiCGSTAB scalars live in a small device array (slots SC_*), so the only
// value the host reads per iteration is (R,R) for the stop test.
enum { SC_S = 0, SC_W1, SC_ALPHA, SC_W1B, SC_W2, SC_OMEGA, SC_XN, SC_W1C, SC_BETA, SC_BN, SC_N,
       SC_TRR = 20, SC_TBB = 21 };
enum { RED_PLAIN = 0, RED_ALPHA, RED_OMEGA, RED_BETA };
constexpr int VEC_BLOCKS = 2048;

// [n0, n1) of the active cells; rng must hold {nc, 0} on entry
__global__ void __launch_bounds__(256) k_active_range(long long nc, const int *__restrict__ indc,
                                                      unsigned long long *rng)
{
    unsigned long long lo = ~0ull, hi = 0;
    for (long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x; n < nc;
         n += (long long)gridDim.x * blockDim.x) {
        if (indc[n] != -1) {
            lo = min(lo, (unsigned long long)n);
            hi = max(hi, (unsigned long long)n + 1);
        }
    }
    __shared__ unsigned long long slo[256], shi[256];
    slo[threadIdx.x] = lo;
    shi[threadIdx.x] = hi;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            slo[threadIdx.x] = min(slo[threadIdx.x], slo[threadIdx.x + w]);
            shi[threadIdx.x] = max(shi[threadIdx.x], shi[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0 && shi[0] > 0) {
        atomicMin(&rng[0], slo[0]);
        atomicMax(&rng[1], shi[0]);
    }
}

__global__ void k_range_init(unsigned long long *rng, long long nc)
{
    rng[0] = (unsigned long long)nc;
    rng[1] = 0;
}

#define ACTIVE_LOOP(n, rng, nblocks)                                                   \
    const long long n0_ = (long long)(rng)[0], n1_ = (long long)(rng)[1];           \
    for (long long n = n0_ + (long long)blockIdx.x * blockDim.x + threadIdx.x; n < n1_; \
         n += (long long)(nblocks) * blockDim.x)

// y = A x on active cells, vf_mzax.f:42-56 (active cells are interior, so all
// seven neighbours exist)
__global__ void __launch_bounds__(256) k_mzax(Dims d, const unsigned long long *rng,
                                              const double *__restrict__ ad,
                                              const double *__restrict__ ali,
                                              const double *__restrict__ alj,
                                              const double *__restrict__ alk,
                                              const double *__restrict__ aui,
                                              const double *__restrict__ auj,
                                              const double *__restrict__ auk,
                                              const int *__restrict__ indc,
                                              const double *__restrict__ x, double *__restrict__ y)
{
    const long long ni = d.ni, nij = d.nij;
    ACTIVE_LOOP(n, rng, VEC_BLOCKS) {
        if (indc[n] == -1) continue;
        y[n] = ad[n] * x[n] + ali[n] * x[n - 1] + alj[n] * x[n - ni] + alk[n] * x[n - nij]
             + aui[n] * x[n + 1] + auj[n] * x[n + ni] + auk[n] * x[n + nij];
    }
}

// Partial sums of up to three products over active cells in a fixed
// block-strided order; k_reduce_final adds the partials in block order.
template <int NPROD>
__device__ __forceinline__ void block_reduce_store(double (&acc)[3], double *partial)
{
    __shared__ double sh[3][REDUCE_THREADS];
#pragma unroll
    for (int q = 0; q < NPROD; q++) sh[q][threadIdx.x] = acc[q];
    __syncthreads();
    for (int off = REDUCE_THREADS / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
#pragma unroll
            for (int q = 0; q < NPROD; q++) sh[q][threadIdx.x] += sh[q][threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NPROD; q++) partial[q * REDUCE_BLOCKS + blockIdx.x] = sh[q][0];
    }
}

// (x, y) over active cells, vf_mzip.f:32-43
__global__ void __launch_bounds__(REDUCE_THREADS) k_dot(const unsigned long long *rng,
                                                        const int *__restrict__ mask,
                                                        const double *__restrict__ x,
                                                        const double *__restrict__ y,
                                                        double *partial)
{
    double acc[3] = {0.0, 0.0, 0.0};
    ACTIVE_LOOP(n, rng, REDUCE_BLOCKS) {
        if (mask[n] != -1) acc[0] += x[n] * y[n];
    }
    block_reduce_store<1>(acc, partial);
}

// omega numerator and denominator: (r, ae) and (ae, ae), vf_m1bcgs.f:199-203
__global__ void __launch_bounds__(REDUCE_THREADS) k_dot2(const unsigned long long *rng,
                                                         const int *__restrict__ mask,
                                                         const double *__restrict__ r,
                                                         const double *__restrict__ ae,
                                                         double *partial)
{
    double acc[3] = {0.0, 0.0, 0.0};
    ACTIVE_LOOP(n, rng, REDUCE_BLOCKS) {
        if (mask[n] != -1) {
            acc[0] += r[n] * ae[n];
            acc[1] += ae[n] * ae[n];
        }
    }
    block_reduce_store<2>(acc, partial);
}

// x += alpha p + omega r ; r -= omega ae ; partials of (r, r) and (r0, r),
// vf_m1bcgs.f:205-227
__global__ void __launch_bounds__(REDUCE_THREADS) k_update_xr(const unsigned long long *rng,
                                                              const int *__restrict__ mask,
                                                              double *__restrict__ x,
                                                              double *__restrict__ r,
                                                              const double *__restrict__ p,
                                                              const double *__restrict__ ae,
                                                              const double *__restrict__ r0,
                                                              const double *__restrict__ sc,
                                                              double *partial)
{
    const double alpha = sc[SC_ALPHA], omega = sc[SC_OMEGA];
    double acc[3] = {0.0, 0.0, 0.0};
    ACTIVE_LOOP(n, rng, REDUCE_BLOCKS) {
        if (mask[n] != -1) {
            x[n] = x[n] + (alpha * p[n] + omega * r[n]);
            double rn = r[n] - omega * ae[n];
            r[n] = rn;
            acc[0] += rn * rn;
            acc[1] += r0[n] * rn;
        }
    }
    block_reduce_store<2>(acc, partial);
}

// Adds the partials of nprod products in block order into sc[slot..], then
// forms the BiCGSTAB scalar the next kernel needs (vf_m1bcgs.f:179, 204, 229):
//   RED_ALPHA: alpha = s / w1          RED_OMEGA: omega = w1 / w2
//   RED_BETA:  beta = w1 alpha / (s omega), then s = w1
__global__ void __launch_bounds__(REDUCE_THREADS) k_reduce_final(const double *__restrict__ partial,
                                                                 int nprod, double *sc, int slot,
                                                                 int mode)
{
    __shared__ double sh[REDUCE_THREADS];
    for (int q = 0; q < nprod; q++) {
        double a = 0.0;
        for (int b = threadIdx.x; b < REDUCE_BLOCKS; b += REDUCE_THREADS) a += partial[q * REDUCE_BLOCKS + b];
        sh[threadIdx.x] = a;
        __syncthreads();
        for (int off = REDUCE_THREADS / 2; off > 0; off >>= 1) {
            if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
            __syncthreads();
        }
        if (threadIdx.x == 0) sc[slot + q] = sh[0];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (mode == RED_ALPHA) {
            sc[SC_ALPHA] = sc[SC_S] / sc[SC_W1];
        } else if (mode == RED_OMEGA) {
            sc[SC_OMEGA] = sc[SC_W1B] / sc[SC_W2];
        } else if (mode == RED_BETA) {
            sc[SC_BETA] = sc[SC_W1C] * sc[SC_ALPHA] / (sc[SC_S] * sc[SC_OMEGA]);
            sc[SC_S] = sc[SC_W1C];
        }
    }
}

// ------------------------------------------------------------- vector ops
__global__ void k_fill(long long nc, double *a, double v)
{
    for (long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x; n < nc;
         n += (long long)gridDim.x * blockDim.x)
        a[n] = v;
}

__global__ void k_copy(long long nc, const double *__restrict__ a, double *__restrict__ b)
{
    for (long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x; n < nc;
         n += (long long)gridDim.x * blockDim.x)
        b[n] = a[n];
}

// y = b - y (active), vf_m1bcgs.f:122-130
__global__ void __launch_bounds__(256) k_bminus(const unsigned long long *rng, const int *__restrict__ mask,
                                                const double *__restrict__ b, double *y)
{
    ACTIVE_LOOP(n, rng, VEC_BLOCKS) {
        if (mask[n] != -1) y[n] = b[n] - y[n];
    }
}

// rr = r0, pp = r0 (active), vf_m1bcgs.f:147-155
__global__ void __launch_bounds__(256) k_init_rp(const unsigned long long *rng, const int *__restrict__ mask,
                                                 const double *__restrict__ r0, double *rr, double *pp)
{
    ACTIVE_LOOP(n, rng, VEC_BLOCKS) {
        if (mask[n] != -1) {
            rr[n] = r0[n];
            pp[n] = r0[n];
        }
    }
}

// r -= alpha ap (active), vf_m1bcgs.f:181-189
__global__ void __launch_bounds__(256) k_axpy_neg(const unsigned long long *rng, const int *__restrict__ mask,
                                                  const double *__restrict__ sc,
                                                  const double *__restrict__ ap, double *r)
{
    const double alpha = sc[SC_ALPHA];
    ACTIVE_LOOP(n, rng, VEC_BLOCKS) {
        if (mask[n] != -1) r[n] = r[n] - alpha * ap[n];
    }
}

// p = r + beta (p - omega ap) (active), vf_m1bcgs.f:230-238
__global__ void __launch_bounds__(256) k_update_p(const unsigned long long *rng, const int *__restrict__ mask,
                                                  const double *__restrict__ sc,
                                                  const double *__restrict__ r,
                                                  const double *__restrict__ ap, double *p)
{
    const double beta = sc[SC_BETA], omega = sc[SC_OMEGA];
    ACTIVE_LOOP(n, rng, VEC_BLOCKS) {
        if (mask[n] != -1) p[n] = r[n] + beta * (p[n] - omega * ap[n]);
    }
}

// x0 = xw (active), vf_m1bcgs.f:248-257
__global__ void __launch_bounds__(256) k_copy_active(const unsigned long long *rng, const int *__restrict__ mask,
                                                     const double *__restrict__ a, double *b)
{
    ACTIVE_LOOP(n, rng, VEC_BLOCKS) {
        if (mask[n] != -1) b[n] = a[n];
    }
}

__global__ void k_iset1(int *p, int v) { *p = v; }

inline dim3 stream_grid(long long nc)
{
    long long b = (nc + 255) / 256;
    if (b > 8192) b = 8192;
    return dim3((unsigned)b);
}

}  // namespace

// ------------------------------------------------------------ host drivers
int cad_launch_vpcoef(cad_ctx *c)
{
    Static s = cad_static(c);
    Fields &f = c->f;
    hipLaunchKernelGGL(k_iset1, dim3(1), dim3(1), 0, c->stream, c->icount, 0);
    hipLaunchKernelGGL(k_vpcoef, cell_grid(c->d), cell_block(), 0, c->stream, c->d, s, f.uu, f.vv,
                       f.ww, f.pp, f.ff, f.nf, f.indc, c->ad, c->ali, c->alj, c->alk, c->aui,
                       c->auj, c->auk, c->bb, c->prm, c->icount);
    CAD_CHECK(hipGetLastError());
    CAD_CHECK(hipMemcpyAsync(c->host_cnt, c->icount, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    CAD_CHECK(hipStreamSynchronize(c->stream));
    if (c->host_cnt[0] != 0) {
        cad_set_error("VF_VPCOEF: invalid face index or cell flag (code %d)", c->host_cnt[0]);
        return CAD_ERR_STATE;
    }
    return CAD_OK;
}

static int hyper_launch_range(const Dims &d, int h, Plane &pl, unsigned &blocks)
{
    // interior i in [2,ni-1], j in [2,nj-1], k in [2,nk-1]; i = h - j - k
    int kmin = h - (d.ni - 1) - (d.nj - 1);
    if (kmin < 2) kmin = 2;
    int kmax = h - 4;
    if (kmax > d.nk - 1) kmax = d.nk - 1;
    if (kmax < kmin) return 0;
    pl.h = h;
    pl.k0 = kmin;
    long long threads = (long long)(kmax - kmin + 1) * (d.nj - 2);
    blocks = (unsigned)((threads + 255) / 256);
    return 1;
}

static int sweep_dcmp(cad_ctx *c, double u, int *npos_host)
{
    if (c->sweep_schedule != CAD_SWEEP_HYPERPLANE) return cad_sweep_dcmp(c, u, npos_host);
    const Dims &d = c->d;
    hipLaunchKernelGGL(k_iset1, dim3(1), dim3(1), 0, c->stream, c->icount + 1, 1);
    const int hmin = 6, hmax = (d.ni - 1) + (d.nj - 1) + (d.nk - 1);
    for (int h = hmin; h <= hmax; h++) {
        Plane pl;
        unsigned nb;
        if (!hyper_launch_range(d, h, pl, nb)) continue;
        hipLaunchKernelGGL(k_mzdcmp_plane, dim3(nb), dim3(256), 0, c->stream, d, pl, c->ad, c->ali,
                           c->alj, c->alk, c->aui, c->auj, c->auk, c->f.indc, c->dd, u,
                           c->icount + 1);
    }
    CAD_CHECK(hipGetLastError());
    CAD_CHECK(hipMemcpyAsync(c->host_cnt + 1, c->icount + 1, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    CAD_CHECK(hipStreamSynchronize(c->stream));
    *npos_host = c->host_cnt[1];
    return CAD_OK;
}

// x = M^-1 y  (vf_mzminv.f:38-41)
static int minv_launch(cad_ctx *c, double *x, const double *y)
{
    if (c->sweep_schedule != CAD_SWEEP_HYPERPLANE) return cad_sweep_minv(c, x, y);
    const Dims &d = c->d;
    const int hmin = 6, hmax = (d.ni - 1) + (d.nj - 1) + (d.nk - 1);
    for (int h = hmin; h <= hmax; h++) {
        Plane pl;
        unsigned nb;
        if (!hyper_launch_range(d, h, pl, nb)) continue;
        hipLaunchKernelGGL(k_frwd_plane, dim3(nb), dim3(256), 0, c->stream, d, pl, c->dd, c->ali,
                           c->alj, c->alk, c->f.indc, x, y);
    }
    for (int h = hmax; h >= hmin; h--) {
        Plane pl;
        unsigned nb;
        if (!hyper_launch_range(d, h, pl, nb)) continue;
        hipLaunchKernelGGL(k_bkwd_plane, dim3(nb), dim3(256), 0, c->stream, d, pl, c->dd, c->aui,
                           c->auj, c->auk, c->f.indc, x);
    }
    CAD_CHECK(hipGetLastError());
    return CAD_OK;
}

static int apply_minv(cad_ctx *c, double *x, const double *y)
{
    int rc = cad_kt_begin(c, CAD_KT_MINV);
    if (!rc) rc = minv_launch(c, x, y);
    if (!rc) rc = cad_kt_end(c, CAD_KT_MINV, 0.0);
    return rc;
}

#define RET_IF_(x)             \
    do {                       \
        int rc_ = (x);         \
        if (rc_) return rc_;   \
    } while (0)

static int apply_a(cad_ctx *c, const double *x, double *y)
{
    hipLaunchKernelGGL(k_mzax, dim3(VEC_BLOCKS), dim3(256), 0, c->stream, c->d, c->act_rng, c->ad,
                       c->ali, c->alj, c->alk, c->aui, c->auj, c->auk, c->f.indc, x, y);
    CAD_CHECK(hipGetLastError());
    return CAD_OK;
}

// partials of nprod products in red_partial -> sc[slot..] (+ scalar mode)
static int reduce_to(cad_ctx *c, int nprod, int slot, int mode)
{
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(REDUCE_THREADS), 0, c->stream, c->red_partial,
                       nprod, c->ksc, slot, mode);
    CAD_CHECK(hipGetLastError());
    return CAD_OK;
}

static int dot_to(cad_ctx *c, const double *x, const double *y, int slot, int mode)
{
    hipLaunchKernelGGL(k_dot, dim3(REDUCE_BLOCKS), dim3(REDUCE_THREADS), 0, c->stream, c->act_rng,
                       c->f.indc, x, y, c->red_partial);
    CAD_CHECK(hipGetLastError());
    return reduce_to(c, 1, slot, mode);
}

// n scalars sc[slot..] -> host
static int fetch_sc(cad_ctx *c, int slot, int n, double *out)
{
    CAD_CHECK(hipMemcpyAsync(c->host_red, c->ksc + slot, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CAD_CHECK(hipStreamSynchronize(c->stream));
    for (int q = 0; q < n; q++) out[q] = c->host_red[q];
    return CAD_OK;
}

// Generic reduction used outside the solver (VOF sums, CDIV00): partials in
// red_partial -> n values on the host.
int cad_reduce_fetch(cad_ctx *c, int n, double *host_out)
{
    RET_IF_(reduce_to(c, n, 0, RED_PLAIN));
    return fetch_sc(c, 0, n, host_out);
}

// VF_M1BCGS on the context's matrix with solution vector x0 (device),
// vf_m1bcgs.f:78-262.  Same algebra, scalar recurrences and stop test as the
// reference; the dot products are deterministic two-stage sums.
int cad_solve_m1bcgs(cad_ctx *c, double *x0, double *bnorm_out, double *xnorm_out, int *iter_out)
{
    const Dims &d = c->d;
    const long long nc = d.nc;
    const int *mask = c->f.indc;
    const unsigned long long *rng = c->act_rng;
    hipStream_t st = c->stream;
    const dim3 sg = stream_grid(nc), sb(256), vg(VEC_BLOCKS);
    const dim3 rgd(REDUCE_BLOCKS), rbd(REDUCE_THREADS);
    // XW = X0; DD, YY, R0, RR, PP, AP, AE = 0  (vf_m1bcgs.f:82-99)
    hipLaunchKernelGGL(k_copy, sg, sb, 0, st, nc, x0, c->xw);
    double *zero_vecs[7] = {c->dd, c->yv, c->r0, c->rr, c->pv, c->ap, c->ae};
    for (double *z : zero_vecs) hipLaunchKernelGGL(k_fill, sg, sb, 0, st, nc, z, 0.0);
    hipLaunchKernelGGL(k_range_init, dim3(1), dim3(1), 0, st, c->act_rng, nc);
    hipLaunchKernelGGL(k_active_range, dim3(1024), dim3(256), 0, st, nc, mask, c->act_rng);
    CAD_CHECK(hipGetLastError());
    int iter = -1;
    double bnorm = 0.0, xnorm = 0.0;
    const double u = c->prm.icgtyp != 0 ? c->prm.cgpara : 0.0;
    int npos = 1;
    RET_IF_(sweep_dcmp(c, u, &npos));
    if (npos == 0) {
        *bnorm_out = 0.0;
        *xnorm_out = 0.0;
        *iter_out = -1;
        return CAD_OK;
    }
    const double epsa2 = c->prm.cgepsa * c->prm.cgepsa;
    const double epsr2 = c->prm.cgepsr * c->prm.cgepsr;
    const double div2 = c->prm.cgdiv * c->prm.cgdiv;
    RET_IF_(apply_minv(c, c->yv, c->bb));
    RET_IF_(dot_to(c, c->yv, c->yv, SC_BN, RED_PLAIN));
    iter = 0;
    RET_IF_(apply_a(c, c->xw, c->yv));
    hipLaunchKernelGGL(k_bminus, vg, sb, 0, st, rng, mask, c->bb, c->yv);
    RET_IF_(apply_minv(c, c->r0, c->yv));
    RET_IF_(dot_to(c, c->r0, c->r0, SC_S, RED_PLAIN));
    double init[2];
    RET_IF_(fetch_sc(c, SC_S, 1, &init[0]));
    RET_IF_(fetch_sc(c, SC_BN, 1, &init[1]));
    xnorm = init[0];
    bnorm = init[1];
    // optional second stop condition: the true relative residual
    // ||BB - A XW|| / ||BB|| at most c->true_stop (no counterpart in the
    // reference; holds the plain form to the residual the other forms stop on)
    const bool true_stop = c->true_stop > 0.0;
    double bb2 = 0.0;
    if (true_stop) {
        RET_IF_(dot_to(c, c->bb, c->bb, SC_TBB, RED_PLAIN));
        RET_IF_(fetch_sc(c, SC_TBB, 1, &bb2));
    }
    auto true_ok = [&](bool *ok) -> int {
        RET_IF_(apply_a(c, c->xw, c->yv));
        hipLaunchKernelGGL(k_bminus, vg, sb, 0, st, rng, mask, c->bb, c->yv);
        CAD_CHECK(hipGetLastError());
        RET_IF_(dot_to(c, c->yv, c->yv, SC_TRR, RED_PLAIN));
        double rr2 = 0.0;
        RET_IF_(fetch_sc(c, SC_TRR, 1, &rr2));
        *ok = rr2 <= c->true_stop * c->true_stop * bb2;
        return CAD_OK;
    };
    bool done = xnorm <= bnorm * epsr2 + epsa2;
    if (done && true_stop) RET_IF_(true_ok(&done));
    if (!done) {
        hipLaunchKernelGGL(k_init_rp, vg, sb, 0, st, rng, mask, c->r0, c->rr, c->pv);
        for (;;) {
            iter = iter + 1;
            if (iter > c->prm.icgmax) break;
            RET_IF_(apply_a(c, c->pv, c->yv));
            RET_IF_(apply_minv(c, c->ap, c->yv));
            RET_IF_(dot_to(c, c->r0, c->ap, SC_W1, RED_ALPHA));
            hipLaunchKernelGGL(k_axpy_neg, vg, sb, 0, st, rng, mask, c->ksc, c->ap, c->rr);
            RET_IF_(apply_a(c, c->rr, c->yv));
            RET_IF_(apply_minv(c, c->ae, c->yv));
            hipLaunchKernelGGL(k_dot2, rgd, rbd, 0, st, rng, mask, c->rr, c->ae, c->red_partial);
            RET_IF_(reduce_to(c, 2, SC_W1B, RED_OMEGA));
            hipLaunchKernelGGL(k_update_xr, rgd, rbd, 0, st, rng, mask, c->xw, c->rr, c->pv, c->ae,
                               c->r0, c->ksc, c->red_partial);
            // (R,R) -> SC_XN, (R0,R) -> SC_W1C; beta and the new s on the device
            RET_IF_(reduce_to(c, 2, SC_XN, RED_BETA));
            RET_IF_(fetch_sc(c, SC_XN, 1, &xnorm));
            if (xnorm <= bnorm * epsr2 + epsa2) {
                if (!true_stop) break;
                bool ok = false;
                RET_IF_(true_ok(&ok));
                if (ok) break;
            }
            if (xnorm > div2) {
                iter = -2;
                break;
            }
            hipLaunchKernelGGL(k_update_p, vg, sb, 0, st, rng, mask, c->ksc, c->rr, c->ap, c->pv);
        }
    }
    bnorm = std::sqrt(bnorm);
    xnorm = std::sqrt(xnorm);
    if (iter >= 0) hipLaunchKernelGGL(k_copy_active, vg, sb, 0, st, rng, mask, c->xw, x0);
    CAD_CHECK(hipGetLastError());
    if (c->sweep_schedule != CAD_SWEEP_HYPERPLANE) {
        int code = 0;
        RET_IF_(cad_sweep_status(c, &code));
        if (code != 0) {
            cad_set_error("MILU sweep schedule failed (code %d)", code);
            return CAD_ERR_HIP;
        }
    }
    *bnorm_out = bnorm;
    *xnorm_out = xnorm;
    *iter_out = iter;
    return CAD_OK;
}

// The true relative residual ||BB - A PT||_2 / ||BB||_2 over the active cells,
// in FP64, of the context's matrix and potential (after a solve: its result).
int cad_poisson_residual_impl(cad_ctx *c, double *rel)
{
    const dim3 sb(256), vg(VEC_BLOCKS);
    RET_IF_(apply_a(c, c->pt, c->yv));
    hipLaunchKernelGGL(k_bminus, vg, sb, 0, c->stream, c->act_rng, c->f.indc, c->bb, c->yv);
    CAD_CHECK(hipGetLastError());
    RET_IF_(dot_to(c, c->yv, c->yv, SC_S, RED_PLAIN));
    RET_IF_(dot_to(c, c->bb, c->bb, SC_BN, RED_PLAIN));
    double rr = 0.0, bb = 0.0;
    RET_IF_(fetch_sc(c, SC_S, 1, &rr));
    RET_IF_(fetch_sc(c, SC_BN, 1, &bb));
    *rel = bb > 0.0 ? std::sqrt(rr) / std::sqrt(bb) : std::sqrt(rr);
    return CAD_OK;
}

// Iterations the mixed and learned forms take before handing a solve to
// MILU-BiCGSTAB: cad_set_krylov_cap's value, else 200, at most ICGMAX.
int cad_krylov_cap(const cad_ctx *c)
{
    const int cap = c->krylov_cap > 0 ? c->krylov_cap : 200;
    return cap < c->prm.icgmax ? cap : c->prm.icgmax;
}

// The true relative residual ||BB - A x|| / ||BB|| of a device vector x.
int cad_poisson_residual_x(cad_ctx *c, const double *x, double *rel)
{
    const dim3 sb(256), vg(VEC_BLOCKS);
    RET_IF_(apply_a(c, x, c->yv));
    hipLaunchKernelGGL(k_bminus, vg, sb, 0, c->stream, c->act_rng, c->f.indc, c->bb, c->yv);
    CAD_CHECK(hipGetLastError());
    RET_IF_(dot_to(c, c->yv, c->yv, SC_S, RED_PLAIN));
    RET_IF_(dot_to(c, c->bb, c->bb, SC_BN, RED_PLAIN));
    double rr = 0.0, bb = 0.0;
    RET_IF_(fetch_sc(c, SC_S, 1, &rr));
    RET_IF_(fetch_sc(c, SC_BN, 1, &bb));
    *rel = bb > 0.0 ? std::sqrt(rr) / std::sqrt(bb) : std::sqrt(rr);
    return CAD_OK;
}

// MILU factorisation of the context's matrix on its own (cad_milu_factor).
int cad_milu_factor_impl(cad_ctx *c, int *npos)
{
    const long long nc = c->d.nc;
    hipLaunchKernelGGL(k_fill, stream_grid(nc), dim3(256), 0, c->stream, nc, c->dd, 0.0);
    CAD_CHECK(hipGetLastError());
    const double u = c->prm.icgtyp != 0 ? c->prm.cgpara : 0.0;
    return sweep_dcmp(c, u, npos);
}

// x = M^-1 y for a host y (cad_milu_apply); uses YV and AE as device buffers.
int cad_milu_apply_impl(cad_ctx *c, const double *y, double *x)
{
    const long long nc = c->d.nc;
    CAD_CHECK(hipMemcpyAsync(c->yv, y, nc * sizeof(double), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_fill, stream_grid(nc), dim3(256), 0, c->stream, nc, c->ae, 0.0);
    CAD_CHECK(hipGetLastError());
    RET_IF_(apply_minv(c, c->ae, c->yv));
    CAD_CHECK(hipMemcpyAsync(x, c->ae, nc * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CAD_CHECK(hipStreamSynchronize(c->stream));
    if (c->sweep_schedule != CAD_SWEEP_HYPERPLANE) {
        int code = 0;
        RET_IF_(cad_sweep_status(c, &code));
        if (code != 0) {
            cad_set_error("MILU sweep schedule failed (code %d)", code);
            return CAD_ERR_HIP;
        }
    }
    return CAD_OK;
}

// Average duration of one M^-1 application (forward + backward MILU sweeps of
// vf_mzminv.f) on the context's current factorisation, timed with HIP events on
// the library's stream.  Used by the benchmark to report the dominant kernel.
int cad_time_minv_impl(cad_ctx *c, int reps, float *ms_avg)
{
    if (reps < 1) reps = 1;
    hipEvent_t a, b;
    CAD_CHECK(hipEventCreate(&a));
    CAD_CHECK(hipEventCreate(&b));
    CAD_CHECK(hipEventRecord(a, c->stream));
    for (int r = 0; r < reps; r++) RET_IF_(apply_minv(c, c->ae, c->yv));
    CAD_CHECK(hipEventRecord(b, c->stream));
    CAD_CHECK(hipEventSynchronize(b));
    if (c->sweep_schedule != CAD_SWEEP_HYPERPLANE) {
        int code = 0;
        RET_IF_(cad_sweep_status(c, &code));
        if (code != 0) {
            cad_set_error("MILU sweep schedule failed (code %d)", code);
            return CAD_ERR_HIP;
        }
    }
    float ms = 0.0f;
    CAD_CHECK(hipEventElapsedTime(&ms, a, b));
    hipEventDestroy(a);
    hipEventDestroy(b);
    *ms_avg = ms / reps;
    return CAD_OK;
}

// ---------------------------------------------------------------------------
// Pieces of the solve shared with the learned form (cad_learned.hip).

// The active index range [n0, n1) of the context's INDC into act_rng and hr.
int cad_ext_active_range(cad_ctx *c, unsigned long long hr[2])
{
    hipLaunchKernelGGL(k_range_init, dim3(1), dim3(1), 0, c->stream, c->act_rng, c->d.nc);
    hipLaunchKernelGGL(k_active_range, dim3(1024), dim3(256), 0, c->stream, c->d.nc, c->f.indc,
                       c->act_rng);
    CAD_CHECK(hipGetLastError());
    CAD_CHECK(hipMemcpyAsync(hr, c->act_rng, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                             c->stream));
    CAD_CHECK(hipStreamSynchronize(c->stream));
    return CAD_OK;
}

// y = A x (active cells)
int cad_ext_apply_a(cad_ctx *c, const double *x, double *y) { return apply_a(c, x, y); }

// y = b - A x (active cells)
int cad_ext_residual(cad_ctx *c, const double *x, double *y)
{
    RET_IF_(apply_a(c, x, y));
    hipLaunchKernelGGL(k_bminus, dim3(VEC_BLOCKS), dim3(256), 0, c->stream, c->act_rng, c->f.indc, c->bb, y);
    CAD_CHECK(hipGetLastError());
    return CAD_OK;
}

// (x, y) over the active cells, to the host
int cad_ext_dot(cad_ctx *c, const double *x, const double *y, double *out)
{
    RET_IF_(dot_to(c, x, y, SC_S, RED_PLAIN));
    return fetch_sc(c, SC_S, 1, out);
}
