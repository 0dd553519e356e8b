// cad_refine.hip -- the mixed form's pressure solve: FP64 iterative
// refinement around an FP32 Krylov solve.
//
// The system is VF_VPCOEF's (the FP64 matrix AD..AUK and BB that VF_M1BCGS
// solves, vf_m1bcgs.f:101-262).  Outer iteration, in FP64:
//   r = BB - A x,  stop when ||r|| <= CGEPSR ||BB|| + CGEPSA  (the true residual)
//   solve A d = r / ||r|| approximately in FP32
//   x = x + ||r|| d
// Inner solve, in FP32: BiCGSTAB right-preconditioned by one multigrid V-cycle
// (cad_mg.hip); the stencil matvecs read the multigrid's level-0 rows (the
// matrix rounded to FP32, in the grid's order), every vector is FP32, and the
// dot products are
// accumulated in FP64.  The scalars alpha, omega, rho stay on the device; an
// iteration reads one pair of values on the host, (r, r) and a breakdown
// flag.  The inner solve stops at a relative residual eta = max(1e-6, 0.1
// tol / ||r||): about what FP32 can resolve, and on the last outer step a
// tenth of what is left to the stop test, so that the answer lands below it
// with room to spare.  A solve takes two or three outer steps.  An outer step that does not halve the true residual, or more
// than the iteration cap (cad_krylov_cap), hands the solve to VF_M1BCGS from
// the current x with the same true-residual stop.
#include "cad_common.h"

#include <cmath>

namespace {

constexpr int VB = 2048;
enum { S_RHO1 = 0, S_RHO, S_ALPHA, S_OM, S_RV, S_RR, S_FLAG };   // slots of ksc

#define RANGE_LOOP(n, rng)                                                             \
    const long long n0_ = (long long)(rng)[0], n1_ = (long long)(rng)[1];           \
    for (long long n = n0_ + (long long)blockIdx.x * blockDim.x + threadIdx.x; n < n1_; \
         n += (long long)gridDim.x * blockDim.x)

__device__ __forceinline__ double wsum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// block sums of up to two products into partial[q * REDUCE_BLOCKS + block]
template <int N>
__device__ __forceinline__ void block_sums(const double (&a)[2], double *partial)
{
    __shared__ double sh[2][REDUCE_THREADS / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < N; q++) {
        const double s = wsum(a[q]);
        if (lane == 0) sh[q][w] = s;
    }
    __syncthreads();
    if (threadIdx.x < N) {
        double s = 0.0;
        for (int i = 0; i < REDUCE_THREADS / 64; i++) s += sh[threadIdx.x][i];
        partial[threadIdx.x * REDUCE_BLOCKS + blockIdx.x] = s;
    }
}

// b = r = r0 = (float)(r64 / rn); p = v = d = 0 (active rows)
__global__ void __launch_bounds__(256) k_r_init(const unsigned long long *rng, MgLevel L,
                                                const double *__restrict__ r64, double inv,
                                                float *__restrict__ r, float *__restrict__ r0,
                                                float *__restrict__ p, float *__restrict__ v,
                                                float *__restrict__ d)
{
    RANGE_LOOP(n, rng) {
        const float b = L.nact[n] ? (float)(r64[n] * inv) : 0.0f;
        r[n] = b;
        r0[n] = b;
        p[n] = 0.0f;
        v[n] = 0.0f;
        d[n] = 0.0f;
    }
}

__global__ void k_r_scal_init(double *sc)
{
    sc[S_RHO] = 1.0;
    sc[S_ALPHA] = 1.0;
    sc[S_OM] = 1.0;
    sc[S_FLAG] = 0.0;
}

// partials of (x, y)
__global__ void __launch_bounds__(REDUCE_THREADS) k_r_dot(const unsigned long long *rng, MgLevel L,
                                                          const float *__restrict__ x,
                                                          const float *__restrict__ y, double *partial)
{
    double a[2] = {0.0, 0.0};
    RANGE_LOOP(n, rng) {
        if (L.nact[n]) a[0] += (double)x[n] * (double)y[n];
    }
    block_sums<1>(a, partial);
}

// p = r + beta (p - om v),  beta = (rho1 / rho) (alpha / om)
__global__ void __launch_bounds__(256) k_r_p(const unsigned long long *rng, MgLevel L,
                                             const double *__restrict__ sc, const float *__restrict__ r,
                                             const float *__restrict__ v, float *__restrict__ p)
{
    const float beta = (float)((sc[S_RHO1] / sc[S_RHO]) * (sc[S_ALPHA] / sc[S_OM]));
    const float om = (float)sc[S_OM];
    RANGE_LOOP(n, rng) {
        if (L.nact[n]) p[n] = r[n] + beta * (p[n] - om * v[n]);
    }
}

__device__ __forceinline__ float mv32(const MgLevel &L, const float *__restrict__ x, long long n)
{
    const long long ni = L.d.ni, nij = L.d.nij;
    return L.nad[n] * x[n] + L.nali[n] * x[n - 1] + L.naui[n] * x[n + 1] + L.nalj[n] * x[n - ni]
         + L.nauj[n] * x[n + ni] + L.nalk[n] * x[n - nij] + L.nauk[n] * x[n + nij];
}

// y = A x (FP32 rows of level 0) with the partials of (w, y) and, N = 2, (y, y)
template <int N>
__global__ void __launch_bounds__(REDUCE_THREADS) k_r_mv(const unsigned long long *rng, MgLevel L,
                                                         const float *__restrict__ x,
                                                         const float *__restrict__ w,
                                                         float *__restrict__ y, double *partial)
{
    double a[2] = {0.0, 0.0};
    RANGE_LOOP(n, rng) {
        if (!L.nact[n]) continue;
        const float t = mv32(L, x, n);
        y[n] = t;
        a[0] += (double)w[n] * (double)t;
        if (N == 2) a[1] += (double)t * (double)t;
    }
    block_sums<N>(a, partial);
}

// s = r - alpha v
__global__ void __launch_bounds__(256) k_r_s(const unsigned long long *rng, MgLevel L,
                                             const double *__restrict__ sc, const float *__restrict__ r,
                                             const float *__restrict__ v, float *__restrict__ s)
{
    const float alpha = (float)sc[S_ALPHA];
    RANGE_LOOP(n, rng) {
        if (L.nact[n]) s[n] = r[n] - alpha * v[n];
    }
}

// d += alpha ph + om sh; r = s - om t; partials of (r, r) and (r0, r)
__global__ void __launch_bounds__(REDUCE_THREADS) k_r_xr(const unsigned long long *rng, MgLevel L,
                                                         const double *__restrict__ sc,
                                                         const float *__restrict__ ph,
                                                         const float *__restrict__ shv,
                                                         const float *__restrict__ s,
                                                         const float *__restrict__ t,
                                                         const float *__restrict__ r0,
                                                         float *__restrict__ d, float *__restrict__ r,
                                                         double *partial)
{
    const float alpha = (float)sc[S_ALPHA], om = (float)sc[S_OM];
    double a[2] = {0.0, 0.0};
    RANGE_LOOP(n, rng) {
        if (!L.nact[n]) continue;
        d[n] += alpha * ph[n] + om * shv[n];
        const float rn = s[n] - om * t[n];
        r[n] = rn;
        a[0] += (double)rn * (double)rn;
        a[1] += (double)r0[n] * (double)rn;
    }
    block_sums<2>(a, partial);
}

enum { F_RHO1 = 0, F_ALPHA, F_OMEGA, F_XR };

// the partials' sums in block order, and the scalar step of the mode
__global__ void __launch_bounds__(256) k_r_final(const double *__restrict__ partial, int nprod, double *sc,
                                                 int mode)
{
    __shared__ double sh[2][256];
    for (int q = 0; q < nprod; q++) {
        double a = 0.0;
        for (int b = threadIdx.x; b < REDUCE_BLOCKS; b += 256) a += partial[q * REDUCE_BLOCKS + b];
        sh[q][threadIdx.x] = a;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double s[2] = {0.0, 0.0};
    for (int q = 0; q < nprod; q++)
        for (int i = 0; i < 256; i++) s[q] += sh[q][i];
    const bool ok0 = std::isfinite(s[0]) && s[0] != 0.0;
    switch (mode) {
    case F_RHO1:
        sc[S_RHO1] = s[0];
        if (!ok0) sc[S_FLAG] = 1.0;
        break;
    case F_ALPHA:
        sc[S_RV] = s[0];
        if (ok0)
            sc[S_ALPHA] = sc[S_RHO1] / s[0];
        else {
            sc[S_ALPHA] = 0.0;
            sc[S_FLAG] = 1.0;
        }
        break;
    case F_OMEGA:
        if (std::isfinite(s[1]) && s[1] > 0.0)
            sc[S_OM] = s[0] / s[1];
        else {
            sc[S_OM] = 0.0;
            sc[S_FLAG] = 1.0;
        }
        break;
    case F_XR:
        sc[S_RR] = s[0];
        sc[S_RHO] = sc[S_RHO1];
        sc[S_RHO1] = s[1];
        if (!(std::isfinite(s[1]) && s[1] != 0.0) || !(sc[S_OM] != 0.0)) sc[S_FLAG] = 1.0;
        break;
    }
}

// x += scale d (active rows), FP64
__global__ void __launch_bounds__(256) k_r_update(const unsigned long long *rng, MgLevel L, double scale,
                                                  const float *__restrict__ d, double *__restrict__ x)
{
    RANGE_LOOP(n, rng) {
        if (L.nact[n]) x[n] += scale * (double)d[n];
    }
}

// b = a on the active rows, FP64
__global__ void __launch_bounds__(256) k_r_copy(const unsigned long long *rng, MgLevel L,
                                                const double *__restrict__ a, double *__restrict__ b)
{
    RANGE_LOOP(n, rng) {
        if (L.nact[n]) b[n] = a[n];
    }
}

__global__ void k_r_fill(long long n, double *a)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        a[i] = 0.0;
}

#define RET_IF(x)              \
    do {                       \
        int rc_ = (x);         \
        if (rc_) return rc_;   \
    } while (0)

}  // namespace

int cad_solve_mixed(cad_ctx *c, double *x0, double *bnorm_out, double *xnorm_out, int *iter_out)
{
    const Dims &d = c->d;
    const long long nc = d.nc;
    hipStream_t st = c->stream;
    const dim3 vb(VB), tb(256), rb(REDUCE_BLOCKS), rt(REDUCE_THREADS);
    // FP64: x (XW), the outer residual (RR); FP32 vectors two to an FP64 buffer
    double *x = c->xw, *r64 = c->rr;
    double *fbuf[5] = {c->r0, c->pv, c->ap, c->yv, c->ae};
    for (double *b : fbuf) hipLaunchKernelGGL(k_r_fill, vb, tb, 0, st, nc, b);
    hipLaunchKernelGGL(k_r_fill, vb, tb, 0, st, nc, x);
    CAD_CHECK(hipGetLastError());
    float *r = reinterpret_cast<float *>(c->r0), *r0 = r + nc;
    float *p = reinterpret_cast<float *>(c->pv), *v = p + nc;
    float *s = reinterpret_cast<float *>(c->ap), *t = s + nc;
    float *ph = reinterpret_cast<float *>(c->yv), *shv = ph + nc;
    float *dv = reinterpret_cast<float *>(c->ae);
    unsigned long long hr[2];
    RET_IF(cad_ext_active_range(c, hr));
    if (hr[0] >= hr[1]) {
        *bnorm_out = 0.0;
        *xnorm_out = 0.0;
        *iter_out = 0;
        return CAD_OK;
    }
    const unsigned long long *rng = c->act_rng;
    c->mg_span = hr[1] - hr[0];
    c->mg_k0 = (int)(hr[0] / (unsigned long long)d.nij) + 1;
    c->mg_k1 = (int)((hr[1] - 1) / (unsigned long long)d.nij) + 1;
    RET_IF(cad_mg_setup(c));
    const MgLevel &L0 = c->mg[0];
    // x = x0 on the active rows
    hipLaunchKernelGGL(k_r_copy, vb, tb, 0, st, rng, L0, x0, x);
    CAD_CHECK(hipGetLastError());
    double bn2 = 0.0;
    RET_IF(cad_ext_dot(c, c->bb, c->bb, &bn2));
    const double bnorm = std::sqrt(bn2);
    // the reference's stop on the true residual, held to the context's true
    // relative residual when one is set
    double tol = c->prm.cgepsr * bnorm + c->prm.cgepsa;
    if (c->true_stop > 0.0) tol = std::fmin(tol, c->true_stop * bnorm);
    const int cap = cad_krylov_cap(c);
    int inner_total = 0, outer = 0;
    double rn = 0.0, prev = -1.0;
    bool fail = false;
    double *sc = c->ksc;
    double *hb = c->host_red;
    for (;;) {
        RET_IF(cad_ext_residual(c, x, r64));
        double rr64 = 0.0;
        RET_IF(cad_ext_dot(c, r64, r64, &rr64));
        rn = std::sqrt(rr64);
        if (!std::isfinite(rn)) {
            fail = true;
            break;
        }
        if (rn <= tol) break;
        if (inner_total >= cap || outer >= 20 || (prev > 0.0 && !(rn < 0.5 * prev))) {
            fail = true;
            break;
        }
        prev = rn;
        const double eta = std::fmax(1.0e-6, 0.1 * tol / rn);
        hipLaunchKernelGGL(k_r_init, vb, tb, 0, st, rng, L0, r64, 1.0 / rn, r, r0, p, v, dv);
        hipLaunchKernelGGL(k_r_scal_init, dim3(1), dim3(1), 0, st, sc);
        hipLaunchKernelGGL(k_r_dot, rb, rt, 0, st, rng, L0, r0, r, c->red_partial);
        hipLaunchKernelGGL(k_r_final, dim3(1), dim3(256), 0, st, c->red_partial, 1, sc, (int)F_RHO1);
        CAD_CHECK(hipGetLastError());
        for (int it = 0; it < 60 && inner_total < cap; it++) {
            hipLaunchKernelGGL(k_r_p, vb, tb, 0, st, rng, L0, sc, r, v, p);
            CAD_CHECK(hipGetLastError());
            RET_IF(cad_mg_apply32(c, p, ph, c->mg_k0, c->mg_k1));
            hipLaunchKernelGGL(k_r_mv<1>, rb, rt, 0, st, rng, L0, ph, r0, v, c->red_partial);
            hipLaunchKernelGGL(k_r_final, dim3(1), dim3(256), 0, st, c->red_partial, 1, sc, (int)F_ALPHA);
            hipLaunchKernelGGL(k_r_s, vb, tb, 0, st, rng, L0, sc, r, v, s);
            CAD_CHECK(hipGetLastError());
            RET_IF(cad_mg_apply32(c, s, shv, c->mg_k0, c->mg_k1));
            hipLaunchKernelGGL(k_r_mv<2>, rb, rt, 0, st, rng, L0, shv, s, t, c->red_partial);
            hipLaunchKernelGGL(k_r_final, dim3(1), dim3(256), 0, st, c->red_partial, 2, sc, (int)F_OMEGA);
            hipLaunchKernelGGL(k_r_xr, rb, rt, 0, st, rng, L0, sc, ph, shv, s, t, r0, dv, r, c->red_partial);
            hipLaunchKernelGGL(k_r_final, dim3(1), dim3(256), 0, st, c->red_partial, 2, sc, (int)F_XR);
            CAD_CHECK(hipGetLastError());
            // one host read per iteration: (r, r) and the breakdown flag
            CAD_CHECK(hipMemcpyAsync(hb, sc + S_RR, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
            CAD_CHECK(hipStreamSynchronize(st));
            inner_total++;
            if (hb[1] != 0.0 || !(std::sqrt(hb[0]) > eta)) break;
        }
        hipLaunchKernelGGL(k_r_update, vb, tb, 0, st, rng, L0, rn, dv, x);
        CAD_CHECK(hipGetLastError());
        outer++;
    }
    if (fail && std::isfinite(rn)) {
        // the refinement stalled: the reference's MILU-BiCGSTAB finishes the
        // solve from the current x, held to the same true relative residual
        hipLaunchKernelGGL(k_r_copy, vb, tb, 0, st, rng, L0, x, x0);
        CAD_CHECK(hipGetLastError());
        const double keep_stop = c->true_stop;
        c->true_stop = keep_stop > 0.0 ? std::fmin(keep_stop, c->prm.cgepsr) : c->prm.cgepsr;
        double bn = 0.0, xn = 0.0;
        int fit = 0;
        const int rc = cad_solve_m1bcgs(c, x0, &bn, &xn, &fit);
        c->true_stop = keep_stop;
        RET_IF(rc);
        c->mixed_fallbacks++;
        c->mixed_fallback_iters = fit;
        *bnorm_out = bnorm;
        if (fit >= 0 && fit <= c->prm.icgmax) {
            double rel = 0.0;
            RET_IF(cad_poisson_residual_x(c, x0, &rel));
            *xnorm_out = rel;
            *iter_out = inner_total + fit;
        } else {
            *xnorm_out = bnorm > 0.0 ? rn / bnorm : rn;
            *iter_out = fit < 0 ? fit : c->prm.icgmax + 1;
        }
        return CAD_OK;
    }
    if (!fail) {
        hipLaunchKernelGGL(k_r_copy, vb, tb, 0, st, rng, L0, x, x0);
        CAD_CHECK(hipGetLastError());
    }
    *bnorm_out = bnorm;
    *xnorm_out = bnorm > 0.0 ? rn / bnorm : rn;
    *iter_out = fail ? -2 : inner_total;
    return CAD_OK;
}
