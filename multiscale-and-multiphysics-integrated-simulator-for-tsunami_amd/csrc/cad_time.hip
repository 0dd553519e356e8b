// cad_time.hip -- the automatic time step VF_CDTCAL (Source code/vf_cdtcal.f)
// on the device-resident fields.
//
// The reference folds each non-obstacle cell's limit into DT one cell at a
// time, DT = V/MAX(FL1..FL6, V/DT) then DT = Vd/MAX(nu, Vd/DT): DT becomes
// the cell's limit when that is smaller and otherwise stays DT up to the
// rounding of V/(V/DT).  The device takes the exact minimum of the same
// per-cell quotients over all cells (a min is order-independent, so the
// result is deterministic); it differs from the serial fold only by that
// per-cell rounding, a few ulps at most.
#include "cad_common.h"

namespace {

__global__ void __launch_bounds__(REDUCE_THREADS) k_cdtcal(Dims d, Static s, const double *__restrict__ u,
                                                          const double *__restrict__ v,
                                                          const double *__restrict__ w,
                                                          const double *__restrict__ f,
                                                          const double *__restrict__ bcf,
                                                          const int *__restrict__ nf, double *partial)
{
    __shared__ double sh[REDUCE_THREADS];
    double lim = INFINITY;
    for (long long n = (long long)blockIdx.x * REDUCE_THREADS + threadIdx.x; n < d.nc;
         n += (long long)REDUCE_BLOCKS * REDUCE_THREADS) {
        const int k = (int)(n / d.nij) + 1;
        const long long r = n - (long long)(k - 1) * d.nij;
        const int j = (int)(r / d.ni) + 1;
        const int i = (int)(r - (long long)(j - 1) * d.ni) + 1;
        if (i < 2 || i > d.ni - 1 || j < 2 || j > d.nj - 1 || k < 2 || k > d.nk - 1) continue;
        if (nf[n] == -1) continue;
        // vf_cdtcal.f: cell sizes, face areas, volume
        const double dx = met<0>(s, 2, i), dy = met<1>(s, 2, j), dz = met<2>(s, 2, k);
        const double dxy = dx * dy, dxz = dx * dz, dyz = dy * dz;
        const double fc = f[n];
        const double vol = s.ggv[n] * dx * dy * dz;
        const long long ni = d.ni, nij = d.nij;
        const long long face[6] = {n, n + 1, n, n + ni, n, n + nij};
        const double *g[3] = {s.ggx, s.ggy, s.ggz};
        const double *q[3] = {u, v, w};
        const int *ind[3] = {s.indx, s.indy, s.indz};
        const double ar[3] = {dyz, dxz, dxy};
        double m = 0.0;
#pragma unroll
        for (int e = 0; e < 6; e++) {
            const int a = e / 2;
            const long long fi = face[e];
            double fl = fabs(ar[a] * g[a][fi] * q[a][fi]);
            const int b = ind[a][fi];
            // no flux limit on a boundary face with no water on either side
            if (b >= 1 && dmaxd(fc, bcf[b - 1]) <= 0.0) fl = 0.0;
            m = dmaxd(m, fl);
        }
        if (m > 0.0) lim = dmind(lim, vol / m);
        const double vd = 0.5 / (1.0 / (dx * dx) + 1.0 / (dy * dy) + 1.0 / (dz * dz));
        const double nu = s.anu[n];
        if (nu > 0.0) lim = dmind(lim, vd / nu);
    }
    sh[threadIdx.x] = lim;
    __syncthreads();
    for (int st = REDUCE_THREADS / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) sh[threadIdx.x] = dmind(sh[threadIdx.x], sh[threadIdx.x + st]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(REDUCE_THREADS) k_min_final(const double *__restrict__ partial, int n,
                                                             double *out)
{
    __shared__ double sh[REDUCE_THREADS];
    double m = INFINITY;
    for (int t = threadIdx.x; t < n; t += REDUCE_THREADS) m = dmind(m, partial[t]);
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int st = REDUCE_THREADS / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) sh[threadIdx.x] = dmind(sh[threadIdx.x], sh[threadIdx.x + st]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}

}  // namespace

// VF_CDTCAL: DT = MAX(MIN(limit*DTSAFE, 1.2*W), DTMIN) with the limit no
// larger than DTMAX/DTSAFE and W the last step's DTNOW (DTINIT at NNOW=0).
int cad_cdtcal_impl(cad_ctx *c, int nnow, double dtinit, double dtsafe, double dtmin, double dtmax,
                    double *dt)
{
    Static s = cad_static(c);
    hipStream_t st = c->stream;
    hipLaunchKernelGGL(k_cdtcal, dim3(REDUCE_BLOCKS), dim3(REDUCE_THREADS), 0, st, c->d, s, c->f.uu, c->f.vv,
                       c->f.ww, c->f.ff, c->f.bcf, c->f.nf, c->red_partial);
    hipLaunchKernelGGL(k_min_final, dim3(1), dim3(REDUCE_THREADS), 0, st, c->red_partial, REDUCE_BLOCKS,
                       c->red_out);
    CAD_CHECK(hipGetLastError());
    CAD_CHECK(hipMemcpyAsync(c->host_red, c->red_out, sizeof(double), hipMemcpyDeviceToHost, st));
    CAD_CHECK(hipStreamSynchronize(st));
    double lim = dtmax / dtsafe;
    if (c->host_red[0] < lim) lim = c->host_red[0];
    const double w = nnow == 0 ? dtinit : c->prm.dtnow;
    double t = lim * dtsafe;
    if (1.20 * w < t) t = 1.20 * w;
    *dt = t > dtmin ? t : dtmin;
    return CAD_OK;
}
