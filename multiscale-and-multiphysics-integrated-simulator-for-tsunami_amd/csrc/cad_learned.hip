// cad_learned.hip -- the learned form's pressure solve: flexible GMRES in FP64
// on the same matrix and right-hand side as VF_M1BCGS (vf_m1bcgs.f:101-262),
// preconditioned by the 3-D U-Net of cad_unet.hip, with the reference's
// MILU-BiCGSTAB as the fallback when the Krylov residual stalls.
//
// Preconditioner.  One application to a vector v (mode 1, the default):
//   z1 = V(v)                   one FP32 multigrid V-cycle (cad_mg.hip)
//   r1 = v - A z1               FP64
//   z  = z1 + s N(r1 / s)       s = RMS of r1 over the active cells
// where N is the network (BF16 on the matrix cores, FP32 accumulation): it
// maps a residual, with the porosity GGV, F and the cell type, to the
// correction the V-cycle left out.  Mode 0 is the network alone, z = s N(v/s).
// The network is nonlinear, hence flexible GMRES, which keeps the
// preconditioned vectors Z beside the Arnoldi basis V.  Orthogonalisation is
// classical Gram-Schmidt done twice with the second pass delayed by one step:
// per step one reduction and one update pass over the basis, one host read.
// The stop test is on the true relative residual
// ||BB - A x|| / ||BB|| <= CGEPSR, recomputed in FP64 at every restart.  A
// cycle that does not halve the residual, or more than ICGMAX iterations,
// hands over to VF_M1BCGS from the current x with the same true-residual stop.
//
// Net box.  Voxel (x, y, z) of the network is cell (I, J, K) = (x+2, y+2,
// z+kb): the interior cells in I and J, and in K the planes from the first
// that holds an active cell, rounded up to a multiple of 4; voxels outside
// the grid are inactive.  Input channels: the scaled residual, GGV, F, the
// cell type one-hot (obstacle, fluid, surface, gas) and the activity.
#include "cad_common.h"
#include "cad_unet.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int MAXM = 32;            // largest restart length
constexpr int VBLK = 2048;          // blocks of the vector kernels
// columns of a block's Gram-Schmidt partials: 2 (MAXM-1) + 3 sums, and the
// step's 2 (MAXM-1) + 6 coefficients
constexpr int GS_STRIDE = 2 * MAXM + 4;

}  // namespace

struct Learned {
    cad_unet *net;
    int nx, ny, nzmax;              // the network's box
    int mode;                       // 1: V-cycle then network; 0: network alone
    int m;                          // restart length
    double *vecs;                   // V (m+1), Z (m), then 6 scratch vectors
    int nvec;
    double *part;                   // REDUCE_BLOCKS x GS_STRIDE partial sums
    double *hdev;                   // a step's Gram-Schmidt sums, then its coefficients: 2 GS_STRIDE
    double *scal;                   // [0] s, [1] 1/s, [2] active count
    double *hpin;                   // pinned host mirror of hdev
    hipEvent_t gs_read;             // the step's sums and coefficients are in hpin
    float *y;                       // network output
    unsigned long long prev_rng[2]; // active range of the previous solve
    int kb, nzr;                    // the current box placement
    bool xin_static;                // the network input holds this solve's flag channels
    long long last_iters, last_fallback_iters, solves, fallbacks, total_iters;
    double last_resid;
};

namespace {

__device__ __forceinline__ long long uvox(const UGeo &g, int x, int y, int z)
{
    return (long long)(x + 1) + (long long)g.px * ((y + 1) + (long long)g.py * (z + 1));
}

// network input of the box from the residual r (scaled by scal[1]); plane nzr
// cleared (it is the halo of the run)
__global__ void k_features(Dims d, UGeo g, int kb, int nzr, const double *__restrict__ r,
                           const double *__restrict__ scal, const double *__restrict__ ggv,
                           const double *__restrict__ ff, const int *__restrict__ nf,
                           const int *__restrict__ indc, bf16 *__restrict__ x)
{
    const long long total = (long long)g.nx * g.ny * (nzr + 1);
    const double inv = scal[1];
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int xq = (int)(e % g.nx);
        const int yq = (int)((e / g.nx) % g.ny);
        const int zq = (int)(e / ((long long)g.nx * g.ny));
        bf16x8 o;
#pragma unroll
        for (int c = 0; c < 8; c++) o[c] = (bf16)0.f;
        const int i = xq + 2, j = yq + 2, k = zq + kb;
        if (zq < nzr && i <= d.ni - 1 && j <= d.nj - 1 && k <= d.nk - 1) {
            const long long n = cid(d, i, j, k);
            const int t = nf[n];
            const bool act = indc[n] != -1;
            o[0] = (bf16)(act ? (float)(r[n] * inv) : 0.f);
            o[1] = (bf16)(float)ggv[n];
            o[2] = (bf16)(float)ff[n];
            o[3] = (bf16)(t == -1 ? 1.f : 0.f);
            o[4] = (bf16)(t == 0 ? 1.f : 0.f);
            o[5] = (bf16)(t >= 1 && t <= 6 ? 1.f : 0.f);
            o[6] = (bf16)(t >= 7 ? 1.f : 0.f);
            o[7] = (bf16)(act ? 1.f : 0.f);
        }
        *reinterpret_cast<bf16x8 *>(x + uvox(g, xq, yq, zq) * 8) = o;
    }
}

// channel 0 of the network input (the scaled residual) on the box's active
// voxels; the other channels, and the inactive voxels' zeros, are the ones
// k_features wrote at the start of the solve (the flags do not change within
// it).  Same value as k_features writes.
__global__ void k_features_r(Dims d, UGeo g, int kb, int nzr, const double *__restrict__ r,
                             const double *__restrict__ scal, const int *__restrict__ indc,
                             bf16 *__restrict__ x)
{
    const long long total = (long long)g.nx * g.ny * nzr;
    const double inv = scal[1];
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int xq = (int)(e % g.nx);
        const int yq = (int)((e / g.nx) % g.ny);
        const int zq = (int)(e / ((long long)g.nx * g.ny));
        const int i = xq + 2, j = yq + 2, k = zq + kb;
        if (i > d.ni - 1 || j > d.nj - 1 || k > d.nk - 1) continue;
        const long long n = cid(d, i, j, k);
        if (indc[n] == -1) continue;
        x[uvox(g, xq, yq, zq) * 8] = (bf16)(float)(r[n] * inv);
    }
}

__device__ __forceinline__ bool box_voxel(const Dims &d, long long n, int nx, int ny, int kb, int nzr,
                                          long long &u)
{
    const int i = (int)(n % d.ni) + 1;
    const int j = (int)((n / d.ni) % d.nj) + 1;
    const int k = (int)(n / d.nij) + 1;
    const int xq = i - 2, yq = j - 2, zq = k - kb;
    if (xq < 0 || xq >= nx || yq < 0 || yq >= ny || zq < 0 || zq >= nzr) return false;
    u = xq + (long long)nx * (yq + (long long)ny * zq);
    return true;
}

// z = z1 + s y over the active range (z1 may be null), 0 on inactive cells
__global__ void __launch_bounds__(256) k_scatter(Dims d, const unsigned long long *rng, int nx, int ny, int kb,
                                                 int nzr, const int *__restrict__ indc,
                                                 const double *__restrict__ z1,
                                                 const double *__restrict__ scal,
                                                 const float *__restrict__ y, double *__restrict__ z)
{
    const long long n0 = (long long)rng[0], n1 = (long long)rng[1];
    const double s = scal[0];
    for (long long n = n0 + blockIdx.x * (long long)blockDim.x + threadIdx.x; n < n1;
         n += (long long)gridDim.x * blockDim.x) {
        if (indc[n] == -1) {
            z[n] = 0.0;
            continue;
        }
        long long u;
        const double c = box_voxel(d, n, nx, ny, kb, nzr, u) ? s * (double)y[u] : 0.0;
        z[n] = (z1 ? z1[n] : 0.0) + c;
    }
}

// sampling: target e/s and activity mask in the box's voxel order
__global__ void k_box_target(Dims d, int nx, int ny, int kb, int nzr, const int *__restrict__ indc,
                             const double *__restrict__ e, const double *__restrict__ scal,
                             float *__restrict__ t, float *__restrict__ m)
{
    const long long total = (long long)nx * ny * nzr;
    const double inv = scal[1];
    for (long long u = blockIdx.x * (long long)blockDim.x + threadIdx.x; u < total;
         u += (long long)gridDim.x * blockDim.x) {
        const int xq = (int)(u % nx), yq = (int)((u / nx) % ny), zq = (int)(u / ((long long)nx * ny));
        const int i = xq + 2, j = yq + 2, k = zq + kb;
        float tv = 0.f, mv = 0.f;
        if (i <= d.ni - 1 && j <= d.nj - 1 && k <= d.nk - 1) {
            const long long n = cid(d, i, j, k);
            if (indc[n] != -1) {
                tv = (float)(e[n] * inv);
                mv = 1.f;
            }
        }
        t[u] = tv;
        m[u] = mv;
    }
}

// the matrix rows of the box's voxels, FP32, coefficient-major: AD, ALI,
// AUI, ALJ, AUJ, ALK, AUK (out[c * voxels + u]); outside the grid AD = 1 and
// no couplings
__global__ void k_box_matrix(Dims d, int nx, int ny, int kb, int nzr, const double *__restrict__ ad,
                             const double *__restrict__ ali, const double *__restrict__ aui,
                             const double *__restrict__ alj, const double *__restrict__ auj,
                             const double *__restrict__ alk, const double *__restrict__ auk,
                             float *__restrict__ out)
{
    const long long total = (long long)nx * ny * nzr;
    for (long long u = blockIdx.x * (long long)blockDim.x + threadIdx.x; u < total;
         u += (long long)gridDim.x * blockDim.x) {
        const int xq = (int)(u % nx), yq = (int)((u / nx) % ny), zq = (int)(u / ((long long)nx * ny));
        const int i = xq + 2, j = yq + 2, k = zq + kb;
        const bool in = i <= d.ni - 1 && j <= d.nj - 1 && k <= d.nk - 1;
        const long long n = in ? cid(d, i, j, k) : 0;
        const double *a[7] = {ad, ali, aui, alj, auj, alk, auk};
#pragma unroll
        for (int q = 0; q < 7; q++) out[q * total + u] = in ? (float)a[q][n] : (q == 0 ? 1.f : 0.f);
    }
}

__device__ __forceinline__ double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// The delayed Gram-Schmidt's reduction, one pass over the active cells: with
// the final basis V_0..V_{nv-1}, the candidate u (the previous step's new
// vector, orthogonalised once) and w = A z, the block partials of
//   (V_i, u) in columns 0..nv-1, (V_i, w) in nv..2nv-1, (u, u), (u, w), (w, w)
// in 2nv..2nv+2 (row stride GS_STRIDE).  NV >= nv bounds the registers.
template <int NV>
__global__ void __launch_bounds__(256) k_dgs_reduce(const unsigned long long *rng, const int *__restrict__ mask,
                                                    const double *__restrict__ V, long long nc, int nv,
                                                    const double *__restrict__ u, const double *__restrict__ w,
                                                    double *__restrict__ part)
{
    double as[NV], ap[NV], uu = 0.0, uw = 0.0, ww = 0.0;
#pragma unroll
    for (int i = 0; i < NV; i++) as[i] = ap[i] = 0.0;
    const long long n0 = (long long)rng[0], n1 = (long long)rng[1];
    for (long long n = n0 + blockIdx.x * (long long)blockDim.x + threadIdx.x; n < n1;
         n += (long long)gridDim.x * blockDim.x) {
        if (mask[n] == -1) continue;
        const double un = u[n], wn = w[n];
        uu += un * un;
        uw += un * wn;
        ww += wn * wn;
#pragma unroll
        for (int i = 0; i < NV; i++)
            if (i < nv) {
                const double vi = V[i * nc + n];
                as[i] += vi * un;
                ap[i] += vi * wn;
            }
    }
    __shared__ double sh[4][2 * NV + 3];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    auto put = [&](int col, double v) {
        const double t = wave_sum(v);
        if (lane == 0) sh[wv][col] = t;
    };
#pragma unroll
    for (int i = 0; i < NV; i++)
        if (i < nv) {
            put(i, as[i]);
            put(nv + i, ap[i]);
        }
    put(2 * nv, uu);
    put(2 * nv + 1, uw);
    put(2 * nv + 2, ww);
    __syncthreads();
    if ((int)threadIdx.x < 2 * nv + 3)
        part[(long long)blockIdx.x * GS_STRIDE + threadIdx.x] =
            (sh[0][threadIdx.x] + sh[1][threadIdx.x]) + (sh[2][threadIdx.x] + sh[3][threadIdx.x]);
}

// out[i] = sum over the nblk block partials of column i (row stride
// stride): block i of the grid sums column i, each thread a fixed stride of
// partials, then a fixed tree
__global__ void __launch_bounds__(256) k_cols_final(const double *__restrict__ part, int nblk, int ncol, int stride,
                                                    double *__restrict__ out)
{
    const int i = blockIdx.x;
    if (i >= ncol) return;
    double s = 0.0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) s += part[(long long)b * stride + i];
    __shared__ double sh[4];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[i] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// The step's coefficients from the reduction's sums (s_i, p_i, uu, uw, ww),
// one thread, fixed order.  The candidate u = V s + alpha v with
// alpha^2 = uu - |s|^2; w's coefficient on the new v is h = (uw - s.p) / alpha
// and its remainder's norm beta^2 = ww - |p|^2 - h^2 (Pythagoras).  cf:
// s_i / alpha, p_i / beta, 1 / alpha, h / beta, 1 / beta (the update pass's
// factors), then alpha, h, beta.  A non-positive alpha^2 or beta^2 (the
// candidate or the remainder in the basis's span) gives 0 there and in the
// factors that divide by it.
__global__ void k_dgs_coef(int nv, const double *__restrict__ sum, double *__restrict__ cf)
{
    if (threadIdx.x != 0) return;
    double ss = 0.0, sp = 0.0, pp = 0.0;
    for (int i = 0; i < nv; i++) {
        ss += sum[i] * sum[i];
        sp += sum[i] * sum[nv + i];
        pp += sum[nv + i] * sum[nv + i];
    }
    const double a2 = sum[2 * nv] - ss;
    const double al = a2 > 0.0 ? sqrt(a2) : 0.0;
    const double h = al > 0.0 ? (sum[2 * nv + 1] - sp) / al : 0.0;
    const double b2 = sum[2 * nv + 2] - pp - h * h;
    const double be = b2 > 0.0 ? sqrt(b2) : 0.0;
    const double ia = al > 0.0 ? 1.0 / al : 0.0, ib = be > 0.0 ? 1.0 / be : 0.0;
    for (int i = 0; i < nv; i++) {
        cf[i] = sum[i] * ia;
        cf[nv + i] = sum[nv + i] * ib;
    }
    cf[2 * nv] = ia;
    cf[2 * nv + 1] = h * ib;
    cf[2 * nv + 2] = ib;
    cf[2 * nv + 3] = al;
    cf[2 * nv + 4] = h;
    cf[2 * nv + 5] = be;
}

// The delayed Gram-Schmidt's update, one pass over the active cells: the
// candidate becomes the final basis vector v = (u - sum_i s_i V_i) / alpha (in
// place of u), and w the next candidate (w - sum_i p_i V_i - h v) / beta (in
// place of w), with the factors of k_dgs_coef.
template <int NV>
__global__ void __launch_bounds__(256) k_dgs_update(const unsigned long long *rng, const int *__restrict__ mask,
                                                    const double *__restrict__ V, long long nc, int nv,
                                                    const double *__restrict__ cf, double *__restrict__ u,
                                                    double *__restrict__ w)
{
    __shared__ double c[2 * NV + 3];
    for (int i = threadIdx.x; i < 2 * nv + 3; i += blockDim.x) c[i] = cf[i];
    __syncthreads();
    const double ia = c[2 * nv], hb = c[2 * nv + 1], ib = c[2 * nv + 2];
    const long long n0 = (long long)rng[0], n1 = (long long)rng[1];
    for (long long n = n0 + blockIdx.x * (long long)blockDim.x + threadIdx.x; n < n1;
         n += (long long)gridDim.x * blockDim.x) {
        if (mask[n] == -1) continue;
        double vs = 0.0, vp = 0.0;
#pragma unroll
        for (int i = 0; i < NV; i++)
            if (i < nv) {
                const double vi = V[i * nc + n];
                vs += c[i] * vi;
                vp += c[nv + i] * vi;
            }
        const double v = u[n] * ia - vs;
        u[n] = v;
        w[n] = w[n] * ib - vp - hb * v;
    }
}

// w += sign * sum_i h_i V_i (active cells)
__global__ void __launch_bounds__(256) k_maxpy(const unsigned long long *rng, const int *__restrict__ mask,
                                               const double *__restrict__ V, long long nc, int nv,
                                               const double *__restrict__ h, double sign,
                                               double *__restrict__ w)
{
    __shared__ double hs[MAXM + 1];
    if ((int)threadIdx.x < nv) hs[threadIdx.x] = sign * h[threadIdx.x];
    __syncthreads();
    const long long n0 = (long long)rng[0], n1 = (long long)rng[1];
    for (long long n = n0 + blockIdx.x * (long long)blockDim.x + threadIdx.x; n < n1;
         n += (long long)gridDim.x * blockDim.x) {
        if (mask[n] == -1) continue;
        double s = w[n];
        for (int i = 0; i < nv; i++) s += hs[i] * V[i * nc + n];
        w[n] = s;
    }
}

// w *= a (active cells)
__global__ void k_vscale(const unsigned long long *rng, const int *__restrict__ mask, double a,
                         double *__restrict__ w)
{
    const long long n0 = (long long)rng[0], n1 = (long long)rng[1];
    for (long long n = n0 + blockIdx.x * (long long)blockDim.x + threadIdx.x; n < n1;
         n += (long long)gridDim.x * blockDim.x)
        if (mask[n] != -1) w[n] *= a;
}

// a = b - a, or a = b - c into a (active cells)
__global__ void k_vsub(const unsigned long long *rng, const int *__restrict__ mask,
                       const double *__restrict__ b, const double *__restrict__ c,
                       double *__restrict__ a)
{
    const long long n0 = (long long)rng[0], n1 = (long long)rng[1];
    for (long long n = n0 + blockIdx.x * (long long)blockDim.x + threadIdx.x; n < n1;
         n += (long long)gridDim.x * blockDim.x)
        if (mask[n] != -1) a[n] = b[n] - c[n];
}

// a = b (active cells; every cell of the range when mask is null)
__global__ void k_vcopy(const unsigned long long *rng, const int *__restrict__ mask,
                        const double *__restrict__ b, double *__restrict__ a)
{
    const long long n0 = (long long)rng[0], n1 = (long long)rng[1];
    for (long long n = n0 + blockIdx.x * (long long)blockDim.x + threadIdx.x; n < n1;
         n += (long long)gridDim.x * blockDim.x)
        if (!mask || mask[n] != -1) a[n] = b[n];
}

__global__ void k_vzero(long long n0, long long n1, double *__restrict__ a)
{
    for (long long n = n0 + blockIdx.x * (long long)blockDim.x + threadIdx.x; n < n1;
         n += (long long)gridDim.x * blockDim.x)
        a[n] = 0.0;
}

// partial sums of r^2 and of the active count
__global__ void __launch_bounds__(256) k_sumsq(const unsigned long long *rng, const int *__restrict__ mask,
                                               const double *__restrict__ r, double *__restrict__ part)
{
    double a = 0.0, c = 0.0;
    const long long n0 = (long long)rng[0], n1 = (long long)rng[1];
    for (long long n = n0 + blockIdx.x * (long long)blockDim.x + threadIdx.x; n < n1;
         n += (long long)gridDim.x * blockDim.x)
        if (mask[n] != -1) {
            a += r[n] * r[n];
            c += 1.0;
        }
    __shared__ double sh[2][4];
    a = wave_sum(a);
    c = wave_sum(c);
    if ((threadIdx.x & 63) == 0) {
        sh[0][threadIdx.x >> 6] = a;
        sh[1][threadIdx.x >> 6] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[(long long)blockIdx.x * (MAXM + 1)] = sh[0][0] + sh[0][1] + sh[0][2] + sh[0][3];
        part[(long long)blockIdx.x * (MAXM + 1) + 1] = sh[1][0] + sh[1][1] + sh[1][2] + sh[1][3];
    }
}

// s = sqrt(sum r^2 / count), 1/s (1 when r is zero)
__global__ void __launch_bounds__(256) k_rms_final(const double *__restrict__ part, int nblk,
                                                   double *__restrict__ scal)
{
    double a = 0.0, c = 0.0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
        a += part[(long long)b * (MAXM + 1)];
        c += part[(long long)b * (MAXM + 1) + 1];
    }
    __shared__ double sh[2][4];
    a = wave_sum(a);
    c = wave_sum(c);
    if ((threadIdx.x & 63) == 0) {
        sh[0][threadIdx.x >> 6] = a;
        sh[1][threadIdx.x >> 6] = c;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    a = (sh[0][0] + sh[0][1]) + (sh[0][2] + sh[0][3]);
    c = (sh[1][0] + sh[1][1]) + (sh[1][2] + sh[1][3]);
    const double s = c > 0.0 ? sqrt(a / c) : 0.0;
    scal[0] = s;
    scal[1] = s > 0.0 ? 1.0 / s : 1.0;
    scal[2] = c;
}

#define RET_IF(x)              \
    do {                       \
        int rc_ = (x);         \
        if (rc_) return rc_;   \
    } while (0)

int lalloc(void **p, size_t bytes)
{
    if (hipMalloc(p, bytes) != hipSuccess) {
        cad_set_error("learned form: hipMalloc of %zu bytes failed", bytes);
        *p = nullptr;
        return CAD_ERR_HIP;
    }
    return CAD_OK;
}

int learned_init(cad_ctx *c)
{
    if (c->lrn) return CAD_OK;
    Learned *L = new Learned();
    std::memset((void *)L, 0, sizeof(Learned));
    c->lrn = L;
    const Dims &d = c->d;
    L->nx = ((d.ni - 2) + 63) / 64 * 64;
    L->ny = ((d.nj - 2) + 3) / 4 * 4;
    L->nzmax = ((d.nk - 2) + 3) / 4 * 4;
    L->mode = 1;
    L->m = 30;
    CAD_CHECK(hipSetDevice(c->device));
    RET_IF(cad_unet_create(L->nx, L->ny, L->nzmax, 1, 0, c->stream, &L->net));
    L->nvec = 2 * MAXM + 1 + 6;
    RET_IF(lalloc((void **)&L->vecs, (size_t)L->nvec * d.nc * sizeof(double)));
    RET_IF(lalloc((void **)&L->part, (size_t)REDUCE_BLOCKS * GS_STRIDE * sizeof(double)));
    RET_IF(lalloc((void **)&L->hdev, 2 * GS_STRIDE * sizeof(double)));
    RET_IF(lalloc((void **)&L->scal, 4 * sizeof(double)));
    RET_IF(lalloc((void **)&L->y, (size_t)L->nx * L->ny * L->nzmax * sizeof(float)));
    if (hipHostMalloc((void **)&L->hpin, 2 * GS_STRIDE * sizeof(double), hipHostMallocDefault) != hipSuccess) {
        cad_set_error("learned form: hipHostMalloc failed");
        return CAD_ERR_HIP;
    }
    CAD_CHECK(hipEventCreateWithFlags(&L->gs_read, hipEventDisableTiming));
    hipLaunchKernelGGL(k_vzero, dim3(VBLK), dim3(256), 0, c->stream, 0LL, (long long)L->nvec * d.nc, L->vecs);
    CAD_CHECK(hipGetLastError());
    // weights until the caller sets trained ones: the library's initialiser
    std::vector<float> p(cad_unet_param_count());
    RET_IF(cad_unet_init_params(0, p.data()));
    RET_IF(cad_unet_set_params(L->net, p.data()));
    L->prev_rng[0] = 0;
    L->prev_rng[1] = (unsigned long long)d.nc;
    return CAD_OK;
}

double *vec(cad_ctx *c, int i) { return c->lrn->vecs + (long long)i * c->d.nc; }
// V_j = vec(j), Z_j = vec(MAXM+1+j), scratch S_q = vec(2*MAXM+1+q)
double *Vv(cad_ctx *c, int j) { return vec(c, j); }
double *Zv(cad_ctx *c, int j) { return vec(c, MAXM + 1 + j); }
double *Sv(cad_ctx *c, int q) { return vec(c, 2 * MAXM + 1 + q); }

int rms(cad_ctx *c, const double *r)
{
    Learned *L = c->lrn;
    hipLaunchKernelGGL(k_sumsq, dim3(REDUCE_BLOCKS), dim3(256), 0, c->stream, c->act_rng, c->f.indc, r, L->part);
    hipLaunchKernelGGL(k_rms_final, dim3(1), dim3(256), 0, c->stream, L->part, REDUCE_BLOCKS, L->scal);
    CAD_CHECK(hipGetLastError());
    return CAD_OK;
}

// features of r into the destination (network input or a sample buffer),
// scaled by scal (set by rms, or from partials already in L->part)
int features_scaled(cad_ctx *c, const double *r, bf16 *dst, const UGeo &g)
{
    Learned *L = c->lrn;
    hipLaunchKernelGGL(k_features, dim3(4096), dim3(256), 0, c->stream, c->d, g, L->kb, L->nzr, r, L->scal,
                       c->ggv, c->f.ff, c->f.nf, c->f.indc, dst);
    CAD_CHECK(hipGetLastError());
    return CAD_OK;
}

int features(cad_ctx *c, const double *r, bf16 *dst, const UGeo &g)
{
    RET_IF(rms(c, r));
    return features_scaled(c, r, dst, g);
}

// z = P(v)
int precond(cad_ctx *c, const double *v, double *z)
{
    Learned *L = c->lrn;
    const double *z1 = nullptr;
    UGeo g = cad_unet_geo(L->net);
    g.nz = L->nzr;
    bf16 *xin = reinterpret_cast<bf16 *>(cad_unet_input(L->net));
    if (L->mode == 1) {
        // t0 = V(v); t1 = v - A t0 on the V-cycle's FP32 level-0 matrix (the
        // network's input needs no more), with its sum of squares
        double *t0 = Sv(c, 0), *t1 = Sv(c, 1);
        RET_IF(cad_mg_apply_resid(c, v, t0, t1, L->part, MAXM + 1, c->mg_k0, c->mg_k1));
        hipLaunchKernelGGL(k_rms_final, dim3(1), dim3(256), 0, c->stream, L->part, REDUCE_BLOCKS, L->scal);
        CAD_CHECK(hipGetLastError());
        if (L->xin_static) {
            hipLaunchKernelGGL(k_features_r, dim3(4096), dim3(256), 0, c->stream, c->d, g, L->kb, L->nzr, t1,
                               L->scal, c->f.indc, xin);
            CAD_CHECK(hipGetLastError());
        } else {
            RET_IF(features_scaled(c, t1, xin, g));
            L->xin_static = true;
        }
        z1 = t0;
    } else {
        RET_IF(features(c, v, xin, g));
        L->xin_static = true;
    }
    RET_IF(cad_kt_begin(c, CAD_KT_UNET));
    RET_IF(cad_unet_forward_impl(L->net));
    RET_IF(cad_kt_end(c, CAD_KT_UNET, 2.0 * cad_unet_mac_per_voxel() * (double)L->nx * L->ny * L->nzr));
    hipLaunchKernelGGL(k_scatter, dim3(VBLK), dim3(256), 0, c->stream, c->d, c->act_rng, L->nx, L->ny, L->kb,
                       L->nzr, c->f.indc, z1, L->scal, cad_unet_output(L->net), z);
    CAD_CHECK(hipGetLastError());
    return CAD_OK;
}

// the solve's set-up: active range, multigrid levels, the network's box
int setup(cad_ctx *c, bool *empty)
{
    Learned *L = c->lrn;
    const Dims &d = c->d;
    unsigned long long hr[2];
    RET_IF(cad_ext_active_range(c, hr));
    *empty = hr[0] >= hr[1];
    // vectors read by the matvec must be 0 outside the new active range.
    // Every solve writes its whole range (0 on the inactive cells in it) and
    // nothing outside it, so only the cells of the previous range that fall
    // outside the new one can hold values: clear those.
    const long long z0 = (long long)L->prev_rng[0], z1 = (long long)L->prev_rng[1];
    const long long n0 = (long long)hr[0], n1 = (long long)hr[1];
    const long long cut[2][2] = {{z0, std::min(z1, n0 < n1 ? n0 : z1)}, {std::max(z0, n0 < n1 ? n1 : z1), z1}};
    for (const auto &iv : cut) {
        if (iv[1] <= iv[0]) continue;
        for (int j = 0; j < MAXM; j++)
            hipLaunchKernelGGL(k_vzero, dim3(VBLK), dim3(256), 0, c->stream, iv[0], iv[1], Zv(c, j));
        for (int q = 0; q < 6; q++)
            hipLaunchKernelGGL(k_vzero, dim3(VBLK), dim3(256), 0, c->stream, iv[0], iv[1], Sv(c, q));
    }
    CAD_CHECK(hipGetLastError());
    L->prev_rng[0] = hr[0];
    L->prev_rng[1] = hr[1];
    if (*empty) return CAD_OK;
    c->mg_span = hr[1] - hr[0];
    const int k0 = (int)(hr[0] / (unsigned long long)d.nij) + 1;
    const int k1 = (int)((hr[1] - 1) / (unsigned long long)d.nij) + 1;
    c->mg_k0 = k0;
    c->mg_k1 = k1;
    if (L->mode == 1) RET_IF(cad_mg_setup(c));
    L->kb = k0;
    L->nzr = std::min(((k1 - k0 + 1) + 3) / 4 * 4, L->nzmax);
    RET_IF(cad_unet_set_depth(L->net, L->nzr));
    L->xin_static = false;   // new flags: the network input is written whole once
    return CAD_OK;
}

}  // namespace

void cad_learned_free(cad_ctx *c)
{
    Learned *L = c->lrn;
    if (!L) return;
    if (L->net) cad_unet_destroy(L->net);
    void *p[] = {L->vecs, L->part, L->hdev, L->scal, L->y};
    for (void *q : p)
        if (q) (void)hipFree(q);
    if (L->hpin) (void)hipHostFree(L->hpin);
    if (L->gs_read) (void)hipEventDestroy(L->gs_read);
    delete L;
    c->lrn = nullptr;
}

int cad_solve_learned(cad_ctx *c, double *x0, double *bnorm_out, double *xnorm_out, int *iter_out)
{
    RET_IF(learned_init(c));
    Learned *L = c->lrn;
    const Dims &d = c->d;
    const long long nc = d.nc;
    const int *mask = c->f.indc;
    hipStream_t st = c->stream;
    bool empty = false;
    RET_IF(setup(c, &empty));
    L->last_iters = 0;
    L->last_fallback_iters = 0;
    L->solves++;
    if (empty) {
        *bnorm_out = 0.0;
        *xnorm_out = 0.0;
        *iter_out = 0;
        return CAD_OK;
    }
    double bn2 = 0.0;
    RET_IF(cad_ext_dot(c, c->bb, c->bb, &bn2));
    const double bnorm = std::sqrt(bn2);
    // the reference's stop on the true residual, held to the context's true
    // relative residual when one is set
    double tol = c->prm.cgepsr * bnorm + c->prm.cgepsa;
    if (c->true_stop > 0.0) tol = std::fmin(tol, c->true_stop * bnorm);
    double *x = Sv(c, 2), *r = Sv(c, 3);
    // x is x0 on the whole range, the inactive cells' zeros included: the
    // matvec reads them as neighbours, so the residual the solve stops on is
    // then the one of the x0 it returns (the masked copy back leaves those
    // cells of x0 as they are)
    hipLaunchKernelGGL(k_vcopy, dim3(VBLK), dim3(256), 0, st, c->act_rng, nullptr, x0, x);
    CAD_CHECK(hipGetLastError());
    const int m = L->m;
    const int cap = cad_krylov_cap(c);
    int iters = 0;
    double rnorm = 0.0;
    bool done = false, stalled = false;
    // R: the rotated Hessenberg columns (R[i * m + k]); Hr: the unrotated
    // columns awaiting their final entries (Hr[k * (m + 1) + i])
    std::vector<double> H((size_t)(m + 1) * m), Hr((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), y(m),
        col(m + 1);
    double prev = -1.0;
    // column k of R from col (rows 0..k+1, unrotated): the final rotations
    // G_0..G_{k-1}, then G_k from the column itself; returns |g_{k+1}| as it
    // would be, and applies G_k to g only if commit
    auto rotate = [&](int k, double sub, bool commit) {
        for (int i = 0; i < k; i++) {
            const double a = cs[i] * col[i] + sn[i] * col[i + 1];
            col[i + 1] = -sn[i] * col[i] + cs[i] * col[i + 1];
            col[i] = a;
        }
        const double den = std::hypot(col[k], sub);
        const double ck = den > 0.0 ? col[k] / den : 1.0, sk = den > 0.0 ? sub / den : 0.0;
        const double est = std::fabs(sk * g[k]);
        if (commit) {
            cs[k] = ck;
            sn[k] = sk;
            col[k] = den;
            for (int i = 0; i <= k; i++) H[(size_t)i * m + k] = col[i];
            g[k + 1] = -sk * g[k];
            g[k] = ck * g[k];
        }
        return est;
    };
    while (!done && !stalled) {
        // r = BB - A x into V_0
        RET_IF(cad_ext_residual(c, x, Vv(c, 0)));
        double rr = 0.0;
        RET_IF(cad_ext_dot(c, Vv(c, 0), Vv(c, 0), &rr));
        rnorm = std::sqrt(rr);
        if (!(rnorm > tol) || !std::isfinite(rnorm)) {
            done = std::isfinite(rnorm);
            stalled = !done;
            break;
        }
        // a cycle that did not halve the true residual, or the iteration cap,
        // hands over to VF_M1BCGS
        if (iters >= cap || (prev > 0.0 && !(rnorm < 0.5 * prev))) {
            stalled = true;
            break;
        }
        prev = rnorm;
        hipLaunchKernelGGL(k_vscale, dim3(VBLK), dim3(256), 0, st, c->act_rng, mask, 1.0 / rnorm, Vv(c, 0));
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = rnorm;
        // Arnoldi with classical Gram-Schmidt done twice, the second pass
        // delayed by one step (DCGS2: Swirydowicz, Langou, Ananthan, Yang and
        // Thomas, Numer. Linear Algebra Appl. 28, 2021): step j preconditions
        // the candidate u_j in slot j (V_0 at j = 0), orthogonal to V_0..V_{j-1}
        // after one pass, and one reduction over the basis gives both its
        // second pass (s = V^T u_j) and w's first (p = V^T w).  The update pass
        // then writes the final v_j = (u_j - V s) / alpha and the next
        // candidate.  Column j-1 of the Hessenberg matrix gets its last terms
        // here (u_j = V s + alpha v_j), column j is provisional until the next
        // step; one read-back per step.
        int j = 0;            // columns in R
        double bprev = 0.0;   // the previous column's provisional subdiagonal
        for (;;) {
            const int nv = j;
            RET_IF(precond(c, Vv(c, j), Zv(c, j)));
            double *w = Vv(c, j + 1);
            RET_IF(cad_ext_apply_a(c, Zv(c, j), w));
            double *sum = L->hdev, *cf = L->hdev + GS_STRIDE;
            const bool last = j + 1 == m || iters + 1 >= cap;
            // register arrays sized to the basis: 8, 16, 24 or 32 vectors
#define CAD_GS(NVB)                                                                                              \
    do {                                                                                                         \
        hipLaunchKernelGGL(k_dgs_reduce<NVB>, dim3(REDUCE_BLOCKS), dim3(256), 0, st, c->act_rng, mask, Vv(c, 0),   \
                           nc, nv, Vv(c, j), w, L->part);                                                        \
        hipLaunchKernelGGL(k_cols_final, dim3(2 * nv + 3), dim3(256), 0, st, L->part, REDUCE_BLOCKS, 2 * nv + 3,  \
                           GS_STRIDE, sum);                                                                      \
        hipLaunchKernelGGL(k_dgs_coef, dim3(1), dim3(64), 0, st, nv, sum, cf);                                   \
        CAD_CHECK(hipMemcpyAsync(L->hpin, L->hdev, 2 * GS_STRIDE * sizeof(double), hipMemcpyDeviceToHost, st)); \
        CAD_CHECK(hipEventRecord(L->gs_read, st));                                                               \
        if (!last)                                                                                               \
            hipLaunchKernelGGL(k_dgs_update<NVB>, dim3(VBLK), dim3(256), 0, st, c->act_rng, mask, Vv(c, 0), nc,  \
                               nv, cf, Vv(c, j), w);                                                             \
    } while (0)
            if (nv <= 8)
                CAD_GS(8);
            else if (nv <= 16)
                CAD_GS(16);
            else if (nv <= 24)
                CAD_GS(24);
            else
                CAD_GS(MAXM);
#undef CAD_GS
            CAD_CHECK(hipGetLastError());
            // the update pass (if any) runs on while the host reads
            CAD_CHECK(hipEventSynchronize(L->gs_read));
            iters++;
            const double *hs = L->hpin, *hc = L->hpin + GS_STRIDE;
            const double al = hc[2 * nv + 3], hj = hc[2 * nv + 4], be = hc[2 * nv + 5];
            if (j == 0) {
                g[0] = rnorm * al;   // r = rnorm u_0 = rnorm alpha v_0
            } else {
                // column j-1 final: A z_{j-1} = ... + bprev u_j, u_j = V s + alpha v_j
                double *hk = &Hr[(size_t)(j - 1) * (m + 1)];
                for (int i = 0; i < j; i++) col[i] = hk[i] + bprev * hs[i];
                rotate(j - 1, bprev * al, true);
                if (!(al > 0.0)) break;   // u_j in the span: the space is invariant
            }
            // column j, provisional: p_i, h_j and the remainder's norm
            for (int i = 0; i < j; i++) col[i] = hs[nv + i];
            col[j] = hj;
            double *hk = &Hr[(size_t)j * (m + 1)];
            for (int i = 0; i <= j; i++) hk[i] = col[i];
            const double est = rotate(j, be, false);
            if (!(be > 0.0) || !(est > tol) || !std::isfinite(est) || last) {
                for (int i = 0; i <= j; i++) col[i] = hk[i];
                rotate(j, be, true);
                j++;
                break;
            }
            bprev = be;
            j++;
        }
        // y = H^-1 g (upper triangular), x += Z y
        for (int i = j - 1; i >= 0; i--) {
            double s = g[i];
            for (int k = i + 1; k < j; k++) s -= H[(size_t)i * m + k] * y[k];
            y[i] = H[(size_t)i * m + i] != 0.0 ? s / H[(size_t)i * m + i] : 0.0;
        }
        CAD_CHECK(hipMemcpyAsync(L->hdev, y.data(), j * sizeof(double), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_maxpy, dim3(VBLK), dim3(256), 0, st, c->act_rng, mask, Zv(c, 0), nc, j, L->hdev,
                           1.0, x);
        CAD_CHECK(hipGetLastError());
    }
    L->last_iters = iters;
    L->total_iters += iters;
    int it = iters;
    double xrel = bnorm > 0.0 ? rnorm / bnorm : rnorm;
    if (!done) {
        // fallback: the reference's MILU-BiCGSTAB from the current x, held to
        // the same true relative residual
        L->fallbacks++;
        const double keep_stop = c->true_stop;
        c->true_stop = keep_stop > 0.0 ? std::fmin(keep_stop, c->prm.cgepsr) : c->prm.cgepsr;
        hipLaunchKernelGGL(k_vcopy, dim3(VBLK), dim3(256), 0, st, c->act_rng, mask, x, r);
        CAD_CHECK(hipGetLastError());
        double bn = 0.0, xn = 0.0;
        int fit = 0;
        const int rc = cad_solve_m1bcgs(c, r, &bn, &xn, &fit);
        c->true_stop = keep_stop;
        RET_IF(rc);
        L->last_fallback_iters = fit;
        if (fit < 0 || fit > c->prm.icgmax) {
            *bnorm_out = bnorm;
            *xnorm_out = xrel;
            *iter_out = fit < 0 ? fit : c->prm.icgmax + 1;
            return CAD_OK;
        }
        hipLaunchKernelGGL(k_vcopy, dim3(VBLK), dim3(256), 0, st, c->act_rng, mask, r, x);
        CAD_CHECK(hipGetLastError());
        RET_IF(cad_ext_residual(c, x, Sv(c, 4)));
        double rr = 0.0;
        RET_IF(cad_ext_dot(c, Sv(c, 4), Sv(c, 4), &rr));
        xrel = bnorm > 0.0 ? std::sqrt(rr) / bnorm : std::sqrt(rr);
        it = iters + fit;
    }
    L->last_resid = xrel;
    hipLaunchKernelGGL(k_vcopy, dim3(VBLK), dim3(256), 0, st, c->act_rng, mask, x, x0);
    CAD_CHECK(hipGetLastError());
    *bnorm_out = bnorm;
    *xnorm_out = xrel;
    *iter_out = it;
    return CAD_OK;
}

extern "C" {

int cad_set_learned_weights(cad_ctx *c, const float *params, int n)
{
    if (!c || !params || n != cad_unet_param_count()) {
        cad_set_error("cad_set_learned_weights: %d parameters, the network has %d", n, cad_unet_param_count());
        return CAD_ERR_ARG;
    }
    RET_IF(learned_init(c));
    return cad_unet_set_params(c->lrn->net, params);
}

int cad_set_learned_mode(cad_ctx *c, int mode, int restart)
{
    if (!c || mode < 0 || mode > 1 || restart < 1 || restart > MAXM) {
        cad_set_error("cad_set_learned_mode: mode %d (0 network, 1 V-cycle + network), restart %d (1..%d)",
                      mode, restart, MAXM);
        return CAD_ERR_ARG;
    }
    RET_IF(learned_init(c));
    c->lrn->mode = mode;
    c->lrn->m = restart;
    return CAD_OK;
}

// Average duration of one forward pass of the network on the box of the
// context's current flags (input: the right-hand side's features), timed with
// HIP events on the context's stream, and its multiply-add work counted as 2
// FLOP per MAC (the 3x3x3 layers at their levels and the output layer).
int cad_time_learned(cad_ctx *c, int reps, float *ms_avg, double *flops)
{
    if (!c || !ms_avg || !flops) return CAD_ERR_ARG;
    if (reps < 1) reps = 1;
    RET_IF(learned_init(c));
    Learned *L = c->lrn;
    bool empty = false;
    RET_IF(setup(c, &empty));
    if (empty) {
        cad_set_error("cad_time_learned: no active cell");
        return CAD_ERR_STATE;
    }
    UGeo g = cad_unet_geo(L->net);
    g.nz = L->nzr;
    RET_IF(features(c, c->bb, reinterpret_cast<bf16 *>(cad_unet_input(L->net)), g));
    RET_IF(cad_unet_forward_impl(L->net));
    hipEvent_t a, b;
    CAD_CHECK(hipEventCreate(&a));
    CAD_CHECK(hipEventCreate(&b));
    CAD_CHECK(hipEventRecord(a, c->stream));
    for (int r = 0; r < reps; r++) RET_IF(cad_unet_forward_impl(L->net));
    CAD_CHECK(hipEventRecord(b, c->stream));
    CAD_CHECK(hipEventSynchronize(b));
    float ms = 0.f;
    CAD_CHECK(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *ms_avg = ms / reps;
    *flops = 2.0 * cad_unet_mac_per_voxel() * (double)L->nx * L->ny * L->nzr;
    return CAD_OK;
}

int cad_learned_stats(cad_ctx *c, long long *out)
{
    if (!c || !out) return CAD_ERR_ARG;
    const Learned *L = c->lrn;
    if (!L) {
        for (int i = 0; i < 6; i++) out[i] = 0;
        return CAD_OK;
    }
    out[0] = L->last_iters;
    out[1] = L->last_fallback_iters;
    out[2] = L->solves;
    out[3] = L->fallbacks;
    out[4] = L->total_iters;
    out[5] = (long long)L->nzr;
    return CAD_OK;
}

// Box of the network for the context's current flags: nx, ny, nz (the run's
// depth) and kb (the first K plane).
int cad_learned_box(cad_ctx *c, int *box)
{
    if (!c || !box) return CAD_ERR_ARG;
    RET_IF(learned_init(c));
    bool empty = false;
    RET_IF(setup(c, &empty));
    box[0] = c->lrn->nx;
    box[1] = c->lrn->ny;
    box[2] = c->lrn->nzr;
    box[3] = c->lrn->kb;
    return CAD_OK;
}

// The context's current matrix on the network's box (k_box_matrix), for
// training losses that apply it; out holds 7 * nx*ny*nz floats.
int cad_learned_box_matrix(cad_ctx *c, float *out)
{
    if (!c || !out) {
        cad_set_error("cad_learned_box_matrix: null argument");
        return CAD_ERR_ARG;
    }
    RET_IF(learned_init(c));
    bool empty = false;
    RET_IF(setup(c, &empty));
    const Learned *L = c->lrn;
    hipLaunchKernelGGL(k_box_matrix, dim3(4096), dim3(256), 0, c->stream, c->d, L->nx, L->ny, L->kb, L->nzr,
                       c->ad, c->ali, c->aui, c->alj, c->auj, c->alk, c->auk, out);
    CAD_CHECK(hipGetLastError());
    CAD_CHECK(hipStreamSynchronize(c->stream));
    return CAD_OK;
}

// Training pairs from the plain solver on the context's current system (the
// matrix and right-hand side of its last step): x* from a full MILU-BiCGSTAB
// solve, and for each k of iters the k-th iterate x_k of the same solve from
// 0; the pair is the residual r = BB - A x_k and the correction e = x* - x_k,
// in mode 1 after one V-cycle z1 = V(r): r - A z1 and e - z1.  Written per
// pair: the network input (box of cad_learned_box, one-voxel zero halo, 8
// BF16 channels, (nz+2)(ny+2)(nx+2) voxels), the target e / s and the
// activity mask (FP32, nx*ny*nz voxels, x fastest).
int cad_learned_sample(cad_ctx *c, int mode, const int *iters, int n, void *feat, float *target, float *mask)
{
    if (!c || !iters || n < 1 || !feat || !target || !mask || mode < 0 || mode > 1) {
        cad_set_error("cad_learned_sample: bad argument");
        return CAD_ERR_ARG;
    }
    RET_IF(learned_init(c));
    Learned *L = c->lrn;
    const int keep_mode = L->mode;
    L->mode = mode;
    bool empty = false;
    int rc = setup(c, &empty);
    L->mode = keep_mode;
    RET_IF(rc);
    if (empty) {
        cad_set_error("cad_learned_sample: no active cell");
        return CAD_ERR_STATE;
    }
    const long long nc = c->d.nc;
    hipStream_t st = c->stream;
    const int *msk = c->f.indc;
    double *xs = Sv(c, 0), *xk = Sv(c, 1), *r = Sv(c, 2), *e = Sv(c, 3), *z1 = Sv(c, 4), *t = Sv(c, 5);
    const cad_params keep = c->prm;
    const double keep_stop = c->true_stop;
    c->prm.cgepsr = 1.0e-10;
    c->prm.cgepsa = 0.0;
    c->prm.icgmax = 2000;
    c->true_stop = 1.0e-10;
    double bn = 0.0, xn = 0.0;
    int it = 0;
    hipLaunchKernelGGL(k_vzero, dim3(VBLK), dim3(256), 0, st, 0LL, nc, xs);
    rc = cad_solve_m1bcgs(c, xs, &bn, &xn, &it);
    c->true_stop = keep_stop;
    if (!rc && (it < 0 || it > c->prm.icgmax)) {
        cad_set_error("cad_learned_sample: the reference solve did not converge (ITER %d)", it);
        rc = CAD_ERR_STATE;
    }
    UGeo g;
    g.nx = L->nx;
    g.ny = L->ny;
    g.nz = L->nzr;
    g.nb = 1;
    g.px = g.nx + 2;
    g.py = g.ny + 2;
    g.pz = g.nz + 2;
    const long long fstride = (long long)g.px * g.py * g.pz * 8;
    const long long tstride = (long long)g.nx * g.ny * g.nz;
    for (int q = 0; q < n && !rc; q++) {
        hipLaunchKernelGGL(k_vzero, dim3(VBLK), dim3(256), 0, st, 0LL, nc, xk);
        c->prm.icgmax = iters[q];
        rc = cad_solve_m1bcgs(c, xk, &bn, &xn, &it);
        if (rc) break;
        hipLaunchKernelGGL(k_vsub, dim3(VBLK), dim3(256), 0, st, c->act_rng, msk, xs, xk, e);
        CAD_CHECK(hipGetLastError());
        if ((rc = cad_ext_residual(c, xk, r))) break;
        if (mode == 1) {
            if ((rc = cad_mg_apply(c, r, z1, c->mg_k0, c->mg_k1))) break;
            if ((rc = cad_ext_apply_a(c, z1, t))) break;
            hipLaunchKernelGGL(k_vsub, dim3(VBLK), dim3(256), 0, st, c->act_rng, msk, r, t, r);
            hipLaunchKernelGGL(k_vsub, dim3(VBLK), dim3(256), 0, st, c->act_rng, msk, e, z1, e);
            CAD_CHECK(hipGetLastError());
        }
        bf16 *fd = reinterpret_cast<bf16 *>(feat) + q * fstride;
        hipLaunchKernelGGL(k_vzero, dim3(64), dim3(256), 0, st, 0LL, fstride / 4,
                           reinterpret_cast<double *>(fd));
        if ((rc = features(c, r, fd, g))) break;
        hipLaunchKernelGGL(k_box_target, dim3(4096), dim3(256), 0, st, c->d, g.nx, g.ny, L->kb, L->nzr, msk, e,
                           L->scal, target + q * tstride, mask + q * tstride);
        CAD_CHECK(hipGetLastError());
    }
    c->prm = keep;
    RET_IF(rc);
    CAD_CHECK(hipStreamSynchronize(st));
    return CAD_OK;
}

}  // extern "C"
