// cad_mg.hip -- FP32 aggregation multigrid, the preconditioner of the mixed form.
//
// The mixed form solves the pressure system of VF_VPCOEF (the same FP64
// matrix AD..AUK, BB as the plain form) with an FP64 right-preconditioned
// BiCGSTAB (cad_poisson.hip) whose preconditioner is one V-cycle of this
// multigrid in FP32.  It replaces the reference's MILU preconditioner
// (vf_mzdcmp.f, vf_mzminv.f), whose triangular sweeps are a serial chain:
// every multigrid pass is a data-parallel 7-point stencil bound by HBM
// bandwidth.
//
// Levels.  Level 0 is the pressure grid with the matrix rounded to FP32;
// inactive rows (INDC = -1) get AD = 1 and no couplings.  A coarse cell
// aggregates the 2x2x2 block of fine cells (fine interior index 2I-2, 2I-1 of
// coarse interior index I, 1-based with one ghost layer), and its row is the
// Galerkin product with piecewise-constant transfer over the active fine
// cells: the coarse diagonal sums the fine diagonals and the couplings
// between cells of the block, a coarse coupling sums the fine couplings that
// cross the block face.  The structure stays 7-point on every level.  A
// coarse row without active fine cells, or with a diagonal that is not
// positive, is inactive.
//
// V-cycle.  Red-black Gauss-Seidel (one sweep = the cells with I+J+K even,
// then odd) before and after the coarse correction; the residual of the fine
// cells is summed into the coarse right-hand side in one kernel; the coarse
// correction is added scaled by MG_SCALE (piecewise-constant prolongation
// under-corrects smooth errors); the coarsest level is smoothed MG_COARSE
// times.  Every level starts from x = 0, so the first (even) half of its
// first sweep is x = b/AD, which the kernel that sets b writes at once.  The
// levels of at most MG_TAIL_CELLS cells run as one workgroup (k_mg_tail), so
// the small grids cost one launch instead of five per level.
#include "cad_common.h"

#define RET_IF(x)              \
    do {                       \
        const int rc_ = (x);   \
        if (rc_ != CAD_OK)     \
            return rc_;        \
    } while (0)

namespace {

constexpr float MG_SCALE = 1.8f;
constexpr int MG_COARSE = 12;
constexpr long long MG_TAIL_CELLS = 4096;

__device__ __forceinline__ long long mid(const Dims &d, int i, int j, int k)
{
    return (long long)(i - 1) + (long long)d.ni * (j - 1) + d.nij * (k - 1);
}

// level 0 from the FP64 matrix
__global__ void __launch_bounds__(256) k_mg_fine(Dims d, const double *__restrict__ ad,
                                                 const double *__restrict__ ali,
                                                 const double *__restrict__ alj,
                                                 const double *__restrict__ alk,
                                                 const double *__restrict__ aui,
                                                 const double *__restrict__ auj,
                                                 const double *__restrict__ auk,
                                                 const int *__restrict__ indc, MgLevel L)
{
    const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= d.nc) return;
    const bool a = indc[n] != -1 && ad[n] > 0.0;
    L.act[n] = a ? 1 : 0;
    L.ad[n] = a ? (float)ad[n] : 1.0f;
    L.ali[n] = a ? (float)ali[n] : 0.0f;
    L.alj[n] = a ? (float)alj[n] : 0.0f;
    L.alk[n] = a ? (float)alk[n] : 0.0f;
    L.aui[n] = a ? (float)aui[n] : 0.0f;
    L.auj[n] = a ? (float)auj[n] : 0.0f;
    L.auk[n] = a ? (float)auk[n] : 0.0f;
}

// Galerkin coarse row of coarse interior cell (I,J,K); one thread per cell
__global__ void __launch_bounds__(256) k_mg_coarsen(MgLevel F, MgLevel C)
{
    const int I = blockIdx.x * blockDim.x + threadIdx.x + 2;
    const int J = blockIdx.y * blockDim.y + threadIdx.y + 2;
    const int K = blockIdx.z * blockDim.z + threadIdx.z + 2;
    if (I > C.d.ni - 1 || J > C.d.nj - 1 || K > C.d.nk - 1) return;
    const Dims &fd = F.d;
    float s_ad = 0.0f, s_li = 0.0f, s_lj = 0.0f, s_lk = 0.0f, s_ui = 0.0f, s_uj = 0.0f, s_uk = 0.0f;
    bool any = false;
    for (int c = 0; c < 8; c++) {
        const int oi = c & 1, oj = (c >> 1) & 1, ok = c >> 2;
        const int i = 2 * I - 2 + oi, j = 2 * J - 2 + oj, k = 2 * K - 2 + ok;
        if (i > fd.ni - 1 || j > fd.nj - 1 || k > fd.nk - 1) continue;
        const long long n = mid(fd, i, j, k);
        if (!F.act[n]) continue;
        any = true;
        s_ad += F.ad[n];
        // -I: inside the block when oi = 1
        if (F.ali[n] != 0.0f && F.act[n - 1]) {
            if (oi) s_ad += F.ali[n]; else s_li += F.ali[n];
        }
        if (F.aui[n] != 0.0f && F.act[n + 1]) {
            if (!oi && i + 1 <= fd.ni - 1) s_ad += F.aui[n]; else s_ui += F.aui[n];
        }
        if (F.alj[n] != 0.0f && F.act[n - fd.ni]) {
            if (oj) s_ad += F.alj[n]; else s_lj += F.alj[n];
        }
        if (F.auj[n] != 0.0f && F.act[n + fd.ni]) {
            if (!oj && j + 1 <= fd.nj - 1) s_ad += F.auj[n]; else s_uj += F.auj[n];
        }
        if (F.alk[n] != 0.0f && F.act[n - fd.nij]) {
            if (ok) s_ad += F.alk[n]; else s_lk += F.alk[n];
        }
        if (F.auk[n] != 0.0f && F.act[n + fd.nij]) {
            if (!ok && k + 1 <= fd.nk - 1) s_ad += F.auk[n]; else s_uk += F.auk[n];
        }
    }
    const long long m = mid(C.d, I, J, K);
    const bool a = any && s_ad > 0.0f;
    C.act[m] = a ? 1 : 0;
    C.ad[m] = a ? s_ad : 1.0f;
    C.ali[m] = a ? s_li : 0.0f;
    C.alj[m] = a ? s_lj : 0.0f;
    C.alk[m] = a ? s_lk : 0.0f;
    C.aui[m] = a ? s_ui : 0.0f;
    C.auj[m] = a ? s_uj : 0.0f;
    C.auk[m] = a ? s_uk : 0.0f;
}

__device__ __forceinline__ float row_off(const MgLevel &L, long long n)
{
    const long long ni = L.d.ni, nij = L.d.nij;
    const float *x = L.x;
    return L.ali[n] * x[n - 1] + L.aui[n] * x[n + 1] + L.alj[n] * x[n - ni] + L.auj[n] * x[n + ni]
         + L.alk[n] * x[n - nij] + L.auk[n] * x[n + nij];
}

// one colour of a Gauss-Seidel sweep: the cells with (I+J+K) % 2 == color,
// thread x covering every second I of its (J,K) line, K in [k0, k1]
__global__ void __launch_bounds__(256) k_mg_smooth(MgLevel L, int color, int k0)
{
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 2;
    const int k = blockIdx.z + k0;
    if (j > L.d.nj - 1 || k > L.d.nk - 1) return;
    const int i = 2 + ((color + j + k) & 1) + 2 * (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i > L.d.ni - 1) return;
    const long long n = mid(L.d, i, j, k);
    if (!L.act[n]) return;
    L.x[n] = (L.b[n] - row_off(L, n)) / L.ad[n];
}

// coarse right-hand side = sum of the fine residuals of the block; coarse x = 0
__global__ void __launch_bounds__(256) k_mg_restrict(MgLevel F, MgLevel C)
{
    const int I = blockIdx.x * blockDim.x + threadIdx.x + 2;
    const int J = blockIdx.y * blockDim.y + threadIdx.y + 2;
    const int K = blockIdx.z * blockDim.z + threadIdx.z + 2;
    if (I > C.d.ni - 1 || J > C.d.nj - 1 || K > C.d.nk - 1) return;
    const Dims &fd = F.d;
    float s = 0.0f;
    for (int c = 0; c < 8; c++) {
        const int i = 2 * I - 2 + (c & 1), j = 2 * J - 2 + ((c >> 1) & 1), k = 2 * K - 2 + (c >> 2);
        if (i > fd.ni - 1 || j > fd.nj - 1 || k > fd.nk - 1) continue;
        const long long n = mid(fd, i, j, k);
        if (!F.act[n]) continue;
        s += F.b[n] - F.ad[n] * F.x[n] - row_off(F, n);
    }
    const long long m = mid(C.d, I, J, K);
    const float bc = C.act[m] ? s : 0.0f;
    C.b[m] = bc;
    // from x = 0 the even half of the first sweep is x = b/AD
    C.x[m] = C.act[m] && ((I + J + K) & 1) == 0 ? bc / C.ad[m] : 0.0f;
}

// x += MG_SCALE * coarse x of the block (active fine cells)
__global__ void __launch_bounds__(256) k_mg_prolong(MgLevel F, MgLevel C, int k0)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 2;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 2;
    const int k = blockIdx.z + k0;
    if (i > F.d.ni - 1 || j > F.d.nj - 1 || k > F.d.nk - 1) return;
    const long long n = mid(F.d, i, j, k);
    if (!F.act[n]) return;
    const long long m = mid(C.d, (i + 2) / 2, (j + 2) / 2, (k + 2) / 2);
    F.x[n] += MG_SCALE * C.x[m];
}

// level-0 right-hand side from the FP64 vector over the active index range
// [rng[0], rng[1]) (rows outside it are inactive: their x stays 0); from
// x = 0 the even half of the first sweep is x = b/AD
__global__ void __launch_bounds__(256) k_mg_in(const unsigned long long *rng, const double *__restrict__ y,
                                               MgLevel L)
{
    const long long n = (long long)rng[0] + (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= (long long)rng[1]) return;
    const bool a = L.act[n] != 0;
    const float b = a ? (float)y[n] : 0.0f;
    L.b[n] = b;
    const long long ni = L.d.ni;
    const long long i = n % ni, j = (n / ni) % L.d.nj, k = n / L.d.nij;
    L.x[n] = a && ((i + j + k + 3) & 1) == 0 ? b / L.ad[n] : 0.0f;
}

// FP64 result from level 0 over the active index range (0 on inactive rows)
__global__ void __launch_bounds__(256) k_mg_out(const unsigned long long *rng, const MgLevel L,
                                                double *__restrict__ x)
{
    const long long n = (long long)rng[0] + (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= (long long)rng[1]) return;
    x[n] = L.act[n] ? (double)L.x[n] : 0.0;
}

// The V-cycle below level t0 inside one workgroup: for each level down to the
// coarsest the odd half-sweep (the even half was set with b), the coarse
// right-hand side; MG_COARSE sweeps on the coarsest; then up: correction and
// one sweep.  Levels are a kernel argument (at most CAD_MG_MAX).
struct MgTail {
    MgLevel L[CAD_MG_MAX];
};

__device__ void tail_half(const MgLevel &L, int color)
{
    const int ni = L.d.ni - 2, nj = L.d.nj - 2, nk = L.d.nk - 2;
    const int cells = ni * nj * nk;
    for (int q = threadIdx.x; q < cells; q += blockDim.x) {
        const int i = q % ni + 2, j = (q / ni) % nj + 2, k = q / (ni * nj) + 2;
        if (((i + j + k) & 1) != color) continue;
        const long long n = mid(L.d, i, j, k);
        if (L.act[n]) L.x[n] = (L.b[n] - row_off(L, n)) / L.ad[n];
    }
    __syncthreads();
}

__global__ void __launch_bounds__(1024) k_mg_tail(MgTail T, int t0, int nl)
{
    for (int l = t0; l < nl - 1; l++) {
        const MgLevel &F = T.L[l], &C = T.L[l + 1];
        tail_half(F, 1);
        const int ni = C.d.ni - 2, nj = C.d.nj - 2, nk = C.d.nk - 2;
        for (int q = threadIdx.x; q < ni * nj * nk; q += blockDim.x) {
            const int I = q % ni + 2, J = (q / ni) % nj + 2, K = q / (ni * nj) + 2;
            float s = 0.0f;
            for (int c = 0; c < 8; c++) {
                const int i = 2 * I - 2 + (c & 1), j = 2 * J - 2 + ((c >> 1) & 1), k = 2 * K - 2 + (c >> 2);
                if (i > F.d.ni - 1 || j > F.d.nj - 1 || k > F.d.nk - 1) continue;
                const long long n = mid(F.d, i, j, k);
                if (!F.act[n]) continue;
                s += F.b[n] - F.ad[n] * F.x[n] - row_off(F, n);
            }
            const long long m = mid(C.d, I, J, K);
            const float bc = C.act[m] ? s : 0.0f;
            C.b[m] = bc;
            C.x[m] = C.act[m] && ((I + J + K) & 1) == 0 ? bc / C.ad[m] : 0.0f;
        }
        __syncthreads();
    }
    const MgLevel &B = T.L[nl - 1];
    tail_half(B, 1);
    for (int s = 1; s < MG_COARSE; s++) {
        tail_half(B, 0);
        tail_half(B, 1);
    }
    for (int l = nl - 2; l >= t0; l--) {
        const MgLevel &F = T.L[l], &C = T.L[l + 1];
        const int ni = F.d.ni - 2, nj = F.d.nj - 2, nk = F.d.nk - 2;
        for (int q = threadIdx.x; q < ni * nj * nk; q += blockDim.x) {
            const int i = q % ni + 2, j = (q / ni) % nj + 2, k = q / (ni * nj) + 2;
            const long long n = mid(F.d, i, j, k);
            if (F.act[n]) F.x[n] += MG_SCALE * C.x[mid(C.d, (i + 2) / 2, (j + 2) / 2, (k + 2) / 2)];
        }
        __syncthreads();
        tail_half(F, 0);
        tail_half(F, 1);
    }
}

dim3 lvl_grid(const Dims &d, int k0, int k1)
{
    return dim3((unsigned)((d.ni - 2 + 63) / 64), (unsigned)((d.nj - 2 + 3) / 4), (unsigned)(k1 - k0 + 1));
}

dim3 half_grid(const Dims &d, int k0, int k1)
{
    const int half = (d.ni - 2 + 1) / 2;
    return dim3((unsigned)((half + 63) / 64), (unsigned)((d.nj - 2 + 3) / 4), (unsigned)(k1 - k0 + 1));
}

// half-sweeps of colours c0..1 on K planes k0..k1
int smooth(cad_ctx *c, const MgLevel &L, int k0, int k1, int c0)
{
    const dim3 b(64, 4, 1), g = half_grid(L.d, k0, k1);
    for (int color = c0; color <= 1; color++)
        hipLaunchKernelGGL(k_mg_smooth, g, b, 0, c->stream, L, color, k0);
    CAD_CHECK(hipGetLastError());
    return CAD_OK;
}

}  // namespace

void cad_mg_free(cad_ctx *c)
{
    for (int l = 0; l < c->mg_n; l++) {
        MgLevel &L = c->mg[l];
        void *p[] = {L.ad, L.ali, L.alj, L.alk, L.aui, L.auj, L.auk, L.x, L.b, L.act};
        for (void *q : p)
            if (q) (void)hipFree(q);
    }
    c->mg_n = 0;
}

int cad_mg_alloc(cad_ctx *c)
{
    if (c->mg_n > 0) return CAD_OK;
    Dims d = c->d;
    int n = 0;
    for (;;) {
        if (n == CAD_MG_MAX) break;
        MgLevel &L = c->mg[n];
        L = MgLevel();
        L.d = d;
        c->mg_n = n + 1;   // counted before its arrays, so a failure frees what was allocated
        const size_t nc = (size_t)d.nc;
        float **f[] = {&L.ad, &L.ali, &L.alj, &L.alk, &L.aui, &L.auj, &L.auk, &L.x, &L.b};
        for (float **p : f) {
            CAD_CHECK(hipMalloc((void **)p, nc * sizeof(float)));
            CAD_CHECK(hipMemset(*p, 0, nc * sizeof(float)));
        }
        CAD_CHECK(hipMalloc((void **)&L.act, nc));
        CAD_CHECK(hipMemset(L.act, 0, nc));
        n++;
        // coarsen while every interior extent is at least 4
        const int ei = d.ni - 2, ej = d.nj - 2, ek = d.nk - 2;
        if (ei < 4 || ej < 4 || ek < 4) break;
        Dims cd;
        cd.ni = (ei + 1) / 2 + 2;
        cd.nj = (ej + 1) / 2 + 2;
        cd.nk = (ek + 1) / 2 + 2;
        cd.nij = (long long)cd.ni * cd.nj;
        cd.nc = cd.nij * cd.nk;
        cd.numb = 0;
        d = cd;
    }
    const MgLevel &top = c->mg[c->mg_n - 1];
    if ((long long)(top.d.ni - 2) * (top.d.nj - 2) * (top.d.nk - 2) > (1 << 20)) {
        cad_set_error("multigrid: coarsest level %d x %d x %d too large", top.d.ni - 2, top.d.nj - 2,
                      top.d.nk - 2);
        return CAD_ERR_ARG;
    }
    CAD_CHECK(hipDeviceSynchronize());
    return CAD_OK;
}

// the hierarchy of the context's current matrix
int cad_mg_setup(cad_ctx *c)
{
    RET_IF(cad_mg_alloc(c));
    const Dims &d = c->d;
    hipLaunchKernelGGL(k_mg_fine, dim3((unsigned)((d.nc + 255) / 256)), dim3(256), 0, c->stream, d,
                       c->ad, c->ali, c->alj, c->alk, c->aui, c->auj, c->auk, c->f.indc, c->mg[0]);
    for (int l = 1; l < c->mg_n; l++) {
        const Dims &cd = c->mg[l].d;
        hipLaunchKernelGGL(k_mg_coarsen, lvl_grid(cd, 2, cd.nk - 1), dim3(64, 4, 1), 0, c->stream,
                           c->mg[l - 1], c->mg[l]);
    }
    CAD_CHECK(hipGetLastError());
    return CAD_OK;
}

// x = M^-1 y: one V-cycle from x = 0; k0..k1 limits level 0 to the K planes
// that hold active cells
int cad_mg_apply(cad_ctx *c, const double *y, double *x, int k0, int k1)
{
    const dim3 b(64, 4, 1);
    const int nl = c->mg_n;
    const unsigned long long span = c->mg_span;
    const unsigned vblocks = (unsigned)((span + 255) / 256);
    hipLaunchKernelGGL(k_mg_in, dim3(vblocks), dim3(256), 0, c->stream, c->act_rng, y, c->mg[0]);
    CAD_CHECK(hipGetLastError());
    // the first level small enough for one workgroup
    int t0 = nl - 1;
    while (t0 > 0 && (long long)(c->mg[t0 - 1].d.ni - 2) * (c->mg[t0 - 1].d.nj - 2) *
                             (c->mg[t0 - 1].d.nk - 2) <= MG_TAIL_CELLS)
        t0--;
    // down: the odd half-sweep (the even half came with b), the coarse b
    for (int l = 0; l < t0; l++) {
        const MgLevel &F = c->mg[l], &C = c->mg[l + 1];
        const int a0 = l == 0 ? k0 : 2, a1 = l == 0 ? k1 : F.d.nk - 1;
        RET_IF(smooth(c, F, a0, a1, 1));
        hipLaunchKernelGGL(k_mg_restrict, lvl_grid(C.d, 2, C.d.nk - 1), b, 0, c->stream, F, C);
        CAD_CHECK(hipGetLastError());
    }
    MgTail T;
    for (int l = 0; l < nl; l++) T.L[l] = c->mg[l];
    hipLaunchKernelGGL(k_mg_tail, dim3(1), dim3(1024), 0, c->stream, T, t0, nl);
    CAD_CHECK(hipGetLastError());
    // up: correct, then one sweep
    for (int l = t0 - 1; l >= 0; l--) {
        const MgLevel &F = c->mg[l], &C = c->mg[l + 1];
        const int a0 = l == 0 ? k0 : 2, a1 = l == 0 ? k1 : F.d.nk - 1;
        hipLaunchKernelGGL(k_mg_prolong, lvl_grid(F.d, a0, a1), b, 0, c->stream, F, C, a0);
        CAD_CHECK(hipGetLastError());
        RET_IF(smooth(c, F, a0, a1, 0));
    }
    hipLaunchKernelGGL(k_mg_out, dim3(vblocks), dim3(256), 0, c->stream, c->act_rng, c->mg[0], x);
    CAD_CHECK(hipGetLastError());
    return CAD_OK;
}

// One V-cycle on the hierarchy and K range of the last mixed solve, timed
// with HIP events on the library's stream (input BB, output AE), and its
// algorithmic bytes: per cell of a level below the coarsest, three
// half-sweeps of 20.5 B (one colour's 7 coefficients, b, activity and x
// written, the other colour's x read), the residual-restriction 38 B (7
// coefficients, b, x, activity; the coarse b and x) and the prolongation
// 9.5 B; the coarsest level's sweeps 41 B each; and per cell of the active
// index range the FP64 -> FP32 input 21 B and the FP32 -> FP64 output 13 B.
int cad_mg_time(cad_ctx *c, int reps, float *ms_avg, double *bytes)
{
    if (c->mg_n <= 0 || c->mg_span == 0) {
        cad_set_error("cad_time_precond: no mixed solve has run on this context");
        return CAD_ERR_STATE;
    }
    if (reps < 1) reps = 1;
    hipEvent_t a, b;
    CAD_CHECK(hipEventCreate(&a));
    CAD_CHECK(hipEventCreate(&b));
    CAD_CHECK(hipEventRecord(a, c->stream));
    for (int r = 0; r < reps; r++) RET_IF(cad_mg_apply(c, c->bb, c->ae, c->mg_k0, c->mg_k1));
    CAD_CHECK(hipEventRecord(b, c->stream));
    CAD_CHECK(hipEventSynchronize(b));
    float ms = 0.0f;
    CAD_CHECK(hipEventElapsedTime(&ms, a, b));
    hipEventDestroy(a);
    hipEventDestroy(b);
    *ms_avg = ms / reps;
    double by = 34.0 * (double)c->mg_span;
    for (int l = 0; l < c->mg_n; l++) {
        const Dims &d = c->mg[l].d;
        const int planes = l == 0 ? c->mg_k1 - c->mg_k0 + 1 : d.nk - 2;
        const double cells = (double)(d.ni - 2) * (double)(d.nj - 2) * (double)planes;
        by += l == c->mg_n - 1 ? 41.0 * MG_COARSE * cells : (3.0 * 20.5 + 38.0 + 9.5) * cells;
    }
    *bytes = by;
    return CAD_OK;
}
