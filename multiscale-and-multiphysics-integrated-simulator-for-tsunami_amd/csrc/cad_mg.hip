// cad_mg.hip -- aggregation multigrid, the preconditioner of the mixed and
// learned forms, in FP32 (and in FP64 for measuring what FP32 buys).
//
// The mixed form solves the pressure system of VF_VPCOEF (the same FP64
// matrix AD..AUK, BB as the plain form) by FP64 iterative refinement around
// an FP32 BiCGSTAB (cad_refine.hip) whose preconditioner is one V-cycle of
// this multigrid.  It replaces the reference's MILU preconditioner
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
// Storage.  Each level stores the two colours of the red-black ordering
// apart: a (J,K) line of ld cells (NUMI rounded up to even) holds first its
// cells with I+J+K even, then those with I+J+K odd, each half in ascending
// I.  A half-sweep then reads and writes whole lines of its own colour and
// reads the other colour's x at the same positions of the neighbouring lines,
// so a thread's neighbours are its neighbour threads' cells and no line is
// fetched for half of its cells.  Level 0 also keeps its rows in the grid's
// natural order (for the FP32 matvec of cad_refine.hip).
//
// V-cycle.  Red-black Gauss-Seidel (one sweep = the cells with I+J+K even,
// then odd) before and after the coarse correction; the residual of the fine
// cells is summed into the coarse right-hand side in one kernel (after the
// odd half-sweep only the even cells' residuals are not zero up to rounding);
// the coarse correction is added scaled by MG_SCALE (piecewise-constant
// prolongation under-corrects smooth errors), on the fly inside the even
// half-sweep that follows it; the coarsest level is smoothed MG_COARSE
// times.  Every level starts from x = 0, so the first (even) half of its
// first sweep is x = b/AD, which the kernel that sets b writes at once.  The
// levels of at most MG_TAIL_CELLS cells run as one workgroup (k_mg_tail), so
// the small grids cost one launch instead of five per level.
#include "cad_common.h"

#include <algorithm>
#include <climits>

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

// position of cell (i,j,k) (1-based, ghosts included) in a level's arrays
template <typename L>
__device__ __forceinline__ long long pos(const L &v, int i, int j, int k)
{
    const long long line = (long long)v.ld * ((j - 1) + (long long)v.d.nj * (k - 1));
    return line + (long long)((i + j + k) & 1) * (v.ld >> 1) + ((i - 1) >> 1);
}

__device__ __forceinline__ long long nat(const Dims &d, int i, int j, int k)
{
    return (long long)(i - 1) + (long long)d.ni * (j - 1) + d.nij * (k - 1);
}

// level 0 from the FP64 matrix: the colour-split rows and, where the level
// has them, the natural-order copy
template <typename V>
__global__ void __launch_bounds__(256) k_mg_fine(Dims d, const double *__restrict__ ad,
                                                 const double *__restrict__ ali,
                                                 const double *__restrict__ alj,
                                                 const double *__restrict__ alk,
                                                 const double *__restrict__ aui,
                                                 const double *__restrict__ auj,
                                                 const double *__restrict__ auk,
                                                 const int *__restrict__ indc, MgLevelT<V> L,
                                                 long long n0, long long n1)
{
    const long long n = n0 + (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= n1) return;
    const int i = (int)(n % d.ni) + 1, j = (int)((n / d.ni) % d.nj) + 1, k = (int)(n / d.nij) + 1;
    const long long m = pos(L, i, j, k);
    const bool a = indc[n] != -1 && ad[n] > 0.0;
    const V r[7] = {a ? (V)ad[n] : (V)1, a ? (V)ali[n] : (V)0, a ? (V)alj[n] : (V)0, a ? (V)alk[n] : (V)0,
                    a ? (V)aui[n] : (V)0, a ? (V)auj[n] : (V)0, a ? (V)auk[n] : (V)0};
    L.act[m] = a ? 1 : 0;
    L.ad[m] = r[0];
    L.ali[m] = r[1];
    L.alj[m] = r[2];
    L.alk[m] = r[3];
    L.aui[m] = r[4];
    L.auj[m] = r[5];
    L.auk[m] = r[6];
    if (L.nad) {
        L.nact[n] = a ? 1 : 0;
        L.nad[n] = r[0];
        L.nali[n] = r[1];
        L.nalj[n] = r[2];
        L.nalk[n] = r[3];
        L.naui[n] = r[4];
        L.nauj[n] = r[5];
        L.nauk[n] = r[6];
    }
}

// Galerkin coarse row of coarse interior cell (I,J,K), K from k0; one thread
// per cell
template <typename V>
__global__ void __launch_bounds__(256) k_mg_coarsen(MgLevelT<V> F, MgLevelT<V> C, int k0)
{
    const int I = blockIdx.x * blockDim.x + threadIdx.x + 2;
    const int J = blockIdx.y * blockDim.y + threadIdx.y + 2;
    const int K = blockIdx.z + k0;
    if (I > C.d.ni - 1 || J > C.d.nj - 1 || K > C.d.nk - 1) return;
    const Dims &fd = F.d;
    V s_ad = 0, s_li = 0, s_lj = 0, s_lk = 0, s_ui = 0, s_uj = 0, s_uk = 0;
    bool any = false;
    for (int c = 0; c < 8; c++) {
        const int oi = c & 1, oj = (c >> 1) & 1, ok = c >> 2;
        const int i = 2 * I - 2 + oi, j = 2 * J - 2 + oj, k = 2 * K - 2 + ok;
        if (i > fd.ni - 1 || j > fd.nj - 1 || k > fd.nk - 1) continue;
        const long long n = pos(F, i, j, k);
        if (!F.act[n]) continue;
        any = true;
        s_ad += F.ad[n];
        // -I: inside the block when oi = 1
        if (F.ali[n] != (V)0 && F.act[pos(F, i - 1, j, k)]) {
            if (oi) s_ad += F.ali[n]; else s_li += F.ali[n];
        }
        if (F.aui[n] != (V)0 && F.act[pos(F, i + 1, j, k)]) {
            if (!oi && i + 1 <= fd.ni - 1) s_ad += F.aui[n]; else s_ui += F.aui[n];
        }
        if (F.alj[n] != (V)0 && F.act[pos(F, i, j - 1, k)]) {
            if (oj) s_ad += F.alj[n]; else s_lj += F.alj[n];
        }
        if (F.auj[n] != (V)0 && F.act[pos(F, i, j + 1, k)]) {
            if (!oj && j + 1 <= fd.nj - 1) s_ad += F.auj[n]; else s_uj += F.auj[n];
        }
        if (F.alk[n] != (V)0 && F.act[pos(F, i, j, k - 1)]) {
            if (ok) s_ad += F.alk[n]; else s_lk += F.alk[n];
        }
        if (F.auk[n] != (V)0 && F.act[pos(F, i, j, k + 1)]) {
            if (!ok && k + 1 <= fd.nk - 1) s_ad += F.auk[n]; else s_uk += F.auk[n];
        }
    }
    const long long m = pos(C, I, J, K);
    const bool a = any && s_ad > (V)0;
    C.act[m] = a ? 1 : 0;
    C.ad[m] = a ? s_ad : (V)1;
    C.ali[m] = a ? s_li : (V)0;
    C.alj[m] = a ? s_lj : (V)0;
    C.alk[m] = a ? s_lk : (V)0;
    C.aui[m] = a ? s_ui : (V)0;
    C.auj[m] = a ? s_uj : (V)0;
    C.auk[m] = a ? s_uk : (V)0;
}

// the off-diagonal part of row (i,j,k) at position n
template <typename V>
__device__ __forceinline__ V row_off(const MgLevelT<V> &L, long long n, int i, int j, int k)
{
    const V *x = L.x;
    return L.ali[n] * x[pos(L, i - 1, j, k)] + L.aui[n] * x[pos(L, i + 1, j, k)]
         + L.alj[n] * x[pos(L, i, j - 1, k)] + L.auj[n] * x[pos(L, i, j + 1, k)]
         + L.alk[n] * x[pos(L, i, j, k - 1)] + L.auk[n] * x[pos(L, i, j, k + 1)];
}

// one colour of a Gauss-Seidel sweep: thread x covers position q of its
// (J,K) line's half of that colour (cell I = 2q+1 or 2q+2), K in [k0, k1]
template <typename V>
__global__ void __launch_bounds__(256) k_mg_smooth(MgLevelT<V> L, int color, int k0)
{
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 2;
    const int k = blockIdx.z + k0;
    if (j > L.d.nj - 1 || k > L.d.nk - 1) return;
    const int q = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int i = 2 * q + 1 + ((color + 1 + j + k) & 1);
    if (i < 2 || i > L.d.ni - 1) return;
    const long long n = pos(L, i, j, k);
    if (!L.act[n]) return;
    L.x[n] = (L.b[n] - row_off(L, n, i, j, k)) / L.ad[n];
}

// Coarse right-hand side = sum of the fine residuals of the block; coarse
// x = 0.  It follows the odd half-sweep, after which an odd cell's residual
// is zero up to rounding (its x was just solved from its row), so only the
// block's four even cells are summed.
template <typename V>
__device__ __forceinline__ V block_residual(const MgLevelT<V> &F, int I, int J, int K)
{
    const Dims &fd = F.d;
    V s = 0;
    for (int c = 0; c < 8; c++) {
        const int i = 2 * I - 2 + (c & 1), j = 2 * J - 2 + ((c >> 1) & 1), k = 2 * K - 2 + (c >> 2);
        if (((i + j + k) & 1) != 0 || i > fd.ni - 1 || j > fd.nj - 1 || k > fd.nk - 1) continue;
        const long long n = pos(F, i, j, k);
        if (!F.act[n]) continue;
        s += F.b[n] - F.ad[n] * F.x[n] - row_off(F, n, i, j, k);
    }
    return s;
}

template <typename V>
__global__ void __launch_bounds__(256) k_mg_restrict(MgLevelT<V> F, MgLevelT<V> C, int k0)
{
    const int I = blockIdx.x * blockDim.x + threadIdx.x + 2;
    const int J = blockIdx.y * blockDim.y + threadIdx.y + 2;
    const int K = blockIdx.z + k0;
    if (I > C.d.ni - 1 || J > C.d.nj - 1 || K > C.d.nk - 1) return;
    const V s = block_residual(F, I, J, K);
    const long long m = pos(C, I, J, K);
    const V bc = C.act[m] ? s : (V)0;
    C.b[m] = bc;
    // from x = 0 the even half of the first sweep is x = b/AD
    C.x[m] = C.act[m] && ((I + J + K) & 1) == 0 ? bc / C.ad[m] : (V)0;
}

// The coarse correction and the even half-sweep after it, in one pass.  The
// correction x += MG_SCALE * coarse x of the block matters only at the odd
// cells: the even half-sweep solves each even cell's x from its row, which
// reads the odd neighbours only.  So each even cell reads its odd
// neighbours' x plus their correction (what a separate correction pass would
// have stored), and the odd cells, which the odd half-sweep that follows
// solves from their rows in turn, are not corrected in memory.
template <typename V>
__device__ __forceinline__ V corrected(const MgLevelT<V> &F, const MgLevelT<V> &C, int i, int j, int k)
{
    const long long n = pos(F, i, j, k);
    return F.act[n] ? F.x[n] + (V)MG_SCALE * C.x[pos(C, (i + 2) / 2, (j + 2) / 2, (k + 2) / 2)] : F.x[n];
}

template <typename V>
__global__ void __launch_bounds__(256) k_mg_correct_even(MgLevelT<V> F, MgLevelT<V> C, int k0)
{
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 2;
    const int k = blockIdx.z + k0;
    if (j > F.d.nj - 1 || k > F.d.nk - 1) return;
    const int q = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int i = 2 * q + 1 + ((1 + j + k) & 1);
    if (i < 2 || i > F.d.ni - 1) return;
    const long long n = pos(F, i, j, k);
    if (!F.act[n]) return;
    const V off = F.ali[n] * corrected(F, C, i - 1, j, k) + F.aui[n] * corrected(F, C, i + 1, j, k)
                + F.alj[n] * corrected(F, C, i, j - 1, k) + F.auj[n] * corrected(F, C, i, j + 1, k)
                + F.alk[n] * corrected(F, C, i, j, k - 1) + F.auk[n] * corrected(F, C, i, j, k + 1);
    F.x[n] = (F.b[n] - off) / F.ad[n];
}

// level-0 right-hand side from a vector of the grid (natural order) on the K
// planes k0..k1 (the rows outside them are inactive: their x stays 0); from
// x = 0 the even half of the first sweep is x = b/AD
template <typename V, typename T>
__global__ void __launch_bounds__(256) k_mg_in(MgLevelT<V> L, const T *__restrict__ y, int k0)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 2;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 2;
    const int k = blockIdx.z + k0;
    if (i > L.d.ni - 1 || j > L.d.nj - 1) return;
    const long long m = pos(L, i, j, k);
    const bool a = L.act[m] != 0;
    const V b = a ? (V)y[nat(L.d, i, j, k)] : (V)0;
    L.b[m] = b;
    L.x[m] = a && ((i + j + k) & 1) == 0 ? b / L.ad[m] : (V)0;
}

// the result from level 0 into a vector of the grid on the K planes k0..k1
// (0 on inactive rows)
template <typename V, typename T>
__global__ void __launch_bounds__(256) k_mg_out(const MgLevelT<V> L, T *__restrict__ x, int k0)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 2;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 2;
    const int k = blockIdx.z + k0;
    if (i > L.d.ni - 1 || j > L.d.nj - 1) return;
    const long long m = pos(L, i, j, k);
    x[nat(L.d, i, j, k)] = L.act[m] ? (T)L.x[m] : (T)0;
}

// k_mg_out and, on the same rows, the level's residual b - A x in V (the
// learned preconditioner's input), with the block partials of its sum of
// squares and of the row count at part[b * stride] and part[b * stride + 1]
// (grid-stride over the K planes k0..k1, gridDim.x blocks of 256)
template <typename V, typename T>
__global__ void __launch_bounds__(256) k_mg_out_resid(const MgLevelT<V> L, T *__restrict__ x, T *__restrict__ r,
                                                      int k0, int k1, double *__restrict__ part, int stride)
{
    const int ni = L.d.ni - 2, nj = L.d.nj - 2;
    const long long total = (long long)ni * nj * (k1 - k0 + 1);
    double a = 0.0, cnt = 0.0;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(e % ni) + 2;
        const long long q = e / ni;
        const int j = (int)(q % nj) + 2, k = (int)(q / nj) + k0;
        const long long m = pos(L, i, j, k), n = nat(L.d, i, j, k);
        if (!L.act[m]) {
            x[n] = (T)0;
            r[n] = (T)0;
            continue;
        }
        const V xm = L.x[m];
        const V rv = L.b[m] - L.ad[m] * xm - row_off(L, m, i, j, k);
        x[n] = (T)xm;
        r[n] = (T)rv;
        a += (double)rv * (double)rv;
        cnt += 1.0;
    }
    __shared__ double sh[2][4];
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_down(a, o, 64);
        cnt += __shfl_down(cnt, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        sh[0][threadIdx.x >> 6] = a;
        sh[1][threadIdx.x >> 6] = cnt;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[(long long)blockIdx.x * stride] = sh[0][0] + sh[0][1] + sh[0][2] + sh[0][3];
        part[(long long)blockIdx.x * stride + 1] = sh[1][0] + sh[1][1] + sh[1][2] + sh[1][3];
    }
}

// The V-cycle below level t0 inside one workgroup: for each level down to the
// coarsest the odd half-sweep (the even half was set with b), the coarse
// right-hand side; MG_COARSE sweeps on the coarsest; then up: the even
// half-sweep with the correction, and the odd one.  The coarsest level, whose
// sweeps are most of the tail's phases, is swept in LDS: its x and b are
// copied in, each thread keeps its cell's row in registers, and x goes back
// at the end, so those phases touch no memory.  Levels are a kernel argument
// (at most CAD_MG_MAX).
constexpr int TAIL_THREADS = 1024;
constexpr int TAIL_LDS_CELLS = 8192;         // positions of the coarsest level held in LDS

template <typename V>
struct MgTail {
    MgLevelT<V> L[CAD_MG_MAX];
};

template <typename V>
__device__ void tail_half(const MgLevelT<V> &L, int color)
{
    const int ni = L.d.ni - 2, nj = L.d.nj - 2, nk = L.d.nk - 2;
    const int cells = ni * nj * nk;
    for (int q = threadIdx.x; q < cells; q += blockDim.x) {
        const int i = q % ni + 2, j = (q / ni) % nj + 2, k = q / (ni * nj) + 2;
        if (((i + j + k) & 1) != color) continue;
        const long long n = pos(L, i, j, k);
        if (L.act[n]) L.x[n] = (L.b[n] - row_off(L, n, i, j, k)) / L.ad[n];
    }
    __syncthreads();
}

// the even half-sweep of F with the correction from C (k_mg_correct_even)
template <typename V>
__device__ void tail_correct_even(const MgLevelT<V> &F, const MgLevelT<V> &C)
{
    const int ni = F.d.ni - 2, nj = F.d.nj - 2, nk = F.d.nk - 2;
    for (int q = threadIdx.x; q < ni * nj * nk; q += blockDim.x) {
        const int i = q % ni + 2, j = (q / ni) % nj + 2, k = q / (ni * nj) + 2;
        if (((i + j + k) & 1) != 0) continue;
        const long long n = pos(F, i, j, k);
        if (!F.act[n]) continue;
        const V off = F.ali[n] * corrected(F, C, i - 1, j, k) + F.aui[n] * corrected(F, C, i + 1, j, k)
                    + F.alj[n] * corrected(F, C, i, j - 1, k) + F.auj[n] * corrected(F, C, i, j + 1, k)
                    + F.alk[n] * corrected(F, C, i, j, k - 1) + F.auk[n] * corrected(F, C, i, j, k + 1);
        F.x[n] = (F.b[n] - off) / F.ad[n];
    }
    __syncthreads();
}

// MG_COARSE sweeps of the coarsest level B in LDS (the even half of the
// first came with b); at most TAIL_THREADS cells and TAIL_LDS_CELLS
// positions (the host sweeps larger coarsest levels in memory)
template <typename V>
__device__ void tail_coarsest(const MgLevelT<V> &B)
{
    __shared__ V xs[TAIL_LDS_CELLS], bs[TAIL_LDS_CELLS];
    const int ns = B.ld * B.d.nj * B.d.nk;
    for (int p = threadIdx.x; p < ns; p += blockDim.x) {
        xs[p] = B.x[p];
        bs[p] = B.b[p];
    }
    const int ni = B.d.ni - 2, nj = B.d.nj - 2, nk = B.d.nk - 2;
    const int q = threadIdx.x;
    const bool on = q < ni * nj * nk;
    const int i = on ? q % ni + 2 : 2, j = on ? (q / ni) % nj + 2 : 2, k = on ? q / (ni * nj) + 2 : 2;
    const long long n = pos(B, i, j, k);
    const bool act = on && B.act[n];
    const bool odd = ((i + j + k) & 1) != 0;
    const V ad = act ? B.ad[n] : (V)1, li = act ? B.ali[n] : (V)0, ui = act ? B.aui[n] : (V)0;
    const V lj = act ? B.alj[n] : (V)0, uj = act ? B.auj[n] : (V)0, lk = act ? B.alk[n] : (V)0;
    const V uk = act ? B.auk[n] : (V)0;
    const long long pi0 = pos(B, i - 1, j, k), pi1 = pos(B, i + 1, j, k), pj0 = pos(B, i, j - 1, k);
    const long long pj1 = pos(B, i, j + 1, k), pk0 = pos(B, i, j, k - 1), pk1 = pos(B, i, j, k + 1);
    __syncthreads();
    for (int h = 1; h < 2 * MG_COARSE; h++) {
        // h odd: the odd half, h even: the even half
        if (act && odd == ((h & 1) != 0))
            xs[n] = (bs[n] - (li * xs[pi0] + ui * xs[pi1] + lj * xs[pj0] + uj * xs[pj1] + lk * xs[pk0]
                              + uk * xs[pk1])) / ad;
        __syncthreads();
    }
    if (on) B.x[n] = xs[n];
    __syncthreads();
}

template <typename V>
__global__ void __launch_bounds__(TAIL_THREADS) k_mg_tail(MgTail<V> T, int t0, int nl, int lds_coarsest)
{
    for (int l = t0; l < nl - 1; l++) {
        const MgLevelT<V> &F = T.L[l], &C = T.L[l + 1];
        tail_half(F, 1);
        const int ni = C.d.ni - 2, nj = C.d.nj - 2, nk = C.d.nk - 2;
        for (int q = threadIdx.x; q < ni * nj * nk; q += blockDim.x) {
            const int I = q % ni + 2, J = (q / ni) % nj + 2, K = q / (ni * nj) + 2;
            const V s = block_residual(F, I, J, K);
            const long long m = pos(C, I, J, K);
            const V bc = C.act[m] ? s : (V)0;
            C.b[m] = bc;
            C.x[m] = C.act[m] && ((I + J + K) & 1) == 0 ? bc / C.ad[m] : (V)0;
        }
        __syncthreads();
    }
    const MgLevelT<V> &B = T.L[nl - 1];
    if (lds_coarsest) {
        tail_coarsest(B);
    } else {
        tail_half(B, 1);
        for (int s = 1; s < MG_COARSE; s++) {
            tail_half(B, 0);
            tail_half(B, 1);
        }
    }
    for (int l = nl - 2; l >= t0; l--) {
        tail_correct_even(T.L[l], T.L[l + 1]);
        tail_half(T.L[l], 1);
    }
}

dim3 lvl_grid(const Dims &d, int k0, int k1)
{
    return dim3((unsigned)((d.ni - 2 + 63) / 64), (unsigned)((d.nj - 2 + 3) / 4), (unsigned)(k1 - k0 + 1));
}

// threads over one colour's half lines: positions 0 .. NUMI/2
dim3 half_grid(const Dims &d, int k0, int k1)
{
    const int half = d.ni / 2 + 1;
    return dim3((unsigned)((half + 63) / 64), (unsigned)((d.nj - 2 + 3) / 4), (unsigned)(k1 - k0 + 1));
}

// half-sweeps of colours c0..1 on K planes k0..k1
template <typename V>
int smooth(cad_ctx *c, const MgLevelT<V> &L, int k0, int k1, int c0)
{
    const dim3 b(64, 4, 1), g = half_grid(L.d, k0, k1);
    for (int color = c0; color <= 1; color++)
        hipLaunchKernelGGL(k_mg_smooth<V>, g, b, 0, c->stream, L, color, k0);
    CAD_CHECK(hipGetLastError());
    return CAD_OK;
}

template <typename V>
void free_levels(MgLevelT<V> *mg, int &n)
{
    for (int l = 0; l < n; l++) {
        MgLevelT<V> &L = mg[l];
        void *p[] = {L.ad, L.ali, L.alj, L.alk, L.aui, L.auj, L.auk, L.x, L.b, L.act, L.nad, L.nact};
        for (void *q : p)
            if (q) (void)hipFree(q);
        L = MgLevelT<V>();
    }
    n = 0;
}

// The levels of a grid: colour-split arrays of ld x NUMJ x NUMK per level
// (and level 0's natural-order copy when natural is set), each level counted
// before its arrays so that a failed allocation frees what was allocated.
template <typename V>
int alloc_levels(const Dims &d0, MgLevelT<V> *mg, int &nl, bool natural)
{
    if (nl > 0) return CAD_OK;
    Dims d = d0;
    int n = 0;
    for (;;) {
        if (n == CAD_MG_MAX) break;
        MgLevelT<V> &L = mg[n];
        L = MgLevelT<V>();
        L.d = d;
        L.ld = (d.ni + 1) / 2 * 2;
        nl = n + 1;
        const size_t ns = (size_t)L.ld * d.nj * d.nk;
        V **f[] = {&L.ad, &L.ali, &L.alj, &L.alk, &L.aui, &L.auj, &L.auk, &L.x, &L.b};
        for (V **p : f) {
            CAD_CHECK(hipMalloc((void **)p, ns * sizeof(V)));
            CAD_CHECK(hipMemset(*p, 0, ns * sizeof(V)));
        }
        CAD_CHECK(hipMalloc((void **)&L.act, ns));
        CAD_CHECK(hipMemset(L.act, 0, ns));
        if (n == 0 && natural) {
            const size_t nc = (size_t)d.nc;
            CAD_CHECK(hipMalloc((void **)&L.nad, 7 * nc * sizeof(V)));
            CAD_CHECK(hipMalloc((void **)&L.nact, nc));
            CAD_CHECK(hipMemset(L.nad, 0, 7 * nc * sizeof(V)));
            CAD_CHECK(hipMemset(L.nact, 0, nc));
            L.nali = L.nad + nc;
            L.nalj = L.nali + nc;
            L.nalk = L.nalj + nc;
            L.naui = L.nalk + nc;
            L.nauj = L.naui + nc;
            L.nauk = L.nauj + nc;
        }
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
    const MgLevelT<V> &top = mg[nl - 1];
    if ((long long)(top.d.ni - 2) * (top.d.nj - 2) * (top.d.nk - 2) > (1 << 20)) {
        cad_set_error("multigrid: coarsest level %d x %d x %d too large", top.d.ni - 2, top.d.nj - 2,
                      top.d.nk - 2);
        return CAD_ERR_ARG;
    }
    CAD_CHECK(hipDeviceSynchronize());
    return CAD_OK;
}

// The levels' rows from the context's matrix on level 0's K planes lo..hi
// and, level by level, the coarse planes whose rows read those (coarse plane
// K reads fine planes 2K-3 .. 2K); every other row keeps what it holds.
template <typename V>
int setup_levels(cad_ctx *c, MgLevelT<V> *mg, int nl, int lo, int hi)
{
    const Dims &d = c->d;
    const long long n0 = (long long)(lo - 1) * d.nij, n1 = (long long)hi * d.nij;
    hipLaunchKernelGGL(k_mg_fine<V>, dim3((unsigned)((n1 - n0 + 255) / 256)), dim3(256), 0, c->stream, d,
                       c->ad, c->ali, c->alj, c->alk, c->aui, c->auj, c->auk, c->f.indc, mg[0], n0, n1);
    for (int l = 1; l < nl; l++) {
        const Dims &cd = mg[l].d;
        lo = std::max(2, (lo + 1) / 2);
        hi = std::min(cd.nk - 1, (hi + 3) / 2);
        if (lo > hi) break;
        hipLaunchKernelGGL(k_mg_coarsen<V>, lvl_grid(cd, lo, hi), dim3(64, 4, 1), 0, c->stream, mg[l - 1],
                           mg[l], lo);
    }
    CAD_CHECK(hipGetLastError());
    return CAD_OK;
}

// x = M^-1 y: one V-cycle from x = 0; k0..k1 limits level 0 to the K planes
// that hold active cells.  With r set, also r = y - A x on level 0's matrix
// and the partials of its sum of squares (k_mg_out_resid, REDUCE_BLOCKS blocks).
template <typename V, typename T>
int vcycle(cad_ctx *c, const MgLevelT<V> *mg, int nl, const T *y, T *x, int k0, int k1, T *r = nullptr,
           double *part = nullptr, int stride = 0)
{
    const dim3 b(64, 4, 1);
    hipLaunchKernelGGL((k_mg_in<V, T>), lvl_grid(mg[0].d, k0, k1), b, 0, c->stream, mg[0], y, k0);
    CAD_CHECK(hipGetLastError());
    // the first level small enough for one workgroup
    int t0 = nl - 1;
    while (t0 > 0 && (long long)(mg[t0 - 1].d.ni - 2) * (mg[t0 - 1].d.nj - 2) * (mg[t0 - 1].d.nk - 2) <=
                         MG_TAIL_CELLS)
        t0--;
    // Each level's passes run on the K planes that can hold its active rows:
    // level 0's k0..k1, and on a coarser level the planes of their parents
    // (fine plane k lies in coarse plane (k + 2) / 2).  Rows outside are
    // inactive, and no active row couples to them (coarsening sums only the
    // couplings between active rows), so what they hold is never used.
    int lo[CAD_MG_MAX], hi[CAD_MG_MAX];
    lo[0] = k0;
    hi[0] = k1;
    for (int l = 1; l < nl; l++) {
        lo[l] = std::max(2, (lo[l - 1] + 2) / 2);
        hi[l] = std::min(mg[l].d.nk - 1, (hi[l - 1] + 2) / 2);
    }
    // down: the odd half-sweep (the even half came with b), the coarse b
    for (int l = 0; l < t0; l++) {
        const MgLevelT<V> &F = mg[l], &C = mg[l + 1];
        RET_IF(smooth(c, F, lo[l], hi[l], 1));
        hipLaunchKernelGGL(k_mg_restrict<V>, lvl_grid(C.d, lo[l + 1], hi[l + 1]), b, 0, c->stream, F, C,
                           lo[l + 1]);
        CAD_CHECK(hipGetLastError());
    }
    MgTail<V> tl;
    for (int l = 0; l < nl; l++) tl.L[l] = mg[l];
    const MgLevelT<V> &top = mg[nl - 1];
    const int lds = (long long)(top.d.ni - 2) * (top.d.nj - 2) * (top.d.nk - 2) <= TAIL_THREADS &&
                    (long long)top.ld * top.d.nj * top.d.nk <= TAIL_LDS_CELLS;
    hipLaunchKernelGGL(k_mg_tail<V>, dim3(1), dim3(TAIL_THREADS), 0, c->stream, tl, t0, nl, lds);
    CAD_CHECK(hipGetLastError());
    // up: correct, then one sweep (the correction rides on the even half)
    for (int l = t0 - 1; l >= 0; l--) {
        const MgLevelT<V> &F = mg[l], &C = mg[l + 1];
        hipLaunchKernelGGL(k_mg_correct_even<V>, half_grid(F.d, lo[l], hi[l]), b, 0, c->stream, F, C, lo[l]);
        CAD_CHECK(hipGetLastError());
        RET_IF(smooth(c, F, lo[l], hi[l], 1));
    }
    if (r)
        hipLaunchKernelGGL((k_mg_out_resid<V, T>), dim3(REDUCE_BLOCKS), dim3(256), 0, c->stream, mg[0], x, r, k0, k1,
                           part, stride);
    else
        hipLaunchKernelGGL((k_mg_out<V, T>), lvl_grid(mg[0].d, k0, k1), b, 0, c->stream, mg[0], x, k0);
    CAD_CHECK(hipGetLastError());
    return CAD_OK;
}

// Algorithmic bytes of one V-cycle with vs-byte values: per cell of a level
// below the coarsest, three half-sweeps (each cell once in one of them: 7
// coefficients, b, x written and the activity byte; x of the other colour's
// neighbours read once: 10 values + 1 byte per cell and colour pair), the
// residual-restriction of the even cells (7 coefficients, b, x, activity,
// and a share of the coarse b and x) and the coarse x that the corrected
// even half-sweep reads; on the coarsest level MG_COARSE full sweeps; per cell of level
// 0's K range the FP64 input read with b and x written (AD and the activity
// read), and x and the activity read with the FP64 output written.
double vcycle_bytes(const cad_ctx *c, const Dims *dims, int nl, int vs)
{
    const double half = (10.0 * vs + 1.0) / 2.0;
    const double restr = (9.0 * vs + 1.0) / 2.0 + 2.0 * vs / 8.0;
    const double prol = vs / 8.0;
    double by = 0.0;
    for (int l = 0; l < nl; l++) {
        const Dims &d = dims[l];
        const int planes = l == 0 ? c->mg_k1 - c->mg_k0 + 1 : d.nk - 2;
        const double cells = (double)(d.ni - 2) * (double)(d.nj - 2) * (double)planes;
        if (l == 0) by += ((8.0 + 3.0 * vs + 1.0) + (vs + 8.0 + 1.0)) * cells;
        by += l == nl - 1 ? 2.0 * half * MG_COARSE * cells : (3.0 * half + restr + prol) * cells;
    }
    return by;
}

}  // namespace

void cad_mg_free(cad_ctx *c)
{
    c->mg_set_k0 = c->mg_set_k1 = 0;
    free_levels(c->mg, c->mg_n);
    free_levels(c->mg64, c->mg64_n);
}

int cad_mg_alloc(cad_ctx *c) { return alloc_levels(c->d, c->mg, c->mg_n, true); }

// the hierarchy of the context's current matrix.  Outside the K planes that
// hold active cells every row is inactive (INDC = -1 there), as a freshly
// allocated level's are, so the rows are rebuilt only on the planes of this
// solve and of the last build (where rows may have been active); the rows
// are those of a build over every plane.
int cad_mg_setup(cad_ctx *c)
{
    RET_IF(cad_mg_alloc(c));
    int lo = c->mg_k0, hi = c->mg_k1;
    if (c->mg_set_k0 > 0) {
        lo = std::min(lo, c->mg_set_k0);
        hi = std::max(hi, c->mg_set_k1);
    }
    c->mg_set_k0 = c->mg_k0;
    c->mg_set_k1 = c->mg_k1;
    return setup_levels(c, c->mg, c->mg_n, lo, hi);
}

int cad_mg_apply(cad_ctx *c, const double *y, double *x, int k0, int k1)
{
    return vcycle(c, c->mg, c->mg_n, y, x, k0, k1);
}

int cad_mg_apply_resid(cad_ctx *c, const double *y, double *x, double *r, double *part, int stride, int k0, int k1)
{
    return vcycle(c, c->mg, c->mg_n, y, x, k0, k1, r, part, stride);
}

int cad_mg_apply32(cad_ctx *c, const float *y, float *x, int k0, int k1)
{
    RET_IF(cad_kt_begin(c, CAD_KT_VCYCLE));
    RET_IF(vcycle(c, c->mg, c->mg_n, y, x, k0, k1));
    double by = 0.0;
    if (c->kt && c->kt->on) {
        Dims dims[CAD_MG_MAX];
        for (int l = 0; l < c->mg_n; l++) dims[l] = c->mg[l].d;
        by = vcycle_bytes(c, dims, c->mg_n, 4);
    }
    return cad_kt_end(c, CAD_KT_VCYCLE, by);
}

// One V-cycle on the hierarchy and K range of the last mixed or learned
// solve, timed with HIP events on the library's stream (input BB, output AE),
// with its algorithmic bytes (vcycle_bytes); fp64 != 0: the same V-cycle with
// every level in FP64 (a hierarchy of its own, built from the same matrix),
// for comparison.
int cad_mg_time(cad_ctx *c, int reps, int fp64, float *ms_avg, double *bytes)
{
    if (c->mg_n <= 0 || c->mg_span == 0) {
        cad_set_error("cad_time_precond: no mixed solve has run on this context");
        return CAD_ERR_STATE;
    }
    if (reps < 1) reps = 1;
    if (fp64) {
        RET_IF(alloc_levels(c->d, c->mg64, c->mg64_n, false));
        RET_IF(setup_levels(c, c->mg64, c->mg64_n, 1, c->d.nk));
    }
    hipEvent_t a, b;
    CAD_CHECK(hipEventCreate(&a));
    CAD_CHECK(hipEventCreate(&b));
    CAD_CHECK(hipEventRecord(a, c->stream));
    for (int r = 0; r < reps; r++) {
        if (fp64)
            RET_IF(vcycle(c, c->mg64, c->mg64_n, c->bb, c->ae, c->mg_k0, c->mg_k1));
        else
            RET_IF(vcycle(c, c->mg, c->mg_n, c->bb, c->ae, c->mg_k0, c->mg_k1));
    }
    CAD_CHECK(hipEventRecord(b, c->stream));
    CAD_CHECK(hipEventSynchronize(b));
    float ms = 0.0f;
    CAD_CHECK(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *ms_avg = ms / reps;
    Dims dims[CAD_MG_MAX];
    const int nl = fp64 ? c->mg64_n : c->mg_n;
    for (int l = 0; l < nl; l++) dims[l] = fp64 ? c->mg64[l].d : c->mg[l].d;
    *bytes = vcycle_bytes(c, dims, nl, fp64 ? 8 : 4);
    return CAD_OK;
}
