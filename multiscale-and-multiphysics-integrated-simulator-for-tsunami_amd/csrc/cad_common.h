// cad_common.h -- shared definitions of the gfx950 VOF-SMAC step library.
//
// Field layout in HBM is the reference's own: every field is one contiguous
// (NUMI,NUMJ,NUMK) array with I fastest (Source code/vf_a1main.f:150-151),
// FP64 for real fields and int32 for flags, so that a field moves between host
// and device with one copy and a wavefront of 64 consecutive I reads 512
// contiguous bytes.  Metric tables XX(6,NUMI), YY(6,NUMJ), ZZ(6,NUMK) keep the
// reference's (6,N) shape (vf_cgrid.f:44-67).  Indices inside kernels are
// 1-based like the reference so each expression can be checked against the
// Fortran line it restates.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

#include "../../include/cadmas_hip.h"

#define CAD_CHECK(x)                                                              \
    do {                                                                          \
        hipError_t e_ = (x);                                                      \
        if (e_ != hipSuccess) {                                                   \
            cad_set_error("%s:%d %s: %s", __FILE__, __LINE__, #x,                 \
                          hipGetErrorString(e_));                                 \
            return CAD_ERR_HIP;                                                   \
        }                                                                         \
    } while (0)

void cad_set_error(const char *fmt, ...);

// Grid geometry passed by value to every kernel.
struct Dims {
    int ni, nj, nk;      // NUMI, NUMJ, NUMK (cells + 2 ghosts)
    long long nij;       // NUMI*NUMJ
    long long nc;        // NUMI*NUMJ*NUMK
    int numb;            // boundary faces
};

// Dupuit-Forchheimer drag table of OPTION DRAG-DF (vf_iiopt.f:86-108),
// passed by value to the predictor: n rows (0 = the CD drag form), YU and
// per row the diameter DR and the coefficients AP, BT.
#define CAD_DRAG_ROWS 10
struct DragDF {
    int n;
    double yu;
    double dr[CAD_DRAG_ROWS], ap[CAD_DRAG_ROWS], bt[CAD_DRAG_ROWS];
};

// Damping zones of MODEL DAMP (vf_iimdl.f:216-258) as VF_VEULER uses them
// (vf_veuler.f:75-113), passed by value to the predictor: per side x-, x+,
// y-, y+ the degree (-1 = off), the width and the coefficients
// P = (degree+1) PARAM sqrt(GRZ0/DEPTH) for u, v (PARAM-XY) and w (PARAM-Z);
// the grid's extent (vf_iigrid.f:124-131).
struct Damp {
    int deg[4];
    double w[4], px[4], pz[4];
    double xmin, xmax, ymin, ymax;
};

// Read-only set-up arrays (device pointers).
struct Static {
    const double *xx, *yy, *zz;
    const double *anu, *cd0;
    const double *ggv, *ggx, *ggy, *ggz;
    const double *glv, *glx, *gly, *glz;
    const int *indx, *indy, *indz;
    const int *indb;     // INDB(4,NUMB)
    int lvo;             // sides (bits 0..3 = X-, X+, Y-, Y+) driven by a LEVEL-ONLY table
};

// The boundary type of face is (1-based) in VF_BWUWN's list pass, VF_VPCOEF,
// VF_VMODIF and VF_BWUWT: a wave-maker face (5) on a side with a fixed level
// (LEVEL-ONLY table) is taken there as a free face (4) (vf_bwuwn.f:161-211,
// vf_vpcoef.f:162-175, vf_vmodif.f:87-110, vf_bwuwt.f:80-116).
__device__ __forceinline__ int face_type(const int *indb, int lvo, int is)
{
    const int *e = indb + (long long)(is - 1) * 4;
    const int ib = e[2], ns = e[1];
    return (ib == 5 && ns >= 1 && ns <= 4 && ((lvo >> (ns - 1)) & 1)) ? 4 : ib;
}

__device__ __forceinline__ long long cid(const Dims &d, int i, int j, int k)
{
    return (long long)(i - 1) + (long long)d.ni * (j - 1) + d.nij * (k - 1);
}

// metric entry m (1..6) of axis a at 1-based index i
template <int A>
__device__ __forceinline__ double met(const Static &s, int m, int i)
{
    const double *t = A == 0 ? s.xx : (A == 1 ? s.yy : s.zz);
    return t[(long long)(i - 1) * 6 + (m - 1)];
}

__device__ __forceinline__ double metr(const Static &s, int a, int m, int i)
{
    const double *t = a == 0 ? s.xx : (a == 1 ? s.yy : s.zz);
    return t[(long long)(i - 1) * 6 + (m - 1)];
}

__device__ __forceinline__ long long stride(const Dims &d, int a)
{
    return a == 0 ? 1LL : (a == 1 ? (long long)d.ni : d.nij);
}

__device__ __forceinline__ double dmaxd(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double dmind(double a, double b) { return a < b ? a : b; }

// Thread -> cell mapping of the full-array kernels: x = I (64 wide), y = J, z = K.
__device__ __forceinline__ bool cell_of_thread(const Dims &d, int &i, int &j, int &k)
{
    i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    k = blockIdx.z * blockDim.z + threadIdx.z + 1;
    return i <= d.ni && j <= d.nj && k <= d.nk;
}

inline dim3 cell_block() { return dim3(64, 4, 1); }
inline dim3 cell_grid(const Dims &d)
{
    return dim3((unsigned)((d.ni + 63) / 64), (unsigned)((d.nj + 3) / 4), (unsigned)d.nk);
}

// Deterministic FP64 reduction: fixed grid of REDUCE_BLOCKS blocks, each block
// writes one partial, one block sums the partials in index order.
constexpr int REDUCE_BLOCKS = 1024;
constexpr int REDUCE_THREADS = 256;

// One level of the multigrid (cad_mg.hip): a grid with one ghost layer, its
// 7-point rows (inactive rows AD = 1, no couplings), the level's unknowns and
// right-hand side, and the row activity, stored colour by colour in lines of
// ld cells; level 0 of the FP32 hierarchy also holds its rows and activity in
// the grid's natural order (nad .. nauk, nact; null elsewhere).
constexpr int CAD_MG_MAX = 12;
template <typename V>
struct MgLevelT {
    Dims d;
    int ld;
    V *ad, *ali, *alj, *alk, *aui, *auj, *auk;
    V *x, *b;
    unsigned char *act;
    V *nad, *nali, *nalj, *nalk, *naui, *nauj, *nauk;
    unsigned char *nact;
};
typedef MgLevelT<float> MgLevel;

// Device-side context of one grid (all fields HBM resident).
struct Fields {
    double *uu, *vv, *ww, *pp, *ff;
    double *bcu, *bcv, *bcw, *bcp, *bcf;
    int *nf, *indc, *nlim;
};

struct cad_ctx {
    int device;
    int ncu;   // compute units of the device
    hipStream_t stream;
    Dims d;
    cad_params prm;
    // set-up arrays
    double *xx, *yy, *zz, *anu, *cd0, *ggv, *ggx, *ggy, *ggz, *glv, *glx, *gly, *glz, *bcvi;
    int *indx, *indy, *indz, *indb;
    Fields f;
    // spare velocity buffers for the out-of-place predictor
    double *u2, *v2, *w2;
    // Poisson matrix (AD..BB) and potential PT
    double *ad, *ali, *alj, *alk, *aui, *auj, *auk, *bb, *pt;
    // BiCGSTAB work vectors (vf_m1bcgs.f: DD, XW, YY, R0, RR, PP, AP, AE)
    double *dd, *xw, *yv, *r0, *rr, *pv, *ap, *ae;
    // MILU products ALI*DD(n-1), ALJ*DD(n-NI), ALK*DD(n-NIJ) of the pipelined sweeps
    double *lpi, *lpj, *lpk;
    // pipelined sweep schedule (cad_sweep.hip); tile shape 0 = 8x8 lines, 1 = 16x4
    int sweep_schedule;                  // CAD_SWEEP_PIPELINED or CAD_SWEEP_HYPERPLANE
    int sw_shape;
    int sw_ntj[2], sw_ntk[2], sw_ntiles[2];
    int *sw_order[2], *sw_range[2], *sw_err;
    void *sw_mb;
    unsigned long long sw_mb_bytes;
    unsigned long long *sw_ticket;
    unsigned long long sw_base;
    unsigned sw_seq;
    // VOF work: F fluxes
    double *flf[3];
    // scratch for reductions and counters
    double *red_partial;    // REDUCE_BLOCKS * 4
    double *red_out;        // 16 results
    double *ksc;            // BiCGSTAB scalars on the device (32 slots)
    unsigned long long *act_rng;   // [n0, n1) of the active cells of the current solve
    int *icount;            // 16 int counters
    double *host_red;       // pinned mirror of red_out
    int *host_cnt;          // pinned mirror of icount
    // timers
    hipEvent_t ev[16];
    bool have_static, have_state;
    // the surface-cell count of the last flagging pass
    int nums;
    // cap on VF_FNFINI's flag-fix passes (0: interior cells + 1, a bound no
    // valid state reaches); cad_set_flag_pass_limit
    long long flag_pass_limit;
    // the plain solve's extra stop on the true relative residual (0: the
    // reference's stop test alone); cad_set_true_residual_stop
    double true_stop;
    // TimerDoor models (cad_tdoor.hip), allocated on first use: droplet
    // fields DROPTX..DROPWW, per-line droplet flags, the level mask of the
    // bubble timers, counters and their pinned mirror; TBUB(NUMK) and the
    // ZZ table on the host
    double *drop[6];
    int *td_flag, *td_cnt, *td_host;
    unsigned char *td_lvl;
    double *tbub, *zz_host;
    bool td_alloc, td_full;
    // matrix-table wave-maker (cad_wave.hip): the table's heights and
    // velocities on the device (ZZ, UN, UT), its times and levels on the
    // host, the time search start MTBNOW, the sides and an error flag
    // the wave-maker tables: FUNC MATRIX (<name>.mtb) and MATRIX2
    // (<name>.mtb2), each with its own kind, search start and phase origin
    struct WaveTable {
        double *tab;        // device: ZZ(NZ) | UN(NZ,NT) | UT(NZ,NT)
        double *tt, *hh;    // host: times and levels
        int nz, nt, mtbnow; // MTBZZ, MTBTT, MTBNOW
        double t0;          // DMTBT0
        int off, only;      // LEVEL-OFF, LEVEL-ONLY
    } wt[2];
    double wv_wvlvl;
    int wv_on[4], wv_l1[4], wv_l2[4];
    double wv_d[4], wv_t[4], wv_a[4], wv_level[4];
    int *wv_err, *wv_err_host;
    int n_wave_faces;   // boundary faces of type 5 in INDB
    // the theory sides (cad_wave_theory.hip): per side the function
    // (-3 / -4 the first / second table, -2 Stokes, -1 cnoidal), height, phase origin and the
    // persistent secant step of VF_STK0; per side the columns' level scales
    // and face velocities (device and pinned host)
    int wv_kind[4];
    double wv_h[4], wv_t0s[4];
    double wv_wl[4];        // wave length of VF_CWMAK0 at set-up (BCTYP(4,side))
    double wv_ang[4][3];    // ANGLE: degrees and the ray's origin SX, SY (BCTYP(9..11,side))
    float wv_stk_dddk;
    double *wv_vws, *wv_vws_host, *wv_uv, *wv_uv_host, *wv_ph_host;
    size_t wv_cols;     // columns per side in the buffers (max(NUMI, NUMJ))
    // radiation sides (cad_wave.hip): on, depth, period; the x and y metric
    // tables on the host for their cell widths
    int op_on[4];
    double op_d[4], op_t[4];
    int n_open_faces;   // boundary faces of type 7 in INDB
    double *xx_host, *yy_host;
    DragDF drag;        // OPTION DRAG-DF (n = 0: the CD form)
    // MODEL DAMP per side: on, DEGREE, PARAM-XY, PARAM-Z, WIDTH, DEPTH
    int dm_on[4], dm_deg[4];
    double dm_par[4][4];
    // the mixed form's multigrid levels (allocated on first use)
    MgLevel mg[CAD_MG_MAX];
    int mg_n;
    // an FP64 hierarchy of the same matrix, built only to time an FP64 V-cycle
    MgLevelT<double> mg64[CAD_MG_MAX];
    int mg64_n;
    unsigned long long mg_span;   // cells in the active index range of the current solve
    int mg_k0, mg_k1;             // K planes of level 0's passes in the current solve
    int mg_set_k0, mg_set_k1;     // K planes level 0's rows were last built on (0: none yet)
    // the learned form's network, Krylov basis and counters (cad_learned.hip),
    // allocated when its weights are set
    struct Learned *lrn;
    // the mixed form's solves that stalled and were finished by MILU-BiCGSTAB,
    // and the BiCGSTAB iterations of the last such solve
    long long mixed_fallbacks;
    int mixed_fallback_iters;
    // cap on the mixed and learned forms' own iterations (0: 200)
    int krylov_cap;
    // in-solve timing of the forms' dominant launches (cad_set_kernel_timing)
    struct cad_ktimer *kt;
};

// ---- in-solve kernel timing --------------------------------------------
// While on, every n-th launch of a form's dominant kernel inside the solves
// (the plain form's M^-1, the mixed form's FP32 V-cycle, the learned form's
// network pass) is bracketed by a pair of HIP events on the library stream,
// with the launch's algorithmic work (bytes or FLOP; 0 where the caller
// supplies it).  The events carry no system-scope fence, so recording them
// flushes no cache.  Pairs sit in a ring per kernel; a slot's elapsed time is
// taken when the slot comes round again, and all are taken when read.
enum { CAD_KT_MINV = 0, CAD_KT_VCYCLE = 1, CAD_KT_UNET = 2, CAD_KT_N = 3 };
struct cad_ktimer {
    static constexpr int PAIRS = 256;
    bool on;
    int every;                                   // time launches 0, n, 2n, ...
    long long seen[CAD_KT_N];                    // launches since turned on
    bool open[CAD_KT_N];                         // a pair's start is recorded
    hipEvent_t ev[CAD_KT_N][2 * PAIRS];
    double work[CAD_KT_N][PAIRS];
    long long next[CAD_KT_N], done[CAD_KT_N];   // pairs recorded, pairs taken
    double ms[CAD_KT_N], total_work[CAD_KT_N];
};
extern "C" {
int cad_kt_begin(cad_ctx *c, int which);
int cad_kt_end(cad_ctx *c, int which, double work);
void cad_kt_free(cad_ctx *c);
}

// ---- launchers implemented in the phase files ---------------------------
int cad_launch_predictor(cad_ctx *c);
// initial: VF_CINIT's pass (NNOW = 0), else the pass after the predictor
int cad_launch_bwuwn(cad_ctx *c, bool initial = false);
int cad_launch_vpcoef(cad_ctx *c);
int cad_solve_m1bcgs(cad_ctx *c, double *x0, double *bnorm, double *xnorm, int *iter);
int cad_launch_vmodif(cad_ctx *c);
int cad_launch_bsuwt(cad_ctx *c);
int cad_launch_bsuwt2(cad_ctx *c);
int cad_launch_bsuwn(cad_ctx *c);
int cad_launch_bwuwt(cad_ctx *c);
int cad_launch_bwpp(cad_ctx *c);
int cad_vof_phase(cad_ctx *c, double *fsum, double *fcut);
int cad_reduce_dot(cad_ctx *c, const double *x, const double *y, const int *mask, double *out_dev);
int cad_reduce_fetch(cad_ctx *c, int n, double *host_out);
int cad_cdiv(cad_ctx *c, double *dv);
int cad_sweep_alloc(cad_ctx *c);
void cad_sweep_free(cad_ctx *c);
unsigned long long cad_sweep_bytes(const cad_ctx *c);
int cad_sweep_bind(cad_ctx *c, char *base, unsigned long long bytes, unsigned seq);
int cad_sweep_dcmp(cad_ctx *c, double u, int *npos_host);
int cad_sweep_minv(cad_ctx *c, double *x, const double *y);
int cad_sweep_status(cad_ctx *c, int *code);
int cad_milu_factor_impl(cad_ctx *c, int *npos);
int cad_poisson_residual_impl(cad_ctx *c, double *rel);
int cad_poisson_residual_x(cad_ctx *c, const double *x, double *rel);
int cad_krylov_cap(const cad_ctx *c);
int cad_solve_mixed(cad_ctx *c, double *x0, double *bnorm, double *xnorm, int *iter);
int cad_mg_alloc(cad_ctx *c);
void cad_mg_free(cad_ctx *c);
int cad_mg_setup(cad_ctx *c);
int cad_mg_apply(cad_ctx *c, const double *y, double *x, int k0, int k1);
int cad_mg_apply32(cad_ctx *c, const float *y, float *x, int k0, int k1);
// cad_mg_apply and r = y - A x on the V-cycle's FP32 level-0 matrix, with the
// block partials of sum r^2 and of the row count at part[b * stride] and
// part[b * stride + 1], b < REDUCE_BLOCKS
int cad_mg_apply_resid(cad_ctx *c, const double *y, double *x, double *r, double *part, int stride, int k0, int k1);
int cad_mg_time(cad_ctx *c, int reps, int fp64, float *ms_avg, double *bytes);
int cad_milu_apply_impl(cad_ctx *c, const double *y, double *x);
Static cad_static(const cad_ctx *c);
int cad_fnfini(cad_ctx *c);
int cad_tdoor_alloc(cad_ctx *c);
void cad_tdoor_free(cad_ctx *c);
int cad_tdoor_init(cad_ctx *c);
int cad_tdoor_phase(cad_ctx *c);
int cad_wave_table(cad_ctx *c, int which, int nz, int nt, const double *zz, const double *tt,
                   const double *hh, const double *un, const double *ut, double t0);
int cad_wave_side(cad_ctx *c, int which, int side, int l1, int l2, double depth, double period,
                  double ampl, double wvlvl);
int cad_wave_apply(cad_ctx *c, bool initial);
int cad_open_side(cad_ctx *c, int side, double depth, double period);
int cad_wave_check(cad_ctx *c);
int cad_launch_bwffsf(cad_ctx *c);
int cad_level_sides(const cad_ctx *c);
int cad_wave_theory_side(cad_ctx *c, int side, int func, int l1, int l2, double depth, double height,
                         double period, double ampl, double wvlvl, double angle, double ax, double ay);
int cad_wave_theory_apply(cad_ctx *c, int jd);
void cad_wave_theory_free(cad_ctx *c);
void cad_wave_free(cad_ctx *c);
int cad_ext_active_range(cad_ctx *c, unsigned long long hr[2]);
int cad_ext_apply_a(cad_ctx *c, const double *x, double *y);
int cad_ext_residual(cad_ctx *c, const double *x, double *y);
int cad_ext_dot(cad_ctx *c, const double *x, const double *y, double *out);
int cad_solve_learned(cad_ctx *c, double *x0, double *bnorm, double *xnorm, int *iter);
void cad_learned_free(cad_ctx *c);
