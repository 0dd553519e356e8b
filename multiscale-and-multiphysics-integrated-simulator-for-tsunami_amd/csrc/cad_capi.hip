// cad_capi.hip -- extern "C" entry points, device context and the step driver.
//
// The step order follows the reference's time loop for one sub-loop
// (Source code/vf_a1main.f:1044-1114): VF_V1CAL (phases a, b, d) then VF_F1CAL
// (phase c).  All fields live in HBM between steps; the host touches them only
// through cad_upload_* / cad_download_*.
#include "cad_common.h"

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

static thread_local char g_err[1024] = "";

void cad_set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

extern "C" const char *cad_last_error(void) { return g_err; }

Static cad_static(const cad_ctx *c)
{
    Static s;
    s.xx = c->xx; s.yy = c->yy; s.zz = c->zz;
    s.anu = c->anu; s.cd0 = c->cd0;
    s.ggv = c->ggv; s.ggx = c->ggx; s.ggy = c->ggy; s.ggz = c->ggz;
    s.glv = c->glv; s.glx = c->glx; s.gly = c->gly; s.glz = c->glz;
    s.indx = c->indx; s.indy = c->indy; s.indz = c->indz;
    s.indb = c->indb;
    s.lvo = cad_level_sides(c);
    return s;
}

int cad_flag_cells(cad_ctx *c, int *nums);
int cad_init_phase(cad_ctx *c, double wvlvl, const double *uvw, int flags, double *fsum, double *fcut);
int cad_time_minv_impl(cad_ctx *c, int reps, float *ms_avg);
int cad_cdtcal_impl(cad_ctx *c, int nnow, double dtinit, double dtsafe, double dtmin, double dtmax,
                    double *dt);
int cad_launch_bsuwn3(cad_ctx *c);

namespace {

__global__ void k_cdiv(Dims d, Static s, const double *__restrict__ u, const double *__restrict__ v,
                       const double *__restrict__ w, const int *__restrict__ indc, double *partial)
{
    __shared__ double sh[REDUCE_THREADS];
    double acc = 0.0;
    for (long long n = (long long)blockIdx.x * REDUCE_THREADS + threadIdx.x; n < d.nc;
         n += (long long)REDUCE_BLOCKS * REDUCE_THREADS) {
        const int k = (int)(n / d.nij) + 1;
        const long long r = n - (long long)(k - 1) * d.nij;
        const int j = (int)(r / d.ni) + 1;
        const int i = (int)(r - (long long)(j - 1) * d.ni) + 1;
        if (i < 2 || i > d.ni - 1 || j < 2 || j > d.nj - 1 || k < 2 || k > d.nk - 1) continue;
        if (indc[n] == -1) continue;
        const long long ni = d.ni, nij = d.nij;
        // vf_cdiv00.f:41-47
        double vd = met<1>(s, 2, j) * met<2>(s, 2, k) * (s.ggx[n + 1] * u[n + 1] - s.ggx[n] * u[n])
                  + met<0>(s, 2, i) * met<2>(s, 2, k) * (s.ggy[n + ni] * v[n + ni] - s.ggy[n] * v[n])
                  + met<0>(s, 2, i) * met<1>(s, 2, j) * (s.ggz[n + nij] * w[n + nij] - s.ggz[n] * w[n]);
        acc += vd * vd;
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int off = REDUCE_THREADS / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

__global__ void k_fill0(long long nc, double *a)
{
    for (long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x; n < nc;
         n += (long long)gridDim.x * blockDim.x)
        a[n] = 0.0;
}

template <typename T>
int dalloc(T **p, size_t n)
{
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void **)p, n * sizeof(T));
    if (e != hipSuccess) {
        cad_set_error("hipMalloc of %zu bytes failed: %s", n * sizeof(T), hipGetErrorString(e));
        *p = nullptr;
        return CAD_ERR_HIP;
    }
    return CAD_OK;
}

template <typename T>
int h2d(cad_ctx *c, T *dst, const T *src, size_t n)
{
    if (!src) {
        cad_set_error("null host pointer");
        return CAD_ERR_ARG;
    }
    CAD_CHECK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, c->stream));
    return CAD_OK;
}

template <typename T>
int d2h(cad_ctx *c, T *dst, const T *src, size_t n)
{
    if (!dst) return CAD_OK;  // caller does not want this array
    CAD_CHECK(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToHost, c->stream));
    return CAD_OK;
}

#define RET_IF(x)              \
    do {                       \
        int rc_ = (x);         \
        if (rc_) return rc_;   \
    } while (0)

}  // namespace

int cad_cdiv(cad_ctx *c, double *dv)
{
    Static s = cad_static(c);
    hipLaunchKernelGGL(k_cdiv, dim3(REDUCE_BLOCKS), dim3(REDUCE_THREADS), 0, c->stream, c->d, s,
                       c->f.uu, c->f.vv, c->f.ww, c->f.indc, c->red_partial);
    CAD_CHECK(hipGetLastError());
    double v;
    RET_IF(cad_reduce_fetch(c, 1, &v));
    *dv = std::sqrt(v);
    return CAD_OK;
}

extern "C" {

int cad_create(int numi, int numj, int numk, int numb, int device, cad_ctx **out)
{
    if (!out || numi < 4 || numj < 4 || numk < 4 || numb < 0) {
        cad_set_error("cad_create: invalid sizes %d x %d x %d, numb %d", numi, numj, numk, numb);
        return CAD_ERR_ARG;
    }
    *out = nullptr;
    CAD_CHECK(hipSetDevice(device));
    cad_ctx *c = new cad_ctx();
    std::memset((void *)c, 0, sizeof(cad_ctx));
    c->device = device;
    if (hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
        c->ncu < 1)
        c->ncu = 256;
    c->d.ni = numi;
    c->d.nj = numj;
    c->d.nk = numk;
    c->d.nij = (long long)numi * numj;
    c->d.nc = c->d.nij * numk;
    c->d.numb = numb;
    const size_t nc = (size_t)c->d.nc, nb = (size_t)(numb > 0 ? numb : 1);
    int rc = 0;
    double **f64[] = {&c->f.uu, &c->f.vv, &c->f.ww, &c->f.pp, &c->f.ff, &c->u2, &c->v2, &c->w2,
                      &c->anu, &c->cd0, &c->ggv, &c->ggx, &c->ggy, &c->ggz, &c->glv, &c->glx,
                      &c->gly, &c->glz, &c->ad, &c->ali, &c->alj, &c->alk, &c->aui, &c->auj,
                      &c->auk, &c->bb, &c->pt, &c->dd, &c->xw, &c->yv, &c->r0, &c->rr, &c->pv,
                      &c->ap, &c->ae, &c->flf[0], &c->flf[1], &c->flf[2], &c->lpi, &c->lpj,
                      &c->lpk};
    for (double **p : f64) rc = rc ? rc : dalloc(p, nc);
    int **i32[] = {&c->f.nf, &c->f.indc, &c->f.nlim, &c->indx, &c->indy, &c->indz};
    for (int **p : i32) rc = rc ? rc : dalloc(p, nc);
    double **bnd[] = {&c->f.bcu, &c->f.bcv, &c->f.bcw, &c->f.bcp, &c->f.bcf, &c->bcvi};
    for (double **p : bnd) rc = rc ? rc : dalloc(p, nb);
    rc = rc ? rc : dalloc(&c->indb, 4 * nb);
    rc = rc ? rc : dalloc(&c->xx, 6 * (size_t)numi);
    rc = rc ? rc : dalloc(&c->yy, 6 * (size_t)numj);
    rc = rc ? rc : dalloc(&c->zz, 6 * (size_t)numk);
    rc = rc ? rc : dalloc(&c->red_partial, 4 * (size_t)REDUCE_BLOCKS);
    rc = rc ? rc : dalloc(&c->red_out, 16);
    rc = rc ? rc : dalloc(&c->ksc, 32);
    rc = rc ? rc : dalloc(&c->act_rng, 2);
    // 16 scalar counters, then the surface-orientation count slots (2 x 1024)
    rc = rc ? rc : dalloc(&c->icount, 16 + 2 * 1024);
    rc = rc ? rc : cad_sweep_alloc(c);
    if (rc) {
        cad_destroy(c);
        return rc;
    }
    c->zz_host = new double[6 * (size_t)numk]();
    c->xx_host = new double[6 * (size_t)numi]();
    c->yy_host = new double[6 * (size_t)numj]();
    if (hipHostMalloc((void **)&c->host_red, 16 * sizeof(double), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&c->host_cnt, 16 * sizeof(int), hipHostMallocDefault) != hipSuccess) {
        cad_set_error("hipHostMalloc failed");
        cad_destroy(c);
        return CAD_ERR_HIP;
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        cad_set_error("hipStreamCreate failed");
        cad_destroy(c);
        return CAD_ERR_HIP;
    }
    for (auto &e : c->ev) hipEventCreate(&e);
    // defaults of vf_a2dflt.f
    cad_params p;
    p.rho0 = 1.0e3; p.grz0 = 9.8; p.anu0 = 1.0e-6; p.dtnow = 1.0e-3;
    p.scmvp = 1.0; p.vvmax = 30.0; p.flower = 1.0e-6; p.fupper = 1.0 - 1.0e-6; p.plower = 1.0e-4;
    p.cgpara = 0.95; p.cgepsa = 1.0e-15; p.cgepsr = 1.0e-12; p.cgdiv = 1.0e30;
    p.icgtyp = 1; p.icgmax = 500; p.ibsuw0 = 0; p.loops = 1; p.precision = CAD_PRECISION_FP64;
    p.idrop = 1; p.wbub = 0.2; p.tnow = 0.0;
    c->prm = p;
    *out = c;
    return CAD_OK;
}

void cad_destroy(cad_ctx *c)
{
    if (!c) return;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    void *ptrs[] = {c->f.uu, c->f.vv, c->f.ww, c->f.pp, c->f.ff, c->u2, c->v2, c->w2, c->anu, c->cd0,
                    c->ggv, c->ggx, c->ggy, c->ggz, c->glv, c->glx, c->gly, c->glz, c->ad, c->ali,
                    c->alj, c->alk, c->aui, c->auj, c->auk, c->bb, c->pt, c->dd, c->xw, c->yv,
                    c->r0, c->rr, c->pv, c->ap, c->ae, c->flf[0], c->flf[1], c->flf[2], c->f.nf,
                    c->f.indc, c->f.nlim, c->indx, c->indy, c->indz, c->f.bcu, c->f.bcv, c->f.bcw,
                    c->f.bcp, c->f.bcf, c->bcvi, c->indb, c->xx, c->yy, c->zz, c->red_partial,
                    c->red_out, c->icount, c->lpi, c->lpj, c->lpk, c->ksc, c->act_rng};
    for (void *p : ptrs)
        if (p) hipFree(p);
    cad_sweep_free(c);
    cad_tdoor_free(c);
    cad_wave_free(c);
    cad_wave_theory_free(c);
    cad_mg_free(c);
    cad_learned_free(c);
    cad_kt_free(c);
    delete[] c->zz_host;
    delete[] c->xx_host;
    delete[] c->yy_host;
    if (c->host_red) hipHostFree(c->host_red);
    if (c->host_cnt) hipHostFree(c->host_cnt);
    for (auto &e : c->ev)
        if (e) hipEventDestroy(e);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

int cad_set_params(cad_ctx *c, const cad_params *p)
{
    if (!c || !p) {
        cad_set_error("cad_set_params: null argument");
        return CAD_ERR_ARG;
    }
    if (p->loops < 1) {
        cad_set_error("OPTION SUB-LOOP %d: LOOPS must be >= 1 (vf_iiopt.f:44)", p->loops);
        return CAD_ERR_ARG;
    }
    if (p->precision != CAD_PRECISION_FP64 && p->precision != CAD_PRECISION_MIXED &&
        p->precision != CAD_PRECISION_LEARNED) {
        cad_set_error("precision mode %d is not available in this build", p->precision);
        return CAD_ERR_UNSUPPORTED;
    }
    if (!(p->dtnow > 0.0) || !(p->rho0 > 0.0)) {
        cad_set_error("cad_set_params: DT and RHO0 must be positive");
        return CAD_ERR_ARG;
    }
    if (p->wbub < 0.0 || p->idrop < 0 || p->idrop > 1) {
        cad_set_error("cad_set_params: WBUB %g / IDROP %d invalid (vf_iiopt.f:60-75)", p->wbub, p->idrop);
        return CAD_ERR_ARG;
    }
    c->prm = *p;
    if (p->wbub >= 1.0e-20 || p->idrop >= 1) {
        CAD_CHECK(hipSetDevice(c->device));
        RET_IF(cad_tdoor_alloc(c));
    }
    return CAD_OK;
}

int cad_get_params(cad_ctx *c, cad_params *p)
{
    if (!c || !p) {
        cad_set_error("cad_get_params: null argument");
        return CAD_ERR_ARG;
    }
    *p = c->prm;
    return CAD_OK;
}

int cad_upload_tdoor(cad_ctx *c, const double *tbub, const double *droptx, const double *dropty,
                     const double *droptz, const double *dropuu, const double *dropvv,
                     const double *dropww)
{
    if (!c || !tbub || !droptx || !dropty || !droptz || !dropuu || !dropvv || !dropww) {
        cad_set_error("cad_upload_tdoor: null argument");
        return CAD_ERR_ARG;
    }
    CAD_CHECK(hipSetDevice(c->device));
    RET_IF(cad_tdoor_alloc(c));
    const size_t nc = (size_t)c->d.nc;
    const double *src[6] = {droptx, dropty, droptz, dropuu, dropvv, dropww};
    for (int a = 0; a < 6; a++) RET_IF(h2d(c, c->drop[a], src[a], nc));
    CAD_CHECK(hipStreamSynchronize(c->stream));
    for (int k = 0; k < c->d.nk; k++) c->tbub[k] = tbub[k];
    // the next droplet pass clears every cell without a droplet itself
    c->td_full = true;
    return CAD_OK;
}

int cad_download_tdoor(cad_ctx *c, double *tbub, double *droptx, double *dropty, double *droptz,
                       double *dropuu, double *dropvv, double *dropww)
{
    if (!c || !tbub || !droptx || !dropty || !droptz || !dropuu || !dropvv || !dropww) {
        cad_set_error("cad_download_tdoor: null argument");
        return CAD_ERR_ARG;
    }
    CAD_CHECK(hipSetDevice(c->device));
    RET_IF(cad_tdoor_alloc(c));
    const size_t nc = (size_t)c->d.nc;
    double *dst[6] = {droptx, dropty, droptz, dropuu, dropvv, dropww};
    for (int a = 0; a < 6; a++) RET_IF(d2h(c, dst[a], c->drop[a], nc));
    CAD_CHECK(hipStreamSynchronize(c->stream));
    for (int k = 0; k < c->d.nk; k++) tbub[k] = c->tbub[k];
    return CAD_OK;
}

int cad_upload_static(cad_ctx *c, const double *xx, const double *yy, const double *zz,
                      const double *anu, const double *cd0, const double *ggv, const double *ggx,
                      const double *ggy, const double *ggz, const double *glv, const double *glx,
                      const double *gly, const double *glz, const int *indx, const int *indy,
                      const int *indz, const int *indb, const double *bcvi)
{
    if (!c) return CAD_ERR_ARG;
    CAD_CHECK(hipSetDevice(c->device));
    const size_t nc = (size_t)c->d.nc, nb = (size_t)c->d.numb;
    // boundary types this build implements (vf_bwuwn.f, vf_bwuwt.f, vf_bwff.f)
    for (size_t l = 0; l < nb; l++) {
        const int tv = indb[4 * l + 2], tf = indb[4 * l + 3];
        // wave-maker (5) and radiation (7) faces carry the type for both on
        // an x or y side (vf_csetup.f:146-205)
        const bool side = (tv == 5 || tv == 7) && tf == tv && indb[4 * l + 1] <= 4;
        if (!side && (!(tv == 1 || tv == 2 || tv == 3 || tv == 4) || !(tf == 1 || tf == 2))) {
            cad_set_error("boundary face %zu: velocity type %d / F type %d not implemented "
                          "(log-law boundaries; wave-maker and radiation faces are type 5 / 7 "
                          "for both on an x or y side)", l + 1, tv, tf);
            return CAD_ERR_UNSUPPORTED;
        }
        if (indb[4 * l] < 1 || (long long)indb[4 * l] > c->d.nc || indb[4 * l + 1] < 1 ||
            indb[4 * l + 1] > 6) {
            cad_set_error("boundary face %zu: invalid INDB entry", l + 1);
            return CAD_ERR_ARG;
        }
    }
    // face indices: -1 (between obstacles), 0 (normal) or a boundary entry 1..NUMB
    const int *inds3[3] = {indx, indy, indz};
    for (int a = 0; a < 3; a++) {
        if (!inds3[a]) {
            cad_set_error("null face-index array");
            return CAD_ERR_ARG;
        }
        for (size_t n = 0; n < nc; n++) {
            const int v = inds3[a][n];
            if (v < -1 || v > c->d.numb) {
                cad_set_error("face index %d at linear position %zu outside [-1, NUMB=%d]", v, n,
                              c->d.numb);
                return CAD_ERR_ARG;
            }
        }
    }
    c->n_wave_faces = c->n_open_faces = 0;
    for (size_t l = 0; l < nb; l++) {
        c->n_wave_faces += indb[4 * l + 2] == 5;
        c->n_open_faces += indb[4 * l + 2] == 7;
    }
    RET_IF(h2d(c, c->xx, xx, 6 * (size_t)c->d.ni));
    RET_IF(h2d(c, c->yy, yy, 6 * (size_t)c->d.nj));
    RET_IF(h2d(c, c->zz, zz, 6 * (size_t)c->d.nk));
    std::memcpy(c->zz_host, zz, 6 * (size_t)c->d.nk * sizeof(double));
    std::memcpy(c->xx_host, xx, 6 * (size_t)c->d.ni * sizeof(double));
    std::memcpy(c->yy_host, yy, 6 * (size_t)c->d.nj * sizeof(double));
    RET_IF(h2d(c, c->anu, anu, nc));
    RET_IF(h2d(c, c->cd0, cd0, nc));
    RET_IF(h2d(c, c->ggv, ggv, nc));
    RET_IF(h2d(c, c->ggx, ggx, nc));
    RET_IF(h2d(c, c->ggy, ggy, nc));
    RET_IF(h2d(c, c->ggz, ggz, nc));
    RET_IF(h2d(c, c->glv, glv, nc));
    RET_IF(h2d(c, c->glx, glx, nc));
    RET_IF(h2d(c, c->gly, gly, nc));
    RET_IF(h2d(c, c->glz, glz, nc));
    RET_IF(h2d(c, c->indx, indx, nc));
    RET_IF(h2d(c, c->indy, indy, nc));
    RET_IF(h2d(c, c->indz, indz, nc));
    if (nb > 0) {
        RET_IF(h2d(c, c->indb, indb, 4 * nb));
        if (bcvi) RET_IF(h2d(c, c->bcvi, bcvi, nb));
    }
    CAD_CHECK(hipStreamSynchronize(c->stream));
    c->have_static = true;
    return CAD_OK;
}

int cad_upload_state(cad_ctx *c, const double *uu, const double *vv, const double *ww,
                     const double *pp, const double *ff, const double *bcu, const double *bcv,
                     const double *bcw, const double *bcp, const double *bcf, const int *nf,
                     const int *indc)
{
    if (!c) return CAD_ERR_ARG;
    CAD_CHECK(hipSetDevice(c->device));
    const size_t nc = (size_t)c->d.nc, nb = (size_t)c->d.numb;
    // every kernel relies on the ghost layer being an inactive obstacle layer
    // (VF_CNFDFL / VF_CINDC): check it before anything is launched
    if (!nf || !indc) {
        cad_set_error("null flag array");
        return CAD_ERR_ARG;
    }
    const Dims &d = c->d;
    for (long long n = 0; n < d.nc; n++) {
        const int k = (int)(n / d.nij) + 1;
        const long long r = n - (long long)(k - 1) * d.nij;
        const int j = (int)(r / d.ni) + 1;
        const int i = (int)(r - (long long)(j - 1) * d.ni) + 1;
        const bool ghost = i == 1 || i == d.ni || j == 1 || j == d.nj || k == 1 || k == d.nk;
        if (nf[n] < -1 || nf[n] > 8) {
            cad_set_error("NF=%d at (%d,%d,%d) is not a cell type", nf[n], i, j, k);
            return CAD_ERR_ARG;
        }
        if (ghost && (nf[n] != -1 || indc[n] != -1)) {
            cad_set_error("ghost cell (%d,%d,%d) must have NF=-1 and INDC=-1", i, j, k);
            return CAD_ERR_ARG;
        }
    }
    RET_IF(h2d(c, c->f.uu, uu, nc));
    RET_IF(h2d(c, c->f.vv, vv, nc));
    RET_IF(h2d(c, c->f.ww, ww, nc));
    RET_IF(h2d(c, c->f.pp, pp, nc));
    RET_IF(h2d(c, c->f.ff, ff, nc));
    RET_IF(h2d(c, c->f.nf, nf, nc));
    RET_IF(h2d(c, c->f.indc, indc, nc));
    if (nb > 0) {
        RET_IF(h2d(c, c->f.bcu, bcu, nb));
        RET_IF(h2d(c, c->f.bcv, bcv, nb));
        RET_IF(h2d(c, c->f.bcw, bcw, nb));
        RET_IF(h2d(c, c->f.bcp, bcp, nb));
        RET_IF(h2d(c, c->f.bcf, bcf, nb));
    }
    CAD_CHECK(hipStreamSynchronize(c->stream));
    c->have_state = true;
    return CAD_OK;
}

int cad_download_state(cad_ctx *c, double *uu, double *vv, double *ww, double *pp, double *ff,
                       double *bcu, double *bcv, double *bcw, double *bcp, double *bcf, int *nf,
                       int *indc)
{
    if (!c) return CAD_ERR_ARG;
    CAD_CHECK(hipSetDevice(c->device));
    const size_t nc = (size_t)c->d.nc, nb = (size_t)c->d.numb;
    RET_IF(d2h(c, uu, c->f.uu, nc));
    RET_IF(d2h(c, vv, c->f.vv, nc));
    RET_IF(d2h(c, ww, c->f.ww, nc));
    RET_IF(d2h(c, pp, c->f.pp, nc));
    RET_IF(d2h(c, ff, c->f.ff, nc));
    RET_IF(d2h(c, nf, c->f.nf, nc));
    RET_IF(d2h(c, indc, c->f.indc, nc));
    if (nb > 0) {
        RET_IF(d2h(c, bcu, c->f.bcu, nb));
        RET_IF(d2h(c, bcv, c->f.bcv, nb));
        RET_IF(d2h(c, bcw, c->f.bcw, nb));
        RET_IF(d2h(c, bcp, c->f.bcp, nb));
        RET_IF(d2h(c, bcf, c->f.bcf, nb));
    }
    CAD_CHECK(hipStreamSynchronize(c->stream));
    return CAD_OK;
}

static int ready(cad_ctx *c)
{
    if (!c) {
        cad_set_error("null context");
        return CAD_ERR_ARG;
    }
    if (!c->have_static || !c->have_state) {
        cad_set_error("set-up arrays and state must be uploaded first");
        return CAD_ERR_STATE;
    }
    if (c->n_wave_faces > 0 && !(c->wv_on[0] || c->wv_on[1] || c->wv_on[2] || c->wv_on[3])) {
        cad_set_error("%d wave-maker faces (type 5) but no wave-maker side: call "
                      "cad_set_wave_table and cad_set_wave", c->n_wave_faces);
        return CAD_ERR_STATE;
    }
    if (c->n_open_faces > 0 && !(c->op_on[0] || c->op_on[1] || c->op_on[2] || c->op_on[3])) {
        cad_set_error("%d radiation faces (type 7) but no radiation side: call cad_set_open",
                      c->n_open_faces);
        return CAD_ERR_STATE;
    }
    CAD_CHECK(hipSetDevice(c->device));
    return CAD_OK;
}

int cad_predictor(cad_ctx *c)
{
    RET_IF(ready(c));
    RET_IF(cad_launch_predictor(c));
    RET_IF(cad_launch_bwuwn(c));
    CAD_CHECK(hipStreamSynchronize(c->stream));
    return cad_wave_check(c);
}

int cad_vpcoef(cad_ctx *c)
{
    RET_IF(ready(c));
    return cad_launch_vpcoef(c);
}

int cad_vpsol(cad_ctx *c, cad_step_info *info)
{
    RET_IF(ready(c));
    // VF_VPSOL: PT = 0 before the solve (vf_vpsol.f:55-61)
    hipLaunchKernelGGL(k_fill0, dim3(4096), dim3(256), 0, c->stream, c->d.nc, c->pt);
    CAD_CHECK(hipGetLastError());
    double bn = 0.0, xn = 0.0;
    int it = 0;
    if (c->prm.precision == CAD_PRECISION_MIXED)
        RET_IF(cad_solve_mixed(c, c->pt, &bn, &xn, &it));
    else if (c->prm.precision == CAD_PRECISION_LEARNED)
        RET_IF(cad_solve_learned(c, c->pt, &bn, &xn, &it));
    else
        RET_IF(cad_solve_m1bcgs(c, c->pt, &bn, &xn, &it));
    CAD_CHECK(hipStreamSynchronize(c->stream));
    if (info) {
        info->cgbnrm = bn;
        info->cgxnrm = xn;
        info->icgitr = it;
    }
    if (it > c->prm.icgmax) {
        cad_set_error("VF_VPSOL: NOT CONVERGED (%d iterations)", it);
        return CAD_ERR_NOT_CONVERGED;
    }
    if (it == -2) {
        cad_set_error("VF_VPSOL: INSTABILITY");
        return CAD_ERR_INSTABILITY;
    }
    if (it == -1) {
        cad_set_error("VF_VPSOL: CAN NOT DECOMPOSITION");
        return CAD_ERR_DECOMPOSITION;
    }
    return CAD_OK;
}

int cad_correction(cad_ctx *c)
{
    RET_IF(ready(c));
    RET_IF(cad_launch_vmodif(c));
    RET_IF(cad_launch_bsuwt(c));
    RET_IF(cad_launch_bsuwt2(c));
    RET_IF(cad_launch_bsuwn(c));
    RET_IF(cad_launch_bwuwt(c));
    RET_IF(cad_launch_bwpp(c));
    CAD_CHECK(hipStreamSynchronize(c->stream));
    return CAD_OK;
}

int cad_v1cal(cad_ctx *c, cad_step_info *info)
{
    RET_IF(ready(c));
    cad_step_info loc;
    std::memset(&loc, 0, sizeof loc);
    hipEventRecord(c->ev[0], c->stream);
    RET_IF(cad_launch_predictor(c));
    RET_IF(cad_launch_bwuwn(c));
    hipEventRecord(c->ev[1], c->stream);
    RET_IF(cad_launch_vpcoef(c));
    int rc = cad_vpsol(c, &loc);
    hipEventRecord(c->ev[2], c->stream);
    if (rc) {
        if (info) *info = loc;
        return rc;
    }
    RET_IF(cad_launch_vmodif(c));
    RET_IF(cad_launch_bsuwt(c));
    RET_IF(cad_launch_bsuwt2(c));
    RET_IF(cad_launch_bsuwn(c));
    RET_IF(cad_launch_bwuwt(c));
    RET_IF(cad_launch_bwpp(c));
    hipEventRecord(c->ev[3], c->stream);
    CAD_CHECK(hipStreamSynchronize(c->stream));
    RET_IF(cad_wave_check(c));
    hipEventElapsedTime(&loc.ms_predictor, c->ev[0], c->ev[1]);
    hipEventElapsedTime(&loc.ms_poisson, c->ev[1], c->ev[2]);
    hipEventElapsedTime(&loc.ms_correction, c->ev[2], c->ev[3]);
    if (info) {
        info->cgbnrm = loc.cgbnrm;
        info->cgxnrm = loc.cgxnrm;
        info->icgitr = loc.icgitr;
        info->ms_predictor = loc.ms_predictor;
        info->ms_poisson = loc.ms_poisson;
        info->ms_correction = loc.ms_correction;
    }
    return CAD_OK;
}

int cad_f1cal(cad_ctx *c, cad_step_info *info)
{
    RET_IF(ready(c));
    double fsum = 0.0, fcut = 0.0;
    hipEventRecord(c->ev[4], c->stream);
    RET_IF(cad_vof_phase(c, &fsum, &fcut));
    hipEventRecord(c->ev[5], c->stream);
    CAD_CHECK(hipStreamSynchronize(c->stream));
    if (info) {
        info->fsum = fsum;
        info->fcut = fcut;
        info->nums = c->nums;
        hipEventElapsedTime(&info->ms_vof, c->ev[4], c->ev[5]);
    }
    return CAD_OK;
}

int cad_step(cad_ctx *c, cad_step_info *info)
{
    cad_step_info loc;
    std::memset(&loc, 0, sizeof loc);
    RET_IF(ready(c));
    // DT = DTNOW: LOOPS passes of TNOW += DT/LOOPS and VF_V1CAL at
    // DTNOW = DT/LOOPS (passes 2.. first refresh the boundary values,
    // vf_v1cal.f:115-118), then DTNOW = DT and VF_F1CAL
    // (vf_a1main.f:1029-1104)
    const double dt = c->prm.dtnow;
    c->prm.dtnow = dt / (double)c->prm.loops;
    int rc = CAD_OK;
    for (int l = 1; l <= c->prm.loops && rc == CAD_OK; l++) {
        c->prm.tnow = c->prm.tnow + c->prm.dtnow;
        if (l >= 2) {
            rc = cad_launch_bsuwn3(c);
            if (rc == CAD_OK) rc = cad_launch_bwuwt(c);
        }
        if (rc == CAD_OK) rc = cad_v1cal(c, &loc);
    }
    c->prm.dtnow = dt;
    if (rc == CAD_OK) rc = cad_f1cal(c, &loc);
    if (info) *info = loc;
    return rc;
}

int cad_cdtcal(cad_ctx *c, int nnow, double dtinit, double dtsafe, double dtmin, double dtmax,
               double *dt)
{
    RET_IF(ready(c));
    if (!dt || !(dtsafe > 0.0) || !(dtinit >= 0.0) || !(dtmin >= 0.0) || !(dtmax >= dtmin)) {
        cad_set_error("cad_cdtcal: invalid TIME AUTO / LIMIT values (vf_iitime.f:37-52)");
        return CAD_ERR_ARG;
    }
    return cad_cdtcal_impl(c, nnow, dtinit, dtsafe, dtmin, dtmax, dt);
}

int cad_init_state(cad_ctx *c, double wvlvl, double uini, double vini, double wini, int flags,
                   cad_step_info *info)
{
    RET_IF(ready(c));
    const double uvw[3] = {uini, vini, wini};
    double fsum = 0.0, fcut = 0.0;
    RET_IF(cad_init_phase(c, wvlvl, uvw, flags, &fsum, &fcut));
    RET_IF(cad_wave_check(c));
    // TimerDoor state (vf_cinit.f:412-421)
    if (c->td_alloc) RET_IF(cad_tdoor_init(c));
    if (info) {
        std::memset(info, 0, sizeof *info);
        info->fsum = fsum;
        info->fcut = fcut;
        info->nums = c->nums;
    }
    return CAD_OK;
}

int cad_set_wave_table(cad_ctx *c, int nz, int nt, const double *zz, const double *tt,
                       const double *hh, const double *un, const double *ut, double t0)
{
    if (!c) return CAD_ERR_ARG;
    return cad_wave_table(c, 0, nz, nt, zz, tt, hh, un, ut, t0);
}

int cad_set_wave_table2(cad_ctx *c, int nz, int nt, const double *zz, const double *tt,
                        const double *hh, const double *un, const double *ut, double t0)
{
    if (!c) return CAD_ERR_ARG;
    return cad_wave_table(c, 1, nz, nt, zz, tt, hh, un, ut, t0);
}

int cad_set_wave(cad_ctx *c, int side, int l1, int l2, double depth, double period, double ampl,
                 double wvlvl)
{
    if (!c) return CAD_ERR_ARG;
    return cad_wave_side(c, 0, side, l1, l2, depth, period, ampl, wvlvl);
}

int cad_set_wave2(cad_ctx *c, int side, int l1, int l2, double depth, double period, double ampl,
                  double wvlvl)
{
    if (!c) return CAD_ERR_ARG;
    return cad_wave_side(c, 1, side, l1, l2, depth, period, ampl, wvlvl);
}

int cad_set_wave_theory(cad_ctx *c, int side, int func, int l1, int l2, double depth,
                        double height, double period, double ampl, double wvlvl, double angle,
                        double ax, double ay)
{
    if (!c) return CAD_ERR_ARG;
    return cad_wave_theory_side(c, side, func, l1, l2, depth, height, period, ampl, wvlvl, angle,
                                ax, ay);
}

int cad_set_open(cad_ctx *c, int side, double depth, double period)
{
    if (!c) return CAD_ERR_ARG;
    return cad_open_side(c, side, depth, period);
}

int cad_set_drag_df(cad_ctx *c, double yu, int n, const double *dr, const double *ap,
                    const double *bt)
{
    if (!c) return CAD_ERR_ARG;
    if (n == 0) {
        c->drag.n = 0;
        return CAD_OK;
    }
    // the checks of VF_IIOPT (vf_iiopt.f:89-106)
    if (n < 0 || n > CAD_DRAG_ROWS || !dr || !ap || !bt) {
        cad_set_error("cad_set_drag_df: 0..%d rows and non-null tables", CAD_DRAG_ROWS);
        return CAD_ERR_ARG;
    }
    if (!(yu > 0.0)) {
        cad_set_error("cad_set_drag_df: YU must be positive");
        return CAD_ERR_ARG;
    }
    for (int l = 0; l < n; l++) {
        if (!(dr[l] > 0.0) || !(ap[l] >= 0.0) || !(bt[l] >= 0.0)) {
            cad_set_error("cad_set_drag_df: row %d needs DR > 0, AP >= 0, BT >= 0", l + 1);
            return CAD_ERR_ARG;
        }
    }
    DragDF t = {};
    t.n = n;
    t.yu = yu;
    for (int l = 0; l < n; l++) {
        t.dr[l] = dr[l];
        t.ap[l] = ap[l];
        t.bt[l] = bt[l];
    }
    c->drag = t;
    return CAD_OK;
}

int cad_set_damp(cad_ctx *c, int side, int degree, double param_xy, double param_z, double width,
                 double depth)
{
    if (!c) return CAD_ERR_ARG;
    if (side < 1 || side > 4) {
        cad_set_error("cad_set_damp: side must be 1..4 (X-, X+, Y-, Y+), got %d", side);
        return CAD_ERR_ARG;
    }
    const int q = side - 1;
    if (degree < 0) {
        c->dm_on[q] = 0;
        return CAD_OK;
    }
    // the checks of VF_IIMDL (vf_iimdl.f:238-257) and VF_CSETUP (vf_csetup.f:221-234)
    if (!(param_xy >= 1.0e-20) || !(param_z >= 1.0e-20) || !(width >= 1.0e-20) ||
        !(depth >= 1.0e-20)) {
        cad_set_error("cad_set_damp: PARAM-XY, PARAM-Z, WIDTH and DEPTH must be positive");
        return CAD_ERR_ARG;
    }
    c->dm_on[q] = 1;
    c->dm_deg[q] = degree;
    c->dm_par[q][0] = param_xy;
    c->dm_par[q][1] = param_z;
    c->dm_par[q][2] = width;
    c->dm_par[q][3] = depth;
    return CAD_OK;
}

int cad_get_wave_level(cad_ctx *c, int side, double *level)
{
    if (!c || !level || side < 1 || side > 4) {
        cad_set_error("cad_get_wave_level: side 1..4 and a non-null result");
        return CAD_ERR_ARG;
    }
    if (!c->wv_on[side - 1]) {
        cad_set_error("cad_get_wave_level: side %d has no wave-maker", side);
        return CAD_ERR_STATE;
    }
    *level = c->wv_level[side - 1];
    return CAD_OK;
}

int cad_time_precond(cad_ctx *c, int reps, float *ms_avg, double *bytes)
{
    RET_IF(ready(c));
    if (!ms_avg || !bytes) return CAD_ERR_ARG;
    if (c->prm.precision != CAD_PRECISION_MIXED) {
        RET_IF(cad_time_minv_impl(c, reps, ms_avg));
        long long active = 0;
        RET_IF(cad_count_active(c, &active));
        *bytes = 96.0 * (double)active;
        return CAD_OK;
    }
    return cad_mg_time(c, reps, 0, ms_avg, bytes);
}

int cad_time_vcycle(cad_ctx *c, int reps, int fp64, float *ms_avg, double *bytes)
{
    RET_IF(ready(c));
    if (!ms_avg || !bytes) return CAD_ERR_ARG;
    return cad_mg_time(c, reps, fp64, ms_avg, bytes);
}

int cad_set_krylov_cap(cad_ctx *c, int cap)
{
    if (!c || cap < 0) {
        cad_set_error("cad_set_krylov_cap: cap %d (0: the default)", cap);
        return CAD_ERR_ARG;
    }
    c->krylov_cap = cap;
    return CAD_OK;
}

int cad_mixed_stats(cad_ctx *c, long long *out)
{
    if (!c || !out) return CAD_ERR_ARG;
    out[0] = c->mixed_fallbacks;
    out[1] = c->mixed_fallback_iters;
    return CAD_OK;
}

int cad_poisson_residual(cad_ctx *c, double *rel)
{
    RET_IF(ready(c));
    if (!rel) return CAD_ERR_ARG;
    return cad_poisson_residual_impl(c, rel);
}

int cad_time_minv(cad_ctx *c, int reps, float *ms_avg)
{
    RET_IF(ready(c));
    if (!ms_avg) return CAD_ERR_ARG;
    return cad_time_minv_impl(c, reps, ms_avg);
}

// ---- in-solve kernel timing ---------------------------------------------

// takes the oldest recorded pair of kernel w; its end event was recorded at
// least one solve iteration ago, so the wait returns at once in practice
static int kt_take(cad_ktimer *t, int w)
{
    const int s = (int)(t->done[w] % cad_ktimer::PAIRS);
    CAD_CHECK(hipEventSynchronize(t->ev[w][2 * s + 1]));
    float ms = 0.f;
    CAD_CHECK(hipEventElapsedTime(&ms, t->ev[w][2 * s], t->ev[w][2 * s + 1]));
    t->ms[w] += ms;
    t->total_work[w] += t->work[w][s];
    t->done[w]++;
    return CAD_OK;
}

int cad_kt_begin(cad_ctx *c, int w)
{
    cad_ktimer *t = c->kt;
    if (!t || !t->on) return CAD_OK;
    t->open[w] = t->seen[w]++ % t->every == 0;
    if (!t->open[w]) return CAD_OK;
    if (t->next[w] - t->done[w] >= cad_ktimer::PAIRS) RET_IF(kt_take(t, w));
    const int s = (int)(t->next[w] % cad_ktimer::PAIRS);
    CAD_CHECK(hipEventRecord(t->ev[w][2 * s], c->stream));
    return CAD_OK;
}

int cad_kt_end(cad_ctx *c, int w, double work)
{
    cad_ktimer *t = c->kt;
    if (!t || !t->on || !t->open[w]) return CAD_OK;
    t->open[w] = false;
    const int s = (int)(t->next[w] % cad_ktimer::PAIRS);
    CAD_CHECK(hipEventRecord(t->ev[w][2 * s + 1], c->stream));
    t->work[w][s] = work;
    t->next[w]++;
    return CAD_OK;
}

void cad_kt_free(cad_ctx *c)
{
    if (!c->kt) return;
    for (auto &row : c->kt->ev)
        for (auto &e : row)
            if (e) hipEventDestroy(e);
    delete c->kt;
    c->kt = nullptr;
}

int cad_set_kernel_timing(cad_ctx *c, int on)
{
    RET_IF(ready(c));
    if (on < 0) return CAD_ERR_ARG;
    if (on && !c->kt) {
        c->kt = new (std::nothrow) cad_ktimer();
        if (!c->kt) {
            cad_set_error("cad_set_kernel_timing: out of host memory");
            return CAD_ERR_STATE;
        }
        for (auto &row : c->kt->ev)
            for (auto &e : row) {
                const hipError_t err = hipEventCreateWithFlags(&e, hipEventDisableSystemFence);
                if (err != hipSuccess) {
                    cad_kt_free(c);   // no timer with missing events is kept
                    cad_set_error("cad_set_kernel_timing: hipEventCreateWithFlags: %s", hipGetErrorString(err));
                    return CAD_ERR_HIP;
                }
            }
    }
    if (!c->kt) return CAD_OK;
    if (on) {
        CAD_CHECK(hipStreamSynchronize(c->stream));   // no pair of an earlier period is in flight
        for (int w = 0; w < CAD_KT_N; w++) {
            c->kt->next[w] = c->kt->done[w] = c->kt->seen[w] = 0;
            c->kt->open[w] = false;
            c->kt->ms[w] = c->kt->total_work[w] = 0.0;
        }
        c->kt->every = on;
    }
    c->kt->on = on != 0;
    return CAD_OK;
}

int cad_kernel_timing(cad_ctx *c, int which, double *ms_avg, double *work_avg, long long *count)
{
    RET_IF(ready(c));
    if (which < 0 || which >= CAD_KT_N || !ms_avg || !work_avg || !count) return CAD_ERR_ARG;
    *ms_avg = *work_avg = 0.0;
    *count = 0;
    cad_ktimer *t = c->kt;
    if (!t) return CAD_OK;
    while (t->done[which] < t->next[which]) RET_IF(kt_take(t, which));
    const long long n = t->done[which];
    *count = n;
    if (n > 0) {
        *ms_avg = t->ms[which] / (double)n;
        *work_avg = t->total_work[which] / (double)n;
    }
    return CAD_OK;
}

int cad_count_active(cad_ctx *c, long long *active)
{
    RET_IF(ready(c));
    if (!active) return CAD_ERR_ARG;
    std::vector<int> h((size_t)c->d.nc);
    CAD_CHECK(hipMemcpy(h.data(), c->f.indc, h.size() * sizeof(int), hipMemcpyDeviceToHost));
    long long n = 0;
    for (int v : h) n += (v != -1);
    *active = n;
    return CAD_OK;
}

int cad_set_flag_pass_limit(cad_ctx *c, long long passes)
{
    if (!c || passes < 0) {
        cad_set_error("cad_set_flag_pass_limit: a context and passes >= 0");
        return CAD_ERR_ARG;
    }
    c->flag_pass_limit = passes;
    return CAD_OK;
}

int cad_set_true_residual_stop(cad_ctx *c, double rel)
{
    if (!c || !(rel >= 0.0)) {
        cad_set_error("cad_set_true_residual_stop: a context and rel >= 0");
        return CAD_ERR_ARG;
    }
    c->true_stop = rel;
    return CAD_OK;
}

int cad_set_sweep_schedule(cad_ctx *c, int schedule)
{
    if (!c || schedule < CAD_SWEEP_PIPELINED || schedule > CAD_SWEEP_PIPELINED_16X4) {
        cad_set_error("cad_set_sweep_schedule: unknown schedule %d", schedule);
        return CAD_ERR_ARG;
    }
    c->sweep_schedule = schedule;
    return CAD_OK;
}

int cad_milu_factor(cad_ctx *c, int *npos)
{
    if (!c || !npos) return CAD_ERR_ARG;
    CAD_CHECK(hipSetDevice(c->device));
    return cad_milu_factor_impl(c, npos);
}

int cad_milu_apply(cad_ctx *c, const double *y, double *x)
{
    if (!c || !y || !x) return CAD_ERR_ARG;
    CAD_CHECK(hipSetDevice(c->device));
    return cad_milu_apply_impl(c, y, x);
}

int cad_cdiv00(cad_ctx *c, double *dv)
{
    RET_IF(ready(c));
    if (!dv) return CAD_ERR_ARG;
    return cad_cdiv(c, dv);
}

int cad_download_matrix(cad_ctx *c, double *ad, double *ali, double *alj, double *alk, double *aui,
                        double *auj, double *auk, double *bb)
{
    if (!c) return CAD_ERR_ARG;
    const size_t nc = (size_t)c->d.nc;
    RET_IF(d2h(c, ad, c->ad, nc));
    RET_IF(d2h(c, ali, c->ali, nc));
    RET_IF(d2h(c, alj, c->alj, nc));
    RET_IF(d2h(c, alk, c->alk, nc));
    RET_IF(d2h(c, aui, c->aui, nc));
    RET_IF(d2h(c, auj, c->auj, nc));
    RET_IF(d2h(c, auk, c->auk, nc));
    RET_IF(d2h(c, bb, c->bb, nc));
    CAD_CHECK(hipStreamSynchronize(c->stream));
    return CAD_OK;
}

int cad_upload_matrix(cad_ctx *c, const double *ad, const double *ali, const double *alj,
                      const double *alk, const double *aui, const double *auj, const double *auk,
                      const double *bb, const int *indc)
{
    if (!c) return CAD_ERR_ARG;
    CAD_CHECK(hipSetDevice(c->device));
    const size_t nc = (size_t)c->d.nc;
    RET_IF(h2d(c, c->ad, ad, nc));
    RET_IF(h2d(c, c->ali, ali, nc));
    RET_IF(h2d(c, c->alj, alj, nc));
    RET_IF(h2d(c, c->alk, alk, nc));
    RET_IF(h2d(c, c->aui, aui, nc));
    RET_IF(h2d(c, c->auj, auj, nc));
    RET_IF(h2d(c, c->auk, auk, nc));
    RET_IF(h2d(c, c->bb, bb, nc));
    RET_IF(h2d(c, c->f.indc, indc, nc));
    CAD_CHECK(hipStreamSynchronize(c->stream));
    return CAD_OK;
}

int cad_download_pt(cad_ctx *c, double *pt)
{
    if (!c) return CAD_ERR_ARG;
    RET_IF(d2h(c, pt, c->pt, (size_t)c->d.nc));
    CAD_CHECK(hipStreamSynchronize(c->stream));
    return CAD_OK;
}

}  // extern "C"
