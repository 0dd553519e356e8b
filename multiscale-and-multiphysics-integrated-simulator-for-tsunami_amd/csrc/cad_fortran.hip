// cad_fortran.hip -- link-level drop-in of the reference's pressure solve.
//
// The reference's replaceable unit is the Fortran subroutine (Source
// code/Makefile:14-83).  This file exports the two routines of the solve
// under their own link names, with their by-reference argument lists:
//
//   vf_m1bcgs_  VF_M1BCGS (vf_m1bcgs.f:1-5), called by VF_VPSOL at
//               vf_vpsol.f:64-69;
//   vf_vpsol_   VF_VPSOL (vf_vpsol.f:1-3), called by VF_V1CAL at
//               vf_v1cal.f:278-281; it reads its sizes and solver
//               parameters from COMMON /VF_ANUMBI/, /VF_ACOMPR/ and
//               /VF_ACOMPI/ and writes CGBNRM, CGXNRM and ICGITR back there,
//               as the reference's routine does.
//
// Linking the reference with this library in place of vf_m1bcgs.o (or of
// vf_vpsol.o and vf_m1bcgs.o) runs its pressure solves on the GPU with no
// change to any other reference file.  The COMMON blocks are defined here as
// weak symbols with the reference's member order, so that a Fortran program
// that declares them (the reference) supplies the one instance both sides
// use, while a caller without them (a C or Python program) gets this
// library's own zero-initialised blocks.
//
// Single MPI rank: the reference's loops run over MYIS..MYIE, MYJS..MYJE
// (vf_m1bcgs.f:118-120), which at one rank are 2..NI-1, 2..NJ-1; the
// parallel buffer DBUF is not used.  The work arrays DD..AE are scratch in the
// reference (no caller reads them after the solve, vf_v1cal.f:278-297) and are
// left untouched; the device keeps its own.
#include "cad_common.h"

#include <cstdio>
#include <cstring>
#include <vector>

extern "C" {

// COMMON blocks, member order of VF_ACOMPR.h, VF_ACOMPI.h and VF_ANUMBI.h
// (MAXNC = 10, VF_A0PRM.h).
__attribute__((weak)) vf_acompr_block vf_acompr_ = {};
__attribute__((weak)) vf_acompi_block vf_acompi_ = {};
__attribute__((weak)) vf_anumbi_block vf_anumbi_ = {};

// The reference's error exit VF_A2ERR(NAME, MSG) (vf_a2err.f:28-49): closes
// its files and stops the run.  Present when this library is linked into the
// reference; absent (null) for other callers.  Fortran passes the lengths of
// the two CHARACTER arguments after the argument list.
__attribute__((weak)) void vf_a2err_(const char *name, const char *msg, size_t lname, size_t lmsg);

}  // extern "C"

namespace {

cad_ctx *g_solver = nullptr;   // reused while the grid size is unchanged

template <typename T>
void gather(std::vector<T> &out, const T *a, int mi, int mj, int ni, int nj, int nk)
{
    out.resize((size_t)ni * nj * nk);
    for (int k = 0; k < nk; k++)
        for (int j = 0; j < nj; j++)
            std::memcpy(&out[((size_t)k * nj + j) * ni], a + ((size_t)k * mj + j) * mi,
                        (size_t)ni * sizeof(T));
}

// One VF_M1BCGS on arrays of shape (mi,mj,mk) whose loop range is
// (ni,nj,nk): X0 in and out, BNORM, XNORM and ITER as vf_m1bcgs.f returns
// them.  CAD_OK, or a negative code when the library itself fails (the
// message in cad_last_error).
int solve_host(const double *const mat[8], double *x0, const int *indc, int mi, int mj, int mk,
               int ni, int nj, int nk, double cgpara, double cgepsa, double cgepsr, double cgdiv,
               int icgtyp, int icgmax, double *bnorm, double *xnorm, int *iter)
{
    if (ni < 4 || nj < 4 || nk < 4 || mi < ni || mj < nj || mk < nk) {
        cad_set_error("VF_M1BCGS: sizes NI,NJ,NK = %d,%d,%d within MI,MJ,MK = %d,%d,%d are not "
                      "a grid with a ghost layer", ni, nj, nk, mi, mj, mk);
        return CAD_ERR_ARG;
    }
    for (int q = 0; q < 8; q++)
        if (!mat[q]) {
            cad_set_error("VF_M1BCGS: null array argument");
            return CAD_ERR_ARG;
        }
    if (!x0 || !indc) {
        cad_set_error("VF_M1BCGS: null array argument");
        return CAD_ERR_ARG;
    }
    if (g_solver && (g_solver->d.ni != ni || g_solver->d.nj != nj || g_solver->d.nk != nk)) {
        cad_destroy(g_solver);
        g_solver = nullptr;
    }
    if (!g_solver) {
        const int rc = cad_create(ni, nj, nk, 0, 0, &g_solver);
        if (rc != CAD_OK) return rc;
    }
    cad_ctx *c = g_solver;
    CAD_CHECK(hipSetDevice(c->device));
    c->prm.cgpara = cgpara;
    c->prm.cgepsa = cgepsa;
    c->prm.cgepsr = cgepsr;
    c->prm.cgdiv = cgdiv;
    c->prm.icgtyp = icgtyp;
    c->prm.icgmax = icgmax;
    const size_t nc = (size_t)c->d.nc;
    const bool packed = mi == ni && mj == nj && mk == nk;
    std::vector<double> m[8], xb;
    std::vector<int> ib;
    const double *mp[8];
    const int *ip = indc;
    double *xp = x0;
    if (packed) {
        for (int q = 0; q < 8; q++) mp[q] = mat[q];
    } else {
        for (int q = 0; q < 8; q++) {
            gather(m[q], mat[q], mi, mj, ni, nj, nk);
            mp[q] = m[q].data();
        }
        gather(ib, indc, mi, mj, ni, nj, nk);
        gather(xb, x0, mi, mj, ni, nj, nk);
        ip = ib.data();
        xp = xb.data();
    }
    int rc = cad_upload_matrix(c, mp[0], mp[1], mp[2], mp[3], mp[4], mp[5], mp[6], mp[7], ip);
    if (rc != CAD_OK) return rc;
    CAD_CHECK(hipMemcpyAsync(c->pt, xp, nc * sizeof(double), hipMemcpyHostToDevice, c->stream));
    rc = cad_solve_m1bcgs(c, c->pt, bnorm, xnorm, iter);
    if (rc != CAD_OK) return rc;
    CAD_CHECK(hipMemcpyAsync(xp, c->pt, nc * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CAD_CHECK(hipStreamSynchronize(c->stream));
    if (!packed)
        for (int k = 0; k < nk; k++)
            for (int j = 0; j < nj; j++)
                std::memcpy(x0 + ((size_t)k * mj + j) * mi, &xb[((size_t)k * nj + j) * ni],
                            (size_t)ni * sizeof(double));
    return CAD_OK;
}

void report(const char *name, const char *msg)
{
    if (vf_a2err_) vf_a2err_(name, msg, std::strlen(name), std::strlen(msg));
}

}  // namespace

extern "C" {

void vf_m1bcgs_(double *dbuf, const double *ad, const double *ali, const double *alj,
                const double *alk, const double *aui, const double *auj, const double *auk,
                const double *bb, double *x0, double *dd, double *xw, double *yy, double *r0,
                double *rr, double *pp, double *ap, double *ae, const double *cgpara,
                const double *cgepsa, const double *cgepsr, const double *cgdiv, double *bnorm,
                double *xnorm, const int *indc, const int *mi, const int *mj, const int *mk,
                const int *ni, const int *nj, const int *nk, const int *icgtyp,
                const int *icgmax, int *iter)
{
    (void)dbuf; (void)dd; (void)xw; (void)yy; (void)r0; (void)rr; (void)pp; (void)ap; (void)ae;
    const double *mat[8] = {ad, ali, alj, alk, aui, auj, auk, bb};
    *bnorm = 0.0;
    *xnorm = 0.0;
    const int rc = solve_host(mat, x0, indc, *mi, *mj, *mk, *ni, *nj, *nk, *cgpara, *cgepsa,
                              *cgepsr, *cgdiv, *icgtyp, *icgmax, bnorm, xnorm, iter);
    if (rc != CAD_OK) {
        // a failure of the library itself: ITER past ICGMAX, which VF_VPSOL
        // stops on (vf_vpsol.f:70-71); the cause stays in cad_last_error()
        std::fprintf(stderr, "vf_m1bcgs (gfx950 library): %s\n", cad_last_error());
        *iter = *icgmax + 1;
    }
}

void vf_vpsol_(double *dbuf, const double *ad, const double *ali, const double *alj,
               const double *alk, const double *aui, const double *auj, const double *auk,
               const double *bb, double *pt, double *dd, double *xw, double *yy, double *r0,
               double *rr, double *pp, double *ap, double *ae, const int *indc)
{
    (void)dbuf; (void)dd; (void)xw; (void)yy; (void)r0; (void)rr; (void)pp; (void)ap; (void)ae;
    const int ni = vf_anumbi_.numi, nj = vf_anumbi_.numj, nk = vf_anumbi_.numk;
    // PT = 0 (vf_vpsol.f:55-61)
    if (ni > 0 && nj > 0 && nk > 0 && pt)
        std::memset(pt, 0, (size_t)ni * nj * nk * sizeof(double));
    const double *mat[8] = {ad, ali, alj, alk, aui, auj, auk, bb};
    double bn = 0.0, xn = 0.0;
    int it = -1;
    const int icgmax = vf_acompi_.icgmax;
    const int rc = solve_host(mat, pt, indc, ni, nj, nk, ni, nj, nk, vf_acompr_.cgpara,
                              vf_acompr_.cgepsa, vf_acompr_.cgepsr, vf_acompr_.cgdiv,
                              vf_acompi_.icgtyp, icgmax, &bn, &xn, &it);
    if (rc != CAD_OK) {
        std::fprintf(stderr, "vf_vpsol (gfx950 library): %s\n", cad_last_error());
        it = icgmax + 1;
    }
    vf_acompr_.cgbnrm = bn;
    vf_acompr_.cgxnrm = xn;
    vf_acompi_.icgitr = it;
    // the checks of vf_vpsol.f:70-75
    if (it > icgmax) {
        if (rc == CAD_OK) cad_set_error("VF_VPSOL: NOT CONVERGED. (%d iterations)", it);
        report("VF_VPSOL", "NOT CONVERGED.");
    } else if (it == -2) {
        cad_set_error("VF_VPSOL: INSTABILITY.");
        report("VF_VPSOL", "INSTABILITY.");
    } else if (it == -1) {
        cad_set_error("VF_VPSOL: CAN NOT DECOMPOSITION.");
        report("VF_VPSOL", "CAN NOT DECOMPOSITION.");
    }
}

}  // extern "C"
