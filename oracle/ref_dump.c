/*
 * ref_dump.c -- link-time state recorder for the reference CADMAS-SURF/3D build.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is linked into oracle/_ref/cadmas_ref.x
 * (the reference solver compiled from its own Fortran sources by
 * oracle/build_ref.sh).  It is never linked into the product library.
 *
 * build_ref.sh renames the reference's own definitions of VF_V1CAL, VF_VPSOL,
 * VF_F1CAL and VF_CDIV00 to *_orig_ with objcopy; the functions below take
 * their place, write the arrays each phase reads and writes to binary dump
 * files, and call the original.  Argument lists follow the reference:
 *   VF_V1CAL   Source code/vf_v1cal.f:1-9     (called vf_a1main.f:1058)
 *   VF_VPSOL   Source code/vf_vpsol.f:1-3     (called vf_v1cal.f:278)
 *   VF_F1CAL   Source code/vf_f1cal.f:1-12    (called vf_a1main.f:1104);
 *              its records also hold the TimerDoor state TBUB, DROPTX..DROPWW
 *   VF_CDIV00  Source code/vf_cdiv00.f:1      (called vf_a1main.f:954)
 * COMMON block layouts follow Source code/VF_ANUMBI.h, VF_ACOMPR.h,
 * VF_ACOMPI.h, VF_ATIMER.h, VF_ATIMEI.h (leading members only).
 *
 * Dumping is enabled by the environment variable CADMAS_DUMP_DIR; with
 * CADMAS_DUMP_STEPS="a,b,c" only those step numbers are written (the
 * per-step divergence norm of VF_CDIV00 is always written).
 *
 * Record format (little endian): "CDMP", char name[24], char type ('d' or
 * 'i'), int32 ndim, int32 dims[ndim] (Fortran order, fastest first), data.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>

typedef double R8;
typedef int32_t I4;

extern struct { I4 numi, numj, numk, numb, nums, leqk, leqt, leqc, numij, numbx; } vf_anumbi_;
extern struct { R8 cgpara, cgepsa, cgepsr, cgdiv, cgbnrm, cgxnrm, feps, flower, fupper, fsum, fcut, plower, scmvp; } vf_acompr_;
extern struct { I4 icgtyp, icgmax, icgitr, iscmvp, iscmff; } vf_acompi_;
extern struct { R8 tend, tnow, dtnow, dtcnst; } vf_atimer_;
extern struct { I4 nend, nnow, idttyp, loops; } vf_atimei_;

/* ------------------------------------------------------------------ */
static FILE *g_fp = NULL;
static int g_iloop = 1;   /* sub-loop pass of the current VF_V1CAL call */
static int g_static_done = 0;
static R8 *g_u = NULL, *g_v = NULL, *g_w = NULL, *g_p = NULL, *g_f = NULL;
static I4 *g_nf = NULL;

static int step_selected(int step)
{
    const char *s = getenv("CADMAS_DUMP_STEPS");
    if (!s || !*s) return 1;
    const char *p = s;
    while (*p) {
        char *end;
        long v = strtol(p, &end, 10);
        if (end == p) break;
        if (v == step) return 1;
        p = end;
        while (*p == ',' || *p == ' ') p++;
    }
    return 0;
}

/* CADMAS_DUMP_TAGS="f1_out,static": only these records (the per-step "dv"
 * record is always written); unset: every record. */
static int tag_selected(const char *tag)
{
    const char *s = getenv("CADMAS_DUMP_TAGS");
    if (!s || !*s || strcmp(tag, "dv") == 0) return 1;
    size_t n = strlen(tag);
    for (const char *p = s; *p;) {
        const char *e = strchr(p, ',');
        size_t m = e ? (size_t)(e - p) : strlen(p);
        if (m == n && strncmp(p, tag, n) == 0) return 1;
        if (!e) break;
        p = e + 1;
    }
    return 0;
}

static int open_dump(const char *tag)
{
    const char *dir = getenv("CADMAS_DUMP_DIR");
    if (!dir || !*dir) return 0;
    if (!step_selected(vf_atimei_.nnow) && strcmp(tag, "static") != 0 && strcmp(tag, "dv") != 0)
        return 0;
    if (!tag_selected(tag)) return 0;
    char path[4096];
    if (strcmp(tag, "static") == 0)
        snprintf(path, sizeof path, "%s/static.bin", dir);
    else if (g_iloop > 1 && strncmp(tag, "f1", 2) != 0 && strcmp(tag, "dv") != 0)
        /* the second and later VF_V1CAL passes of OPTION SUB-LOOP */
        snprintf(path, sizeof path, "%s/s%05d_%s_l%d.bin", dir, vf_atimei_.nnow, tag, g_iloop);
    else
        snprintf(path, sizeof path, "%s/s%05d_%s.bin", dir, vf_atimei_.nnow, tag);
    g_fp = fopen(path, "wb");
    if (!g_fp) { fprintf(stderr, "ref_dump: cannot open %s\n", path); exit(3); }
    return 1;
}

static void close_dump(void) { if (g_fp) fclose(g_fp); g_fp = NULL; }

static void put(const char *name, char type, int ndim, const int *dims, const void *data)
{
    char nm[24];
    memset(nm, 0, sizeof nm);
    strncpy(nm, name, sizeof nm - 1);
    fwrite("CDMP", 1, 4, g_fp);
    fwrite(nm, 1, 24, g_fp);
    fwrite(&type, 1, 1, g_fp);
    I4 nd = ndim;
    fwrite(&nd, 4, 1, g_fp);
    size_t n = 1;
    for (int i = 0; i < ndim; i++) { I4 d = dims[i]; fwrite(&d, 4, 1, g_fp); n *= (size_t)dims[i]; }
    fwrite(data, type == 'd' ? 8 : 4, n, g_fp);
}

static void f3(const char *name, const R8 *a)
{
    int d[3] = {vf_anumbi_.numi, vf_anumbi_.numj, vf_anumbi_.numk};
    put(name, 'd', 3, d, a);
}
static void i3(const char *name, const I4 *a)
{
    int d[3] = {vf_anumbi_.numi, vf_anumbi_.numj, vf_anumbi_.numk};
    put(name, 'i', 3, d, a);
}
static void fb(const char *name, const R8 *a)
{
    int d[1] = {vf_anumbi_.numb};
    put(name, 'd', 1, d, a);
}
static void sd(const char *name, R8 v) { int d[1] = {1}; put(name, 'd', 1, d, &v); }
static void si(const char *name, I4 v) { int d[1] = {1}; put(name, 'i', 1, d, &v); }

static void scalars(void)
{
    sd("DTNOW", vf_atimer_.dtnow);
    sd("TNOW", vf_atimer_.tnow);
    si("NNOW", vf_atimei_.nnow);
    si("NUMS", vf_anumbi_.nums);
    sd("FSUM", vf_acompr_.fsum);
    sd("FCUT", vf_acompr_.fcut);
    sd("CGBNRM", vf_acompr_.cgbnrm);
    sd("CGXNRM", vf_acompr_.cgxnrm);
    si("ICGITR", vf_acompi_.icgitr);
}

static void inds(const I4 *a)
{
    int d[1] = {vf_anumbi_.nums > 0 ? vf_anumbi_.nums : 1};
    I4 zero = 0;
    put("INDS", 'i', 1, d, vf_anumbi_.nums > 0 ? (const void *)a : (const void *)&zero);
}

/* TimerDoor state (vf_f1cal.f:59-65): TBUB(NUMK) and the droplet fields */
static void tdoor(const R8 *tbub, const R8 *dtx, const R8 *dty, const R8 *dtz,
                  const R8 *duu, const R8 *dvv, const R8 *dww)
{
    int d[1] = {vf_anumbi_.numk};
    put("TBUB", 'd', 1, d, tbub);
    f3("DROPTX", dtx); f3("DROPTY", dty); f3("DROPTZ", dtz);
    f3("DROPUU", duu); f3("DROPVV", dvv); f3("DROPWW", dww);
}

/* ------------------------------------------------------------------ */
#define P R8 *
void vf_v1cal_orig_(I4 *iloop, P xx, P yy, P zz, P uu, P vv, P ww, P pp, P ff, P anu, P cm0, P cd0,
                    P ggv, P ggx, P ggy, P ggz, P glv, P glx, P gly, P glz,
                    P bcu, P bcv, P bcw, P bcp, P bcvi, P ggv0, P glv0, P ak, P tt, P cc,
                    P t1, P z1, P h1, P un1, P ut1, P t2, P z2, P h2, P un2, P ut2,
                    P dbuf, P srcuv, P wk01, P wk02, P wk03, P wk04, P wk05, P wk06, P wk07, P wk08,
                    P wk09, P wk10, P wk11, P wk12, P wk13, P wk14, P wk15, P wk16, P wk17,
                    I4 *nf, I4 *indx, I4 *indy, I4 *indz, I4 *indc, I4 *indb, I4 *inds_);

void vf_v1cal_(I4 *iloop, P xx, P yy, P zz, P uu, P vv, P ww, P pp, P ff, P anu, P cm0, P cd0,
               P ggv, P ggx, P ggy, P ggz, P glv, P glx, P gly, P glz,
               P bcu, P bcv, P bcw, P bcp, P bcvi, P ggv0, P glv0, P ak, P tt, P cc,
               P t1, P z1, P h1, P un1, P ut1, P t2, P z2, P h2, P un2, P ut2,
               P dbuf, P srcuv, P wk01, P wk02, P wk03, P wk04, P wk05, P wk06, P wk07, P wk08,
               P wk09, P wk10, P wk11, P wk12, P wk13, P wk14, P wk15, P wk16, P wk17,
               I4 *nf, I4 *indx, I4 *indy, I4 *indz, I4 *indc, I4 *indb, I4 *inds_)
{
    g_u = uu; g_v = vv; g_w = ww; g_p = pp; g_f = ff; g_nf = nf;
    g_iloop = *iloop;
    if (!g_static_done && open_dump("static")) {
        int dx[2] = {6, vf_anumbi_.numi}, dy[2] = {6, vf_anumbi_.numj}, dz[2] = {6, vf_anumbi_.numk};
        put("XX", 'd', 2, dx, xx); put("YY", 'd', 2, dy, yy); put("ZZ", 'd', 2, dz, zz);
        f3("ANU", anu); f3("CM0", cm0); f3("CD0", cd0);
        f3("GGV", ggv); f3("GGX", ggx); f3("GGY", ggy); f3("GGZ", ggz);
        f3("GLV", glv); f3("GLX", glx); f3("GLY", gly); f3("GLZ", glz);
        /* GGV0/GLV0 are only allocated for time-dependent porosity (IPRNT>1) */
        i3("INDX", indx); i3("INDY", indy); i3("INDZ", indz);
        int db[2] = {4, vf_anumbi_.numb};
        put("INDB", 'i', 2, db, indb);
        fb("BCVI", bcvi);
        sd("CGPARA", vf_acompr_.cgpara); sd("CGEPSA", vf_acompr_.cgepsa);
        sd("CGEPSR", vf_acompr_.cgepsr); sd("CGDIV", vf_acompr_.cgdiv);
        sd("FLOWER", vf_acompr_.flower); sd("FUPPER", vf_acompr_.fupper);
        sd("PLOWER", vf_acompr_.plower); sd("SCMVP", vf_acompr_.scmvp); sd("FEPS", vf_acompr_.feps);
        si("ICGTYP", vf_acompi_.icgtyp); si("ICGMAX", vf_acompi_.icgmax);
        close_dump();
        g_static_done = 1;
    }
    if (open_dump("v1_in")) {
        scalars();
        f3("UU", uu); f3("VV", vv); f3("WW", ww); f3("PP", pp); f3("FF", ff);
        fb("BCU", bcu); fb("BCV", bcv); fb("BCW", bcw); fb("BCP", bcp);
        i3("NF", nf); i3("INDC", indc); inds(inds_);
        close_dump();
    }
    vf_v1cal_orig_(iloop, xx, yy, zz, uu, vv, ww, pp, ff, anu, cm0, cd0, ggv, ggx, ggy, ggz,
                   glv, glx, gly, glz, bcu, bcv, bcw, bcp, bcvi, ggv0, glv0, ak, tt, cc,
                   t1, z1, h1, un1, ut1, t2, z2, h2, un2, ut2, dbuf, srcuv,
                   wk01, wk02, wk03, wk04, wk05, wk06, wk07, wk08, wk09, wk10, wk11, wk12,
                   wk13, wk14, wk15, wk16, wk17, nf, indx, indy, indz, indc, indb, inds_);
    if (open_dump("v1_out")) {
        scalars();
        f3("UU", uu); f3("VV", vv); f3("WW", ww); f3("PP", pp);
        fb("BCU", bcu); fb("BCV", bcv); fb("BCW", bcw); fb("BCP", bcp);
        f3("PT", wk09);
        close_dump();
    }
    g_iloop = 1;
}

void vf_vpsol_orig_(P dbuf, P ad, P ali, P alj, P alk, P aui, P auj, P auk, P bb, P pt,
                    P dd, P xw, P yy, P r0, P rr, P pp, P ap, P ae, I4 *indc);

void vf_vpsol_(P dbuf, P ad, P ali, P alj, P alk, P aui, P auj, P auk, P bb, P pt,
               P dd, P xw, P yy, P r0, P rr, P pp, P ap, P ae, I4 *indc)
{
    if (open_dump("ps_in")) {
        scalars();
        /* the predicted velocity u* is what VF_VPCOEF assembled BB from */
        f3("US", g_u); f3("VS", g_v); f3("WS", g_w);
        f3("AD", ad); f3("ALI", ali); f3("ALJ", alj); f3("ALK", alk);
        f3("AUI", aui); f3("AUJ", auj); f3("AUK", auk); f3("BB", bb);
        i3("INDC", indc);
        close_dump();
    }
    vf_vpsol_orig_(dbuf, ad, ali, alj, alk, aui, auj, auk, bb, pt, dd, xw, yy, r0, rr, pp, ap, ae, indc);
    if (open_dump("ps_out")) {
        scalars();
        f3("PT", pt); f3("DD", dd);
        close_dump();
    }
}

void vf_f1cal_orig_(P xx, P yy, P zz, P uu, P vv, P ww, P pp, P ff, P anu, P ggv, P ggx, P ggy, P ggz,
                    P bcu, P bcv, P bcw, P bcp, P bcf, P bcvi, P tbub,
                    P dtx, P dty, P dtz, P duu, P dvv, P dww,
                    P ggv0, P t1, P h1, P t2, P h2, P dbuf, P srcuv, P pppvc,
                    P flfu, P flfv, P flfw, P qf, P wk05, P wk06, P wk07, P wk08, P wk09, P wk10, P wk11,
                    I4 *nf, I4 *indx, I4 *indy, I4 *indz, I4 *indc, I4 *indb, I4 *inds_,
                    I4 *ipvc, I4 *ibuf, I4 *nlim, P fflxx, P fflxy, P delh, P delh_in);

void vf_f1cal_(P xx, P yy, P zz, P uu, P vv, P ww, P pp, P ff, P anu, P ggv, P ggx, P ggy, P ggz,
               P bcu, P bcv, P bcw, P bcp, P bcf, P bcvi, P tbub,
               P dtx, P dty, P dtz, P duu, P dvv, P dww,
               P ggv0, P t1, P h1, P t2, P h2, P dbuf, P srcuv, P pppvc,
               P flfu, P flfv, P flfw, P qf, P wk05, P wk06, P wk07, P wk08, P wk09, P wk10, P wk11,
               I4 *nf, I4 *indx, I4 *indy, I4 *indz, I4 *indc, I4 *indb, I4 *inds_,
               I4 *ipvc, I4 *ibuf, I4 *nlim, P fflxx, P fflxy, P delh, P delh_in)
{
    if (open_dump("f1_in")) {
        scalars();
        f3("UU", uu); f3("VV", vv); f3("WW", ww); f3("PP", pp); f3("FF", ff);
        fb("BCU", bcu); fb("BCV", bcv); fb("BCW", bcw); fb("BCP", bcp); fb("BCF", bcf);
        i3("NF", nf); i3("INDC", indc); inds(inds_);
        tdoor(tbub, dtx, dty, dtz, duu, dvv, dww);
        close_dump();
    }
    vf_f1cal_orig_(xx, yy, zz, uu, vv, ww, pp, ff, anu, ggv, ggx, ggy, ggz, bcu, bcv, bcw, bcp, bcf,
                   bcvi, tbub, dtx, dty, dtz, duu, dvv, dww, ggv0, t1, h1, t2, h2, dbuf, srcuv, pppvc,
                   flfu, flfv, flfw, qf, wk05, wk06, wk07, wk08, wk09, wk10, wk11,
                   nf, indx, indy, indz, indc, indb, inds_, ipvc, ibuf, nlim,
                   fflxx, fflxy, delh, delh_in);
    if (open_dump("f1_out")) {
        scalars();
        f3("UU", uu); f3("VV", vv); f3("WW", ww); f3("PP", pp); f3("FF", ff);
        fb("BCU", bcu); fb("BCV", bcv); fb("BCW", bcw); fb("BCP", bcp); fb("BCF", bcf);
        i3("NF", nf); i3("INDC", indc); inds(inds_);
        i3("NLIM", nlim);
        tdoor(tbub, dtx, dty, dtz, duu, dvv, dww);
        close_dump();
    }
}

void vf_cdiv00_orig_(P val, P xx, P yy, P zz, P uu, P vv, P ww, P ggx, P ggy, P ggz, I4 *indc);

void vf_cdiv00_(P val, P xx, P yy, P zz, P uu, P vv, P ww, P ggx, P ggy, P ggz, I4 *indc)
{
    vf_cdiv00_orig_(val, xx, yy, zz, uu, vv, ww, ggx, ggy, ggz, indc);
    if (open_dump("dv")) {
        sd("DV", *val);
        si("NNOW", vf_atimei_.nnow);
        /* the step size and time of step NNOW, full precision */
        sd("DTNOW", vf_atimer_.dtnow);
        sd("TNOW", vf_atimer_.tnow);
        /* the volume sums and the solve of the step that produced this
         * state (the list line's FSUM, FCUT, !B!, !R!, ITR), full precision */
        sd("FSUM", vf_acompr_.fsum);
        sd("FCUT", vf_acompr_.fcut);
        sd("CGBNRM", vf_acompr_.cgbnrm);
        sd("CGXNRM", vf_acompr_.cgxnrm);
        si("ICGITR", vf_acompi_.icgitr);
        close_dump();
    }
}
