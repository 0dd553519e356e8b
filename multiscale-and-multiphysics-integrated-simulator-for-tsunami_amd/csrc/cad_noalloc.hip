// cad_noalloc.hip -- the MILU-BiCGSTAB solve on device-resident arrays,
// allocating nothing.
//
// cad_m1bcgs_noalloc is a variant of the solve drop-in vf_m1bcgs_ for callers
// that cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc).  It takes the reference's
// argument list of VF_M1BCGS (vf_m1bcgs.f:1-5) as device pointers: the matrix
// AD..AUK, BB, the solution X0 (in and out), the flags INDC and the
// reference's own work arrays DD, XW, YY, R0, RR, PP, AP, AE, which hold the
// solver's vectors exactly as they do in the reference.  Everything else the
// solve needs -- the MILU products LPI/LPJ/LPK, the reduction partials and
// scalars, the active index range, the counters, the sweeps' tile orders,
// ranges, mailboxes, ticket and error flag -- lives in one file-scope device
// array, `scratch`, of CAD_NOALLOC_SCRATCH_BYTES.  Each GPU the library runs
// on holds its own copy until the process ends.  Its device address is taken
// once per device with hipGetSymbolAddress and used for every access from the
// host; the array's name is never passed to a kernel.
//
// Every call on a GPU shares that GPU's array: two calls that overlap on one
// GPU (on different streams, or from different host threads) race and give
// wrong results.  What must start at zero is cleared on each call by small
// kernels (no hipMemsetAsync): the work vectors, the sweep ticket and error
// flag; the mailboxes are cleared once, on the device's first call, and their
// entries afterwards carry sequence numbers below the next call's, which
// continue per device from call to call.
#include "cad_common.h"

#include <cstring>

namespace {

constexpr int MAX_DEVICES = 64;
constexpr unsigned long long LP_BYTES = 3ull * CAD_NOALLOC_MAX_CELLS * 8ull;
constexpr unsigned long long SMALL_BYTES = 64ull << 10;
constexpr unsigned long long SWEEP_BYTES = CAD_NOALLOC_SCRATCH_BYTES - LP_BYTES - SMALL_BYTES;

__device__ __align__(256) unsigned char scratch[CAD_NOALLOC_SCRATCH_BYTES];

struct DeviceState {
    bool ready;
    char *ptr;              // device address of this device's scratch
    unsigned seq;           // next sweep sequence number on this device
    cad_ctx ctx;            // the solve's view of the caller's arrays and the scratch
    double host_red[16];
    int host_cnt[16];
};
DeviceState g_dev[MAX_DEVICES];

__global__ void k_clear_bytes(unsigned long long *p, unsigned long long n8)
{
    for (unsigned long long t = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; t < n8;
         t += (unsigned long long)gridDim.x * blockDim.x)
        p[t] = 0ull;
}

int prepare(int dev, hipStream_t st)
{
    DeviceState &s = g_dev[dev];
    if (s.ready) return CAD_OK;
    char *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        cad_set_error("cad_m1bcgs_noalloc: hipGetSymbolAddress(scratch) on device %d: %s", dev,
                      hipGetErrorString(e));
        return CAD_ERR_HIP;
    }
    // mailboxes start at sequence number 0 (never a launch's number)
    hipLaunchKernelGGL(k_clear_bytes, dim3(1024), dim3(256), 0, st,
                       reinterpret_cast<unsigned long long *>(ptr + LP_BYTES + SMALL_BYTES),
                       SWEEP_BYTES / 8);
    CAD_CHECK(hipGetLastError());
    CAD_CHECK(hipStreamSynchronize(st));
    s.ptr = ptr;
    s.seq = 1;
    s.ready = true;
    return CAD_OK;
}

constexpr unsigned long long align256(unsigned long long n) { return (n + 255ull) & ~255ull; }

}  // namespace

extern "C" int cad_m1bcgs_noalloc(void *stream, const double *ad, const double *ali,
                                  const double *alj, const double *alk, const double *aui,
                                  const double *auj, const double *auk, const double *bb,
                                  double *x0, double *dd, double *xw, double *yy, double *r0,
                                  double *rr, double *pp, double *ap, double *ae, const int *indc,
                                  int ni, int nj, int nk, double cgpara, double cgepsa,
                                  double cgepsr, double cgdiv, int icgtyp, int icgmax,
                                  double *bnorm, double *xnorm, int *iter)
{
    const void *ptrs[] = {ad, ali, alj, alk, aui, auj, auk, bb, x0, dd, xw, yy, r0, rr, pp, ap, ae,
                          indc, bnorm, xnorm, iter};
    for (const void *q : ptrs)
        if (!q) {
            cad_set_error("cad_m1bcgs_noalloc: null argument");
            return CAD_ERR_ARG;
        }
    if (ni < 4 || nj < 4 || nk < 4) {
        cad_set_error("cad_m1bcgs_noalloc: sizes %d x %d x %d are not a grid with a ghost layer",
                      ni, nj, nk);
        return CAD_ERR_ARG;
    }
    const long long nc = (long long)ni * nj * nk;
    if (nc > (long long)CAD_NOALLOC_MAX_CELLS) {
        cad_set_error("cad_m1bcgs_noalloc: %d x %d x %d = %lld cells exceed the limit of %lld "
                      "(CAD_NOALLOC_MAX_CELLS)", ni, nj, nk, nc, (long long)CAD_NOALLOC_MAX_CELLS);
        return CAD_ERR_ARG;
    }
    int dev = 0;
    CAD_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) {
        cad_set_error("cad_m1bcgs_noalloc: device %d beyond %d", dev, MAX_DEVICES);
        return CAD_ERR_ARG;
    }
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (int rc = prepare(dev, st)) return rc;
    DeviceState &s = g_dev[dev];
    cad_ctx &c = s.ctx;
    std::memset((void *)&c, 0, sizeof(cad_ctx));
    c.device = dev;
    c.stream = st;
    c.d.ni = ni;
    c.d.nj = nj;
    c.d.nk = nk;
    c.d.nij = (long long)ni * nj;
    c.d.nc = nc;
    c.sweep_schedule = CAD_SWEEP_PIPELINED;
    if (cad_sweep_bytes(&c) > SWEEP_BYTES) {
        cad_set_error("cad_m1bcgs_noalloc: the sweeps of %d x %d x %d need %llu bytes of mailboxes, "
                      "more than the %llu reserved", ni, nj, nk, cad_sweep_bytes(&c), SWEEP_BYTES);
        return CAD_ERR_ARG;
    }
    c.prm.cgpara = cgpara;
    c.prm.cgepsa = cgepsa;
    c.prm.cgepsr = cgepsr;
    c.prm.cgdiv = cgdiv;
    c.prm.icgtyp = icgtyp;
    c.prm.icgmax = icgmax;
    c.prm.precision = CAD_PRECISION_FP64;
    c.ad = const_cast<double *>(ad);
    c.ali = const_cast<double *>(ali);
    c.alj = const_cast<double *>(alj);
    c.alk = const_cast<double *>(alk);
    c.aui = const_cast<double *>(aui);
    c.auj = const_cast<double *>(auj);
    c.auk = const_cast<double *>(auk);
    c.bb = const_cast<double *>(bb);
    c.f.indc = const_cast<int *>(indc);
    c.dd = dd;
    c.xw = xw;
    c.yv = yy;
    c.r0 = r0;
    c.rr = rr;
    c.pv = pp;
    c.ap = ap;
    c.ae = ae;
    // scratch: LPI | LPJ | LPK (CAD_NOALLOC_MAX_CELLS each), the small
    // arrays, then the sweeps' control arrays
    char *p = s.ptr;
    c.lpi = reinterpret_cast<double *>(p);
    c.lpj = c.lpi + CAD_NOALLOC_MAX_CELLS;
    c.lpk = c.lpj + CAD_NOALLOC_MAX_CELLS;
    char *q = p + LP_BYTES;
    c.red_partial = reinterpret_cast<double *>(q);
    q += align256(4ull * REDUCE_BLOCKS * sizeof(double));
    c.red_out = reinterpret_cast<double *>(q);
    q += align256(16 * sizeof(double));
    c.ksc = reinterpret_cast<double *>(q);
    q += align256(32 * sizeof(double));
    c.act_rng = reinterpret_cast<unsigned long long *>(q);
    q += align256(2 * sizeof(unsigned long long));
    c.icount = reinterpret_cast<int *>(q);
    q += align256((16 + 2 * 1024) * sizeof(int));
    static_assert(align256(4ull * REDUCE_BLOCKS * 8) + 3 * 256 + align256((16 + 2048) * 4) <=
                      SMALL_BYTES, "small scratch arrays");
    c.host_red = s.host_red;
    c.host_cnt = s.host_cnt;
    int rc = cad_sweep_bind(&c, p + LP_BYTES + SMALL_BYTES, SWEEP_BYTES, s.seq);
    if (rc == CAD_OK) rc = cad_solve_m1bcgs(&c, x0, bnorm, xnorm, iter);
    // the sequence numbers this call used stay behind in the mailboxes
    s.seq = c.sw_seq;
    if (rc != CAD_OK) return rc;
    CAD_CHECK(hipStreamSynchronize(st));
    return CAD_OK;
}
