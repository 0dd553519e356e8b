// cad_sweep.hip -- MILU triangular sweeps as one pipelined launch per sweep.
//
// The three recurrences of the MILU-preconditioned solver
//   factorisation  DD(n) = 1/(AD - ALI*DD(n-1)*AU1 - ALJ*DD(n-NI)*AU2 - ALK*DD(n-NIJ)*AU3)
//                                                         (vf_mzdcmp.f:53-64)
//   forward        X(n)  = Y - ALI*DD(n-1)*X(n-1) - ALJ*DD(n-NI)*X(n-NI) - ALK*DD(n-NIJ)*X(n-NIJ)
//                                                         (vf_mzfrwd.f:37-48)
//   backward       X(n)  = (X - AUI*X(n+1) - AUJ*X(n+NI) - AUK*X(n+NIJ)) * DD(n)
//                                                         (vf_mzbkwd.f:37-47)
// each read the value just produced at the three upstream neighbours.  Every
// cell here evaluates the reference's expression on the same operands in the
// same order, so the results are the serial sweep's bits; only the schedule
// differs.
//
// Schedule.  The (J,K) cross-section is cut into tiles of TJ x TK = 64 lines
// along I (8x8 or 16x4, whichever gives the shorter chain of tiles over the
// active region).  One wavefront owns a tile; lane (jl,kl) owns line (J,K) and
// walks I with a skew of jl+kl steps, so that at every step the J-1 and K-1
// neighbours of its cell were computed by lanes jl-1 / kl-1 one step earlier
// (a DPP row shift and a ds_bpermute hand them over).  Tiles pass their last
// column / row to the next tile through 16-byte mailbox entries {value,
// sequence number}, stored every step and read four steps ahead with
// device-scope (sc1) accesses: one entry is one single-copy-atomic 16-byte
// access, so a consumer that reads the current sequence number also reads the
// value that goes with it, with no fence.  A tile obtains its place in the
// dependency order from an atomic ticket, so a tile only ever waits on tiles
// whose wavefronts are already running, and every wait has a bounded spin
// count (exceeding it sets an error flag that the host reports).  Tiles
// without an active cell exit at once; the others walk only the I range that
// holds active cells.
//
// Memory.  The walk is cut into stages of S steps, and a workgroup is two
// wavefronts: the compute wave walks the tile, the loader wave moves memory.
// While stage b computes, the loader brings stage b+1's operands straight
// from HBM into the other LDS buffer with buffer_load ... lds (each line's
// S-cell window, shifted by the lane's skew, is one contiguous run, so one
// instruction moves 16 contiguous bytes of each of 64 line-pieces) and stores
// stage b-1's results, staged in LDS, as contiguous runs; one barrier per
// stage hands the buffers over.  Keeping the bulk traffic in the other wave
// keeps it out of the compute wave's in-order memory counter, so waiting for
// a mailbox entry never waits for a prefetch.  At the start of a stage every
// compute lane reads its S x NS operands from LDS into registers (the group
// layout makes these reads bank-conflict free), so the steps themselves touch
// no memory but the mailboxes.  The loads into LDS are not tracked by the
// compiler's wait insertion, so the loader waits for them explicitly.  Every
// access goes through a raw buffer resource sized to its array: a load
// outside it returns 0 and a store outside it is dropped.
//
// Inactive cells.  Cells with INDC = -1 are not computed and keep the value
// the output array holds, which for every array swept here is zero (the
// solver clears its work vectors and DD at start and writes active cells
// only), so an inactive neighbour contributes 0 exactly as the serial sweep
// reads it from memory.  Activity is read from an operand stream instead of
// INDC: the precomputed LP*/AU* arrays carry a marker value at inactive cells,
// and DD is zero there and positive at active cells.
#include "cad_common.h"

#include <algorithm>
#include <climits>
#include <vector>

namespace {

constexpr unsigned SPIN_LIMIT = 1u << 20;
constexpr unsigned OOB = 0xFFFFFFF0u;
constexpr int CPOL_DEVICE = 16;      // sc1: device-scope coherent access
constexpr int MB_LINES = 16;         // mailbox lines per tile and side (max of TJ, TK)
// inactive-cell marker: a signalling-NaN pattern that no arithmetic produces
constexpr unsigned long long MARK = 0x7FF4DEADBEEF0001ull;

typedef unsigned v4u __attribute__((ext_vector_type(4)));
typedef unsigned v3u __attribute__((ext_vector_type(3)));
typedef unsigned v2u __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) void lds_void;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void *p, unsigned long long bytes)
{
    const unsigned nrec = bytes > 0xFFFFFFF0ull ? 0xFFFFFFF0u : (unsigned)bytes;
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, (int)nrec, 0x00020000);
}

// 16 bytes per lane from buffer offset off straight into LDS at lds + 16*lane
__device__ __forceinline__ void load_lds16(__amdgpu_buffer_rsrc_t r, void *lds, unsigned off)
{
    __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (lds_void *)lds, 16, off, 0, 0, 0);
}

__device__ __forceinline__ bool is_mark(double x)
{
    return __builtin_bit_cast(unsigned long long, x) == MARK;
}

// {value, sequence number} of a mailbox entry: 12 bytes inside the 16-byte
// entry that one store wrote
__device__ __forceinline__ v3u mb_load(__amdgpu_buffer_rsrc_t r, unsigned off)
{
    return __builtin_amdgcn_raw_buffer_load_b96(r, off, 0, CPOL_DEVICE);
}

__device__ __forceinline__ void mb_store(__amdgpu_buffer_rsrc_t r, unsigned off, double x, unsigned seq)
{
    const v2u b = __builtin_bit_cast(v2u, x);
    const v4u e = {b.x, b.y, seq, seq};
    __builtin_amdgcn_raw_buffer_store_b128(e, r, off, 0, CPOL_DEVICE);
}

// Slow path of a mailbox read: polls the entry until it carries sequence
// number seq, at most SPIN_LIMIT times (then flags err and returns what it
// has).  Once any wait has given up, no wait of this schedule waits again, so
// a failed launch still drains in about one spin limit.
__device__ __attribute__((noinline)) v3u mb_wait(const void *mb, unsigned long long bytes,
                                                 unsigned off, unsigned seq, int *err)
{
    const __amdgpu_buffer_rsrc_t r = rsrc(mb, bytes);
    v3u e;
    unsigned spins = 0;
    do {
        __builtin_amdgcn_s_sleep(1);
        __asm__ __volatile__("" ::: "memory");
        e = mb_load(r, off);
        if ((++spins & 1023u) == 0u) {
            if (spins > SPIN_LIMIT) atomicOr(err, 2);
            if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
        }
    } while (e.z != seq);
    return e;
}

// value of lane-1 (row shift by one inside each row of 16 lanes; lanes with
// jl == 0 take their J neighbour from the mailbox instead)
__device__ __forceinline__ double from_lane_minus1(double x)
{
    v2u b = __builtin_bit_cast(v2u, x);
    v2u r;
    r.x = (unsigned)__builtin_amdgcn_update_dpp(0, (int)b.x, 0x111, 0xF, 0xF, false);
    r.y = (unsigned)__builtin_amdgcn_update_dpp(0, (int)b.y, 0x111, 0xF, 0xF, false);
    return __builtin_bit_cast(double, r);
}

// value of lane src (any lane)
__device__ __forceinline__ double from_lane(double x, int src)
{
    v2u b = __builtin_bit_cast(v2u, x);
    const int a = (src & 63) << 2;
    v2u r;
    r.x = (unsigned)__builtin_amdgcn_ds_bpermute(a, (int)b.x);
    r.y = (unsigned)__builtin_amdgcn_ds_bpermute(a, (int)b.y);
    return __builtin_bit_cast(double, r);
}

// Barrier between the two waves of a workgroup that orders their LDS traffic
// only: a full __syncthreads() would also make the compute wave wait for its
// outstanding mailbox accesses every stage.  The compiler barriers keep LDS
// accesses on their side of it.
__device__ __forceinline__ void lds_barrier()
{
    __asm__ __volatile__("s_waitcnt lgkmcnt(0)" ::: "memory");
    __builtin_amdgcn_s_barrier();
    __asm__ __volatile__("" ::: "memory");
}

// value of lane l (uniform l)
__device__ __forceinline__ double read_lane(double x, int l)
{
    v2u b = __builtin_bit_cast(v2u, x);
    v2u r;
    r.x = (unsigned)__builtin_amdgcn_readlane((int)b.x, l);
    r.y = (unsigned)__builtin_amdgcn_readlane((int)b.y, l);
    return __builtin_bit_cast(double, r);
}

struct SweepGeom {
    int ni, nj, nk;
    int ntj, ntk, ntiles;
    long long nij, nc;
};

struct SweepCtl {
    const int *order;               // tiles in dependency order
    const int2 *range;              // per tile: first, last active I (0-based); x > y: none
    v4u *mb;                        // mailboxes: J side then K side, each [tile][16 lines][I]
    unsigned long long mb_bytes;    // bytes of one side
    unsigned long long *ticket;
    unsigned long long base;
    unsigned seq;
    int *err;
};

// forward (L D^-1 + I) x = y with LPI = ALI*DD(n-1), LPJ = ALJ*DD(n-NI),
// LPK = ALK*DD(n-NIJ) formed once per factorisation (the same products the
// reference forms left to right before multiplying by X); LPI carries MARK at
// inactive cells.
struct FwdOp {
    static constexpr int NS = 4;
    static constexpr int S = 16;
    const double *src[NS];          // Y, LPI, LPJ, LPK
    double *out;                    // X
    int *flag;
    __device__ __forceinline__ static bool active(const double *v) { return !is_mark(v[1]); }
    __device__ __forceinline__ static double eval(const double *v, double xi, double xj, double xk)
    {
        return v[0] - v[1] * xi - v[2] * xj - v[3] * xk;
    }
    static constexpr bool CHECK_PIVOT = false;
};

// backward (D + U) x = x, in place; DD is 0 at inactive cells
struct BwdOp {
    static constexpr int NS = 5;
    static constexpr int S = 16;
    const double *src[NS];          // X, AUI, AUJ, AUK, DD
    double *out;                    // X
    int *flag;
    __device__ __forceinline__ static bool active(const double *v) { return v[4] != 0.0; }
    __device__ __forceinline__ static double eval(const double *v, double xi, double xj, double xk)
    {
        return (v[0] - v[1] * xi - v[2] * xj - v[3] * xk) * v[4];
    }
    static constexpr bool CHECK_PIVOT = false;
};

// MILU diagonal; AU1..AU3 are the neighbour sums of vf_mzdcmp.f:57-59 (AU1
// carries MARK at inactive cells)
struct DcmpOp {
    static constexpr int NS = 7;
    static constexpr int S = 8;
    const double *src[NS];          // AD, ALI, ALJ, ALK, AU1, AU2, AU3
    double *out;                    // DD
    int *flag;                      // cleared when a pivot is not positive
    __device__ __forceinline__ static bool active(const double *v) { return !is_mark(v[4]); }
    __device__ __forceinline__ static double eval(const double *v, double xi, double xj, double xk)
    {
        return 1.0 / (v[0] - v[1] * xi * v[4] - v[2] * xj * v[5] - v[3] * xk * v[6]);
    }
    static constexpr bool CHECK_PIVOT = true;
};

// LDS layout of one operand stream of one stage: S window cells x 64 lines.
// A load instruction q moves LPI lines (16 bytes = 2 cells per lane); lane l
// takes line q*LPI + l%LPI, cells 2*(l/LPI) and 2*(l/LPI)+1, and lands at
// gb(q) + 4*l dwords.  gb(q) mod 64 is 0, 2, 32, 34 for q = 0..3 (mod 4), so
// the 32 lanes of a half-wave reading cell e of their own lines hit 32
// distinct even banks.
template <int S>
struct Lay {
    static constexpr int LPI = 128 / S;      // lines per load instruction
    static constexpr int NG = 64 / LPI;      // load instructions per stream
    static constexpr int GP = 320;           // group pitch (dwords)
    static constexpr int WORDS = NG * GP;    // dwords per stream
    __device__ __forceinline__ static int gb(int q)
    {
        return q * GP + 2 * (q & 1) + 32 * ((q >> 1) & 1);
    }
    __device__ __forceinline__ static int addr(int L, int e)
    {
        return gb(L / LPI) + (e >> 1) * (LPI * 4) + (L % LPI) * 4 + (e & 1) * 2;
    }
};

template <class Op>
struct __attribute__((aligned(16))) Smem {
    static constexpr int S = Op::S, NS = Op::NS;
    unsigned st[2][NS][Lay<S>::WORDS];       // operand streams, two stages
    double xo[2][64][S + 1];                 // results of two stages, per line
    int tile;                                // this workgroup's tile
};

// Per-lane maps of one tile.  Lane = jl + TJ*kl owns logical line L = lane
// and walks it with a skew of jl + kl steps; the maps give its line, the
// offsets of the operand loads and result stores of stage 0 (stage b adds
// b*DIR*S cells), the mailbox entries the lane watches and those it feeds.
template <int S, int DIR, int TJ>
struct TileMap {
    static constexpr int TK = 64 / TJ;
    static constexpr int NG = Lay<S>::NG, LPI = Lay<S>::LPI;
    static constexpr int LPS = 64 / S;       // lines per result store
    static constexpr int HELPER = 1 + TJ;    // lane (1,1): reads lane (0,0)'s K entry
    int ni, nj, nk;
    long long nij;
    int tile, bj, bk, jl, kl, skew, len;
    int2 rg;
    bool valid;
    unsigned ld_off[NG], st_off[S];
    unsigned stage_step;                     // byte step per stage (mod 2^32)
    int hlo, hn, fsk, fn;                    // watched / fed windows of steps
    unsigned h0, f0;

    __device__ __forceinline__ int phys_j(int j) const { return DIR > 0 ? j : TJ - 1 - j; }
    __device__ __forceinline__ int phys_k(int k) const { return DIR > 0 ? k : TK - 1 - k; }
    __device__ __forceinline__ bool line_ok(int L) const
    {
        return 1 + bj * TJ + phys_j(L % TJ) <= nj - 2 && 1 + bk * TK + phys_k(L / TJ) <= nk - 2;
    }
    __device__ __forceinline__ long long line_base(int L) const
    {
        return (long long)(1 + bj * TJ + phys_j(L % TJ)) * ni + (long long)(1 + bk * TK + phys_k(L / TJ)) * nij;
    }
    // first cell (ascending I) of line L's window in stage 0
    __device__ __forceinline__ int win0(int L) const
    {
        const int sk = L % TJ + L / TJ;
        return DIR > 0 ? rg.x - sk : rg.y - (S - 1) + sk;
    }
    __device__ __forceinline__ unsigned mb_off(int tl, int l, int i) const
    {
        return (unsigned)((((long long)tl * MB_LINES + l) * ni + i) * 16);
    }

    __device__ __forceinline__ void init(const SweepGeom &geo, const SweepCtl &c, int t, int lane)
    {
        ni = geo.ni;
        nj = geo.nj;
        nk = geo.nk;
        nij = geo.nij;
        tile = t;
        bj = t % geo.ntj;
        bk = t / geo.ntj;
        rg = c.range[t];
        len = rg.y - rg.x + 1;
        jl = lane % TJ;
        kl = lane / TJ;
        skew = jl + kl;
        valid = line_ok(lane);
#pragma unroll
        for (int q = 0; q < NG; q++) {
            const int L = q * LPI + lane % LPI;
            const long long e = line_base(L) + win0(L) + 2 * (lane / LPI);
            ld_off[q] = line_ok(L) && e >= 0 ? (unsigned)(e * 8) : OOB;
        }
#pragma unroll
        for (int q = 0; q < S; q++) {
            const int L = q * LPS + lane / S;
            const long long e = line_base(L) + win0(L) + lane % S;
            st_off[q] = line_ok(L) && e >= 0 ? (unsigned)(e * 8) : OOB;
        }
        stage_step = (unsigned)(DIR * S * 8);

        // Mailbox reads.  A lane with jl == 0 reads the J-upstream tile's
        // entry for its own line, a lane with kl == 0 (jl > 0) the K-upstream
        // one; lane HELPER, which has no halo of its own, reads lane (0,0)'s
        // K entry.
        const unsigned half_k = (unsigned)c.mb_bytes;
        const int pbj = bj - DIR, pbk = bk - DIR;
        const int ptj = pbj >= 0 && pbj < geo.ntj ? pbj + bk * geo.ntj : -1;
        const int ptk = pbk >= 0 && pbk < geo.ntk ? bj + pbk * geo.ntj : -1;
        const bool has_cj = bj + DIR >= 0 && bj + DIR < geo.ntj;
        const bool has_ck = bk + DIR >= 0 && bk + DIR < geo.ntk;
        int hsrc = -1, hline = 0, hsk = skew;
        bool hval = valid;
        unsigned hbase = 0;
        int2 hr = make_int2(1, 0);
        if (jl == 0) {
            if (ptj >= 0) {
                hsrc = ptj;
                hr = c.range[ptj];
            }
            hline = phys_k(kl);
            if (kl == TK - 1) {
                // the J-upstream corner cell (its lane 63) is published on
                // its K side only
                hbase = half_k;
                hline = phys_j(TJ - 1);
            }
        } else if (kl == 0 || lane == HELPER) {
            if (ptk >= 0) {
                hsrc = ptk;
                hr = c.range[ptk];
            }
            hbase = half_k;
            hline = phys_j(jl);
            if (lane == HELPER) {
                hline = phys_j(0);
                hsk = 0;
                hval = line_ok(0);
            }
        }
        // the watched entry of step t is at h0 + DIR*16*t for t in
        // [hlo, hlo+hn): the steps where the watched cell is walked and lies
        // in the producer's range
        hlo = 0;
        hn = 0;
        h0 = 0;
        if (hsrc >= 0 && hval) {
            // cell_i(t, hsk) in [hr.x, hr.y] and 0 <= t - hsk < len
            const int ta = DIR > 0 ? hr.x - rg.x + hsk : rg.y - hr.y + hsk;
            const int tb = DIR > 0 ? hr.y - rg.x + hsk : rg.y - hr.x + hsk;
            hlo = max(ta, hsk);
            hn = max(min(tb, hsk + len - 1) - hlo + 1, 0);
            h0 = hbase + mb_off(hsrc, hline, 0) + (unsigned)((DIR > 0 ? rg.x - hsk : rg.y + hsk) * 16);
        }
        // Mailbox writes, one per lane and step.  Lanes with kl == TK-1 feed
        // the K-downstream tile; lanes with jl == TJ-1 (kl < TK-1) the
        // J-downstream one.  The corner lane 63 writes its K entry only,
        // which the J-downstream tile reads too, so it writes it when either
        // downstream tile exists.
        bool fok = false;
        int fline = 0;
        unsigned fbase = 0;
        fsk = skew;
        if (kl == TK - 1) {
            fok = valid && (has_ck || (jl == TJ - 1 && has_cj));
            fline = phys_j(jl);
            fbase = half_k;
        } else if (jl == TJ - 1) {
            fok = valid && has_cj;
            fline = phys_k(kl);
        }
        // the fed entry of step t is at f0 + DIR*16*t for t in [fsk, fsk+fn)
        f0 = fbase + mb_off(t, fline, 0) + (unsigned)((DIR > 0 ? rg.x - fsk : rg.y + fsk) * 16);
        fn = fok ? len : 0;
    }
    __device__ __forceinline__ unsigned h_off(int t) const
    {
        return (unsigned)(t - hlo) < (unsigned)hn ? h0 + (unsigned)(DIR * 16 * t) : OOB;
    }
    __device__ __forceinline__ unsigned f_off(int t) const
    {
        return (unsigned)(t - fsk) < (unsigned)fn ? f0 + (unsigned)(DIR * 16 * t) : OOB;
    }
    // offset of result store q of stage b for this lane, OOB outside the
    // tile's active I range
    __device__ __forceinline__ unsigned store_off(int q, int lane, int b) const
    {
        const int L = q * LPS + lane / S, e = lane % S;
        const int skl = L % TJ + L / TJ;
        const int i = (DIR > 0 ? rg.x - skl + S * b : rg.y - (S - 1) + skl - S * b) + e;
        return st_off[q] == OOB || i < rg.x || i > rg.y ? OOB : st_off[q] + stage_step * (unsigned)b;
    }
};

// issue the loads of stage b's operands into LDS buffer b&1
template <class Op, class TM>
__device__ __forceinline__ void load_stage(Smem<Op> &sm, const __amdgpu_buffer_rsrc_t *rs, const TM &m, int b)
{
    using LY = Lay<Op::S>;
    const unsigned adv = m.stage_step * (unsigned)b;
#pragma unroll
    for (int s = 0; s < Op::NS; s++) {
#pragma unroll
        for (int q = 0; q < LY::NG; q++)
            load_lds16(rs[s], &sm.st[b & 1][s][LY::gb(q)], m.ld_off[q] == OOB ? OOB : m.ld_off[q] + adv);
    }
}

// results of stage b (in xo[b&1]): store q covers lines q*LPS .. q*LPS+LPS-1.
// Every cell of the tile's active I range is stored, inactive ones as 0 (the
// value the output array holds there), so that each line's window is one
// contiguous run and HBM sees whole lines rather than masked pieces; cells
// outside the range are not stored.
template <class Op, class TM>
__device__ __forceinline__ void store_stage(const Smem<Op> &sm, __amdgpu_buffer_rsrc_t rout, const TM &m,
                                            int lane, int b)
{
    constexpr int S = Op::S;
#pragma unroll
    for (int q = 0; q < S; q++) {
        double x = sm.xo[b & 1][q * TM::LPS + lane / S][lane % S];
        if (is_mark(x)) x = 0.0;
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(v2u, x), rout, m.store_off(q, lane, b), 0, 0);
    }
}

// The tile of this workgroup from the ticket (thread 0 draws it), or -1.
__device__ __forceinline__ int draw_tile(const SweepGeom &g, const SweepCtl &c, int dir, int *slot)
{
    if (threadIdx.x == 0) {
        const unsigned long long tk = atomicAdd(c.ticket, 1ull);
        const long long tix = (long long)(tk - c.base);
        if (tix < 0 || tix >= g.ntiles) {
            atomicOr(c.err, 1);
            *slot = -1;
        } else {
            *slot = c.order[dir > 0 ? tix : g.ntiles - 1 - tix];
        }
    }
    __syncthreads();
    return __builtin_amdgcn_readfirstlane(*slot);
}

// One step of a lane's walk: the reference's expression on the operands of
// the cell and its three upstream neighbours' new values (xi from the lane's
// previous step, xj / xk from lanes jl-1 / kl-1 one step earlier or, on the
// tile's upstream faces, from the halo values hv and hk0 of lane (0,0)'s K
// neighbour).
template <class Op, int TJ>
__device__ __forceinline__ double sweep_step(const double *v, bool act, double xcur, double hv, double hk0,
                                             int jl, int kl, int lane)
{
    const double sj = from_lane_minus1(xcur);
    const double sk = from_lane(xcur, lane - TJ);
    const double xj = jl == 0 ? hv : sj;
    const double xk = kl == 0 ? (jl == 0 ? hk0 : hv) : sk;
    const double ev = Op::eval(v, xcur, xj, xk);
    return act ? ev : 0.0;
}

template <class Op, int DIR, int TJ>
__global__ void __launch_bounds__(128) k_sweep(SweepGeom g, SweepCtl c, Op op)
{
    constexpr int NS = Op::NS, S = Op::S;
    // mailbox read-ahead (steps); the ring is indexed by p % HD, so HD must
    // divide the stage length
    constexpr int HD = 4;
    static_assert(S % HD == 0, "ring depth must divide the stage length");
    using LY = Lay<S>;
    using TM = TileMap<S, DIR, TJ>;
    __shared__ Smem<Op> sm;
    const int lane = threadIdx.x & 63;
    const bool loader = threadIdx.x >= 64;   // wave 1 moves memory, wave 0 computes
    const int tile = draw_tile(g, c, DIR, &sm.tile);
    if (tile < 0) return;
    if (c.range[tile].x > c.range[tile].y) return;
    TM m;
    m.init(g, c, tile, lane);
    const int nsteps = m.len + (TJ - 1) + (TM::TK - 1);
    const int nstage = (nsteps + S - 1) / S;
    const __amdgpu_buffer_rsrc_t rmb = rsrc(c.mb, 2 * c.mb_bytes);
    const unsigned seq = c.seq;

    if (loader) {
        __amdgpu_buffer_rsrc_t rs[NS];
#pragma unroll
        for (int s = 0; s < NS; s++) rs[s] = rsrc(op.src[s], (unsigned long long)g.nc * 8);
        const __amdgpu_buffer_rsrc_t rout = rsrc(op.out, (unsigned long long)g.nc * 8);
        // stage b+1's operands in while stage b computes; stage b-1's results out
        load_stage(sm, rs, m, 0);
        __asm__ __volatile__("s_waitcnt vmcnt(0)" ::: "memory");
        lds_barrier();
        for (int b = 0; b < nstage; b++) {
            if (b + 1 < nstage) load_stage(sm, rs, m, b + 1);
            if (b > 0) store_stage(sm, rout, m, lane, b - 1);
            __asm__ __volatile__("s_waitcnt vmcnt(0)" ::: "memory");
            lds_barrier();
        }
        store_stage(sm, rout, m, lane, nstage - 1);
        return;
    }

    double xcur = 0.0;
    bool bad_pivot = false;
    v3u ring[HD];
    // Each step issues one mailbox load (HD steps ahead) and one mailbox
    // store.  The loads ahead of the first step are interleaved with stores
    // that fall outside the buffer (dropped), so that the vector-memory
    // instructions in flight when a stage starts follow the same pattern as
    // in the middle of a stage: the wait for a ring entry then counts the
    // HD-1 loads and HD stores issued after it, and the first steps of a
    // stage do not also wait for the loads and stores of the last few steps.
#pragma unroll
    for (int d = 0; d < HD; d++) {
        ring[d] = mb_load(rmb, m.h_off(d));
        mb_store(rmb, OOB, 0.0, seq);
        // the stores have one address: keep each (no store merging)
        __asm__ __volatile__("" ::: "memory");
    }
    lds_barrier();
    for (int b = 0; b < nstage; b++) {
        // this stage's operands: LDS -> registers
        double sv[S][NS];
#pragma unroll
        for (int p = 0; p < S; p++) {
            const int e = DIR > 0 ? p : S - 1 - p;
#pragma unroll
            for (int s = 0; s < NS; s++)
                sv[p][s] = *reinterpret_cast<const double *>(&sm.st[b & 1][s][LY::addr(lane, e)]);
        }
#pragma unroll
        for (int p = 0; p < S; p++) {
            const int t = b * S + p;
            const bool lv = m.valid && (unsigned)(t - m.skew) < (unsigned)m.len;
            const int e = DIR > 0 ? p : S - 1 - p;
            const bool act = lv && Op::active(sv[p]);
            // mailbox entry of this step (read HD steps ago); a stale one is
            // polled for, in a wave-uniform slow path
            v3u he = ring[p % HD];
            const unsigned ho = m.h_off(t);
            const bool act0 = (__builtin_amdgcn_ballot_w64(act) & 1ull) != 0;
            const bool stale = ho != OOB && (lane == TM::HELPER ? act0 : act) && he.z != seq;
            if (__builtin_amdgcn_ballot_w64(stale)) {
                if (stale) he = mb_wait(c.mb, 2 * c.mb_bytes, ho, seq, c.err);
            }
            ring[p % HD] = mb_load(rmb, m.h_off(t + HD));
            const double hv = __builtin_bit_cast(double, v2u{he.x, he.y});
            const double hk0 = read_lane(hv, TM::HELPER);
            const double r = sweep_step<Op, TJ>(sv[p], act, xcur, hv, hk0, m.jl, m.kl, lane);
            if (Op::CHECK_PIVOT) bad_pivot = bad_pivot || (act && r <= 0.0);
            xcur = r;
            sm.xo[b & 1][lane][e] = act ? r : __builtin_bit_cast(double, MARK);
            mb_store(rmb, m.f_off(t), r, seq);
        }
        lds_barrier();
    }
    if (Op::CHECK_PIVOT && bad_pivot) __hip_atomic_store(op.flag, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Active I range of every tile (INDC != -1 on its lines) and the largest
// active K (0-based) in *kmax.
template <int TJ>
__global__ void __launch_bounds__(256) k_tile_range(SweepGeom g, const int *__restrict__ indc,
                                                    int2 *__restrict__ range, int *kmax)
{
    constexpr int TK = 64 / TJ;
    const int tile = blockIdx.x;
    const int bj = tile % g.ntj, bk = tile / g.ntj;
    int lo = INT_MAX, hi = -1, kh = -1;
    for (int l = 0; l < 64; l++) {
        const int j = 1 + bj * TJ + l % TJ, k = 1 + bk * TK + l / TJ;
        if (j > g.nj - 2 || k > g.nk - 2) continue;
        const int *row = indc + (long long)j * g.ni + (long long)k * g.nij;
        for (int i = 1 + threadIdx.x; i <= g.ni - 2; i += blockDim.x) {
            if (row[i] != -1) {
                lo = min(lo, i);
                hi = max(hi, i);
                kh = max(kh, k);
            }
        }
    }
    __shared__ int slo[256], shi[256], skh[256];
    slo[threadIdx.x] = lo;
    shi[threadIdx.x] = hi;
    skh[threadIdx.x] = kh;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            slo[threadIdx.x] = min(slo[threadIdx.x], slo[threadIdx.x + w]);
            shi[threadIdx.x] = max(shi[threadIdx.x], shi[threadIdx.x + w]);
            skh[threadIdx.x] = max(skh[threadIdx.x], skh[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        range[tile] = shi[0] < 0 ? make_int2(1, 0) : make_int2(slo[0], shi[0]);
        if (skh[0] >= 0) atomicMax(kmax, skh[0]);
    }
}

__device__ __forceinline__ bool interior(const Dims &d, int i, int j, int k)
{
    return i >= 2 && j >= 2 && k >= 2 && i <= d.ni - 1 && j <= d.nj - 1 && k <= d.nk - 1;
}

// AU1..AU3 of vf_mzdcmp.f:57-59 on active interior cells; AU1 = MARK on
// inactive interior cells
__global__ void __launch_bounds__(256) k_dcmp_pre(Dims d, const double *__restrict__ aui,
                                                  const double *__restrict__ auj,
                                                  const double *__restrict__ auk,
                                                  const int *__restrict__ indc, double u,
                                                  double *__restrict__ au1, double *__restrict__ au2,
                                                  double *__restrict__ au3)
{
    int i, j, k;
    if (!cell_of_thread(d, i, j, k) || !interior(d, i, j, k)) return;
    const long long n = cid(d, i, j, k);
    if (indc[n] == -1) {
        au1[n] = __builtin_bit_cast(double, MARK);
        return;
    }
    const long long ni = d.ni, nij = d.nij;
    au1[n] = aui[n - 1] + u * (auj[n - 1] + auk[n - 1]);
    au2[n] = auj[n - ni] + u * (aui[n - ni] + auk[n - ni]);
    au3[n] = auk[n - nij] + u * (aui[n - nij] + auj[n - nij]);
}

// LPI = ALI*DD(n-1), LPJ = ALJ*DD(n-NI), LPK = ALK*DD(n-NIJ); LPI = MARK on
// inactive interior cells
__global__ void __launch_bounds__(256) k_lprod(Dims d, const double *__restrict__ ali,
                                               const double *__restrict__ alj,
                                               const double *__restrict__ alk,
                                               const double *__restrict__ dd,
                                               const int *__restrict__ indc, double *__restrict__ lpi,
                                               double *__restrict__ lpj, double *__restrict__ lpk)
{
    int i, j, k;
    if (!cell_of_thread(d, i, j, k) || !interior(d, i, j, k)) return;
    const long long n = cid(d, i, j, k);
    if (indc[n] == -1) {
        lpi[n] = __builtin_bit_cast(double, MARK);
        return;
    }
    lpi[n] = ali[n] * dd[n - 1];
    lpj[n] = alj[n] * dd[n - d.ni];
    lpk[n] = alk[n] * dd[n - d.nij];
}

__global__ void k_set_int(int *p, int v) { *p = v; }

constexpr int SHAPE_TJ[2] = {8, 16};

SweepGeom geom(const cad_ctx *c, int shape)
{
    SweepGeom g;
    g.ni = c->d.ni;
    g.nj = c->d.nj;
    g.nk = c->d.nk;
    g.ntj = c->sw_ntj[shape];
    g.ntk = c->sw_ntk[shape];
    g.ntiles = c->sw_ntiles[shape];
    g.nij = c->d.nij;
    g.nc = c->d.nc;
    return g;
}

SweepCtl ctl(cad_ctx *c, int shape)
{
    SweepCtl s;
    s.order = c->sw_order[shape];
    s.range = reinterpret_cast<const int2 *>(c->sw_range[shape]);
    s.mb = reinterpret_cast<v4u *>(c->sw_mb);
    s.mb_bytes = c->sw_mb_bytes;
    s.ticket = c->sw_ticket;
    s.base = c->sw_base;
    s.seq = c->sw_seq;
    s.err = c->sw_err;
    return s;
}

template <class Op, int DIR>
int launch(cad_ctx *c, const Op &op)
{
    const int shape = c->sw_shape;
    SweepGeom g = geom(c, shape);
    SweepCtl s = ctl(c, shape);
    if (shape == 0) {
        hipLaunchKernelGGL((k_sweep<Op, DIR, 8>), dim3(g.ntiles), dim3(128), 0, c->stream, g, s, op);
    } else {
        hipLaunchKernelGGL((k_sweep<Op, DIR, 16>), dim3(g.ntiles), dim3(128), 0, c->stream, g, s, op);
    }
    CAD_CHECK(hipGetLastError());
    c->sw_base += (unsigned long long)g.ntiles;
    c->sw_seq += 1;
    return CAD_OK;
}

}  // namespace

int cad_sweep_alloc(cad_ctx *c)
{
    const Dims &d = c->d;
    unsigned long long lines = 0;
    for (int sh = 0; sh < 2; sh++) {
        const int tj = SHAPE_TJ[sh], tk = 64 / tj;
        c->sw_ntj[sh] = (d.nj - 2 + tj - 1) / tj;
        c->sw_ntk[sh] = (d.nk - 2 + tk - 1) / tk;
        c->sw_ntiles[sh] = c->sw_ntj[sh] * c->sw_ntk[sh];
        const int nt = c->sw_ntiles[sh], ntj = c->sw_ntj[sh];
        std::vector<int> order(nt);
        for (int t = 0; t < nt; t++) order[t] = t;
        std::stable_sort(order.begin(), order.end(), [ntj](int a, int b) {
            const int da = a % ntj + a / ntj, db = b % ntj + b / ntj;
            if (da != db) return da < db;
            return a % ntj < b % ntj;
        });
        CAD_CHECK(hipMalloc((void **)&c->sw_order[sh], nt * sizeof(int)));
        CAD_CHECK(hipMalloc((void **)&c->sw_range[sh], nt * 2 * sizeof(int)));
        CAD_CHECK(hipMemcpy(c->sw_order[sh], order.data(), nt * sizeof(int), hipMemcpyHostToDevice));
        lines = std::max(lines, (unsigned long long)nt * MB_LINES);
    }
    c->sw_mb_bytes = lines * d.ni * 16;
    if (2 * c->sw_mb_bytes > 0xFFFFFFF0ull || (unsigned long long)d.nc * 8 > 0xFFFFFFF0ull) {
        cad_set_error("grid too large for 32-bit sweep offsets (%d x %d x %d)", d.ni, d.nj, d.nk);
        return CAD_ERR_ARG;
    }
    CAD_CHECK(hipMalloc(&c->sw_mb, 2 * c->sw_mb_bytes));
    CAD_CHECK(hipMalloc((void **)&c->sw_ticket, sizeof(unsigned long long)));
    CAD_CHECK(hipMalloc((void **)&c->sw_err, 2 * sizeof(int)));
    CAD_CHECK(hipMemset(c->sw_mb, 0, 2 * c->sw_mb_bytes));
    CAD_CHECK(hipMemset(c->sw_ticket, 0, sizeof(unsigned long long)));
    CAD_CHECK(hipMemset(c->sw_err, 0, 2 * sizeof(int)));
    CAD_CHECK(hipDeviceSynchronize());
    c->sw_base = 0;
    c->sw_seq = 1;
    c->sw_shape = 0;
    return CAD_OK;
}

namespace {

// Position of each tile in the dependency order that cad_sweep_alloc sorts on
// the host: by anti-diagonal bj + bk, then by bj.
__global__ void k_tile_order(int ntj, int ntk, int *order)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntj * ntk) return;
    const int bj = t % ntj, diag = t % ntj + t / ntj;
    int pos = 0;
    for (int s = 0; s < diag; s++) pos += min(ntj - 1, s) - max(0, s - (ntk - 1)) + 1;
    order[pos + bj - max(0, diag - (ntk - 1))] = t;
}

__global__ void k_clear_u64(unsigned long long *p, int n)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0ull;
}

constexpr unsigned long long align256(unsigned long long n) { return (n + 255ull) & ~255ull; }

}  // namespace

// Bytes of device memory the pipelined schedule's control arrays take for
// the context's grid (tile orders and ranges of both shapes, mailboxes,
// ticket, error flag), each part 256-byte aligned.
unsigned long long cad_sweep_bytes(const cad_ctx *c)
{
    const Dims &d = c->d;
    unsigned long long total = 0, lines = 0;
    for (int sh = 0; sh < 2; sh++) {
        const int tj = SHAPE_TJ[sh], tk = 64 / tj;
        const unsigned long long nt =
            (unsigned long long)((d.nj - 2 + tj - 1) / tj) * ((d.nk - 2 + tk - 1) / tk);
        total += align256(nt * sizeof(int)) + align256(nt * 2 * sizeof(int));
        lines = std::max(lines, nt * MB_LINES);
    }
    return total + align256(2 * lines * d.ni * 16) + 256 + 256;
}

// The same control arrays as cad_sweep_alloc, carved from [base, base +
// bytes) that the caller owns instead of allocated: the tile orders are
// written by a kernel, the ticket and the error flag are cleared by one, all
// on c->stream.  The mailboxes are not cleared: their entries carry the
// sequence numbers of earlier launches, and seq (the first launch's number)
// must exceed every sequence number written into this memory before.
int cad_sweep_bind(cad_ctx *c, char *base, unsigned long long bytes, unsigned seq)
{
    const Dims &d = c->d;
    const unsigned long long need = cad_sweep_bytes(c);
    if (need > bytes) {
        cad_set_error("sweep control arrays need %llu bytes, %llu available", need, bytes);
        return CAD_ERR_ARG;
    }
    char *p = base;
    unsigned long long lines = 0;
    for (int sh = 0; sh < 2; sh++) {
        const int tj = SHAPE_TJ[sh], tk = 64 / tj;
        c->sw_ntj[sh] = (d.nj - 2 + tj - 1) / tj;
        c->sw_ntk[sh] = (d.nk - 2 + tk - 1) / tk;
        c->sw_ntiles[sh] = c->sw_ntj[sh] * c->sw_ntk[sh];
        const int nt = c->sw_ntiles[sh];
        c->sw_order[sh] = reinterpret_cast<int *>(p);
        p += align256((unsigned long long)nt * sizeof(int));
        c->sw_range[sh] = reinterpret_cast<int *>(p);
        p += align256((unsigned long long)nt * 2 * sizeof(int));
        hipLaunchKernelGGL(k_tile_order, dim3((nt + 255) / 256), dim3(256), 0, c->stream,
                           c->sw_ntj[sh], c->sw_ntk[sh], c->sw_order[sh]);
        lines = std::max(lines, (unsigned long long)nt * MB_LINES);
    }
    c->sw_mb_bytes = lines * d.ni * 16;
    if (2 * c->sw_mb_bytes > 0xFFFFFFF0ull || (unsigned long long)d.nc * 8 > 0xFFFFFFF0ull) {
        cad_set_error("grid too large for 32-bit sweep offsets (%d x %d x %d)", d.ni, d.nj, d.nk);
        return CAD_ERR_ARG;
    }
    c->sw_mb = p;
    p += align256(2 * c->sw_mb_bytes);
    c->sw_ticket = reinterpret_cast<unsigned long long *>(p);
    p += 256;
    c->sw_err = reinterpret_cast<int *>(p);
    hipLaunchKernelGGL(k_clear_u64, dim3(1), dim3(64), 0, c->stream, c->sw_ticket, 1);
    hipLaunchKernelGGL(k_clear_u64, dim3(1), dim3(64), 0, c->stream,
                       reinterpret_cast<unsigned long long *>(c->sw_err), 1);
    CAD_CHECK(hipGetLastError());
    c->sw_base = 0;
    c->sw_seq = seq;
    c->sw_shape = 0;
    return CAD_OK;
}

void cad_sweep_free(cad_ctx *c)
{
    void *p[] = {c->sw_order[0], c->sw_order[1], c->sw_range[0], c->sw_range[1], c->sw_mb,
                 c->sw_ticket, c->sw_err};
    for (void *q : p)
        if (q) (void)hipFree(q);
}

// Tile ranges for the current INDC, choice of the tile shape, the MILU
// diagonal DD and the products LPI/LPJ/LPK.  *npos_host = 0 when a pivot is
// not positive (vf_mzdcmp.f:61).
int cad_sweep_dcmp(cad_ctx *c, double u, int *npos_host)
{
    const Dims &d = c->d;
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, c->stream, c->icount + 3, -1);
    for (int sh = 0; sh < 2; sh++) {
        SweepGeom g = geom(c, sh);
        if (sh == 0)
            hipLaunchKernelGGL(k_tile_range<8>, dim3(g.ntiles), dim3(256), 0, c->stream, g, c->f.indc,
                               reinterpret_cast<int2 *>(c->sw_range[sh]), c->icount + 3);
        else
            hipLaunchKernelGGL(k_tile_range<16>, dim3(g.ntiles), dim3(256), 0, c->stream, g, c->f.indc,
                               reinterpret_cast<int2 *>(c->sw_range[sh]), c->icount + 3);
    }
    CAD_CHECK(hipGetLastError());
    CAD_CHECK(hipMemcpyAsync(c->host_cnt + 3, c->icount + 3, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    CAD_CHECK(hipStreamSynchronize(c->stream));
    // tile shape: the one with the shorter chain of tiles over the active K range
    const int kmax = c->host_cnt[3];
    if (c->sweep_schedule == CAD_SWEEP_PIPELINED_8X8) {
        c->sw_shape = 0;
    } else if (c->sweep_schedule == CAD_SWEEP_PIPELINED_16X4) {
        c->sw_shape = 1;
    } else {
        int best = 0, best_hops = INT_MAX;
        for (int sh = 0; sh < 2; sh++) {
            const int tk = 64 / SHAPE_TJ[sh];
            const int hops = c->sw_ntj[sh] + (kmax > 0 ? (kmax + tk - 1) / tk : 1);
            if (hops < best_hops) {
                best_hops = hops;
                best = sh;
            }
        }
        c->sw_shape = best;
    }
    hipLaunchKernelGGL(k_dcmp_pre, cell_grid(d), cell_block(), 0, c->stream, d, c->aui, c->auj,
                       c->auk, c->f.indc, u, c->lpi, c->lpj, c->lpk);
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, c->stream, c->icount + 1, 1);
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, c->stream, c->sw_err, 0);
    CAD_CHECK(hipGetLastError());
    DcmpOp op;
    op.src[0] = c->ad;
    op.src[1] = c->ali;
    op.src[2] = c->alj;
    op.src[3] = c->alk;
    op.src[4] = c->lpi;
    op.src[5] = c->lpj;
    op.src[6] = c->lpk;
    op.out = c->dd;
    op.flag = c->icount + 1;
    if (int rc = launch<DcmpOp, 1>(c, op)) return rc;
    hipLaunchKernelGGL(k_lprod, cell_grid(d), cell_block(), 0, c->stream, d, c->ali, c->alj,
                       c->alk, c->dd, c->f.indc, c->lpi, c->lpj, c->lpk);
    CAD_CHECK(hipGetLastError());
    CAD_CHECK(hipMemcpyAsync(c->host_cnt + 1, c->icount + 1, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    CAD_CHECK(hipMemcpyAsync(c->host_cnt + 2, c->sw_err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    CAD_CHECK(hipStreamSynchronize(c->stream));
    if (c->host_cnt[2] != 0) {
        cad_set_error("MILU sweep schedule failed (code %d)", c->host_cnt[2]);
        return CAD_ERR_HIP;
    }
    *npos_host = c->host_cnt[1];
    return CAD_OK;
}

// x = M^-1 y (vf_mzminv.f:38-41): forward then backward sweep, in place on x.
int cad_sweep_minv(cad_ctx *c, double *x, const double *y)
{
    FwdOp f;
    f.src[0] = y;
    f.src[1] = c->lpi;
    f.src[2] = c->lpj;
    f.src[3] = c->lpk;
    f.out = x;
    f.flag = nullptr;
    if (int rc = launch<FwdOp, 1>(c, f)) return rc;
    BwdOp b;
    b.src[0] = x;
    b.src[1] = c->aui;
    b.src[2] = c->auj;
    b.src[3] = c->auk;
    b.src[4] = c->dd;
    b.out = x;
    b.flag = nullptr;
    return launch<BwdOp, -1>(c, b);
}

// Error flag of the sweeps since the last call (0 = none); clears it.
int cad_sweep_status(cad_ctx *c, int *code)
{
    CAD_CHECK(hipMemcpyAsync(c->host_cnt + 2, c->sw_err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    CAD_CHECK(hipStreamSynchronize(c->stream));
    *code = c->host_cnt[2];
    if (*code) {
        hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, c->stream, c->sw_err, 0);
        CAD_CHECK(hipGetLastError());
    }
    return CAD_OK;
}
