// cad_unet.hip -- the learned preconditioner's network: a 3-level 3-D U-Net
// (3x3x3 convolutions, 16/32/64 channels, 390,913 parameters) with BF16
// weights and activations on the gfx950 matrix cores and FP32 accumulation,
// forward and backward, and an Adam step, all hand-written.
//
// Layout.  Activations are channels-last BF16: voxel v holds its C channels
// contiguously, so one lane's 8 consecutive channels are one 16-byte load.
// A level's volumes (a batch of nb boxes of nx x ny x nz voxels) are stored
// with a one-voxel zero halo on every side of every box, so a 3x3x3
// convolution with zero padding reads its neighbours with no bounds tests and
// boxes of a batch never see each other.  Kernels write interior voxels only;
// halos stay zero from allocation.  A tensor is (pointer, channel stride):
// the decoder's concatenation buffers are written in channel slices (the
// skip connection's encoder output in the upper half, the upsampled coarse
// features in the lower half), so the concatenation costs no copy.
//
// Convolution as implicit GEMM on v_mfma_f32_16x16x32_bf16, marched along z.
// The z taps are folded into the output rows: for each dz in {-1, 0, +1}
//   P_dz (Cout x voxels of plane p) = W_dz (Cout x K) * X_p^T (K x voxels),  K = 9*Cin (dx, dy, channel),
// and input plane p adds P_dz to out(p - dz).
// A block (4 waves, one y row each, NT groups of 16 voxels along x) walks a
// run of z planes; its three FP32 accumulator sets hold out(p+1), out(p) and
// out(p-1) while plane p is read, and out(p-1), then complete, is stored.
// Every input fragment a wave loads feeds the three dz weight tiles, so an
// input plane is read once per block, and the x and y taps are per-lane load
// offsets within the plane, which the block's four rows share in L1.
// A lane holds A[m = l%16][k = 8*(l/16)..+7] (8 weights, packed so that it is
// one 16-byte load, staged in LDS), B[k = 8*(l/16)..+7][n = l%16] (8 channels
// of one (dx,dy) tap of voxel n: one 16-byte load) and D[m = 4*(l/16)+i][n =
// l%16] (4 consecutive output channels of one voxel: one 8-byte store).
// Blocks of neighbouring rows are dealt to the same XCD, so the rows they
// share stay in one L2.  The backward pass for the activations is the same
// kernel with the taps flipped and the weight matrix transposed (packed that
// way), and the weight gradients are an FP32 reduction over voxels.
//
// A convolution's input may come from two buffers (the decoder's skip
// concatenation: channels below `split` from the upsampled coarse features,
// the rest from the encoder's output), and its output may go to two, so
// every buffer holds whole voxel records and no kernel writes part of one.
#include "cad_common.h"
#include "cad_unet.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <type_traits>
#include <vector>

namespace {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ long long vox(const UGeo &g, int x, int y, int z, int b)
{
    return (long long)(x + 1) + (long long)g.px * ((y + 1) + (long long)g.py * ((z + 1) + (long long)g.pz * b));
}

// interior voxel number u (x fastest, then y, z, box) -> padded index
__device__ __forceinline__ long long vox_of(const UGeo &g, long long u)
{
    const int x = (int)(u % g.nx);
    u /= g.nx;
    const int y = (int)(u % g.ny);
    u /= g.ny;
    const int z = (int)(u % g.nz);
    const int b = (int)(u / g.nz);
    return vox(g, x, y, z, b);
}

__device__ __forceinline__ long long tap_off(const UGeo &g, int t)
{
    const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = t / 9 - 1;
    return dx + (long long)g.px * (dy + (long long)g.py * dz);
}

// K steps of 32 of the folded convolution for cin input channels (9 taps)
__host__ __device__ constexpr int fold_kc(int cin) { return (9 * cin + 31) / 32; }
// bytes of A fragments a launch of MT channel tiles stages in LDS, and the cap
constexpr int conv_a_bytes(int cin, int mt) { return 3 * mt * fold_kc(cin) * 64 * 16; }
// blocks stage at most this much in LDS; weights above CONV_A_WIDE take
// blocks of 16 waves (one per CU) so that more rows share one staged copy
constexpr int CONV_A_LDS_MAX = 112 * 1024;
constexpr int CONV_A_WIDE = 40 * 1024;
__host__ __device__ constexpr int conv_wx(int cin, int mt) { return conv_a_bytes(cin, mt) > CONV_A_WIDE ? 4 : 1; }

// input channels [0, split) at p, [split, C) at p + delta, both with records
// of cs channels (delta: the second half of a concatenation, allocated right
// after the first so that one base and 32-bit lane offsets reach both)
struct ConvIn {
    const bf16 *p;
    long long delta;
    int cs, split;
};
// Output channels [0, split) to p0, the rest to p1 (p0 null: not stored).
// Two optional extras of the same BF16 values: up, the 2x nearest-neighbour
// upsampling into the next finer level (padded dims upx, upy, upz of ucs-
// channel records), and y, the network's output layer y = ow[16] + sum_c
// ow[c] h[c] in FP32 at the interior voxel number (a 16-channel output only).
struct ConvOut {
    bf16 *p0, *p1;
    int cs0, cs1, split;
    bf16 *up;
    int ucs, upx, upy, upz;
    const float *ow;
    float *y;
};
__device__ __forceinline__ const bf16 *chan(const ConvIn &t, long long v, int c)
{
    return c < t.split ? t.p + v * t.cs + c : t.p + t.delta + v * t.cs + (c - t.split);
}
__device__ __forceinline__ bf16 *chan(const ConvOut &t, long long v, int c)
{
    return c < t.split ? t.p0 + v * t.cs0 + c : t.p1 + v * t.cs1 + (c - t.split);
}

// v_permlane16_swap on two copies of x: rows (x0, x0, x2, x2) and (x1, x1, x3,
// x3) of the wave's four 16-lane rows.  The second operand is an opaque copy:
// given the same value twice, the compiler takes the two results as equal.
__device__ __forceinline__ void rows_pair(unsigned x, unsigned &even, unsigned &odd)
{
    unsigned y = x;
    __asm__("" : "+v"(y));
    const auto r = __builtin_amdgcn_permlane16_swap(x, y, false, false);
    even = r[0];
    odd = r[1];
}
// v_permlane32_swap on two copies of x: rows (x0, x1, x0, x1) and (x2, x3, x2, x3)
__device__ __forceinline__ void halves_pair(unsigned x, unsigned &lo, unsigned &hi)
{
    unsigned y = x;
    __asm__("" : "+v"(y));
    const auto r = __builtin_amdgcn_permlane32_swap(x, y, false, false);
    lo = r[0];
    hi = r[1];
}

// Stores out(o), complete in the accumulators a of a wave (NT groups of 16
// voxels from x0 = 16 gi0 + n, row y, box b): bias, ReLU, BF16, then the
// outputs out names: EPI 0 p0/p1 only, 1 also the upsampling (p0 may be
// null), 2 also the output layer (p0 may be null).  Called uniformly across
// the wave, so the output layer's channel sum may cross lanes.  ow: the
// output layer's weights of this lane's four channels and its bias (EPI 2).
template <int MT, int NT, bool RELU, int EPI>
__device__ __forceinline__ void conv_store(const UGeo &g, const ConvOut &out, const f4 (&a)[MT][NT],
                                           const float (&bs)[MT][4], const float (&ow)[5], int m0, int kg,
                                           int gi0, int ng, int x0, int y, int o, int b)
{
    const long long v0 = vox(g, x0, y, o, b);
#pragma unroll
    for (int m = 0; m < MT; m++) {
        const int co = 16 * (m0 + m) + 4 * kg;
#pragma unroll
        for (int j = 0; j < NT; j++) {
            if (gi0 + j >= ng) continue;
            bf16x4 ov;
#pragma unroll
            for (int c = 0; c < 4; c++) {
                float val = a[m][j][c] + bs[m][c];
                if (RELU) val = val > 0.f ? val : 0.f;
                ov[c] = (bf16)val;
            }
            if (EPI == 0 || out.p0) *reinterpret_cast<bf16x4 *>(chan(out, v0 + 16 * j, co)) = ov;
            if (EPI == 1) {
                // Lane rows kg and kg ^ 1 hold channels 8 (kg / 2) .. +7 of the
                // voxel between them; rows_pair gives every lane rows (2 r,
                // 2 r + 1) of its pair, so each lane stores those 8
                // channels (16 bytes) into fine voxel 2x + (kg & 1), and a
                // wave's store covers whole fine records.
                const uint2 w = __builtin_bit_cast(uint2, ov);
                uint4 q;
                rows_pair(w.x, q.x, q.z);
                rows_pair(w.y, q.y, q.w);
                const int xf = 2 * (x0 + 16 * j) + 1 + (kg & 1), yf = 2 * y + 1, zf = 2 * o + 1;
                const int cf = 16 * (m0 + m) + 8 * (kg >> 1);
                const long long vf =
                    xf + (long long)out.upx * (yf + (long long)out.upy * (zf + (long long)out.upz * b));
                bf16 *pf = out.up + vf * out.ucs + cf;
                const long long sy = (long long)out.upx * out.ucs, sz = sy * out.upy;   // uniform
                *reinterpret_cast<uint4 *>(pf) = q;
                *reinterpret_cast<uint4 *>(pf + sy) = q;
                *reinterpret_cast<uint4 *>(pf + sz) = q;
                *reinterpret_cast<uint4 *>(pf + sy + sz) = q;
            }
            if (EPI == 2) {
                // channels 4 kg .. 4 kg + 3 here, the other twelve in the other
                // lane rows: rows (0, 1) and (2, 3) summed in pairs, then the
                // pair sums
                float s = 0.f;
#pragma unroll
                for (int c = 0; c < 4; c++) s += ow[c] * (float)ov[c];
                unsigned a, c;
                rows_pair(__builtin_bit_cast(unsigned, s), a, c);
                const float t = __builtin_bit_cast(float, a) + __builtin_bit_cast(float, c);
                halves_pair(__builtin_bit_cast(unsigned, t), a, c);
                s = __builtin_bit_cast(float, a) + __builtin_bit_cast(float, c);
                if (kg == 0)
                    out.y[(x0 + 16 * j) + (long long)g.nx * (y + (long long)g.ny * (o + (long long)g.nz * b))] =
                        s + ow[4];
            }
        }
    }
}

// 3x3x3 convolution, CIN -> 16*MT channels: output channel tiles m0..m0+MT-1
// (wp points at tile m0).  A block is 4 rows (y) by WX wave columns (x) of
// NT groups of 16 voxels, WX = conv_wx(CIN, MT).  Tiles (x, y, z run of zlen
// planes, box) are numbered x fastest; XCD k (blocks b with b % 8 == k) takes
// tiles [k per_xcd, (k+1) per_xcd).  Groups past nx and rows past ny are
// idle.  wp: packed A fragments [MT][3 dz][KC][64 lanes] of 8 BF16.  bias may
// be null.
template <int CIN, int MT, int NT, bool RELU, int EPI>
__global__ void __launch_bounds__(256 * conv_wx(CIN, MT), conv_wx(CIN, MT) == 1 ? 3 : 4)
    k_conv3(UGeo g, ConvIn in, const bf16x8 *__restrict__ wp, const float *__restrict__ bias, ConvOut out, int m0,
            int ntx, int nty, int nzc, int zlen, int per_xcd, int ntiles)
{
    constexpr int KC = fold_kc(CIN);
    constexpr int WX = conv_wx(CIN, MT);
    constexpr int NA = 3 * MT * KC * 64;
    static_assert(conv_a_bytes(CIN, MT) <= CONV_A_LDS_MAX, "weights must fit in LDS");
    __shared__ bf16x8 as[NA];
    const int tile = (int)(blockIdx.x & 7) * per_xcd + (int)(blockIdx.x >> 3);
    if (tile >= ntiles) return;   // the whole block
    const int tx = tile % ntx, ty = (tile / ntx) % nty, zc = (tile / (ntx * nty)) % nzc;
    const int b = tile / (ntx * nty * nzc);
    const int z0 = zc * zlen, z1 = min(g.nz, z0 + zlen);
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int n = lane & 15, kg = lane >> 4;
    const int y = 4 * ty + (wv & 3);
    const int gi0 = (tx * WX + (wv >> 2)) * NT, ng = g.nx >> 4;
    const bool yok = y < g.ny && gi0 < ng;
    const int x0 = 16 * gi0 + n;
    const int np = z1 - z0 + 2;   // input planes z0-1 .. z1
    const int nstep = np * KC;
    // Byte offsets of the K steps' taps for each lane, from the wave's base
    // voxel (x = 16 gi0 - 1, y - 1, plane z0 - 1): the same for every plane.
    // The loads then take a wave-uniform base and a 32-bit lane offset.
    __shared__ unsigned boff[KC][64];
    for (int e = threadIdx.x; e < KC * 64; e += blockDim.x) {
        const int q = e >> 6, l = e & 63;
        const int k0 = 32 * q + 8 * (l >> 4);
        int t = k0 / CIN;
        const int ci0 = k0 % CIN;
        if (t > 8) t = 8;   // K padding: the weights there are zero
        const int rel = (l & 15) + t % 3 + g.px * (t / 3);
        const bool hi = ci0 >= in.split;
        boff[q][l] = (unsigned)(2LL * ((long long)rel * in.cs + (hi ? in.delta + ci0 - in.split : ci0)));
    }
    // Loads go through a buffer resource at the wave's base voxel: a lane
    // offset from the table, a wave-uniform plane and group offset.
    const long long vb = vox(g, 16 * gi0 - 1, y - 1, z0 - 1, b);
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc((void *)(in.p + vb * in.cs), (short)0, -1, 0x00020000);
    const int plane_bytes = 2 * g.px * g.py * in.cs;
    // groups past nx read the wave's first group instead; their sums are not stored
    int gofs[NT];
#pragma unroll
    for (int j = 0; j < NT; j++) gofs[j] = (gi0 + j < ng ? j : 0) * 32 * in.cs;
    auto load_b = [&](int st, bf16x8 *dst) {
        const int pi = st / KC, q = st - pi * KC;
        const int voff = (int)boff[q][lane];
#pragma unroll
        for (int j = 0; j < NT; j++)
            dst[j] = __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(
                                                    rs, voff, pi * plane_bytes + gofs[j], 0));
    };
    for (int e = threadIdx.x; e < NA; e += blockDim.x) as[e] = wp[e];
    __syncthreads();
    if (!yok) return;   // past ny or nx: no loads, no stores, no barrier after this
    constexpr int PD = 2;   // steps in flight ahead of the one computed, in a ring of 3
    bf16x8 ring[3][NT];
#pragma unroll
    for (int p = 0; p < PD; p++)
        if (p < nstep) load_b(p, ring[p]);
    const f4 zero = {0.f, 0.f, 0.f, 0.f};
    f4 acc[3][MT][NT];   // slot (o - z0 + 2) % 3 holds out(o)
#pragma unroll
    for (int sl = 0; sl < 3; sl++)
#pragma unroll
        for (int m = 0; m < MT; m++)
#pragma unroll
            for (int j = 0; j < NT; j++) acc[sl][m][j] = zero;
    float bs[MT][4];
#pragma unroll
    for (int m = 0; m < MT; m++)
#pragma unroll
        for (int i = 0; i < 4; i++) bs[m][i] = bias ? bias[16 * (m0 + m) + 4 * kg + i] : 0.f;
    float ow[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
    if constexpr (EPI == 2) {
#pragma unroll
        for (int c = 0; c < 4; c++) ow[c] = out.ow[4 * kg + c];
        ow[4] = out.ow[16];
    }
#pragma unroll 1
    for (int p0 = 0; p0 < np; p0 += 3) {
#pragma unroll
        for (int u = 0; u < 3; u++) {
            const int pi = p0 + u;   // input plane z0 - 1 + pi
            if (pi >= np) break;
#pragma unroll
            for (int q = 0; q < KC; q++) {
                const int st = pi * KC + q;
                if (st + PD < nstep) load_b(st + PD, ring[(u * KC + q + PD) % 3]);
                const bf16x8 *cur = ring[(u * KC + q) % 3];
#pragma unroll
                for (int m = 0; m < MT; m++)
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        // dz = d - 1 adds to out(p - dz), slot (u + 2 - d) % 3
                        const int fa = ((3 * m + d) * KC + q) * 64 + lane;
                        const bf16x8 af = as[fa];
#pragma unroll
                        for (int j = 0; j < NT; j++)
                            acc[(u + 2 - d) % 3][m][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                                af, cur[j], acc[(u + 2 - d) % 3][m][j], 0, 0, 0);
                    }
            }
            // out(p - 1) is complete: slot u % 3
            const int o = z0 - 2 + pi;
            if (o >= z0 && o < z1)
                conv_store<MT, NT, RELU, EPI>(g, out, acc[u % 3], bs, ow, m0, kg, gi0, ng, x0, y, o, b);
#pragma unroll
            for (int m = 0; m < MT; m++)
#pragma unroll
                for (int j = 0; j < NT; j++) acc[u % 3][m][j] = zero;
        }
    }
}

// bytes of a staged input plane tile: 6 rows (the block's 4 and their y
// neighbours) by the block's 16*NT*WX voxels and their x neighbours
__host__ __device__ constexpr int conv_tile_bytes(int cin, int nt, int wx) { return 6 * (16 * nt * wx + 2) * cin * 2; }
constexpr int CONV_STAGED_MAX = 160 * 1024 - 2048;   // gfx950: 160 KB of LDS per workgroup
// (in blocks of 16 waves, 128 registers a lane: the staging registers fit
// beside 8 accumulator quads, not 12)
__host__ __device__ constexpr bool conv_staged(int cin, int mt, int nt)
{
    return conv_tile_bytes(cin, nt, conv_wx(cin, mt)) + conv_a_bytes(cin, mt) <= CONV_STAGED_MAX &&
           (conv_wx(cin, mt) == 1 || mt * nt <= 2);
}

// k_conv3 with the input staged in LDS: for each input plane the block copies
// its tile (16-byte chunks, each thread a fixed set of them) into LDS, and
// the waves read their B fragments from there; the next plane's chunks are
// loaded into registers while the current plane is computed.  Same tiling,
// numbering and arguments as k_conv3.
template <int CIN, int MT, int NT, bool RELU, int EPI>
__global__ void __launch_bounds__(256 * conv_wx(CIN, MT), conv_wx(CIN, MT) == 1 ? 3 : 4)
    k_conv3s(UGeo g, ConvIn in, const bf16x8 *__restrict__ wp, const float *__restrict__ bias, ConvOut out,
             int m0, int ntx, int nty, int nzc, int zlen, int per_xcd, int ntiles)
{
    constexpr int KC = fold_kc(CIN);
    constexpr int WX = conv_wx(CIN, MT);
    constexpr int NA = 3 * MT * KC * 64;
    constexpr int TW = 16 * NT * WX + 2;   // tile width in voxels
    constexpr int CH = CIN / 8;            // 16-byte chunks per voxel
    constexpr int NCH = 6 * TW * CH;       // chunks per tile
    constexpr int NTHR = 256 * WX;
    constexpr int PER = (NCH + NTHR - 1) / NTHR;
    static_assert(conv_staged(CIN, MT, NT), "tile and weights must fit in LDS");
    __shared__ bf16x8 as[NA];
    __shared__ bf16x8 tl[NCH];
    const int tile = (int)(blockIdx.x & 7) * per_xcd + (int)(blockIdx.x >> 3);
    if (tile >= ntiles) return;   // the whole block
    const int tx = tile % ntx, ty = (tile / ntx) % nty, zc = (tile / (ntx * nty)) % nzc;
    const int b = tile / (ntx * nty * nzc);
    const int z0 = zc * zlen, z1 = min(g.nz, z0 + zlen);
    const int tid = threadIdx.x, lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int n = lane & 15, kg = lane >> 4;
    const int wy = wv & 3, wxi = wv >> 2;
    const int xb = 16 * NT * WX * tx, y0 = 4 * ty;
    const int y = y0 + wy;
    const int gi0 = (xb >> 4) + wxi * NT, ng = g.nx >> 4;
    const bool active = y < g.ny && gi0 < ng;
    const int np = z1 - z0 + 2;   // input planes z0-1 .. z1
    // this thread's chunks: byte offsets from the block's base voxel
    // (xb - 1, y0 - 1, plane z0 - 1); chunks past the box's halo are skipped
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        (void *)(in.p + vox(g, xb - 1, y0 - 1, z0 - 1, b) * in.cs), (short)0, -1, 0x00020000);
    const int plane_bytes = 2 * g.px * g.py * in.cs;
    int goff[PER];
    bool gok[PER];
#pragma unroll
    for (int i = 0; i < PER; i++) {
        const int c = tid + i * NTHR;
        const int r = c / (TW * CH), rem = c - r * (TW * CH);
        const int col = rem / CH, c8 = 8 * (rem - col * CH);
        gok[i] = c < NCH && xb - 1 + col <= g.nx && y0 - 1 + r <= g.ny;
        const bool hi = c8 >= in.split;
        goff[i] = (int)(2LL * ((long long)(r * g.px + col) * in.cs + (hi ? in.delta + c8 - in.split : c8)));
    }
    bf16x8 stg[PER];
    auto fetch = [&](int pi) {
#pragma unroll
        for (int i = 0; i < PER; i++)
            if (gok[i])
                stg[i] = __builtin_bit_cast(bf16x8,
                                            __builtin_amdgcn_raw_buffer_load_b128(rs, goff[i], pi * plane_bytes, 0));
    };
    fetch(0);
    for (int e = tid; e < NA; e += NTHR) as[e] = wp[e];
    const f4 zero = {0.f, 0.f, 0.f, 0.f};
    f4 acc[3][MT][NT];   // slot (o - z0 + 2) % 3 holds out(o)
#pragma unroll
    for (int sl = 0; sl < 3; sl++)
#pragma unroll
        for (int m = 0; m < MT; m++)
#pragma unroll
            for (int j = 0; j < NT; j++) acc[sl][m][j] = zero;
    float bs[MT][4];
#pragma unroll
    for (int m = 0; m < MT; m++)
#pragma unroll
        for (int i = 0; i < 4; i++) bs[m][i] = bias ? bias[16 * (m0 + m) + 4 * kg + i] : 0.f;
    float ow[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
    if constexpr (EPI == 2) {
#pragma unroll
        for (int c = 0; c < 4; c++) ow[c] = out.ow[4 * kg + c];
        ow[4] = out.ow[16];
    }
    // the wave's first B chunk: row wy (tap dy = -1), voxel wxi*16*NT + n (tap dx = -1)
    const int lb = (wy * TW + wxi * 16 * NT + n) * CH;
#pragma unroll 1
    for (int p0 = 0; p0 < np; p0 += 3) {
#pragma unroll
        for (int u = 0; u < 3; u++) {
            const int pi = p0 + u;   // input plane z0 - 1 + pi
            if (pi >= np) break;
            __syncthreads();   // the previous tile is read
#pragma unroll
            for (int i = 0; i < PER; i++)
                if (gok[i]) tl[tid + i * NTHR] = stg[i];
            __syncthreads();
            if (pi + 1 < np) fetch(pi + 1);
            if (active) {
                // Plane pi adds to out(z0 + pi - d) for d = 0, 1, 2; the run's
                // first two and last two planes have taps whose output lies
                // outside it.  Blocks of 16 waves, whose runs are short, skip
                // those products (on runs of two planes, half the work).
                const int zl = z1 - z0;
                // MASK: skip the taps whose output lies outside the run
                auto mac = [&](auto mask) {
#pragma unroll
                    for (int q = 0; q < KC; q++) {
                        const int k0 = 32 * q + 8 * kg;
                        int t = k0 / CIN;
                        if (t > 8) t = 8;   // K padding: the weights there are zero
                        const int bq = lb + ((t / 3) * TW + t % 3) * CH + (k0 % CIN) / 8;
                        bf16x8 bf[NT];
#pragma unroll
                        for (int j = 0; j < NT; j++) bf[j] = tl[bq + 16 * j * CH];
#pragma unroll
                        for (int m = 0; m < MT; m++)
#pragma unroll
                            for (int d = 0; d < 3; d++) {
                                if (decltype(mask)::value && (pi - d < 0 || pi - d >= zl)) continue;
                                // dz = d - 1 adds to out(p - dz), slot (u + 2 - d) % 3
                                const bf16x8 af = as[((3 * m + d) * KC + q) * 64 + lane];
#pragma unroll
                                for (int j = 0; j < NT; j++)
                                    acc[(u + 2 - d) % 3][m][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                                        af, bf[j], acc[(u + 2 - d) % 3][m][j], 0, 0, 0);
                            }
                    }
                };
                if constexpr (WX == 4) {
                    // two copies of the loop: a branch inside it costs more
                    if (pi >= 2 && pi < zl)
                        mac(std::false_type{});
                    else
                        mac(std::true_type{});
                } else {
#pragma unroll
                    for (int q = 0; q < KC; q++) {
                        const int k0 = 32 * q + 8 * kg;
                        int t = k0 / CIN;
                        if (t > 8) t = 8;
                        const int bq = lb + ((t / 3) * TW + t % 3) * CH + (k0 % CIN) / 8;
                        bf16x8 bf[NT];
#pragma unroll
                        for (int j = 0; j < NT; j++) bf[j] = tl[bq + 16 * j * CH];
#pragma unroll
                        for (int m = 0; m < MT; m++)
#pragma unroll
                            for (int d = 0; d < 3; d++) {
                                const bf16x8 af = as[((3 * m + d) * KC + q) * 64 + lane];
#pragma unroll
                                for (int j = 0; j < NT; j++)
                                    acc[(u + 2 - d) % 3][m][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                                        af, bf[j], acc[(u + 2 - d) % 3][m][j], 0, 0, 0);
                            }
                    }
                }
                // out(p - 1) is complete: slot u % 3
                const int o = z0 - 2 + pi;
                if (o >= z0 && o < z1)
                    conv_store<MT, NT, RELU, EPI>(g, out, acc[u % 3], bs, ow, m0, kg, gi0, ng, 16 * gi0 + n, y, o,
                                                  b);
#pragma unroll
                for (int m = 0; m < MT; m++)
#pragma unroll
                    for (int j = 0; j < NT; j++) acc[u % 3][m][j] = zero;
            }
        }
    }
}

// Packs FP32 weights W[t][ci][co] (27 x cin x cout, t = dx+1 + 3(dy+1) + 9(dz+1))
// into the A fragments of k_conv3 for cin -> cout: tile 3m + dzi holds output
// channels 16m..16m+15 of the z tap dzi, its K runs over (dx, dy) taps and
// channels.  flip: the backward-data operator, whose input has cout channels
// and output cin, W'[t][co][ci] = W[26-t][ci][co].
__global__ void k_pack(const float *__restrict__ w, int cin, int cout, int flip, bf16 *__restrict__ wp)
{
    const int ki = flip ? cout : cin;   // input channels of the packed operator
    const int ko = flip ? cin : cout;   // its output channels
    const int kc = fold_kc(ki), mt = ko / 16;
    const long long total = 3LL * mt * kc * 64 * 8;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const int el = (int)(e % 8);
        const int lane = (int)((e / 8) % 64);
        const int q = (int)((e / 512) % kc);
        const int tile = (int)(e / (512LL * kc));
        const int m = tile / 3, dzi = tile % 3;
        const int k = 32 * q + 8 * (lane >> 4) + el;
        const int o = 16 * m + (lane & 15);
        float v = 0.f;
        const int tp = k / ki, c = k % ki;
        if (tp < 9) {
            const int t = tp + 9 * dzi;
            if (!flip)
                v = w[((long long)t * cin + c) * cout + o];
            else
                v = w[((long long)(26 - t) * cin + o) * cout + c];
        }
        wp[e] = (bf16)v;
    }
}

// FP32 weight gradient of one 3x3x3 layer: dW[t][ci][co] += sum_v X[v+off_t][ci] dY[v][co]
// and (tap 13's blocks) db[co] += sum_v dY[v][co].  Block: one tap, a range of
// interior voxels; threads own 4x4 (ci, co) blocks, several thread groups
// split each tile of voxels; partial sums are added atomically.
template <int CIN, int COUT>
__global__ void __launch_bounds__(256) k_conv_dw(UGeo g, ConvIn x,
                                                 const bf16 *__restrict__ dy, int cs_dy,
                                                 long long per_block, float *__restrict__ dw,
                                                 float *__restrict__ db)
{
    constexpr int TV = 32;
    constexpr int NBLK = (CIN / 4) * (COUT / 4);
    constexpr int GROUPS = NBLK >= 256 ? 1 : 256 / NBLK;
    __shared__ float xs[TV][CIN + 1];
    __shared__ float ds[TV][COUT + 1];
    __shared__ float red[GROUPS > 1 ? GROUPS : 1][NBLK > 256 ? 1 : NBLK][17];
    const int t = blockIdx.y;
    const long long off = tap_off(g, t);
    const long long total = (long long)g.nx * g.ny * g.nz * g.nb;
    const long long u0 = blockIdx.x * per_block;
    const long long u1 = u0 + per_block < total ? u0 + per_block : total;
    const int tid = threadIdx.x;
    // thread -> (group, 4x4 block)
    const int grp = tid / (NBLK >= 256 ? 256 : NBLK);
    const int blk0 = tid % (NBLK >= 256 ? 256 : NBLK);
    constexpr int BPT = NBLK > 256 ? NBLK / 256 : 1;   // 4x4 blocks per thread
    float acc[BPT][16];
    float dbacc[BPT][4];
#pragma unroll
    for (int p = 0; p < BPT; p++) {
#pragma unroll
        for (int e = 0; e < 16; e++) acc[p][e] = 0.f;
#pragma unroll
        for (int e = 0; e < 4; e++) dbacc[p][e] = 0.f;
    }
    for (long long ub = u0; ub < u1; ub += TV) {
        __syncthreads();
        for (int e = tid; e < TV * (CIN + COUT); e += 256) {
            const int vi = e / (CIN + COUT), c = e % (CIN + COUT);
            const long long u = ub + vi;
            float val = 0.f;
            if (u < u1) {
                const long long v = vox_of(g, u);
                val = c < CIN ? (float)*chan(x, v + off, c) : (float)dy[v * cs_dy + (c - CIN)];
            }
            if (c < CIN)
                xs[vi][c] = val;
            else
                ds[vi][c - CIN] = val;
        }
        __syncthreads();
        if (grp < GROUPS) {
#pragma unroll
            for (int p = 0; p < BPT; p++) {
                const int blk = blk0 + p * 256;
                const int bi = blk / (COUT / 4), bo = blk % (COUT / 4);
                for (int vi = grp; vi < TV; vi += GROUPS) {
                    float xv[4], dv[4];
#pragma unroll
                    for (int a = 0; a < 4; a++) {
                        xv[a] = xs[vi][4 * bi + a];
                        dv[a] = ds[vi][4 * bo + a];
                    }
#pragma unroll
                    for (int a = 0; a < 4; a++)
#pragma unroll
                        for (int c = 0; c < 4; c++) acc[p][a * 4 + c] += xv[a] * dv[c];
                    if (bi == 0)
#pragma unroll
                        for (int c = 0; c < 4; c++) dbacc[p][c] += dv[c];
                }
            }
        }
    }
    // reduce over the thread groups, then one atomic per weight
    if (GROUPS > 1) {
        __syncthreads();
        if (grp < GROUPS) {
#pragma unroll
            for (int e = 0; e < 16; e++) red[grp][blk0][e] = acc[0][e];
        }
        __syncthreads();
        if (grp == 0) {
            for (int gq = 1; gq < GROUPS; gq++)
#pragma unroll
                for (int e = 0; e < 16; e++) acc[0][e] += red[gq][blk0][e];
        }
        __syncthreads();
        if (grp < GROUPS) {
#pragma unroll
            for (int e = 0; e < 4; e++) red[grp][blk0][e] = dbacc[0][e];
        }
        __syncthreads();
        if (grp == 0) {
            for (int gq = 1; gq < GROUPS; gq++)
#pragma unroll
                for (int e = 0; e < 4; e++) dbacc[0][e] += red[gq][blk0][e];
        }
    }
    if (grp != 0) return;
#pragma unroll
    for (int p = 0; p < BPT; p++) {
        const int blk = blk0 + p * 256;
        const int bi = blk / (COUT / 4), bo = blk % (COUT / 4);
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int c = 0; c < 4; c++)
                atomicAdd(&dw[((long long)t * CIN + 4 * bi + a) * COUT + 4 * bo + c], acc[p][a * 4 + c]);
        if (t == 13 && bi == 0 && db)
#pragma unroll
            for (int c = 0; c < 4; c++) atomicAdd(&db[4 * bo + c], dbacc[p][c]);
    }
}

// Element-wise kernels run over the x rows of a level: block b covers
// elements [256 (b % segs), 256 (b % segs + 1)) of row b / segs, an element
// being (x, channel group of 8) with the group fastest; row = y + ny (z + nz box).
__device__ __forceinline__ bool row_elem(const UGeo &g, int segs, int cg, int &x, int &c0, int &y, int &z,
                                         int &b)
{
    const int seg = (int)(blockIdx.x % (unsigned)segs), row = (int)(blockIdx.x / (unsigned)segs);
    const int t = seg * 256 + (int)threadIdx.x;
    if (t >= g.nx * cg) return false;
    x = t / cg;
    c0 = (t - x * cg) * 8;
    y = row % g.ny;
    z = (row / g.ny) % g.nz;
    b = row / (g.ny * g.nz);
    return true;
}

// 2x2x2 average pooling of C channels (fine level gf -> coarse level gc)
__global__ void k_pool(int segs, UGeo gc, UGeo gf, const bf16 *__restrict__ in, int cs_in, int C,
                       bf16 *__restrict__ out, int cs_out)
{
    int x, c0, y, z, b;
    if (!row_elem(gc, segs, C / 8, x, c0, y, z, b)) return;
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int d = 0; d < 8; d++) {
        const long long vf = vox(gf, 2 * x + (d & 1), 2 * y + ((d >> 1) & 1), 2 * z + (d >> 2), b);
        const bf16x8 a = *reinterpret_cast<const bf16x8 *>(in + vf * cs_in + c0);
#pragma unroll
        for (int i = 0; i < 8; i++) s[i] += (float)a[i];
    }
    bf16x8 o;
#pragma unroll
    for (int i = 0; i < 8; i++) o[i] = (bf16)(0.125f * s[i]);
    *reinterpret_cast<bf16x8 *>(out + vox(gc, x, y, z, b) * cs_out + c0) = o;
}

// backward of k_pool: the fine gradient (accumulated into) += coarse / 8
__global__ void k_pool_bwd(int segs, UGeo gc, UGeo gf, const bf16 *__restrict__ dc, int cs_c, int C,
                           bf16 *__restrict__ df, int cs_f)
{
    int x, c0, y, z, b;
    if (!row_elem(gf, segs, C / 8, x, c0, y, z, b)) return;
    const bf16x8 a = *reinterpret_cast<const bf16x8 *>(dc + vox(gc, x >> 1, y >> 1, z >> 1, b) * cs_c + c0);
    bf16x8 *p = reinterpret_cast<bf16x8 *>(df + vox(gf, x, y, z, b) * cs_f + c0);
    bf16x8 o = *p;
#pragma unroll
    for (int i = 0; i < 8; i++) o[i] = (bf16)((float)o[i] + 0.125f * (float)a[i]);
    *p = o;
}

// backward of the 2x nearest-neighbour upsampling (stored by the coarse
// layer's convolution): coarse gradient = sum of its 8 fine voxels' gradients
__global__ void k_up_bwd(int segs, UGeo gc, UGeo gf, const bf16 *__restrict__ df, int cs_f, int C,
                         bf16 *__restrict__ dc, int cs_c)
{
    int x, c0, y, z, b;
    if (!row_elem(gc, segs, C / 8, x, c0, y, z, b)) return;
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int d = 0; d < 8; d++) {
        const long long vf = vox(gf, 2 * x + (d & 1), 2 * y + ((d >> 1) & 1), 2 * z + (d >> 2), b);
        const bf16x8 a = *reinterpret_cast<const bf16x8 *>(df + vf * cs_f + c0);
#pragma unroll
        for (int i = 0; i < 8; i++) s[i] += (float)a[i];
    }
    bf16x8 o;
#pragma unroll
    for (int i = 0; i < 8; i++) o[i] = (bf16)s[i];
    *reinterpret_cast<bf16x8 *>(dc + vox(gc, x, y, z, b) * cs_c + c0) = o;
}

// dA *= (A > 0) over C channels: the gradient through a ReLU output A
__global__ void k_relu_bwd(int segs, UGeo g, const bf16 *__restrict__ a, int cs_a, int C, bf16 *__restrict__ d,
                           int cs_d)
{
    int x, c0, y, z, b;
    if (!row_elem(g, segs, C / 8, x, c0, y, z, b)) return;
    const long long v = vox(g, x, y, z, b);
    const bf16x8 av = *reinterpret_cast<const bf16x8 *>(a + v * cs_a + c0);
    bf16x8 *p = reinterpret_cast<bf16x8 *>(d + v * cs_d + c0);
    bf16x8 o = *p;
#pragma unroll
    for (int i = 0; i < 8; i++)
        if (!((float)av[i] > 0.f)) o[i] = (bf16)0.f;
    *p = o;
}

// backward of the output layer y = b + sum_c w[c] h[c] (applied by the last
// convolution): dh[v][c] = w[c] dy[u]; dw[c] += sum dy h[c]; db += sum dy
__global__ void __launch_bounds__(256) k_out_bwd(UGeo g, const bf16 *__restrict__ h,
                                                 const float *__restrict__ w,
                                                 const float *__restrict__ dy, bf16 *__restrict__ dh,
                                                 float *__restrict__ dwb)
{
    __shared__ float part[256][17];
    float acc[17];
#pragma unroll
    for (int c = 0; c < 17; c++) acc[c] = 0.f;
    const long long total = (long long)g.nx * g.ny * g.nz * g.nb;
    for (long long u = blockIdx.x * (long long)blockDim.x + threadIdx.x; u < total;
         u += (long long)gridDim.x * blockDim.x) {
        const long long v = vox_of(g, u);
        const float d = dy[u];
        bf16x8 o0, o1;
        const bf16x8 a0 = *reinterpret_cast<const bf16x8 *>(h + v * 16);
        const bf16x8 a1 = *reinterpret_cast<const bf16x8 *>(h + v * 16 + 8);
#pragma unroll
        for (int i = 0; i < 8; i++) {
            o0[i] = (bf16)(w[i] * d);
            o1[i] = (bf16)(w[8 + i] * d);
            acc[i] += d * (float)a0[i];
            acc[8 + i] += d * (float)a1[i];
        }
        acc[16] += d;
        *reinterpret_cast<bf16x8 *>(dh + v * 16) = o0;
        *reinterpret_cast<bf16x8 *>(dh + v * 16 + 8) = o1;
    }
#pragma unroll
    for (int c = 0; c < 17; c++) part[threadIdx.x][c] = acc[c];
    __syncthreads();
    if (threadIdx.x < 17) {
        float s = 0.f;
        for (int t = 0; t < 256; t++) s += part[t][threadIdx.x];
        atomicAdd(&dwb[threadIdx.x], s);
    }
}

// masked relative squared error per box: loss_b = sum m (y - e)^2 / sum m e^2;
// dy = 2 m (y - e) / (sum m e^2 * nb); loss summed into *loss (sum over boxes / nb)
__global__ void __launch_bounds__(256) k_loss(UGeo g, const float *__restrict__ y,
                                              const float *__restrict__ e, const float *__restrict__ m,
                                              const float *__restrict__ esq, float *__restrict__ dy,
                                              float *__restrict__ loss)
{
    __shared__ float part[256];
    float acc = 0.f;
    const long long per = (long long)g.nx * g.ny * g.nz;
    const long long total = per * g.nb;
    for (long long u = blockIdx.x * (long long)blockDim.x + threadIdx.x; u < total;
         u += (long long)gridDim.x * blockDim.x) {
        const int b = (int)(u / per);
        const float s = esq[b] > 0.f ? 1.f / (esq[b] * (float)g.nb) : 0.f;
        const float r = m[u] * (y[u] - e[u]);
        dy[u] = 2.f * r * s;
        acc += r * r * s;
    }
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, part[0]);
}

// per box: sum m e^2
__global__ void __launch_bounds__(256) k_esq(UGeo g, const float *__restrict__ e,
                                             const float *__restrict__ m, float *__restrict__ esq)
{
    __shared__ float part[256];
    const long long per = (long long)g.nx * g.ny * g.nz;
    const int b = blockIdx.y;
    float acc = 0.f;
    for (long long u = blockIdx.x * (long long)blockDim.x + threadIdx.x; u < per;
         u += (long long)gridDim.x * blockDim.x) {
        const float v = e[b * per + u];
        acc += m[b * per + u] * v * v;
    }
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(&esq[b], part[0]);
}

__global__ void k_adam(long long n, float *__restrict__ p, const float *__restrict__ gr,
                       float *__restrict__ m1, float *__restrict__ m2, float lr, float b1, float b2,
                       float c1, float c2, float eps)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const float g = gr[i];
        const float a = b1 * m1[i] + (1.f - b1) * g;
        const float b = b2 * m2[i] + (1.f - b2) * g * g;
        m1[i] = a;
        m2[i] = b;
        p[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
    }
}

// zeroes plane z (interior numbering) of box 0, all cs channels of each voxel
__global__ void k_zero_plane(UGeo g, int z, bf16 *__restrict__ a, int cs)
{
    const long long total = (long long)g.px * g.py * cs;
    const long long base = ((long long)(z + 1) * g.py) * g.px * cs;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x)
        a[base + e] = (bf16)0.f;
}

__global__ void k_zero_f(long long n, float *p)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = 0.f;
}

// ---------------------------------------------------------------- layers
// The network's 3x3x3 layers in order: (level, cin, cout).
struct LayerDef {
    int level, cin, cout;
};
constexpr LayerDef LAYERS[] = {
    {0, 8, 16},  {0, 16, 16},  // e1a, e1b
    {1, 16, 32}, {1, 32, 32},  // e2a, e2b
    {2, 32, 64}, {2, 64, 64},  // b3a, b3b
    {2, 64, 32},               // u3 (coarse conv before upsampling)
    {1, 64, 32}, {1, 32, 32},  // d2a, d2b
    {1, 32, 16},               // u2
    {0, 32, 16}, {0, 16, 16},  // d1a, d1b
};
constexpr int NL = sizeof(LAYERS) / sizeof(LAYERS[0]);
enum { E1A, E1B, E2A, E2B, B3A, B3B, U3, D2A, D2B, U2, D1A, D1B };

unsigned grid_for(long long n) { return (unsigned)std::min<long long>((n + 255) / 256, 8192); }

}  // namespace

struct cad_unet {
    int device;
    int ncu;                        // compute units of the device
    hipStream_t stream;
    bool own_stream;
    UGeo g[3];
    long long wofs[NL], bofs[NL], out_ofs;
    long long nparam;
    float *p, *gr, *m1, *m2;        // parameters, gradients, Adam moments (FP32)
    bf16 *wp[NL], *wt[NL];          // packed forward and backward-data weights
    long long wp_n[NL], wt_n[NL];
    // activations (channels-last BF16 with halos)
    // (cat1a, cat1b) and (cat2a, cat2b): the decoders' concatenated inputs,
    // upsampled coarse features and the encoder's skip output
    bf16 *x, *h1a, *cat1a, *cat1b, *p1, *h2a, *cat2a, *cat2b, *p2, *h3a, *h3b, *h3c, *h2d, *h2e, *h2f, *h1d, *h1e;
    // gradients
    bf16 *dh1e, *dh1d, *dcat1a, *dcat1b, *dh2f, *dh2e, *dh2d, *dcat2a, *dcat2b, *dh3c, *dh3b, *dh3a, *dp2, *dh2a,
        *dp1, *dh1a;
    float *y, *dy, *esq, *loss;
    bool train;
    long long step;
    std::vector<void *> allocs;
};

namespace {

template <typename T>
int ualloc(cad_unet *u, T **p, size_t n)
{
    void *q = nullptr;
    if (hipMalloc(&q, n * sizeof(T) + 256) != hipSuccess) {
        cad_set_error("cad_unet: hipMalloc of %zu bytes failed", n * sizeof(T));
        return CAD_ERR_HIP;
    }
    if (hipMemsetAsync(q, 0, n * sizeof(T) + 256, u->stream) != hipSuccess) {
        cad_set_error("cad_unet: hipMemset failed");
        return CAD_ERR_HIP;
    }
    u->allocs.push_back(q);
    *p = reinterpret_cast<T *>(q);
    return CAD_OK;
}

long long padded(const UGeo &g) { return (long long)g.px * g.py * g.pz * g.nb; }

ConvIn one_in(const bf16 *p, int cs, int c) { return ConvIn{p, 0, cs, c}; }
// a concatenation: channels [0, split) at p, the rest at p + delta, records of cs
ConvIn cat_in(const bf16 *p, long long delta, int cs, int split) { return ConvIn{p, delta, cs, split}; }
ConvOut one_out(bf16 *p, int cs, int c) { return ConvOut{p, nullptr, cs, 0, c}; }

// waves a launch aims at: four per SIMD
int conv_target_waves(const cad_unet *u) { return 16 * u->ncu; }

template <int CIN, int MT, int NT>
int conv_launch(cad_unet *u, const UGeo &g, const ConvIn &in, const bf16 *wp, const float *bias,
                const ConvOut &out, int m0, bool relu)
{
    constexpr int WX = conv_wx(CIN, MT);
    const int ng = g.nx / 16;
    const int ntx = (ng + NT * WX - 1) / (NT * WX), nty = (g.ny + 3) / 4;
    const long long plane_tiles = (long long)ntx * nty * g.nb;
    // z runs: the fewest that give the target wave count (each run reads two
    // planes beyond its own)
    const long long per_plane = plane_tiles * 4 * WX;
    const long long want = (conv_target_waves(u) + per_plane - 1) / per_plane;
    const int nzc0 = (int)std::min<long long>(std::max<long long>(want, 1), g.nz);
    const int zlen = (g.nz + nzc0 - 1) / nzc0;
    const int nzc = (g.nz + zlen - 1) / zlen;
    const long long ntiles = plane_tiles * nzc;
    // the kernel's load offsets (bytes from a wave's base voxel: lane part,
    // plane part and group part) add up in 32 bits
    const long long plane = 2LL * g.px * g.py * in.cs;
    if (ntiles > (1LL << 30) ||
        2 * (in.delta + (long long)(3 * g.px + 64) * in.cs) + (zlen + 2) * plane >= (1LL << 32) ||
        plane >= (1LL << 31)) {
        cad_set_error("cad_unet: box too large for the convolution's tiling and 32-bit load offsets");
        return CAD_ERR_ARG;
    }
    const int per_xcd = (int)((ntiles + 7) / 8);
    const dim3 grid((unsigned)(8 * per_xcd)), blk(256 * WX);
    const bf16x8 *w = reinterpret_cast<const bf16x8 *>(wp);
    // the fused epilogues exist where the network has them: the upsampled
    // layers' (64 and 32 input channels) and the last layer's (16)
    const int epi = out.up ? 1 : out.y ? 2 : 0;
#define CAD_CONV(R, E)                                                                                       \
    do {                                                                                                     \
        if constexpr (conv_staged(CIN, MT, NT))                                                             \
            hipLaunchKernelGGL((k_conv3s<CIN, MT, NT, R, E>), grid, blk, 0, u->stream, g, in, w, bias, out, m0, \
                               ntx, nty, nzc, zlen, per_xcd, (int)ntiles);                                  \
        else                                                                                                 \
            hipLaunchKernelGGL((k_conv3<CIN, MT, NT, R, E>), grid, blk, 0, u->stream, g, in, w, bias, out, m0,  \
                               ntx, nty, nzc, zlen, per_xcd, (int)ntiles);                                  \
    } while (0)
    if (epi == 0) {
        if (relu)
            CAD_CONV(true, 0);
        else
            CAD_CONV(false, 0);
    } else if (!relu || (epi == 1 && CIN != 32 && CIN != 64) || (epi == 2 && CIN != 16)) {
        cad_set_error("cad_unet: no fused output for a layer of %d input channels", CIN);
        return CAD_ERR_ARG;
    } else if (epi == 1) {
        if constexpr (CIN == 32 || CIN == 64) CAD_CONV(true, 1);
    } else {
        if constexpr (CIN == 16) CAD_CONV(true, 2);
    }
#undef CAD_CONV
    CAD_CHECK(hipGetLastError());
    return CAD_OK;
}

// One launch per share of mt output-channel tiles (mt = 2 where the output
// has more than one tile, 1 otherwise), each with the widest x groups (nt)
// that still give the target wave count at one plane per z run.
template <int CIN>
int conv_ci(cad_unet *u, const UGeo &g, const ConvIn &in, const bf16 *wp, const float *bias, const ConvOut &out,
            int cout, bool relu)
{
    const int mtot = cout / 16;
    if (mtot < 1 || cout % 16 || g.nx % 16) {
        cad_set_error("cad_unet: no convolution tiling for %d voxels along x and %d channels", g.nx, cout);
        return CAD_ERR_ARG;
    }
    if (out.y && cout != 16) {
        cad_set_error("cad_unet: the output layer follows a 16-channel layer, not %d", cout);
        return CAD_ERR_ARG;
    }
    const int ng = g.nx / 16;
    auto waves = [&](int nt, int wx) {
        return (long long)((ng + nt * wx - 1) / (nt * wx)) * wx * 4 * ((g.ny + 3) / 4) * g.nb * g.nz;
    };
    const int mt = mtot % 2 == 0 ? 2 : 1;
    const int wx = conv_wx(CIN, mt);
    int nt = mt == 2 || wx == 4 ? 2 : 4;   // 12 accumulator quads at most; 8 in blocks of 16 waves
    while (nt > 1 && waves(nt, wx) < conv_target_waves(u)) nt /= 2;
    // a shape whose input tile and weights fit in LDS is preferred with one x
    // group (measured: faster for 32 -> 32 channels at half resolution; one
    // channel tile per launch for 64 input channels is not)
    if (!conv_staged(CIN, mt, nt) && conv_staged(CIN, mt, 1)) nt = 1;
    const long long tile = 3LL * fold_kc(CIN) * 64 * 8;   // BF16 per channel tile
    for (int m0 = 0; m0 < mtot; m0 += mt) {
        const bf16 *w = wp + m0 * tile;
        int rc;
        if (mt == 2)
            rc = nt == 2 ? conv_launch<CIN, 2, 2>(u, g, in, w, bias, out, m0, relu)
                         : conv_launch<CIN, 2, 1>(u, g, in, w, bias, out, m0, relu);
        else if constexpr (conv_wx(CIN, 1) == 1)
            rc = nt == 4   ? conv_launch<CIN, 1, 4>(u, g, in, w, bias, out, m0, relu)
                 : nt == 2 ? conv_launch<CIN, 1, 2>(u, g, in, w, bias, out, m0, relu)
                           : conv_launch<CIN, 1, 1>(u, g, in, w, bias, out, m0, relu);
        else
            rc = nt == 2 ? conv_launch<CIN, 1, 2>(u, g, in, w, bias, out, m0, relu)
                         : conv_launch<CIN, 1, 1>(u, g, in, w, bias, out, m0, relu);
        if (rc) return rc;
    }
    return CAD_OK;
}

int conv(cad_unet *u, const UGeo &g, int cin, const ConvIn &in, const bf16 *wp, const float *bias,
         const ConvOut &out, int cout, bool relu)
{
    switch (cin) {
    case 8: return conv_ci<8>(u, g, in, wp, bias, out, cout, relu);
    case 16: return conv_ci<16>(u, g, in, wp, bias, out, cout, relu);
    case 32: return conv_ci<32>(u, g, in, wp, bias, out, cout, relu);
    case 64: return conv_ci<64>(u, g, in, wp, bias, out, cout, relu);
    }
    cad_set_error("cad_unet: no convolution for %d input channels", cin);
    return CAD_ERR_ARG;
}

int layer_fwd(cad_unet *u, int l, const ConvIn &in, const ConvOut &out)
{
    const LayerDef &L = LAYERS[l];
    return conv(u, u->g[L.level], L.cin, in, u->wp[l], u->p + u->bofs[l], out, L.cout, true);
}
int layer_fwd(cad_unet *u, int l, const ConvIn &in, bf16 *out, int cs_out)
{
    return layer_fwd(u, l, in, one_out(out, cs_out, LAYERS[l].cout));
}
// an output of c channels stored at p (null: not stored) and upsampled into
// the buffer up of the level below (records of cs_up channels)
ConvOut up_out(const cad_unet *u, int level, bf16 *p, int cs, int c, bf16 *up, int cs_up)
{
    ConvOut o = one_out(p, cs, c);
    const UGeo &gf = u->g[level - 1];
    o.up = up;
    o.ucs = cs_up;
    o.upx = gf.px;
    o.upy = gf.py;
    o.upz = gf.pz;
    return o;
}

// gradient of the layer's input (from dY at its output) and of its weights
int layer_bwd(cad_unet *u, int l, const ConvIn &in, const bf16 *dy, int cs_dy, const ConvOut *dx)
{
    const LayerDef &L = LAYERS[l];
    const UGeo &g = u->g[L.level];
    if (dx) {
        int rc = conv(u, g, L.cout, one_in(dy, cs_dy, L.cout), u->wt[l], nullptr, *dx, L.cin, false);
        if (rc) return rc;
    }
    const long long total = (long long)g.nx * g.ny * g.nz * g.nb;
    const long long per = 4096;
    const dim3 grid((unsigned)((total + per - 1) / per), 27), blk(256);
    float *dw = u->gr + u->wofs[l], *db = u->gr + u->bofs[l];
#define CAD_DW(CI, CO)                                                                                \
    if (L.cin == CI && L.cout == CO)                                                                  \
        hipLaunchKernelGGL((k_conv_dw<CI, CO>), grid, blk, 0, u->stream, g, in, dy, cs_dy, per, dw, db);
    CAD_DW(8, 16) CAD_DW(16, 16) CAD_DW(16, 32) CAD_DW(32, 32) CAD_DW(32, 64) CAD_DW(64, 64)
    CAD_DW(64, 32) CAD_DW(32, 16)
#undef CAD_DW
    CAD_CHECK(hipGetLastError());
    return CAD_OK;
}

int pack_all(cad_unet *u)
{
    for (int l = 0; l < NL; l++) {
        const LayerDef &L = LAYERS[l];
        hipLaunchKernelGGL(k_pack, dim3(grid_for(u->wp_n[l])), dim3(256), 0, u->stream, u->p + u->wofs[l],
                           L.cin, L.cout, 0, u->wp[l]);
        if (u->train && l != E1A)
            hipLaunchKernelGGL(k_pack, dim3(grid_for(u->wt_n[l])), dim3(256), 0, u->stream,
                               u->p + u->wofs[l], L.cin, L.cout, 1, u->wt[l]);
    }
    CAD_CHECK(hipGetLastError());
    return CAD_OK;
}

template <typename K, typename... A>
int ew(cad_unet *u, K kern, long long n, A... args)
{
    hipLaunchKernelGGL(kern, dim3(grid_for(n)), dim3(256), 0, u->stream, args...);
    CAD_CHECK(hipGetLastError());
    return CAD_OK;
}

// an element-wise kernel over the rows of gi, cg elements per voxel
template <typename K, typename... A>
int ewr(cad_unet *u, K kern, const UGeo &gi, int cg, A... args)
{
    const int segs = (gi.nx * cg + 255) / 256;
    const long long nblk = (long long)gi.ny * gi.nz * gi.nb * segs;
    hipLaunchKernelGGL(kern, dim3((unsigned)nblk), dim3(256), 0, u->stream, segs, args...);
    CAD_CHECK(hipGetLastError());
    return CAD_OK;
}

long long nvox(const UGeo &g) { return (long long)g.nx * g.ny * g.nz * g.nb; }

}  // namespace

#define RET_IF(x)              \
    do {                       \
        int rc_ = (x);         \
        if (rc_) return rc_;   \
    } while (0)

int cad_unet_forward_impl(cad_unet *u)
{
    const UGeo *g = u->g;
    RET_IF(layer_fwd(u, E1A, one_in(u->x, 8, 8), u->h1a, 16));
    RET_IF(layer_fwd(u, E1B, one_in(u->h1a, 16, 16), u->cat1b, 16));
    RET_IF(ewr(u, k_pool, g[1], 2, g[1], g[0], u->cat1b, 16, 16, u->p1, 16));
    RET_IF(layer_fwd(u, E2A, one_in(u->p1, 16, 16), u->h2a, 32));
    RET_IF(layer_fwd(u, E2B, one_in(u->h2a, 32, 32), u->cat2b, 32));
    RET_IF(ewr(u, k_pool, g[2], 4, g[2], g[1], u->cat2b, 32, 32, u->p2, 32));
    RET_IF(layer_fwd(u, B3A, one_in(u->p2, 32, 32), u->h3a, 64));
    RET_IF(layer_fwd(u, B3B, one_in(u->h3a, 64, 64), u->h3b, 64));
    // u3 and u2 store their outputs upsampled into the decoders' inputs, and
    // d1b applies the output layer; the coarse outputs and h1e themselves
    // are kept for the backward pass only
    RET_IF(layer_fwd(u, U3, one_in(u->h3b, 64, 64), up_out(u, 2, u->train ? u->h3c : nullptr, 32, 32, u->cat2a, 32)));
    RET_IF(layer_fwd(u, D2A, cat_in(u->cat2a, u->cat2b - u->cat2a, 32, 32), u->h2d, 32));
    RET_IF(layer_fwd(u, D2B, one_in(u->h2d, 32, 32), u->h2e, 32));
    RET_IF(layer_fwd(u, U2, one_in(u->h2e, 32, 32), up_out(u, 1, u->train ? u->h2f : nullptr, 16, 16, u->cat1a, 16)));
    RET_IF(layer_fwd(u, D1A, cat_in(u->cat1a, u->cat1b - u->cat1a, 16, 16), u->h1d, 16));
    ConvOut o1 = one_out(u->train ? u->h1e : nullptr, 16, 16);
    o1.ow = u->p + u->out_ofs;
    o1.y = u->y;
    RET_IF(layer_fwd(u, D1B, one_in(u->h1d, 16, 16), o1));
    return CAD_OK;
}

static int backward(cad_unet *u)
{
    const UGeo *g = u->g;
    float *gout = u->gr + u->out_ofs;
    hipLaunchKernelGGL(k_out_bwd, dim3(512), dim3(256), 0, u->stream, g[0], u->h1e, u->p + u->out_ofs, u->dy,
                       u->dh1e, gout);
    CAD_CHECK(hipGetLastError());
    const ConvOut dh1d = one_out(u->dh1d, 16, 16), dcat1{u->dcat1a, u->dcat1b, 16, 16, 16};
    const ConvOut dh2e = one_out(u->dh2e, 32, 32), dh2d = one_out(u->dh2d, 32, 32);
    const ConvOut dcat2{u->dcat2a, u->dcat2b, 32, 32, 32}, dh3b = one_out(u->dh3b, 64, 64);
    const ConvOut dh3a = one_out(u->dh3a, 64, 64), dp2 = one_out(u->dp2, 32, 32);
    const ConvOut dh2a = one_out(u->dh2a, 32, 32), dp1 = one_out(u->dp1, 16, 16);
    const ConvOut dh1a = one_out(u->dh1a, 16, 16);
    RET_IF(ewr(u, k_relu_bwd, g[0], 2, g[0], u->h1e, 16, 16, u->dh1e, 16));
    RET_IF(layer_bwd(u, D1B, one_in(u->h1d, 16, 16), u->dh1e, 16, &dh1d));
    RET_IF(ewr(u, k_relu_bwd, g[0], 2, g[0], u->h1d, 16, 16, u->dh1d, 16));
    RET_IF(layer_bwd(u, D1A, cat_in(u->cat1a, u->cat1b - u->cat1a, 16, 16), u->dh1d, 16, &dcat1));
    // the skip's first half: the upsampled h2f
    RET_IF(ewr(u, k_up_bwd, g[1], 2, g[1], g[0], u->dcat1a, 16, 16, u->dh2f, 16));
    RET_IF(ewr(u, k_relu_bwd, g[1], 2, g[1], u->h2f, 16, 16, u->dh2f, 16));
    RET_IF(layer_bwd(u, U2, one_in(u->h2e, 32, 32), u->dh2f, 16, &dh2e));
    RET_IF(ewr(u, k_relu_bwd, g[1], 4, g[1], u->h2e, 32, 32, u->dh2e, 32));
    RET_IF(layer_bwd(u, D2B, one_in(u->h2d, 32, 32), u->dh2e, 32, &dh2d));
    RET_IF(ewr(u, k_relu_bwd, g[1], 4, g[1], u->h2d, 32, 32, u->dh2d, 32));
    RET_IF(layer_bwd(u, D2A, cat_in(u->cat2a, u->cat2b - u->cat2a, 32, 32), u->dh2d, 32, &dcat2));
    RET_IF(ewr(u, k_up_bwd, g[2], 4, g[2], g[1], u->dcat2a, 32, 32, u->dh3c, 32));
    RET_IF(ewr(u, k_relu_bwd, g[2], 4, g[2], u->h3c, 32, 32, u->dh3c, 32));
    RET_IF(layer_bwd(u, U3, one_in(u->h3b, 64, 64), u->dh3c, 32, &dh3b));
    RET_IF(ewr(u, k_relu_bwd, g[2], 8, g[2], u->h3b, 64, 64, u->dh3b, 64));
    RET_IF(layer_bwd(u, B3B, one_in(u->h3a, 64, 64), u->dh3b, 64, &dh3a));
    RET_IF(ewr(u, k_relu_bwd, g[2], 8, g[2], u->h3a, 64, 64, u->dh3a, 64));
    RET_IF(layer_bwd(u, B3A, one_in(u->p2, 32, 32), u->dh3a, 64, &dp2));
    // the skip's second half (e2b's output): decoder part + pooling part
    RET_IF(ewr(u, k_pool_bwd, g[1], 4, g[2], g[1], u->dp2, 32, 32, u->dcat2b, 32));
    RET_IF(ewr(u, k_relu_bwd, g[1], 4, g[1], u->cat2b, 32, 32, u->dcat2b, 32));
    RET_IF(layer_bwd(u, E2B, one_in(u->h2a, 32, 32), u->dcat2b, 32, &dh2a));
    RET_IF(ewr(u, k_relu_bwd, g[1], 4, g[1], u->h2a, 32, 32, u->dh2a, 32));
    RET_IF(layer_bwd(u, E2A, one_in(u->p1, 16, 16), u->dh2a, 32, &dp1));
    RET_IF(ewr(u, k_pool_bwd, g[0], 2, g[1], g[0], u->dp1, 16, 16, u->dcat1b, 16));
    RET_IF(ewr(u, k_relu_bwd, g[0], 2, g[0], u->cat1b, 16, 16, u->dcat1b, 16));
    RET_IF(layer_bwd(u, E1B, one_in(u->h1a, 16, 16), u->dcat1b, 16, &dh1a));
    RET_IF(ewr(u, k_relu_bwd, g[0], 2, g[0], u->h1a, 16, 16, u->dh1a, 16));
    RET_IF(layer_bwd(u, E1A, one_in(u->x, 8, 8), u->dh1a, 16, nullptr));
    return CAD_OK;
}

extern "C" {

// multiply-adds per level-1 voxel of one forward pass
double cad_unet_mac_per_voxel(void)
{
    double m = 16.0;
    for (int l = 0; l < NL; l++) m += 27.0 * LAYERS[l].cin * LAYERS[l].cout / (double)(1 << (3 * LAYERS[l].level));
    return m;
}

int cad_unet_param_count(void)
{
    long long n = 17;
    for (int l = 0; l < NL; l++) n += 27LL * LAYERS[l].cin * LAYERS[l].cout + LAYERS[l].cout;
    return (int)n;
}

int cad_unet_create(int nx, int ny, int nz, int nb, int train, void *stream, cad_unet **out)
{
    if (!out || nx < 64 || nx % 64 || nx > 1024 || ny < 4 || ny % 4 || nz < 4 || nz % 4 || nb < 1) {
        cad_set_error("cad_unet_create: box %d x %d x %d (x a multiple of 64 up to 1024, y and z of 4), "
                      "batch %d",
                      nx, ny, nz, nb);
        return CAD_ERR_ARG;
    }
    *out = nullptr;
    cad_unet *u = new cad_unet();
    CAD_CHECK(hipGetDevice(&u->device));
    CAD_CHECK(hipDeviceGetAttribute(&u->ncu, hipDeviceAttributeMultiprocessorCount, u->device));
    if (stream) {
        u->stream = reinterpret_cast<hipStream_t>(stream);
    } else {
        if (hipStreamCreateWithFlags(&u->stream, hipStreamNonBlocking) != hipSuccess) {
            delete u;
            cad_set_error("cad_unet_create: hipStreamCreate failed");
            return CAD_ERR_HIP;
        }
        u->own_stream = true;
    }
    u->train = train != 0;
    for (int l = 0; l < 3; l++) {
        UGeo &g = u->g[l];
        g.nx = nx >> l;
        g.ny = ny >> l;
        g.nz = nz >> l;
        g.nb = nb;
        g.px = g.nx + 2;
        g.py = g.ny + 2;
        g.pz = g.nz + 2;
    }
    long long o = 0;
    for (int l = 0; l < NL; l++) {
        u->wofs[l] = o;
        o += 27LL * LAYERS[l].cin * LAYERS[l].cout;
        u->bofs[l] = o;
        o += LAYERS[l].cout;
    }
    u->out_ofs = o;
    o += 17;
    u->nparam = o;
    int rc = 0;
    auto A = [&](auto **p, size_t n) {
        if (!rc) rc = ualloc(u, p, n);
    };
    A(&u->p, u->nparam);
    for (int l = 0; l < NL; l++) {
        const LayerDef &L = LAYERS[l];
        u->wp_n[l] = 3LL * (L.cout / 16) * fold_kc(L.cin) * 64 * 8;
        A(&u->wp[l], u->wp_n[l]);
        if (u->train && l != E1A) {
            u->wt_n[l] = 3LL * (L.cin / 16) * fold_kc(L.cout) * 64 * 8;
            A(&u->wt[l], u->wt_n[l]);
        }
    }
    const long long v0 = padded(u->g[0]), v1 = padded(u->g[1]), v2 = padded(u->g[2]);
    A(&u->x, v0 * 8);
    A(&u->h1a, v0 * 16);
    A(&u->cat1a, v0 * 32);
    if (!rc) u->cat1b = u->cat1a + v0 * 16;
    A(&u->p1, v1 * 16);
    A(&u->h2a, v1 * 32);
    A(&u->cat2a, v1 * 64);
    if (!rc) u->cat2b = u->cat2a + v1 * 32;
    A(&u->p2, v2 * 32);
    A(&u->h3a, v2 * 64);
    A(&u->h3b, v2 * 64);
    A(&u->h3c, v2 * 32);
    A(&u->h2d, v1 * 32);
    A(&u->h2e, v1 * 32);
    A(&u->h2f, v1 * 16);
    A(&u->h1d, v0 * 16);
    A(&u->h1e, v0 * 16);
    A(&u->y, nvox(u->g[0]));
    if (u->train) {
        A(&u->gr, u->nparam);
        A(&u->m1, u->nparam);
        A(&u->m2, u->nparam);
        A(&u->dh1e, v0 * 16);
        A(&u->dh1d, v0 * 16);
        A(&u->dcat1a, v0 * 16);
        A(&u->dcat1b, v0 * 16);
        A(&u->dh2f, v1 * 16);
        A(&u->dh2e, v1 * 32);
        A(&u->dh2d, v1 * 32);
        A(&u->dcat2a, v1 * 32);
        A(&u->dcat2b, v1 * 32);
        A(&u->dh3c, v2 * 32);
        A(&u->dh3b, v2 * 64);
        A(&u->dh3a, v2 * 64);
        A(&u->dp2, v2 * 32);
        A(&u->dh2a, v1 * 32);
        A(&u->dp1, v1 * 16);
        A(&u->dh1a, v0 * 16);
        A(&u->dy, nvox(u->g[0]));
        A(&u->esq, (size_t)nb);
        A(&u->loss, 1);
    }
    if (!rc && hipStreamSynchronize(u->stream) != hipSuccess) rc = CAD_ERR_HIP;
    if (rc) {
        cad_unet_destroy(u);
        return rc;
    }
    *out = u;
    return CAD_OK;
}

void cad_unet_destroy(cad_unet *u)
{
    if (!u) return;
    (void)hipStreamSynchronize(u->stream);
    for (void *q : u->allocs) (void)hipFree(q);
    if (u->own_stream) (void)hipStreamDestroy(u->stream);
    delete u;
}

int cad_unet_set_params(cad_unet *u, const float *host)
{
    if (!u || !host) {
        cad_set_error("cad_unet_set_params: null argument");
        return CAD_ERR_ARG;
    }
    CAD_CHECK(hipMemcpyAsync(u->p, host, u->nparam * sizeof(float), hipMemcpyHostToDevice, u->stream));
    if (u->train) {
        RET_IF(ew(u, k_zero_f, u->nparam, u->nparam, u->m1));
        RET_IF(ew(u, k_zero_f, u->nparam, u->nparam, u->m2));
        u->step = 0;
    }
    RET_IF(pack_all(u));
    CAD_CHECK(hipStreamSynchronize(u->stream));
    return CAD_OK;
}

int cad_unet_get_params(cad_unet *u, float *host)
{
    if (!u || !host) {
        cad_set_error("cad_unet_get_params: null argument");
        return CAD_ERR_ARG;
    }
    CAD_CHECK(hipMemcpyAsync(host, u->p, u->nparam * sizeof(float), hipMemcpyDeviceToHost, u->stream));
    CAD_CHECK(hipStreamSynchronize(u->stream));
    return CAD_OK;
}

int cad_unet_init_params(int seed, float *host)
{
    if (!host) {
        cad_set_error("cad_unet_init_params: null argument");
        return CAD_ERR_ARG;
    }
    std::mt19937_64 rng((unsigned long long)seed);
    long long o = 0;
    for (int l = 0; l < NL; l++) {
        const LayerDef &L = LAYERS[l];
        std::normal_distribution<float> nd(0.f, std::sqrt(2.f / (27.f * L.cin)));
        for (long long i = 0; i < 27LL * L.cin * L.cout; i++) host[o++] = nd(rng);
        for (int i = 0; i < L.cout; i++) host[o++] = 0.f;
    }
    std::normal_distribution<float> nd(0.f, std::sqrt(1.f / 16.f));
    for (int i = 0; i < 16; i++) host[o++] = nd(rng);
    host[o++] = 0.f;
    return CAD_OK;
}

// Runs the network (one box) on its first nz planes from now on: nz a
// multiple of 4, at most the box's depth; the plane above the run's last one
// is cleared in every buffer a convolution reads, so it acts as the halo.
int cad_unet_set_depth(cad_unet *u, int nz)
{
    if (!u || u->g[0].nb != 1 || nz < 4 || nz % 4 || nz > u->g[0].pz - 2) {
        cad_set_error("cad_unet_set_depth: depth %d (a multiple of 4 within the box, one box)", nz);
        return CAD_ERR_ARG;
    }
    for (int l = 0; l < 3; l++) u->g[l].nz = nz >> l;
    struct B {
        bf16 *a;
        int lvl, cs;
    } bufs[] = {{u->x, 0, 8},     {u->h1a, 0, 16},  {u->cat1a, 0, 16}, {u->cat1b, 0, 16}, {u->h1d, 0, 16},
                {u->h1e, 0, 16},  {u->p1, 1, 16},   {u->h2a, 1, 32},  {u->cat2a, 1, 32}, {u->cat2b, 1, 32},
                {u->h2d, 1, 32},  {u->h2e, 1, 32},  {u->h2f, 1, 16},  {u->p2, 2, 32},
                {u->h3a, 2, 64},  {u->h3b, 2, 64},  {u->h3c, 2, 32}};
    for (const B &b : bufs) {
        const UGeo &g = u->g[b.lvl];
        if (g.nz + 1 > g.pz - 2) continue;   // the box's own halo plane
        hipLaunchKernelGGL(k_zero_plane, dim3(64), dim3(256), 0, u->stream, g, g.nz, b.a, b.cs);
    }
    CAD_CHECK(hipGetLastError());
    return CAD_OK;
}

hipStream_t cad_unet_stream(cad_unet *u) { return u->stream; }
UGeo cad_unet_geo(cad_unet *u) { return u->g[0]; }

void *cad_unet_input(cad_unet *u) { return u ? (void *)u->x : nullptr; }
float *cad_unet_output(cad_unet *u) { return u ? u->y : nullptr; }

// Copies the input features (BF16, 8 channels, padded boxes with zero halos)
// from a device buffer of the caller.
int cad_unet_set_input(cad_unet *u, const void *dev)
{
    if (!u || !dev) {
        cad_set_error("cad_unet_set_input: null argument");
        return CAD_ERR_ARG;
    }
    CAD_CHECK(hipMemcpyAsync(u->x, dev, padded(u->g[0]) * 8 * sizeof(bf16), hipMemcpyDeviceToDevice,
                             u->stream));
    return CAD_OK;
}

// Copies the output (FP32, interior voxels, x fastest) to a device buffer of
// the caller and waits for it.
int cad_unet_get_output(cad_unet *u, float *dev)
{
    if (!u || !dev) {
        cad_set_error("cad_unet_get_output: null argument");
        return CAD_ERR_ARG;
    }
    CAD_CHECK(hipMemcpyAsync(dev, u->y, nvox(u->g[0]) * sizeof(float), hipMemcpyDeviceToDevice, u->stream));
    CAD_CHECK(hipStreamSynchronize(u->stream));
    return CAD_OK;
}

int cad_unet_forward(cad_unet *u)
{
    if (!u) {
        cad_set_error("cad_unet_forward: null network");
        return CAD_ERR_ARG;
    }
    return cad_unet_forward_impl(u);
}

int cad_unet_train_step(cad_unet *u, const float *target, const float *mask, float lr, float *loss_out)
{
    if (!u || !u->train || !target || !mask) {
        cad_set_error("cad_unet_train_step: not a training network or null argument");
        return CAD_ERR_ARG;
    }
    RET_IF(cad_unet_forward_impl(u));
    const UGeo &g = u->g[0];
    RET_IF(ew(u, k_zero_f, u->nparam, u->nparam, u->gr));
    RET_IF(ew(u, k_zero_f, (long long)g.nb, (long long)g.nb, u->esq));
    RET_IF(ew(u, k_zero_f, 1, 1LL, u->loss));
    hipLaunchKernelGGL(k_esq, dim3(64, g.nb), dim3(256), 0, u->stream, g, target, mask, u->esq);
    hipLaunchKernelGGL(k_loss, dim3(grid_for(nvox(g))), dim3(256), 0, u->stream, g, u->y, target, mask,
                       u->esq, u->dy, u->loss);
    CAD_CHECK(hipGetLastError());
    RET_IF(backward(u));
    u->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)u->step), c2 = 1.f - std::pow(b2, (float)u->step);
    RET_IF(ew(u, k_adam, u->nparam, u->nparam, u->p, u->gr, u->m1, u->m2, lr, b1, b2, c1, c2, 1e-8f));
    RET_IF(pack_all(u));
    if (loss_out) {
        CAD_CHECK(hipMemcpyAsync(loss_out, u->loss, sizeof(float), hipMemcpyDeviceToHost, u->stream));
        CAD_CHECK(hipStreamSynchronize(u->stream));
    }
    return CAD_OK;
}

// One Adam step from the gradient dy of a loss with respect to the output
// of the last forward pass (computed by the caller, FP32 in the output's
// order): backward through that pass, Adam, repack.
int cad_unet_train_step_dy(cad_unet *u, const float *dy, float lr)
{
    if (!u || !u->train || !dy) {
        cad_set_error("cad_unet_train_step_dy: not a training network or null argument");
        return CAD_ERR_ARG;
    }
    RET_IF(ew(u, k_zero_f, u->nparam, u->nparam, u->gr));
    CAD_CHECK(hipMemcpyAsync(u->dy, dy, nvox(u->g[0]) * sizeof(float), hipMemcpyDeviceToDevice, u->stream));
    RET_IF(backward(u));
    u->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)u->step), c2 = 1.f - std::pow(b2, (float)u->step);
    RET_IF(ew(u, k_adam, u->nparam, u->nparam, u->p, u->gr, u->m1, u->m2, lr, b1, b2, c1, c2, 1e-8f));
    RET_IF(pack_all(u));
    CAD_CHECK(hipStreamSynchronize(u->stream));
    return CAD_OK;
}

// The gradient of every parameter from the caller's gradient dy with respect
// to the output of the last forward pass (no step).
int cad_unet_gradients_dy(cad_unet *u, const float *dy, float *host_grad)
{
    if (!u || !u->train || !dy || !host_grad) {
        cad_set_error("cad_unet_gradients_dy: not a training network or null argument");
        return CAD_ERR_ARG;
    }
    RET_IF(ew(u, k_zero_f, u->nparam, u->nparam, u->gr));
    CAD_CHECK(hipMemcpyAsync(u->dy, dy, nvox(u->g[0]) * sizeof(float), hipMemcpyDeviceToDevice, u->stream));
    RET_IF(backward(u));
    CAD_CHECK(hipMemcpyAsync(host_grad, u->gr, u->nparam * sizeof(float), hipMemcpyDeviceToHost, u->stream));
    CAD_CHECK(hipStreamSynchronize(u->stream));
    return CAD_OK;
}

int cad_unet_gradients(cad_unet *u, const float *target, const float *mask, float *host_grad, float *loss_out)
{
    if (!u || !u->train || !target || !mask || !host_grad) {
        cad_set_error("cad_unet_gradients: not a training network or null argument");
        return CAD_ERR_ARG;
    }
    RET_IF(cad_unet_forward_impl(u));
    const UGeo &g = u->g[0];
    RET_IF(ew(u, k_zero_f, u->nparam, u->nparam, u->gr));
    RET_IF(ew(u, k_zero_f, (long long)g.nb, (long long)g.nb, u->esq));
    RET_IF(ew(u, k_zero_f, 1, 1LL, u->loss));
    hipLaunchKernelGGL(k_esq, dim3(64, g.nb), dim3(256), 0, u->stream, g, target, mask, u->esq);
    hipLaunchKernelGGL(k_loss, dim3(grid_for(nvox(g))), dim3(256), 0, u->stream, g, u->y, target, mask,
                       u->esq, u->dy, u->loss);
    CAD_CHECK(hipGetLastError());
    RET_IF(backward(u));
    CAD_CHECK(hipMemcpyAsync(host_grad, u->gr, u->nparam * sizeof(float), hipMemcpyDeviceToHost, u->stream));
    if (loss_out) CAD_CHECK(hipMemcpyAsync(loss_out, u->loss, sizeof(float), hipMemcpyDeviceToHost, u->stream));
    CAD_CHECK(hipStreamSynchronize(u->stream));
    return CAD_OK;
}

}  // extern "C"
