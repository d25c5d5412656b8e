// surrogate.hip -- kernels of the 3D convolutional surrogate (no reference
// counterpart: the reference has no learned model).
//
// Activations are channels-last bf16, [batch][z][y][x][channel]; a tensor
// view is (pointer, edge D, channel stride ld, channel offset off), so a
// concatenation of two tensors is one buffer written at two offsets.
// 3^3 convolutions are implicit GEMMs on the bf16 MFMA of gfx950
// (v_mfma_f32_16x16x32_bf16, FP32 accumulation): M = voxels, N = output
// channels, K = 27 taps x input channels, k = tap * Cin + ci with
// tap = (dz+1)*9 + (dy+1)*3 + (dx+1), zero padding outside the volume.
// Weight gradients are GEMMs over the voxels on the same MFMA.  Activation:
// leaky ReLU (slope 0.05), so no unit can die for good during training.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "surrogate.hpp"

namespace gq {
namespace sur {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr float kLeak = 0.05f;  // leaky ReLU slope for negative inputs

__device__ __forceinline__ float bf2f(__bf16 v) { return (float)v; }

__device__ __forceinline__ bf16x8 load8(const __bf16 *p) { return *reinterpret_cast<const bf16x8 *>(p); }

__device__ __forceinline__ bf16x8 zero8()
{
    bf16x8 z;
#pragma unroll
    for (int j = 0; j < 8; j++) z[j] = (__bf16)0.0f;
    return z;
}

// 8 input channels c .. c+7 (c a multiple of 8) of voxel (gx, gy, gz) of item
// bb: below cup from the half-resolution tensor `up` at (gx/2, gy/2, gz/2)
// (nearest 2x upsampling), else channel c - cup of `in` (edge in.D).
__device__ __forceinline__ bf16x8 load_input8(const View &in, const View &up, int cup, int64_t bb, int gx, int gy,
                                              int gz, int c)
{
    if (c < cup) {
        const int64_t Du = up.D;
        const int64_t vu = bb * Du * Du * Du + (gx >> 1) + Du * ((gy >> 1) + Du * (int64_t)(gz >> 1));
        return load8(up.p + vu * up.ld + up.off + c);
    }
    const int64_t D = in.D;
    const int64_t vn = bb * D * D * D + gx + D * (gy + D * (int64_t)gz);
    return load8(in.p + vn * in.ld + in.off + (c - cup));
}

}  // namespace

// 3^3 convolution by voxel bricks: one workgroup per (8x8x4-voxel output
// brick, group of NTG 16-channel output tiles).  The brick's 10x10x6 input halo
// is staged in LDS channels-last, 32 input channels at a time (one 16-byte
// read per lane and tap: no transposition, the GEMM's k runs along the
// channels of a tap), and each wave runs 4 voxel tiles x NTG channel tiles
// over the 27 taps, the weights read once per K-step and reused by the 4
// voxel tiles.
namespace {
// LDS layout of the halo: voxel (hx, hy, hz) at hx * CB_LD + hy * CB_RST +
// hz * CB_PST (elements).  A 16-lane A read covers 8 voxels of each of two
// adjacent rows; at 80-byte voxels (5 x 16 bytes: 8 voxels on 8 distinct
// 16-byte bank groups) a row stride of 8 mod 16 groups puts the second row on
// the other 8, so the read takes one LDS cycle (unpadded rows, 50 groups,
// share two groups and take two).
constexpr int CB_LD = 40;                  // voxel stride (32 channels + 8 pad)
constexpr int CB_RST = 10 * CB_LD + 48;    // row stride: 56 groups of 16 bytes
constexpr int CB_PST = 10 * CB_RST;        // plane stride
constexpr int CB_SIZE = 6 * CB_PST;        // 53,760 bytes: three workgroups per CU
}  // namespace

template <int NTG, bool OUT_BF16>
__global__ __launch_bounds__(256) void k_conv3_brick(ConvArgs a)
{
    __shared__ __attribute__((aligned(16))) __bf16 s_in[CB_SIZE];
    const int D = a.in.D;
    const int nbx = D >> 3, nbz = D >> 2;
    const int b = blockIdx.x;
    const int bx = b % nbx, by = (b / nbx) % nbx, bz = (b / (nbx * nbx)) % nbz, bb = b / (nbx * nbx * nbz);
    const int x0 = 8 * bx, y0 = 8 * by, z0 = 4 * bz;
    const int nt0 = blockIdx.y * NTG;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int li = lane & 15, kq = lane >> 4;
    const int Cin = a.cin, K = 27 * Cin;
    const int64_t D3 = (int64_t)D * D * D;
    f32x4 acc[4][NTG];
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
        for (int n = 0; n < NTG; n++) acc[m][n] = f32x4{0.f, 0.f, 0.f, 0.f};
    // this lane's A rows: voxel tile mt = 4 wave + m covers brick rows 2 mt, 2 mt + 1 (row = y + 8 z)
    int hbase[4];
#pragma unroll
    for (int m = 0; m < 4; m++) {
        const int r = 2 * (4 * wave + m) + (li >> 3);
        hbase[m] = (li & 7) * CB_LD + (r & 7) * CB_RST + (r >> 3) * CB_PST;
    }
    // staging: 600 halo voxels x 4 groups of 8 channels, 10 per thread in
    // batches of NB loads in flight (more at once cost occupancy)
    // Loads are unconditional (a lane with nothing to load reads its item's
    // first voxel and stores zeros), so no branch joins wait on them.
    constexpr int NB = NTG <= 2 ? 5 : NTG == 3 ? 3 : 2;
    const int64_t Du = a.cup > 0 ? a.up.D : 0;
    const __bf16 *in_item = a.in.p + bb * D3 * a.in.ld + a.in.off;
    const __bf16 *up_item = a.cup > 0 ? a.up.p + bb * Du * Du * Du * a.up.ld + a.up.off : in_item;
    for (int c0 = 0; c0 < Cin; c0 += 32) {
        __syncthreads();
#pragma unroll 1
        for (int e0 = t; e0 < 600 * 4; e0 += 256 * NB) {
            bf16x8 sv[NB];
            int so[NB];
            bool ok[NB];
#pragma unroll
            for (int q = 0; q < NB; q++) {
                const int e = e0 + 256 * q;
                const int h = e >> 2, g = e & 3;
                const int hx = h % 10, hy = (h / 10) % 10, hz = h / 100;
                const int gx = x0 + hx - 1, gy = y0 + hy - 1, gz = z0 + hz - 1;
                const int ci = c0 + 8 * g;
                so[q] = e < 600 * 4 ? hx * CB_LD + hy * CB_RST + hz * CB_PST + 8 * g : -1;
                ok[q] = e < 600 * 4 && ci < Cin && gx >= 0 && gx < D && gy >= 0 && gy < D && gz >= 0 && gz < D;
                const int64_t vu = (gx >> 1) + Du * ((gy >> 1) + Du * (int64_t)(gz >> 1));
                const int64_t vn = gx + (int64_t)D * (gy + (int64_t)D * gz);
                const __bf16 *src = !ok[q] ? in_item
                                    : ci < a.cup ? up_item + vu * a.up.ld + ci
                                                 : in_item + vn * a.in.ld + (ci - a.cup);
                sv[q] = load8(src);
            }
#pragma unroll
            for (int q = 0; q < NB; q++)
                if (so[q] >= 0) *reinterpret_cast<bf16x8 *>(&s_in[so[q]]) = ok[q] ? sv[q] : zero8();
        }
        __syncthreads();
        const bool kok = c0 + 8 * kq < Cin;
        // weights of tap + PF are loaded while tap computes (register ring of PF + 1 taps)
        constexpr int PF = 3;
        bf16x8 vb[PF + 1][NTG];
        auto load_w = [&](int tap, bf16x8 *dst) {
#pragma unroll
            for (int n = 0; n < NTG; n++) {
                const int co = 16 * (nt0 + n) + li;
                dst[n] = (kok && co < a.cout) ? load8(a.w + (int64_t)co * K + tap * Cin + c0 + 8 * kq) : zero8();
            }
        };
#pragma unroll
        for (int p = 0; p < PF; p++) load_w(p, vb[p]);
#pragma unroll
        for (int tap = 0; tap < 27; tap++) {
            if (tap + PF < 27) load_w(tap + PF, vb[(tap + PF) % (PF + 1)]);
            const int dx = tap % 3, dy = (tap / 3) % 3, dz = tap / 9;
            const int hoff = dx * CB_LD + dy * CB_RST + dz * CB_PST;
#pragma unroll
            for (int m = 0; m < 4; m++) {
                const bf16x8 va = *reinterpret_cast<const bf16x8 *>(&s_in[hbase[m] + hoff + 8 * kq]);
#pragma unroll
                for (int n = 0; n < NTG; n++)
                    acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(va, vb[tap % (PF + 1)][n], acc[m][n], 0, 0, 0);
            }
        }
    }
    // D row 4 kq + rr = voxel of the tile, column li = output channel of the tile
#pragma unroll
    for (int n = 0; n < NTG; n++) {
        const int co = 16 * (nt0 + n) + li;
        if (co >= a.cout) continue;
        const float bias = a.bias ? a.bias[co] : 0.f;
#pragma unroll
        for (int m = 0; m < 4; m++) {
            // rows 4 kq .. 4 kq + 3 of tile m: voxels x0 + 4 (kq & 1) + rr of one (y, z) row
            const int r = 2 * (4 * wave + m) + (kq >> 1);
            const int64_t v0 =
                bb * D3 + (x0 + 4 * (kq & 1)) + (int64_t)D * ((y0 + (r & 7)) + (int64_t)D * (z0 + (r >> 3)));
            const int64_t o0 = v0 * a.out_ld + a.out_off + co;
#pragma unroll
            for (int rr = 0; rr < 4; rr++) {
                float val = acc[m][n][rr] + bias;
                if (a.relu) val = val > 0.f ? val : kLeak * val;
                const int64_t o = o0 + (int64_t)rr * a.out_ld;
                if (OUT_BF16) a.out_bf[o] = (__bf16)val;
                else a.out_f[o] = a.accumulate ? a.out_f[o] + val : val;
            }
        }
    }
}

// The same convolution for layers with few input channels (CIN = 8, 16, 48):
// the GEMM's k = tap * C + ci runs in steps of 32 across taps.  The input
// channels are staged CP at a time (CP = 8, 16 or 32: passes over channel
// blocks of CP, the same accumulators), the brick's 10x10x6 halo with CP
// channels per voxel in LDS, so a 32-k step covers 32 / CP whole taps: lane
// group kq reads channels 8 * (kq % (CP / 8)) .. +7 of tap
// TPS * st + kq / (CP / 8) (TPS = 32 / CP taps per step).  Every per-step
// offset is then a compile-time constant or a select among TPS of them: the
// weights are read from one per-lane pointer at immediate offsets, and the four
// voxel tiles of a wave lie two halo rows apart, so their LDS reads are one
// address plus immediate offsets (no division or 64-bit address arithmetic in
// the step loop).  CIN = 8: 7 steps; 16: 14; 48 in three passes of 16: 42
// (38 KB of LDS with the row padding, four workgroups per CU).  Only the last step of a pass can
// run past tap 26: its weights are zero there and its reads stay in the halo.
// Same output mapping as k_conv3_brick.
namespace {
// LDS layout of the halo for CP channels per voxel: voxel (hx, hy, hz) at
// hx * LD + hy * RST + hz * PST (elements).  A wave's 16-lane read covers 8
// voxels of each of two rows; with 16 lanes x 16 bytes per LDS cycle the
// rows must fall on complementary halves of the 64 banks: the row stride
// (in 16-byte units) is odd for 32-byte voxels (CP = 8) and 8 mod 16 for
// 48-byte voxels (CP = 16), hence the row padding.
template <int CP>
struct FlatHalo {
    static constexpr int LD = CP + 8;                                        // voxel stride (pad: banks)
    static constexpr int RST = 10 * LD + (CP == 8 ? 8 : CP == 16 ? 80 : 48);  // row stride
    static constexpr int PST = 10 * RST;                                     // plane stride
    static constexpr int SIZE = 6 * PST;
    // offset of tap (dz+1)*9 + (dy+1)*3 + (dx+1) relative to the output voxel's (0,0,0) tap
    __host__ __device__ static constexpr int tap(int t) { return (t % 3) * LD + ((t / 3) % 3) * RST + (t / 9) * PST; }
};
}  // namespace

template <int CIN, int CP, int NTG, bool OUT_BF16>
__global__ __launch_bounds__(256) void k_conv3_flat(ConvArgs a)
{
    static_assert(CIN % CP == 0 && (CP == 8 || CP == 16 || CP == 32), "channel passes of 8, 16 or 32");
    typedef FlatHalo<CP> H;
    constexpr int KT = 27 * CIN;   // the weights' k extent
    constexpr int GPT = CP / 8;    // lane groups per tap
    constexpr int TPS = 4 / GPT;   // taps per step
    constexpr int NST = (27 + TPS - 1) / TPS;
    __shared__ __attribute__((aligned(16))) __bf16 s_in[H::SIZE];
    const int D = a.in.D;
    const int nbx = D >> 3, nbz = D >> 2;
    const int b = blockIdx.x;
    const int bx = b % nbx, by = (b / nbx) % nbx, bz = (b / (nbx * nbx)) % nbz, bb = b / (nbx * nbx * nbz);
    const int x0 = 8 * bx, y0 = 8 * by, z0 = 4 * bz;
    const int nt0 = blockIdx.y * NTG;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int li = lane & 15, kq = lane >> 4;
    const int jt = kq / GPT, ci = 8 * (kq % GPT);  // the lane's tap within a step, its channels within a tap
    const int64_t D3 = (int64_t)D * D * D;
    f32x4 acc[4][NTG];
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
        for (int n = 0; n < NTG; n++) acc[m][n] = f32x4{0.f, 0.f, 0.f, 0.f};
    // voxel tile m of this wave: rows r = 8 * wave + 2 * m + (li >> 3) of the brick's 8 (y) x 4 (z)
    // rows, i.e. halo voxel (li & 7, 2 m + (li >> 3), wave) at tap 0: tiles 2 rows apart
    const int hbase0 = (li & 7) * H::LD + (li >> 3) * H::RST + wave * H::PST;
    // The thread's halo elements (voxel h, 8-channel group g) are the same in
    // every pass: their LDS offsets and voxel indices in `in` and in the
    // half-resolution `up` are computed once (vin < 0: outside the volume).
    constexpr int NE = (600 * GPT + 255) / 256;
    int e_lds[NE], e_g[NE], e_in[NE], e_up[NE];
    const int Du = a.cup > 0 ? a.up.D : 0;
#pragma unroll
    for (int q = 0; q < NE; q++) {
        const int e = t + 256 * q;
        const int h = e / GPT, g = e - h * GPT;
        const int hx = h % 10, hy = (h / 10) % 10, hz = h / 100;
        const int gx = x0 + hx - 1, gy = y0 + hy - 1, gz = z0 + hz - 1;
        const bool inside = gx >= 0 && gx < D && gy >= 0 && gy < D && gz >= 0 && gz < D;
        e_lds[q] = e < 600 * GPT ? hx * H::LD + hy * H::RST + hz * H::PST + 8 * g : -1;
        e_g[q] = 8 * g;
        e_in[q] = inside ? gx + D * (gy + D * gz) : -1;
        e_up[q] = inside ? (gx >> 1) + Du * ((gy >> 1) + Du * (gz >> 1)) : 0;
    }
    const __bf16 *in_item = a.in.p + bb * D3 * a.in.ld + a.in.off;
    const __bf16 *up_item = a.cup > 0 ? a.up.p + bb * (int64_t)Du * Du * Du * a.up.ld + a.up.off : nullptr;
    const __bf16 *sa = s_in + hbase0 + ci;
    // weights PF steps ahead (an L2 hit takes several hundred cycles; a step's
    // MFMAs about 64 per wave), the first PF issued before the halo staging
    constexpr int PF = NTG == 1 ? 6 : 3;
#pragma unroll 1
    for (int cb = 0; cb < CIN; cb += CP) {
        // per lane: weights of channel ci of tap jt of step 0 (tap TPS * st + jt at + st * TPS * CIN)
        const __bf16 *wl[NTG];
#pragma unroll
        for (int n = 0; n < NTG; n++) wl[n] = a.w + (int64_t)(16 * (nt0 + n) + li) * KT + jt * CIN + cb + ci;
        bf16x8 vb[PF + 1][NTG];
        auto load_w = [&](int st, bf16x8 *dst) {
            const bool ok = TPS * st + jt < 27;  // false only in a pass's last step
#pragma unroll
            for (int n = 0; n < NTG; n++) dst[n] = ok ? load8(wl[n] + st * TPS * CIN) : zero8();
        };
#pragma unroll
        for (int p = 0; p < PF; p++)
            if (p < NST) load_w(p, vb[p]);
        if (cb > 0) __syncthreads();  // the previous pass has read the halo
#pragma unroll
        for (int q = 0; q < NE; q++) {
            if (e_lds[q] < 0) continue;
            bf16x8 v = zero8();
            if (e_in[q] >= 0) {
                // channels c .. c+7: below cup from `up` (nearest 2x upsampling), else channel c - cup of `in`
                const int c = cb + e_g[q];
                v = c < a.cup ? load8(up_item + (int64_t)e_up[q] * a.up.ld + c)
                              : load8(in_item + (int64_t)e_in[q] * a.in.ld + (c - a.cup));
            }
            *reinterpret_cast<bf16x8 *>(&s_in[e_lds[q]]) = v;
        }
        __syncthreads();
#pragma unroll
        for (int st = 0; st < NST; st++) {
            if (st + PF < NST) load_w(st + PF, vb[(st + PF) % (PF + 1)]);
            // halo offset of the lane's tap: one of TPS constants (past tap 26: tap 26's, weights zero)
            int ho;
            if constexpr (TPS == 1) {
                ho = H::tap(st);
            } else {
                const int t0 = TPS * st;
                ho = H::tap(t0 < 27 ? t0 : 26);
#pragma unroll
                for (int q = 1; q < TPS; q++)
                    if (jt == q) ho = H::tap(t0 + q < 27 ? t0 + q : 26);
            }
            const __bf16 *sp = sa + ho;
#pragma unroll
            for (int m = 0; m < 4; m++) {
                const bf16x8 va = *reinterpret_cast<const bf16x8 *>(sp + 2 * m * H::RST);
#pragma unroll
                for (int n = 0; n < NTG; n++)
                    acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(va, vb[st % (PF + 1)][n], acc[m][n], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int n = 0; n < NTG; n++) {
        const int co = 16 * (nt0 + n) + li;
        const float bias = a.bias ? a.bias[co] : 0.f;
#pragma unroll
        for (int m = 0; m < 4; m++) {
            // rows 4 kq .. 4 kq + 3 of tile m: voxels x0 + 4 (kq & 1) + rr of one (y, z) row
            const int r = 2 * (4 * wave + m) + (kq >> 1);
            const int64_t v0 =
                bb * D3 + (x0 + 4 * (kq & 1)) + (int64_t)D * ((y0 + (r & 7)) + (int64_t)D * (z0 + (r >> 3)));
            const int64_t o0 = v0 * a.out_ld + a.out_off + co;
#pragma unroll
            for (int rr = 0; rr < 4; rr++) {
                float val = acc[m][n][rr] + bias;
                if (a.relu) val = val > 0.f ? val : kLeak * val;
                const int64_t o = o0 + (int64_t)rr * a.out_ld;
                if (OUT_BF16) a.out_bf[o] = (__bf16)val;
                else a.out_f[o] = a.accumulate ? a.out_f[o] + val : val;
            }
        }
    }
}

// Weight gradient by voxel bricks: dW[co][tap*Cin + ci] = sum_v dZ[v][co] X[v + tap][ci].
// Workgroup (16-channel input tile, output group of <= 32 channels, brick
// group): for each 8x8x4-voxel brick of its group it stages dZ^T [co][voxel]
// and the 10x10x6 halo of X^T for its 16 input channels in LDS -- three
// copies shifted by 0/1/2 in x so that every tap's 8-voxel row is one aligned
// 16-byte read -- and runs all 27 taps on the MFMA (A = dZ^T rows, B = the
// tap-shifted X^T, K = voxels in 32-voxel steps of 4 brick rows).  Partials
// per brick group, reduced in group order afterwards (deterministic).
namespace {
constexpr int WB_ROWS = 61;   // 10 x 6 halo rows + 1 pad row per input channel (LDS banks)
constexpr int WB_DZLD = 264;  // 256 voxels + 8 pad per output channel
}  // namespace

// NCOT: 16-channel output tiles per group (1 or 2); the groups of one launch
// start at output group cog0.  A wave's tiles are tap * NCOT + cot for tap =
// wave / NCOT + (4 / NCOT) j, all with cot = wave % NCOT: one dZ^T read per
// K-step serves all of them, and their X^T offsets are set once.
template <int NCOT>
__global__ __launch_bounds__(256) void k_conv3_wgrad_brick(WgradArgs a, float *bias_partial, int bricks_per_group,
                                                           int cog0)
{
    __shared__ __attribute__((aligned(16))) __bf16 s_x[3 * 16 * WB_ROWS * 8];
    __shared__ __attribute__((aligned(16))) __bf16 s_dz[32 * WB_DZLD];
    __shared__ float s_b[256];
    const int cit = blockIdx.x, cog = cog0 + blockIdx.y, bg = blockIdx.z;
    const int Cin = a.cin, Cout = a.cout;
    constexpr int nco = 16 * NCOT;
    const int D = a.x.D;
    const int nbx = D >> 3, nbz = D >> 2;
    const int64_t D3 = (int64_t)D * D * D;
    const int nbricks = a.batch * nbx * nbx * nbz;
    const int b0 = bg * bricks_per_group;
    const int b1 = b0 + bricks_per_group < nbricks ? b0 + bricks_per_group : nbricks;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int li = lane & 15, kq = lane >> 4;
    const int ci0 = 16 * cit;
    constexpr int MT = (27 * NCOT + 3) / 4;  // tiles per wave
    constexpr int TSTEP = 4 / NCOT;          // tap stride between a wave's tiles
    const int cot = wave % NCOT, tap0 = wave / NCOT;
    f32x4 acc[MT];
    int boff[MT];  // X^T element offset of tile j at brick row 0
#pragma unroll
    for (int j = 0; j < MT; j++) {
        acc[j] = f32x4{0.f, 0.f, 0.f, 0.f};
        const int tap = tap0 + TSTEP * j < 27 ? tap0 + TSTEP * j : 26;
        const int dx = tap % 3, dy = (tap / 3) % 3, dz = tap / 9;
        boff[j] = ((dx * 16 + li) * WB_ROWS + dz * 10 + dy) * 8;
    }
    const int aoff = (16 * cot + li) * WB_DZLD;
    float bsum = 0.f;
    for (int b = b0; b < b1; b++) {
        const int bx = b % nbx, by = (b / nbx) % nbx, bz = (b / (nbx * nbx)) % nbz, bb = b / (nbx * nbx * nbz);
        const int x0 = 8 * bx, y0 = 8 * by, z0 = 4 * bz;
        __syncthreads();  // the previous brick's LDS reads are done
        // X^T halo, three x-shifted copies: copy s, channel c, row (hz*10 + hy), xp <- halo x = xp + s.
        // One thread per (halo row, 8-channel half): the row's 10 voxels are loaded and transposed
        // in registers (bf16 pairs packed per dword), and each copy's channel row of 8 x values is
        // one 16-byte store.
        if (t < 120) {
            const int row = t >> 1, half = t & 1;
            const int hy = row % 10, hz = row / 10;
            const int gy = y0 + hy - 1, gz = z0 + hz - 1;
            const int c0 = ci0 + 8 * half;
            const bool rowin = c0 < Cin && gy >= 0 && gy < D && gz >= 0 && gz < D;
            uint32_t w[10][4];
#pragma unroll
            for (int x = 0; x < 10; x++) {
                const int gx = x0 + x - 1;
                bf16x8 v = zero8();
                if (rowin && gx >= 0 && gx < D) v = load_input8(a.x, a.up, a.cup, bb, gx, gy, gz, c0);
                const uint4 u = __builtin_bit_cast(uint4, v);
                w[x][0] = u.x, w[x][1] = u.y, w[x][2] = u.z, w[x][3] = u.w;
            }
#pragma unroll
            for (int c = 0; c < 8; c++)
#pragma unroll
                for (int sh = 0; sh < 3; sh++) {
                    uint32_t d[4];
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        const uint32_t lo = w[sh + 2 * k][c >> 1], hi = w[sh + 2 * k + 1][c >> 1];
                        d[k] = (c & 1) ? (lo >> 16) | (hi & 0xffff0000u) : (lo & 0xffffu) | (hi << 16);
                    }
                    *reinterpret_cast<uint4 *>(&s_x[((sh * 16 + 8 * half + c) * WB_ROWS + row) * 8]) =
                        uint4{d[0], d[1], d[2], d[3]};
                }
        }
        // dZ^T [co][voxel], voxel = x + 8 (y + 8 z) inside the brick: one thread of waves 2-3 (the
        // halo rows use waves 0-1) per (8-voxel x row, 8-channel group), transposed in registers as
        // above, one 16-byte store per channel
        constexpr int ng = nco >> 3;
        if (t >= 128 && t - 128 < 32 * ng) {
            const int g = (t - 128) % ng, ryz = (t - 128) / ng;
            const int ly = ryz & 7, lz = ryz >> 3;
            uint32_t w[8][4];
#pragma unroll
            for (int x = 0; x < 8; x++) {
                const int64_t vn = bb * D3 + (x0 + x) + (int64_t)D * ((y0 + ly) + (int64_t)D * (z0 + lz));
                const uint4 u = __builtin_bit_cast(uint4, load8(a.dz.p + vn * a.dz.ld + a.dz.off + 32 * cog + 8 * g));
                w[x][0] = u.x, w[x][1] = u.y, w[x][2] = u.z, w[x][3] = u.w;
            }
#pragma unroll
            for (int c = 0; c < 8; c++) {
                uint32_t d[4];
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const uint32_t lo = w[2 * k][c >> 1], hi = w[2 * k + 1][c >> 1];
                    d[k] = (c & 1) ? (lo >> 16) | (hi & 0xffff0000u) : (lo & 0xffffu) | (hi << 16);
                }
                *reinterpret_cast<uint4 *>(&s_dz[(8 * g + c) * WB_DZLD + 8 * ryz]) = uint4{d[0], d[1], d[2], d[3]};
            }
        }
        __syncthreads();
        if (cit == 0 && (t & 31) < nco) {
            const int co = t & 31, q = t >> 5;  // voxels 32q .. 32q+31
            for (int v = 0; v < 32; v++) bsum += bf2f(s_dz[co * WB_DZLD + 32 * q + v]);
        }
#pragma unroll 1
        for (int ks = 0; ks < 8; ks++) {
            const int r = 4 * ks + kq;  // brick row: y = r % 8, z = r / 8
            const int rowoff = ((r >> 3) * 10 + (r & 7)) * 8;
            const bf16x8 va = *reinterpret_cast<const bf16x8 *>(&s_dz[aoff + 8 * r]);
#pragma unroll
            for (int j = 0; j < MT; j++) {
                if (tap0 + TSTEP * j < 27) {  // uniform over the wave
                    const bf16x8 vb = *reinterpret_cast<const bf16x8 *>(&s_x[boff[j] + rowoff]);
                    acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(va, vb, acc[j], 0, 0, 0);
                }
            }
        }
    }
    // D row 4*kq + rr = output channel within the tile, column li = input channel within the tile
    const int Kw = 27 * Cin;
#pragma unroll
    for (int j = 0; j < MT; j++) {
        const int tap = tap0 + TSTEP * j;
        if (tap >= 27) continue;
        const int ci = ci0 + li;
        if (ci >= Cin) continue;
#pragma unroll
        for (int rr = 0; rr < 4; rr++) {
            const int co = 32 * cog + 16 * cot + 4 * kq + rr;
            a.partial[(int64_t)bg * Cout * Kw + (int64_t)co * Kw + tap * Cin + ci] = acc[j][rr];
        }
    }
    if (cit == 0) {
        s_b[t] = bsum;
        __syncthreads();
        if (t < nco) {
            float sb = 0.f;
            for (int q = 0; q < 8; q++) sb += s_b[32 * q + t];
            bias_partial[(int64_t)bg * Cout + 32 * cog + t] = sb;
        }
    }
}

// out[i] = sum over chunks of partial[c*n + i] for few outputs and many
// chunks: one workgroup per output, thread t sums chunks t, t+256, ... in
// order, then a fixed tree over the workgroup (deterministic).
__global__ __launch_bounds__(256) void k_reduce_chunks_few(const float *partial, int nchunks, int64_t n, float *out)
{
    __shared__ float red[4];
    const int64_t i = blockIdx.x;
    float s = 0.f;
    for (int c = threadIdx.x; c < nchunks; c += 256) s += partial[(int64_t)c * n + i];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[i] = (red[0] + red[1]) + (red[2] + red[3]);
}

// out[i] = sum over chunks c (in order) of partial[c*n + i]
__global__ __launch_bounds__(256) void k_reduce_chunks(const float *partial, int nchunks, int64_t n, float *out)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    // chunks added in order; eight loads in flight at a time
    float s = 0.f;
    int c = 0;
    for (; c + 8 <= nchunks; c += 8) {
        float v[8];
#pragma unroll
        for (int j = 0; j < 8; j++) v[j] = partial[(int64_t)(c + j) * n + i];
#pragma unroll
        for (int j = 0; j < 8; j++) s += v[j];
    }
    for (; c < nchunks; c++) s += partial[(int64_t)c * n + i];
    out[i] = s;
}

// Pooling, upsampling and the leaky-ReLU backward move 8 channels per thread
// (one 16-byte bf16 transfer, or two for FP32): edges are powers of two (lg =
// log2 D) and channel counts, strides and offsets multiples of 8, which the
// launchers check.  Per channel the arithmetic and its order are those of a
// one-channel-per-thread loop.
namespace {

__device__ __forceinline__ void f32x8_load(const float *p, float *v)
{
    const float4 a = reinterpret_cast<const float4 *>(p)[0], b = reinterpret_cast<const float4 *>(p)[1];
    v[0] = a.x, v[1] = a.y, v[2] = a.z, v[3] = a.w, v[4] = b.x, v[5] = b.y, v[6] = b.z, v[7] = b.w;
}

__device__ __forceinline__ void f32x8_store(float *p, const float *v)
{
    reinterpret_cast<float4 *>(p)[0] = make_float4(v[0], v[1], v[2], v[3]);
    reinterpret_cast<float4 *>(p)[1] = make_float4(v[4], v[5], v[6], v[7]);
}

// voxel v of a batch of (2^lg)^3 volumes -> item b and coordinates
struct Vox {
    int64_t b;
    int x, y, z;
};
__device__ __forceinline__ Vox vox_of(int64_t v, int lg)
{
    const int m = (1 << lg) - 1;
    const int r = (int)(v & (((int64_t)1 << (3 * lg)) - 1));
    return Vox{v >> (3 * lg), r & m, (r >> lg) & m, r >> (2 * lg)};
}
__device__ __forceinline__ int64_t vox_index(int64_t b, int x, int y, int z, int lg)
{
    return (b << (3 * lg)) + x + ((int64_t)y << lg) + ((int64_t)z << (2 * lg));
}

}  // namespace

// 2x2x2 average pooling (forward) and its adjoint (accumulated into dx).
__global__ __launch_bounds__(256) void k_avgpool(View in, View out, int C, int batch, int lgo)
{
    const int G = C >> 3;
    const int64_t n = ((int64_t)batch << (3 * lgo)) * G;
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    const int64_t v = e / G;
    const int c = 8 * (int)(e - v * G);
    const Vox o = vox_of(v, lgo);
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int64_t vi = vox_index(o.b, 2 * o.x + (k & 1), 2 * o.y + ((k >> 1) & 1), 2 * o.z + (k >> 2), lgo + 1);
        const bf16x8 q = load8(in.p + vi * in.ld + in.off + c);
#pragma unroll
        for (int j = 0; j < 8; j++) s[j] += bf2f(q[j]);
    }
    bf16x8 r;
#pragma unroll
    for (int j = 0; j < 8; j++) r[j] = (__bf16)(0.125f * s[j]);
    *reinterpret_cast<bf16x8 *>(out.p + v * out.ld + out.off + c) = r;
}

__global__ __launch_bounds__(256) void k_avgpool_bwd(const float *dout, int64_t dout_ld, FView din, int C, int batch,
                                                     int lgi)
{
    // din at the fine level; each fine voxel receives dout(parent) / 8
    const int G = C >> 3;
    const int64_t n = ((int64_t)batch << (3 * lgi)) * G;
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    const int64_t v = e / G;
    const int c = 8 * (int)(e - v * G);
    const Vox f = vox_of(v, lgi);
    const int64_t vo = vox_index(f.b, f.x >> 1, f.y >> 1, f.z >> 1, lgi - 1);
    float g[8], d[8];
    f32x8_load(dout + vo * dout_ld + c, g);
    float *p = din.p + v * din.ld + din.off + c;
    f32x8_load(p, d);
#pragma unroll
    for (int j = 0; j < 8; j++) d[j] += 0.125f * g[j];
    f32x8_store(p, d);
}

// Adjoint of nearest-neighbour 2x upsampling (sum of the 8 children): one
// thread per coarse voxel and 8 channels.  (The forward upsampling is folded
// into the convolutions' input reads, load_input8.)
__global__ __launch_bounds__(256) void k_upsample_bwd(FView dout, float *din, int64_t din_ld, int C, int batch,
                                                      int lgi)
{
    const int G = C >> 3;
    const int64_t n = ((int64_t)batch << (3 * lgi)) * G;
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    const int64_t v = e / G;
    const int c = 8 * (int)(e - v * G);
    const Vox q = vox_of(v, lgi);
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int64_t vo = vox_index(q.b, 2 * q.x + (k & 1), 2 * q.y + ((k >> 1) & 1), 2 * q.z + (k >> 2), lgi + 1);
        float g[8];
        f32x8_load(dout.p + vo * dout.ld + dout.off + c, g);
#pragma unroll
        for (int j = 0; j < 8; j++) s[j] += g[j];
    }
    f32x8_store(din + v * din_ld + c, s);
}

// dZ = dY * (Y > 0 ? 1 : leak) -> bf16 (leaky-ReLU backward through the post-activation value)
__global__ __launch_bounds__(256) void k_relu_bwd(FView dy, View y, __bf16 *dz, int C, int64_t nvox)
{
    const int G = C >> 3;
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= nvox * G) return;
    const int64_t v = e / G;
    const int c = 8 * (int)(e - v * G);
    float g[8];
    f32x8_load(dy.p + v * dy.ld + dy.off + c, g);
    const bf16x8 yv = load8(y.p + v * y.ld + y.off + c);
    bf16x8 r;
#pragma unroll
    for (int j = 0; j < 8; j++) r[j] = (__bf16)(bf2f(yv[j]) > 0.f ? g[j] : kLeak * g[j]);
    *reinterpret_cast<bf16x8 *>(dz + v * C + c) = r;
}

// Output head: p = sigmoid(sum_c w[c] h[c] + b); loss = mean (p - t)^2 +
// lam * mean BCE(sigmoid(kappa (p - th)), [t > th]); writes p and, when dh is
// set, dL/dh (FP32) and per-block partials of the loss and of dL/dw, dL/db.
__global__ __launch_bounds__(256) void k_head(HeadArgs a)
{
    __shared__ float red[4][20];
    const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const bool ok = v < a.nvox;
    float h[16];
    float zl = a.b;
    {
        const bf16x8 h0 = ok ? load8(a.h.p + v * a.h.ld + a.h.off) : zero8();
        const bf16x8 h1 = ok ? load8(a.h.p + v * a.h.ld + a.h.off + 8) : zero8();
#pragma unroll
        for (int c = 0; c < 8; c++) {
            h[c] = bf2f(h0[c]);
            h[c + 8] = bf2f(h1[c]);
        }
    }
#pragma unroll
    for (int c = 0; c < 16; c++) zl += a.w[c] * h[c];
    const float p = 1.f / (1.f + __expf(-zl));
    float part[18];
#pragma unroll
    for (int k = 0; k < 18; k++) part[k] = 0.f;
    if (ok) {
        a.pred[v] = p;
        if (a.target) {
            const float tv = a.target[v];
            const float m = tv > a.th ? 1.f : 0.f;
            const float s = 1.f / (1.f + __expf(-a.kappa * (p - a.th)));
            const float sc = fminf(fmaxf(s, 1e-7f), 1.f - 1e-7f);
            const float bce = -(m * __logf(sc) + (1.f - m) * __logf(1.f - sc));
            const float invn = 1.f / (float)a.nvox;
            part[16] = ((p - tv) * (p - tv) + a.lam * bce) * invn;
            const float dldp = (2.f * (p - tv) + a.lam * (s - m) * a.kappa) * invn;
            const float dz = dldp * p * (1.f - p);
            if (a.dh) {
                float4 *dh4 = reinterpret_cast<float4 *>(a.dh + v * 16);
#pragma unroll
                for (int q4 = 0; q4 < 4; q4++)
                    dh4[q4] = make_float4(dz * a.w[4 * q4], dz * a.w[4 * q4 + 1], dz * a.w[4 * q4 + 2], dz * a.w[4 * q4 + 3]);
#pragma unroll
                for (int c = 0; c < 16; c++) part[c] = dz * h[c];
            }
            part[17] = dz;
        }
    }
    if (!a.partial) return;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 18; k++) {
        float s = part[k];
        for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
        if (lane == 0) red[wv][k] = s;
    }
    __syncthreads();
    if (threadIdx.x < 18)
        a.partial[(int64_t)blockIdx.x * 18 + threadIdx.x] =
            ((red[0][threadIdx.x] + red[1][threadIdx.x]) + red[2][threadIdx.x]) + red[3][threadIdx.x];
}

// Input assembly: channels {params[0..3] broadcast, u0, pw, pg, 1}.
__global__ __launch_bounds__(256) void k_assemble_input(const float *params, const float *u0, const float *pw,
                                                        const float *pg, int64_t D3, int batch, __bf16 *x)
{
    const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (v >= D3 * batch) return;
    const int64_t b = v / D3, r = v - b * D3;
    __bf16 *o = x + v * 8;
    for (int c = 0; c < 4; c++) o[c] = (__bf16)params[b * 4 + c];
    o[4] = (__bf16)u0[v];
    o[5] = (__bf16)pw[r];
    o[6] = (__bf16)pg[r];
    o[7] = (__bf16)1.0f;
}

// Adam step on FP32 master weights; refreshes the bf16 forward copy.
__global__ __launch_bounds__(256) void k_adam(float *w, const float *g, float *m, float *v, int64_t n, float lr,
                                              float b1, float b2, float eps, float c1, float c2)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// bf16 copies of a conv's weights: forward layout [co][tap*Cin+ci] and the
// flipped, transposed layout of the data gradient [ci][tap'*Cout+co] with
// tap' = 26 - tap.
__global__ __launch_bounds__(256) void k_pack_weights(const float *w, int cin, int cout, __bf16 *wf, __bf16 *wt)
{
    const int64_t K = 27 * (int64_t)cin, n = (int64_t)cout * K;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int co = (int)(i / K), k = (int)(i % K), tap = k / cin, ci = k % cin;
    const __bf16 b = (__bf16)w[i];
    wf[i] = b;
    wt[(int64_t)ci * 27 * cout + (int64_t)(26 - tap) * cout + co] = b;
}

// Flip-rate log-likelihood of each batch item's prediction against 64^3
// observations: per-block FP64 partials [item][block][3], then an ordered
// per-item reduction (T1, FLAIR, PET terms).
__global__ __launch_bounds__(256) void k_sur_lik_flip(const float *__restrict__ pred, int64_t n,
                                                      const uint8_t *__restrict__ t1, const uint8_t *__restrict__ t2,
                                                      const float *__restrict__ pet, double th1, double th2,
                                                      double lp1, double lp0, double c0, double is2, double *partial)
{
    __shared__ double red[3][4];
    const int b = blockIdx.y;
    const float *p = pred + (int64_t)b * n;
    double acc[3] = {0.0, 0.0, 0.0};
    for (int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x; v < n; v += (int64_t)gridDim.x * 256) {
        const double m = p[v];
        acc[0] += ((m > th1) == (t1[v] != 0)) ? lp1 : lp0;
        acc[1] += ((m > th2) == (t2[v] != 0)) ? lp1 : lp0;
        const float q = pet[v];
        if (!__builtin_isnan(q)) {
            const double r = (double)q - m;
            acc[2] += c0 - r * r * is2;
        }
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        double s = acc[k];
        for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
        if (lane == 0) red[k][wv] = s;
    }
    __syncthreads();
    if (threadIdx.x < 3)
        partial[((int64_t)b * gridDim.x + blockIdx.x) * 3 + threadIdx.x] =
            ((red[threadIdx.x][0] + red[threadIdx.x][1]) + red[threadIdx.x][2]) + red[threadIdx.x][3];
}

__global__ void k_sum_item_partials(const double *partial, int nb, double *out)
{
    const int k = threadIdx.x, b = blockIdx.x;
    if (k >= 3) return;
    double s = 0.0;
    for (int i = 0; i < nb; i++) s += partial[((int64_t)b * nb + i) * 3 + k];
    out[b * 3 + k] = s;
}

hipError_t launch_sur_lik_flip(const float *pred, int64_t n, int batch, const uint8_t *t1, const uint8_t *t2,
                               const float *pet, double th1, double th2, double lp1, double lp0, double c0,
                               double is2, double *partial, int nb, double *out, hipStream_t st)
{
    hipLaunchKernelGGL(k_sur_lik_flip, dim3(nb, batch), dim3(256), 0, st, pred, n, t1, t2, pet, th1, th2, lp1, lp0,
                       c0, is2, partial);
    hipLaunchKernelGGL(k_sum_item_partials, dim3(batch), dim3(64), 0, st, partial, nb, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
hipError_t launch_conv3(const ConvArgs &a, hipStream_t st)
{
    const int D = a.in.D;
    const int nt = (a.cout + 15) / 16;
    if (a.cin % 8 != 0 || a.in.ld % 8 != 0 || a.in.off % 8 != 0 || a.cout % 16 != 0 || nt > 12 || D < 8 ||
        D % 8 != 0)
        return hipErrorInvalidValue;
    if (a.cup < 0 || a.cup % 8 != 0 || a.cup >= a.cin ||
        (a.cup > 0 && (2 * a.up.D != D || a.up.ld % 8 != 0 || a.up.off % 8 != 0)))
        return hipErrorInvalidValue;
    const int ntg = nt <= 4 ? nt : (nt % 4 == 0 ? 4 : 3);  // output tiles per workgroup
    if (nt % ntg != 0) return hipErrorInvalidValue;
    const dim3 grid((unsigned)(a.batch * (D / 8) * (D / 8) * (D / 4)), (unsigned)(nt / ntg));
    const bool bf = a.out_bf != nullptr;
    if (a.cin == 8 || a.cin == 16 || a.cin == 48) {
#define GQ_FLAT(C, N)                                                                                        \
    if (bf) hipLaunchKernelGGL((k_conv3_flat<C, (C > 32 ? 16 : C), N, true>), grid, dim3(256), 0, st, a);  \
    else hipLaunchKernelGGL((k_conv3_flat<C, (C > 32 ? 16 : C), N, false>), grid, dim3(256), 0, st, a);
#define GQ_FLAT_N(C)                    \
    switch (ntg) {                      \
    case 1: GQ_FLAT(C, 1) break;        \
    case 2: GQ_FLAT(C, 2) break;        \
    case 3: GQ_FLAT(C, 3) break;        \
    case 4: GQ_FLAT(C, 4) break;        \
    default: return hipErrorInvalidValue; \
    }
        if (a.cin == 8) { GQ_FLAT_N(8) }
        else if (a.cin == 16) { GQ_FLAT_N(16) }
        else { GQ_FLAT_N(48) }
#undef GQ_FLAT_N
#undef GQ_FLAT
        return hipGetLastError();
    }
#define GQ_CONV(N)                                                                        \
    if (bf) hipLaunchKernelGGL((k_conv3_brick<N, true>), grid, dim3(256), 0, st, a);      \
    else hipLaunchKernelGGL((k_conv3_brick<N, false>), grid, dim3(256), 0, st, a);
    switch (ntg) {
    case 1: GQ_CONV(1) break;
    case 2: GQ_CONV(2) break;
    case 3: GQ_CONV(3) break;
    case 4: GQ_CONV(4) break;
    default: return hipErrorInvalidValue;
    }
#undef GQ_CONV
    return hipGetLastError();
}

int wgrad_groups(int cin, int cout, int D, int batch)
{
    const int nbricks = batch * (D / 8) * (D / 8) * (D / 4);
    const int per_group = ((cin + 15) / 16) * ((cout + 31) / 32);  // workgroups per brick group
    int g = (1024 + per_group - 1) / per_group;                     // about 1024 workgroups in all
    return g < nbricks ? (g > 0 ? g : 1) : nbricks;
}

hipError_t launch_conv3_wgrad_brick(const WgradArgs &a, int ngroups, float *bias_partial, hipStream_t st)
{
    const int D = a.x.D;
    if (a.cout % 16 != 0 || a.cin % 8 != 0 || D < 8 || D % 8 != 0 || a.dz.D != D || ngroups < 1)
        return hipErrorInvalidValue;
    if (a.cup < 0 || a.cup % 8 != 0 || a.cup >= a.cin ||
        (a.cup > 0 && (2 * a.up.D != D || a.up.ld % 8 != 0 || a.up.off % 8 != 0)))
        return hipErrorInvalidValue;
    const int nbricks = a.batch * (D / 8) * (D / 8) * (D / 4);
    const int per = (nbricks + ngroups - 1) / ngroups;
    // output groups of 32 channels, then a last group of 16 when cout is an odd multiple of 16
    const int full = a.cout / 32;
    if (full > 0)
        hipLaunchKernelGGL(k_conv3_wgrad_brick<2>, dim3((unsigned)((a.cin + 15) / 16), (unsigned)full,
                           (unsigned)ngroups), dim3(256), 0, st, a, bias_partial, per, 0);
    if (a.cout % 32 != 0)
        hipLaunchKernelGGL(k_conv3_wgrad_brick<1>, dim3((unsigned)((a.cin + 15) / 16), 1u, (unsigned)ngroups),
                           dim3(256), 0, st, a, bias_partial, per, full);
    return hipGetLastError();
}

hipError_t launch_reduce_chunks(const float *partial, int nchunks, int64_t n, float *out, hipStream_t st)
{
    if (n <= 256 && nchunks >= 64) {  // a single workgroup would add every chunk serially
        hipLaunchKernelGGL(k_reduce_chunks_few, dim3((unsigned)n), dim3(256), 0, st, partial, nchunks, n, out);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_reduce_chunks, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, partial, nchunks, n,
                       out);
    return hipGetLastError();
}

static dim3 g1(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

namespace {
// log2 of a power-of-two edge, or -1
int lg_edge(int D)
{
    if (D < 1 || (D & (D - 1)) != 0) return -1;
    int l = 0;
    while ((1 << l) < D) l++;
    return l;
}
bool ch8(int C, int64_t ld, int64_t off) { return C > 0 && C % 8 == 0 && ld % 8 == 0 && off % 8 == 0 && off + C <= ld; }
}  // namespace

hipError_t launch_avgpool(const View &in, const View &out, int C, int batch, hipStream_t st)
{
    const int lgo = lg_edge(out.D);
    if (lgo < 0 || in.D != 2 * out.D || !ch8(C, in.ld, in.off) || !ch8(C, out.ld, out.off)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_avgpool, g1((int64_t)out.D * out.D * out.D * batch * (C / 8)), dim3(256), 0, st, in, out, C,
                       batch, lgo);
    return hipGetLastError();
}

hipError_t launch_avgpool_bwd(const float *dout, int64_t dout_ld, const FView &din, int C, int batch, hipStream_t st)
{
    const int lgi = lg_edge(din.D);
    if (lgi < 1 || !ch8(C, dout_ld, 0) || !ch8(C, din.ld, din.off)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_avgpool_bwd, g1((int64_t)din.D * din.D * din.D * batch * (C / 8)), dim3(256), 0, st, dout,
                       dout_ld, din, C, batch, lgi);
    return hipGetLastError();
}

hipError_t launch_upsample_bwd(const FView &dout, float *din, int64_t din_ld, int C, int batch, hipStream_t st)
{
    const int64_t Di = dout.D / 2;
    const int lgi = lg_edge((int)Di);
    if (lgi < 0 || !ch8(C, dout.ld, dout.off) || !ch8(C, din_ld, 0)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_upsample_bwd, g1(Di * Di * Di * batch * (C / 8)), dim3(256), 0, st, dout, din, din_ld, C,
                       batch, lgi);
    return hipGetLastError();
}

hipError_t launch_relu_bwd(const FView &dy, const View &y, __bf16 *dz, int C, int64_t nvox, hipStream_t st)
{
    if (!ch8(C, dy.ld, dy.off) || !ch8(C, y.ld, y.off)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_relu_bwd, g1(nvox * (C / 8)), dim3(256), 0, st, dy, y, dz, C, nvox);
    return hipGetLastError();
}

hipError_t launch_head(const HeadArgs &a, hipStream_t st)
{
    if (!ch8(16, a.h.ld, a.h.off)) return hipErrorInvalidValue;  // two 8-channel loads per voxel
    hipLaunchKernelGGL(k_head, g1(a.nvox), dim3(256), 0, st, a);
    return hipGetLastError();
}

int head_blocks(int64_t nvox) { return (int)((nvox + 255) / 256); }

hipError_t launch_assemble_input(const float *params, const float *u0, const float *pw, const float *pg, int64_t D3,
                                 int batch, __bf16 *x, hipStream_t st)
{
    hipLaunchKernelGGL(k_assemble_input, g1(D3 * batch), dim3(256), 0, st, params, u0, pw, pg, D3, batch, x);
    return hipGetLastError();
}

hipError_t launch_adam(float *w, const float *g, float *m, float *v, int64_t n, float lr, float b1, float b2,
                       float eps, float c1, float c2, hipStream_t st)
{
    hipLaunchKernelGGL(k_adam, g1(n), dim3(256), 0, st, w, g, m, v, n, lr, b1, b2, eps, c1, c2);
    return hipGetLastError();
}

hipError_t launch_pack_weights(const float *w, int cin, int cout, __bf16 *wf, __bf16 *wt, hipStream_t st)
{
    hipLaunchKernelGGL(k_pack_weights, g1((int64_t)cout * 27 * cin), dim3(256), 0, st, w, cin, cout, wf, wt);
    return hipGetLastError();
}

}  // namespace sur
}  // namespace gq
