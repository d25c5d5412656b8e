// kernels.hip -- gfx950 kernels of the FK forward-solve + likelihood path.
//
// Data layout in HBM: a pool of 16^3 blocks, one float array per field
// (phi ping-pong, pw, pg), block slot s at [s*4096, s*4096+4096), x fastest.
// Per-slot tables (built on the host from the grid topology) give the
// same-level neighbour of each face/edge/corner region and, for regions
// covered by a coarser or finer leaf, the first row of a CSR table of
// (source cell, weight) terms.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

#include <hip/hip_cooperative_groups.h>

#include <hipcub/device/device_scan.hpp>

namespace gq {

namespace {

constexpr int B = 16;
constexpr int B3 = 4096;
constexpr int PLW = B + 2; // plane width of the column-streaming step

// LDS tile of a step slab: ZT + 2 planes (z0-1 .. z0+ZT) of 18 rows (y = -1
// .. 16) x 16 columns (x = 0 .. 15), then the x halo columns (x = -1, x = 16)
// of the ZT interior planes.  A row stride of 16 floats puts the 64 lanes of
// a wave (4 rows of 16) on distinct banks for every read but the two x halo
// columns (with rows of 18, rows 0 and 1 of a wave shared two banks).
template <int ZT>
struct SlabTile {
    static constexpr int RW = 16, PL = 18 * 16;
    static constexpr int XW = (ZT + 2) * PL, XE = XW + 16 * ZT;
    static constexpr int SIZE = XE + 16 * ZT;
    // x 0..15, y -1..16, plane zl 0..ZT+1 (zl = z - z0 + 1)
    __device__ static int cell(int x, int y, int zl) { return x + RW * (y + 1) + PL * zl; }
    // x = -1 / x = 16 of row y 0..15 of interior plane k 0..ZT-1
    __device__ static int west(int y, int k) { return XW + y + 16 * k; }
    __device__ static int east(int y, int k) { return XE + y + 16 * k; }
};

__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }

// Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a
// contiguous chunk of the Morton-ordered work list so neighbouring blocks
// share an L2.
__device__ __forceinline__ int xcd_remap(int wg, int nwg)
{
    // nwg is a multiple of 8 (the launcher pads the grid)
    const int per = nwg >> 3;
    return (wg & 7) * per + (wg >> 3);
}

// Face plane cell of a neighbour block seen across face f (0:-x 1:+x 2:-y 3:+y 4:-z 5:+z)
// at face coordinates (u, v).
__device__ __forceinline__ int nbr_face_cell(int f, int u, int v)
{
    switch (f) {
    case 0: return (B - 1) + B * (u + B * v);  // neighbour's x = 15, (y,z) = (u,v)
    case 1: return 0 + B * (u + B * v);
    case 2: return u + B * ((B - 1) + B * v);  // neighbour's y = 15, (x,z) = (u,v)
    case 3: return u + B * (0 + B * v);
    case 4: return u + B * (v + B * (B - 1));  // neighbour's z = 15, (x,y) = (u,v)
    default: return u + B * (v + B * 0);
    }
}

// 1/K = (pw+pg) / (pw*Dw + pg*Dg): the reciprocal of the tissue-normalised
// diffusivity of Case2 (Glioma_ReactionDiffusionOperator.h:174-188).  0/0 ->
// NaN (CSF / outside), as in the reference.
// Contraction is off in the per-cell arithmetic and every FMA is explicit, so
// each step kernel (whatever its tiling, unrolling or scheduling) rounds
// exactly alike and the tile choice never changes a bit.
__device__ __forceinline__ float inv_k(float w, float g, float Dw, float Dg)
{
#pragma clang fp contract(off)
    const float b = w + g;
    const float D = __builtin_fmaf(g, Dg, w * Dw);
    return b * frcp(D);
}

// Per-cell step coefficient, precomputed once per anatomy and parameter set:
// the cell's 1/K where that value also settles the tissue test, else the
// marker IK_SLOW (the step then recomputes 1/K and p_w + p_g from pw, pg).
// With p_w, p_g >= 0 and phi >= 0, tissue = p_w + p_g + phi > 0 holds exactly
// when p_w + p_g > 0 (then 1/K is never NaN) or phi > 0 (p_w = p_g = 0, 1/K =
// 0/0 = NaN); cells with a negative p_w or p_g (Interp4 prolongation can make
// them) or with p_w = -p_g != 0 take the marker, and a cell whose phi is
// negative (after a split) recomputes its own sum in the step.  The step's
// results are bit-identical to recomputing everything every step.
constexpr unsigned IK_SLOW = 0xFFFFFFFFu;

__device__ __forceinline__ float ik_code(float w, float g, float Dw, float Dg)
{
#pragma clang fp contract(off)
    const float ik = inv_k(w, g, Dw, Dg);
    const float b = w + g;
    const bool nan = ik != ik;
    return ((b > 0.f && !nan) || (b == 0.f && nan)) ? ik : __uint_as_float(IK_SLOW);
}

// 1/K of a neighbour cell from its code.
__device__ __forceinline__ float ik_of(float code, const float *__restrict__ pw, const float *__restrict__ pg,
                                       size_t i, float Dw, float Dg)
{
    return __float_as_uint(code) != IK_SLOW ? code : inv_k(pw[i], pg[i], Dw, Dg);
}

// 1/K of the updated cell and a stand-in for p_w + p_g that gives the same
// tissue test (bsum + phi > 0) as the true sum.
__device__ __forceinline__ void ik_centre(float code, float pc, const float *__restrict__ pw,
                                          const float *__restrict__ pg, size_t i, float Dw, float Dg, float &ik,
                                          float &bsum)
{
#pragma clang fp contract(off)
    if (__float_as_uint(code) != IK_SLOW && pc >= 0.f) {
        ik = code;
        bsum = code != code ? 0.f : 1.f;
    } else {
        const float w = pw[i], g = pg[i];
        ik = inv_k(w, g, Dw, Dg);
        bsum = w + g;
    }
}

// One cell of the fused RD operator + UpdateTumor (Case2,
// Glioma_ReactionDiffusionOperator.h:150-210, UpdateTumor :224-236): phi of
// the cell and its six face neighbours (-x,+x,-y,+y,-z,+z), their 1/K, and
// p_w + p_g of the cell.
__device__ __forceinline__ float fk_cell(float pc, float pW, float pE, float pS, float pN, float pB, float pF,
                                         float kc, float kW, float kE, float kS, float kN, float kB, float kF,
                                         float bsum, float ih2, float rho, float dt)
{
#pragma clang fp contract(off)
    // phi = 0 on the cell and its six neighbours: every term below is zero
    // (or NaN next to CSF, which the clamp turns into 0) and the update is +0
    // exactly, so waves far from the tumour skip the arithmetic.
    if ((pc == 0.f) & (pW == 0.f) & (pE == 0.f) & (pS == 0.f) & (pN == 0.f) & (pB == 0.f) & (pF == 0.f)) return 0.f;
    float d0 = 0.f, d1 = 0.f, d2 = 0.f, d3 = 0.f, d4 = 0.f, d5 = 0.f;
    if (bsum + pc > 0.f) {  // tissue = p_w + p_g + phi > 0 (:155-157)
        d0 = 2.f * frcp(kc + kW);
        d1 = 2.f * frcp(kc + kE);
        d2 = 2.f * frcp(kc + kS);
        d3 = 2.f * frcp(kc + kN);
        d4 = 2.f * frcp(kc + kB);
        d5 = 2.f * frcp(kc + kF);
        // Neumann doubling, sequential as the reference (:198-203).  It can
        // only act when some coefficient is zero (1/K = inf next to the
        // face); outside the tissue every coefficient is zero and doubling
        // them changes nothing.  A NaN coefficient makes the result 0
        // whatever the doubling does, so min() may ignore NaNs.
        const float mn = __builtin_fminf(__builtin_fminf(__builtin_fminf(__builtin_fabsf(d0), __builtin_fabsf(d1)),
                                                         __builtin_fminf(__builtin_fabsf(d2), __builtin_fabsf(d3))),
                                         __builtin_fminf(__builtin_fabsf(d4), __builtin_fabsf(d5)));
        if (mn == 0.f) {
            if (d0 == 0.f && pW > 0.f) d1 *= 2.f;
            if (d1 == 0.f && pE > 0.f) d0 *= 2.f;
            if (d2 == 0.f && pS > 0.f) d3 *= 2.f;
            if (d3 == 0.f && pN > 0.f) d2 *= 2.f;
            if (d4 == 0.f && pB > 0.f) d5 *= 2.f;
            if (d5 == 0.f && pF > 0.f) d4 *= 2.f;
        }
    }
    const float sdf = d0 + d1 + d2 + d3 + d4 + d5;
    const float reac = rho * pc * (1.f - pc);
    float dphi;
    if (__builtin_isfinite(sdf)) {
        // flux form sum_k df_k (phi_k - phi_c): the same quantity as the
        // reference's in + out, without its cancellation in FP32
        float diff = d0 * (pW - pc);
        diff = __builtin_fmaf(d1, pE - pc, diff);
        diff = __builtin_fmaf(d2, pS - pc, diff);
        diff = __builtin_fmaf(d3, pN - pc, diff);
        diff = __builtin_fmaf(d4, pB - pc, diff);
        diff = __builtin_fmaf(d5, pF - pc, diff);
        dphi = __builtin_fmaf(ih2, diff, reac);
    } else {  // NaN/inf coefficients: the reference's literal form
        float s = d0 * pW;
        s = __builtin_fmaf(d1, pE, s);
        s = __builtin_fmaf(d2, pS, s);
        s = __builtin_fmaf(d3, pN, s);
        s = __builtin_fmaf(d4, pB, s);
        s = __builtin_fmaf(d5, pF, s);
        const float in = ih2 * s;
        const float out = -(sdf * pc * ih2);
        dphi = out + in + reac;
    }
    float pn = __builtin_fmaf(dt, dphi, pc);
    pn = (0.f < pn) ? pn : 0.f;  // std::max((Real)0, phi): NaN -> 0
    pn = (pn < 1.f) ? pn : 1.f;  // std::min((Real)1, phi)
    return pn;
}

}  // namespace

// One cross-level face-ghost row: sum_k w_k phi[src_k] as four interleaved
// FP32 partial sums (terms j, j+4, ...) combined as (s0 + s1) + (s2 + s3) --
// exactly what k_face_ghosts' four lanes and two shuffles compute, so the
// in-kernel evaluation and the separate launch give the same bits.
__device__ __forceinline__ float face_row_sum(const float *__restrict__ phi, const int64_t *__restrict__ ptr,
                                              const int *__restrict__ src, const float *__restrict__ w, int row)
{
#pragma clang fp contract(off)
    float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
    int64_t k = ptr[row];
    const int64_t e = ptr[row + 1];
    // eight terms per round trip: all index/weight loads, then all values,
    // then the FMAs in term order (term j into partial sum j mod 4)
    for (; k < e; k += 8) {
        int si[8];
        float wk[8], v[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int64_t kk = k + j < e ? k + j : k;
            si[j] = src[kk];
            wk[j] = w[kk];
        }
#pragma unroll
        for (int j = 0; j < 8; j++) v[j] = phi[si[j]];
        if (k + 8 <= e) {
            s0 = __builtin_fmaf(wk[0], v[0], s0);
            s1 = __builtin_fmaf(wk[1], v[1], s1);
            s2 = __builtin_fmaf(wk[2], v[2], s2);
            s3 = __builtin_fmaf(wk[3], v[3], s3);
            s0 = __builtin_fmaf(wk[4], v[4], s0);
            s1 = __builtin_fmaf(wk[5], v[5], s1);
            s2 = __builtin_fmaf(wk[6], v[6], s2);
            s3 = __builtin_fmaf(wk[7], v[7], s3);
        } else {
            const int m = (int)(e - k);
#pragma unroll
            for (int j = 0; j < 8; j++) {
                if (j < m) {
                    float &acc = (j & 3) == 0 ? s0 : (j & 3) == 1 ? s1 : (j & 3) == 2 ? s2 : s3;
                    acc = __builtin_fmaf(wk[j], v[j], acc);
                }
            }
        }
    }
    return (s0 + s1) + (s2 + s3);
}

// Cross-level face ghost of phi_in for the step (row index absolute).
__device__ __forceinline__ float step_face_ghost(const StepArgs &a, int row)
{
    return a.gphi ? a.gphi[row] : face_row_sum(a.phi_in, a.fptr, a.fsrc, a.fw, row);
}

// ---------------------------------------------------------------------------
// Fused RD operator + UpdateTumor.  A workgroup of 256 threads updates a slab
// of 16 x 16 x ZT cells of one block (ZT = 16: the whole block); each thread
// owns one (x,y) column of ZT cells.  phi and 1/K of the slab and its 1-cell
// halo are staged in LDS (SlabTile: 16-float rows, x halos apart).  Halo planes in z come
// from the same block when inside it, else from the z-neighbour block (same
// level) or the cross-level ghost rows; x/y halos from the face neighbours.
// ---------------------------------------------------------------------------
template <int ZT>
__device__ __forceinline__ void fk_step_item(const StepArgs &a, int item, float *s_phi, float *s_ik)
{
    constexpr int NS = B / ZT;  // slabs per block
    typedef SlabTile<ZT> T;
    const int w = item / NS;
    const int z0 = (item - w * NS) * ZT;
    const int *rec = a.wrec + (size_t)w * WREC;
    const int slot = rec[0];
    const int t = threadIdx.x;
    const int tx = t & 15, ty = t >> 4;
    const size_t base = (size_t)slot * B3;
    float Dw = a.Dw, Dg = a.Dg, rho = a.rho, dt = a.dt;
    if (a.sprm) {
        const float *q = a.sprm + 4 * rec[14];
        Dw = q[0];
        Dg = q[1];
        rho = q[2];
        dt = q[3];
    }

    float breg[ZT];
    if (a.ikc) {
        float p[ZT], c[ZT];
#pragma unroll
        for (int k = 0; k < ZT; k++) {
            const size_t i = base + (z0 + k) * 256 + t;
            p[k] = a.phi_in[i];
            c[k] = a.ikc[i];
        }
#pragma unroll
        for (int k = 0; k < ZT; k++) {
            const size_t i = base + (z0 + k) * 256 + t;
            const int l = T::cell(tx, ty, k + 1);
            float ik;
            ik_centre(c[k], p[k], a.pw, a.pg, i, Dw, Dg, ik, breg[k]);
            s_phi[l] = p[k];
            s_ik[l] = ik;
        }
    } else {
#pragma unroll
        for (int k = 0; k < ZT; k++) {
            const size_t i = base + (z0 + k) * 256 + t;
            const float p = a.phi_in[i];
            const float wm = a.pw[i];
            const float gm = a.pg[i];
            breg[k] = wm + gm;
            const int l = T::cell(tx, ty, k + 1);
            s_phi[l] = p;
            s_ik[l] = inv_k(wm, gm, Dw, Dg);
        }
    }
    // z halo planes
#pragma unroll
    for (int side = 0; side < 2; side++) {
        const int z = side == 0 ? z0 - 1 : z0 + ZT;
        const int l = T::cell(tx, ty, side == 0 ? 0 : ZT + 1);
        float pv, kv;
        if (z >= 0 && z < B) {
            const size_t i = base + z * 256 + t;
            pv = a.phi_in[i];
            kv = a.ikc ? ik_of(a.ikc[i], a.pw, a.pg, i, Dw, Dg) : inv_k(a.pw[i], a.pg[i], Dw, Dg);
        } else {
            const int f = side == 0 ? 4 : 5;
            const int nb = rec[2 + f];
            if (nb >= 0) pv = a.phi_in[(size_t)nb * B3 + nbr_face_cell(f, tx, ty)];
            else pv = step_face_ghost(a, rec[8 + f] + t);
            // the face table, contiguous per face (measured faster than the
            // neighbour's own 1/K codes, which are strided across its rows)
            kv = a.fik[((size_t)slot * 6 + f) * 256 + t];
        }
        s_phi[l] = pv;
        s_ik[l] = kv;
    }
    // x / y faces of the slab: 4 faces x 16 x ZT cells
    const int fn0 = rec[2], fn1 = rec[3], fn2 = rec[4], fn3 = rec[5];
    const int fr0 = rec[8], fr1 = rec[9], fr2 = rec[10], fr3 = rec[11];
    for (int q = t; q < 4 * 16 * ZT; q += 256) {
        const int f = q / (16 * ZT);
        const int r = q - f * 16 * ZT;
        const int u = r & 15, v = z0 + (r >> 4);
        const int nb = f == 0 ? fn0 : f == 1 ? fn1 : f == 2 ? fn2 : fn3;
        float pv;
        if (nb >= 0) pv = a.phi_in[(size_t)nb * B3 + nbr_face_cell(f, u, v)];
        else pv = step_face_ghost(a, (f == 0 ? fr0 : f == 1 ? fr1 : f == 2 ? fr2 : fr3) + u + 16 * v);
        const float kv = a.fik[((size_t)slot * 6 + f) * 256 + u + 16 * v];
        int l;
        switch (f) {
        case 0: l = T::west(u, v - z0); break;
        case 1: l = T::east(u, v - z0); break;
        case 2: l = T::cell(u, -1, v - z0 + 1); break;
        default: l = T::cell(u, B, v - z0 + 1); break;
        }
        s_phi[l] = pv;
        s_ik[l] = kv;
    }
    __syncthreads();

    const float ih2 = __int_as_float(rec[1]);
#pragma unroll
    for (int k = 0; k < ZT; k++) {
        const int c = T::cell(tx, ty, k + 1);
        const int cw = tx > 0 ? c - 1 : T::west(ty, k);
        const int ce = tx < B - 1 ? c + 1 : T::east(ty, k);
        const float pc = s_phi[c];
        const float pW = s_phi[cw], pE = s_phi[ce];
        const float pS = s_phi[c - T::RW], pN = s_phi[c + T::RW];
        const float pB = s_phi[c - T::PL], pF = s_phi[c + T::PL];
        const float pn = fk_cell(pc, pW, pE, pS, pN, pB, pF, s_ik[c], s_ik[cw], s_ik[ce], s_ik[c - T::RW],
                                 s_ik[c + T::RW], s_ik[c - T::PL], s_ik[c + T::PL], breg[k], ih2, rho, dt);
        a.phi_out[base + (z0 + k) * 256 + t] = pn;
    }
}

template <int ZT>
__global__ __launch_bounds__(256) void k_fk_step(StepArgs a)
{
    constexpr int NS = B / ZT;
    __shared__ float s_phi[SlabTile<ZT>::SIZE];
    __shared__ float s_ik[SlabTile<ZT>::SIZE];
    const int item = xcd_remap(blockIdx.x, gridDim.x);
    if (item >= a.nwork * NS) return;  // padding workgroups (whole workgroup exits)
    fk_step_item<ZT>(a, item, s_phi, s_ik);
}

// ---------------------------------------------------------------------------
// Two fused steps for a block whose 26 neighbours are all leaves of its own
// level (every block of a work list without cross-level faces: a uniform
// grid).  One workgroup of 1024 threads per block stages phi_n and 1/K over
// the block and a two-cell halo ([-2,18)^3, from the 26 neighbours), computes
// phi_{n+1} over [-1,17)^3 (the block and the one-cell halo the second step
// reads, i.e. the neighbours' own first-step values, recomputed here), then
// phi_{n+2} of the block.  Both steps use fk_cell on the same inputs as two
// single-step launches, so the result is the same to the bit; HBM sees one
// read of phi and the 1/K codes and one write per two steps.  phi_in and
// phi_out are distinct buffers (neighbours read phi_in concurrently).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_fk_step2(StepArgs a)
{
    constexpr int R = 20, R2 = R * R;  // phi_n, 1/K over [-2,18)^3
    constexpr int Q = 18, Q2 = Q * Q;  // tissue sums, phi_{n+1} over [-1,17)^3
    __shared__ float s_p0[R * R2];
    __shared__ float s_ik[R * R2];
    __shared__ float s_bs[Q * Q2];
    __shared__ float s_p1[Q * Q2];
    __shared__ int s_nb[27];
    const int item = xcd_remap(blockIdx.x, gridDim.x);
    if (item >= a.nwork) return;  // padding workgroups (whole workgroup exits)
    const int *rec = a.wrec + (size_t)item * WREC;
    const int slot = rec[0];
    const int t = threadIdx.x;
    float Dw = a.Dw, Dg = a.Dg, rho = a.rho, dt = a.dt;
    if (a.sprm) {
        const float *q = a.sprm + 4 * rec[14];
        Dw = q[0];
        Dg = q[1];
        rho = q[2];
        dt = q[3];
    }
    const float ih2 = __int_as_float(rec[1]);
    if (t < 27) {
        // region (dx,dy,dz) in {-1,0,1}^3 -> t = (dx+1) + 3(dy+1) + 9(dz+1); the
        // table leaves the centre out.  A missing neighbour (never launched so)
        // falls back to the block itself rather than an invalid slot.
        int nb = slot;
        if (t != 13) nb = a.nbr26[(size_t)slot * 26 + (t < 13 ? t : t - 1)];
        s_nb[t] = nb >= 0 ? nb : slot;
    }
    __syncthreads();
    // all loads of the staging issued before any is used
    constexpr int NST = (R * R2 + 1023) / 1024;
    float pv[NST], cv[NST];
    size_t jv[NST];
#pragma unroll
    for (int r = 0; r < NST; r++) {
        const int i = t + 1024 * r;
        jv[r] = 0;
        pv[r] = cv[r] = 0.f;
        if (i < R * R2) {
            const int x = i % R - 2, y = (i / R) % R - 2, z = i / R2 - 2;
            const int rx = x < 0 ? 0 : (x < B ? 1 : 2);
            const int ry = y < 0 ? 0 : (y < B ? 1 : 2);
            const int rz = z < 0 ? 0 : (z < B ? 1 : 2);
            jv[r] = (size_t)s_nb[rx + 3 * ry + 9 * rz] * B3 + ((x + B) & 15) + B * ((y + B) & 15) +
                    256 * ((z + B) & 15);
            pv[r] = a.phi_in[jv[r]];
            cv[r] = a.ikc[jv[r]];
        }
    }
#pragma unroll
    for (int r = 0; r < NST; r++) {
        const int i = t + 1024 * r;
        if (i < R * R2) {
            const int lx = i % R, ly = (i / R) % R, lz = i / R2;
            float ik, bs;
            ik_centre(cv[r], pv[r], a.pw, a.pg, jv[r], Dw, Dg, ik, bs);
            s_p0[i] = pv[r];
            s_ik[i] = ik;
            if (lx >= 1 && lx <= Q && ly >= 1 && ly <= Q && lz >= 1 && lz <= Q)
                s_bs[(lx - 1) + Q * (ly - 1) + Q2 * (lz - 1)] = bs;
        }
    }
    __syncthreads();
    for (int i = t; i < Q * Q2; i += 1024) {
        const int qx = i % Q, qy = (i / Q) % Q, qz = i / Q2;
        const int c = (qx + 1) + R * (qy + 1) + R2 * (qz + 1);
        s_p1[i] = fk_cell(s_p0[c], s_p0[c - 1], s_p0[c + 1], s_p0[c - R], s_p0[c + R], s_p0[c - R2], s_p0[c + R2],
                          s_ik[c], s_ik[c - 1], s_ik[c + 1], s_ik[c - R], s_ik[c + R], s_ik[c - R2], s_ik[c + R2],
                          s_bs[i], ih2, rho, dt);
    }
    __syncthreads();
    const size_t base = (size_t)slot * B3;
#pragma unroll
    for (int r = 0; r < B3 / 1024; r++) {
        const int i = t + 1024 * r;
        const int x = i & 15, y = (i >> 4) & 15, z = i >> 8;
        const int c1 = (x + 1) + Q * (y + 1) + Q2 * (z + 1);
        const int c0 = (x + 2) + R * (y + 2) + R2 * (z + 2);
        a.phi_out[base + i] =
            fk_cell(s_p1[c1], s_p1[c1 - 1], s_p1[c1 + 1], s_p1[c1 - Q], s_p1[c1 + Q], s_p1[c1 - Q2], s_p1[c1 + Q2],
                    s_ik[c0], s_ik[c0 - 1], s_ik[c0 + 1], s_ik[c0 - R], s_ik[c0 + R], s_ik[c0 - R2], s_ik[c0 + R2],
                    s_bs[c1], ih2, rho, dt);
    }
}

// Cross-level face ghosts of phi for the step: four lanes per row, partial
// sums combined with lane shuffles.
// g: global lane (four per row); every lane of a wave calls it (shuffles).
__device__ __forceinline__ void face_ghost_lanes(const float *__restrict__ phi, const int64_t *__restrict__ ptr,
                                                 const int *__restrict__ src, const float *__restrict__ w,
                                                 float *__restrict__ out, int nrows, int g)
{
#pragma clang fp contract(off)
    const int r = g >> 2, lane = g & 3;
    float acc = 0.f;
    if (r < nrows) {
        // terms lane, lane+4, ... in order; four of them loaded per round trip
        const int64_t e = ptr[r + 1];
        for (int64_t k = ptr[r] + lane; k < e; k += 16) {
            int si[4];
            float wk[4], v[4];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int64_t kk = k + 4 * j < e ? k + 4 * j : k;
                si[j] = src[kk];
                wk[j] = w[kk];
            }
#pragma unroll
            for (int j = 0; j < 4; j++) v[j] = phi[si[j]];
#pragma unroll
            for (int j = 0; j < 4; j++)
                if (k + 4 * j < e) acc = __builtin_fmaf(wk[j], v[j], acc);
        }
    }
    acc += __shfl_xor(acc, 1, 64);
    acc += __shfl_xor(acc, 2, 64);
    if (r < nrows && lane == 0) out[r] = acc;
}

__global__ __launch_bounds__(256) void k_face_ghosts(const float *__restrict__ phi, const int64_t *__restrict__ ptr,
                                                     const int *__restrict__ src, const float *__restrict__ w,
                                                     float *__restrict__ out, int nrows)
{
    face_ghost_lanes(phi, ptr, src, w, out, nrows, blockIdx.x * blockDim.x + threadIdx.x);
}

// ---------------------------------------------------------------------------
// Persistent chunk of steps: one cooperative launch advances m steps.  Each
// step is two phases separated by grid-wide barriers (cooperative groups):
// the cross-level face ghosts of phi_in (four lanes per row, as
// k_face_ghosts), then the step of every work item of that step's prefix
// (fk_step_item, as k_fk_step<ZT>), so every value equals the per-step
// launches' to the bit.  The workgroups stay resident for the whole chunk
// (the launch fails rather than oversubscribe the device); each XCD walks a
// contiguous range of the work items, the same range every step.
// ---------------------------------------------------------------------------
template <int ZT>
__global__ __launch_bounds__(256) void k_fk_persistent(PersistArgs p)
{
    constexpr int NS = B / ZT;
    __shared__ float s_phi[SlabTile<ZT>::SIZE];
    __shared__ float s_ik[SlabTile<ZT>::SIZE];
    cooperative_groups::grid_group grid = cooperative_groups::this_grid();
    const int per = gridDim.x >> 3;  // workgroups per XCD (the launcher uses a multiple of 8)
    const int xcd = blockIdx.x & 7, rank = blockIdx.x >> 3;
    int c = p.c0;
    for (int j = 0; j < p.m; j++) {
        const int nrows = p.nrows[j];
        if (nrows > 0 && p.a.gphi) {  // uniform over the grid
            for (int g0 = blockIdx.x * 256; g0 < 4 * nrows; g0 += gridDim.x * 256)
                face_ghost_lanes(p.phi[c], p.a.fptr, p.a.fsrc, p.a.fw, p.gout, nrows, g0 + (int)threadIdx.x);
            __threadfence();
            grid.sync();
        }
        StepArgs a = p.a;
        a.phi_in = p.phi[c];
        a.phi_out = p.phi[c ^ 1];
        a.nwork = p.nwork[j];
        const int items = a.nwork * NS;
        const int lo = (int)((int64_t)items * xcd / 8), hi = (int)((int64_t)items * (xcd + 1) / 8);
        for (int it = lo + rank; it < hi; it += per) {
            fk_step_item<ZT>(a, it, s_phi, s_ik);
            __syncthreads();  // the LDS tiles are reused by the next item
        }
        __threadfence();
        grid.sync();
        c ^= 1;
    }
}

// 1/K of every face-halo cell and, when ikc is set, the per-cell step
// coefficients of the block (static between topology / parameter changes).
__global__ __launch_bounds__(256) void k_face_prep(const float *__restrict__ pw, const float *__restrict__ pg,
                                                   const int *__restrict__ work, int nwork,
                                                   const int *__restrict__ fnbr, const int *__restrict__ fbase,
                                                   const int *__restrict__ sfbase,
                                                   const int *__restrict__ slot_sample,
                                                   const int64_t *__restrict__ ptr, const int *__restrict__ src,
                                                   const float *__restrict__ w, float *__restrict__ fik,
                                                   float *__restrict__ ikc, float Dw, float Dg,
                                                   const float *__restrict__ sprm)
{
    // blockIdx.y: 0-5 the face halos, 6 the block's codes (independent
    // chains of loads in separate workgroups)
    const int wi = blockIdx.x;
    if (wi >= nwork) return;
    const int slot = work[wi];
    if (sprm) {
        Dw = sprm[4 * slot_sample[slot]];
        Dg = sprm[4 * slot_sample[slot] + 1];
    }
    const int t = threadIdx.x;
    const int f = blockIdx.y;
    if (f < 6) {
        const int nb = fnbr[slot * 6 + f];
        float wm, gm;
        if (nb >= 0) {
            const size_t i = (size_t)nb * B3 + nbr_face_cell(f, t & 15, t >> 4);
            wm = pw[i];
            gm = pg[i];
        } else {
            const int r = sfbase[slot_sample[slot]] + fbase[slot * 6 + f] + t;
            wm = 0.f;
            gm = 0.f;
            // terms in order, eight loaded per round trip
            const int64_t e = ptr[r + 1];
            for (int64_t k = ptr[r]; k < e; k += 8) {
                int si[8];
                float wk[8], a[8], b[8];
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const int64_t kk = k + j < e ? k + j : k;
                    si[j] = src[kk];
                    wk[j] = w[kk];
                }
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    a[j] = pw[si[j]];
                    b[j] = pg[si[j]];
                }
#pragma unroll
                for (int j = 0; j < 8; j++)
                    if (k + j < e) {
                        wm = __builtin_fmaf(wk[j], a[j], wm);
                        gm = __builtin_fmaf(wk[j], b[j], gm);
                    }
            }
        }
        fik[((size_t)slot * 6 + f) * 256 + t] = inv_k(wm, gm, Dw, Dg);
    } else if (ikc) {
        const size_t b0 = (size_t)slot * B3;
        float wv[16], gv[16];
#pragma unroll
        for (int k = 0; k < 16; k++) {
            wv[k] = pw[b0 + k * 256 + t];
            gv[k] = pg[b0 + k * 256 + t];
        }
#pragma unroll
        for (int k = 0; k < 16; k++) ikc[b0 + k * 256 + t] = ik_code(wv[k], gv[k], Dw, Dg);
    }
}

// ---------------------------------------------------------------------------
// The same fused step streaming z: one workgroup per block, thread = one (x,y)
// column of 16 cells.  The column's phi and 1/K of planes z-1, z, z+1 live in
// registers; only the current plane (with its x/y halo) is staged in LDS,
// double-buffered (one barrier per plane), so a workgroup needs 5 KB of LDS
// and many blocks stay in flight.  The arithmetic per cell is k_fk_step's.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_fk_step_col(StepArgs a)
{
    __shared__ float s_phi[2][PLW * PLW];
    __shared__ float s_ik[2][PLW * PLW];
    const int item = xcd_remap(blockIdx.x, gridDim.x);
    if (item >= a.nwork) return;
    const int *rec = a.wrec + (size_t)item * WREC;
    const int slot = rec[0];
    const int t = threadIdx.x;
    const int tx = t & 15, ty = t >> 4;
    const size_t base = (size_t)slot * B3;
    float Dw = a.Dw, Dg = a.Dg, rho = a.rho, dt = a.dt;
    if (a.sprm) {
        const float *q = a.sprm + 4 * rec[14];
        Dw = q[0];
        Dg = q[1];
        rho = q[2];
        dt = q[3];
    }
    // halo role of this thread for the x/y faces: t < 64 -> face t/16, coordinate t%16
    const int hf = t >> 4, hu = t & 15;
    int hnb = -1, hrow = 0;
    if (t < 64) {
        hnb = hf == 0 ? rec[2] : hf == 1 ? rec[3] : hf == 2 ? rec[4] : rec[5];
        if (hnb < 0) hrow = hf == 0 ? rec[8] : hf == 1 ? rec[9] : hf == 2 ? rec[10] : rec[11];
    }
    // z-1 plane of this column
    float pB, kB;
    {
        const int nb = rec[6];
        if (nb >= 0) pB = a.phi_in[(size_t)nb * B3 + nbr_face_cell(4, tx, ty)];
        else pB = step_face_ghost(a, rec[12] + t);
        kB = a.fik[((size_t)slot * 6 + 4) * 256 + t];
    }
    // 1/K and p_w + p_g of plane z of this column, from the precomputed
    // coefficients (code = the loaded value) or from pw, pg (w, g)
    const bool pre = a.ikc != nullptr;
    auto resolve = [&](int z, float p, float cw, float g, float &ik, float &bs) {
        if (pre) ik_centre(cw, p, a.pw, a.pg, base + z * 256 + t, Dw, Dg, ik, bs);
        else {
            ik = inv_k(cw, g, Dw, Dg);
            bs = cw + g;
        }
    };
    float pC = a.phi_in[base + t];
    float bC;
    float kC;
    if (pre) resolve(0, pC, a.ikc[base + t], 0.f, kC, bC);
    else resolve(0, pC, a.pw[base + t], a.pg[base + t], kC, bC);
    // z+1 plane (register prefetch): wF holds the code when the coefficients are precomputed
    float pF = a.phi_in[base + 256 + t];
    float wF = pre ? a.ikc[base + 256 + t] : a.pw[base + 256 + t];
    float gF = pre ? 0.f : a.pg[base + 256 + t];
    const float ih2 = __int_as_float(rec[1]);
    for (int k = 0; k < B; k++) {
        const int buf = k & 1;
        float *sp = s_phi[buf];
        float *sk = s_ik[buf];
        const int c = (tx + 1) + PLW * (ty + 1);
        sp[c] = pC;
        sk[c] = kC;
        if (t < 64) {
            float pv;
            if (hnb >= 0) pv = a.phi_in[(size_t)hnb * B3 + nbr_face_cell(hf, hu, k)];
            else pv = step_face_ghost(a, hrow + hu + 16 * k);
            const float kv = a.fik[((size_t)slot * 6 + hf) * 256 + hu + 16 * k];
            int l;
            switch (hf) {
            case 0: l = 0 + PLW * (hu + 1); break;
            case 1: l = (B + 1) + PLW * (hu + 1); break;
            case 2: l = (hu + 1); break;
            default: l = (hu + 1) + PLW * (B + 1); break;
            }
            sp[l] = pv;
            sk[l] = kv;
        }
        // plane k+1's 1/K for this column (pF, wF, gF were loaded one plane ahead)
        float kF, bF = 0.f;
        if (k + 1 < B) resolve(k + 1, pF, wF, gF, kF, bF);
        else {
            const int nb = rec[7];
            if (nb >= 0) pF = a.phi_in[(size_t)nb * B3 + nbr_face_cell(5, tx, ty)];
            else pF = step_face_ghost(a, rec[13] + t);
            kF = a.fik[((size_t)slot * 6 + 5) * 256 + t];
        }
        // prefetch plane k+2
        float pN = 0.f, wN = 0.f, gN = 0.f;
        if (k + 2 < B) {
            const size_t i = base + (k + 2) * 256 + t;
            pN = a.phi_in[i];
            if (pre) wN = a.ikc[i];
            else {
                wN = a.pw[i];
                gN = a.pg[i];
            }
        }
        __syncthreads();
        const float pc = pC;
        const float pW = sp[c - 1], pE = sp[c + 1];
        const float pS = sp[c - PLW], pN2 = sp[c + PLW];
        const float pn = fk_cell(pc, pW, pE, pS, pN2, pB, pF, kC, sk[c - 1], sk[c + 1], sk[c - PLW], sk[c + PLW],
                                 kB, kF, bC, ih2, rho, dt);
        a.phi_out[base + k * 256 + t] = pn;
        // shift the column window
        pB = pC;
        kB = kC;
        pC = pF;
        kC = kF;
        bC = bF;
        pF = pN;
        wF = wN;
        gF = gN;
    }
}

// ---------------------------------------------------------------------------
// Ghost points across levels, evaluated on the device.
//
// A point p (global index at level lam, vertex-centred) takes the value MRAG's
// ghost machinery assigns it (MRAG_BBInfoCreator.inl, MRAG_SmartBlockFinder.h):
// the leaf whose region holds it is looked up; same level -> its value;
// finer leaf -> the level lam+1 point 2p (Ha = {1}); coarser leaf ->
// sum_i Hs(2i - p) * value(lam-1, i), tensor product, zero weights skipped.
// The recursion is walked depth-first with the synthesis children in
// increasing (z, y, x) order and child weight w * ((hx*hy)*hz); values are
// accumulated acc + w*v in that order (FP64 round-to-nearest, no
// contraction), the same order as the CPU oracle.
// ---------------------------------------------------------------------------
namespace {
constexpr int MAXD = 9;

__device__ __forceinline__ double hs_w(int order, int i)
{
    if (order == 2) return (i < -1 || i > 1) ? 0.0 : (i == 0 ? 1.0 : 0.5);
    switch (i) {
    case -3: case 3: return -1.0 / 16.0;
    case -1: case 1: return 9.0 / 16.0;
    case 0: return 1.0;
    default: return 0.0;
    }
}

// Look-up entry of the leaf holding the level-lam point p: its slot and its
// level in one word (LUT_LEVEL_SHIFT), so resolving a point costs one load.
__device__ __forceinline__ int lying_entry(const LabArgs &g, const int *__restrict__ lut, int lam, int64_t px,
                                           int64_t py, int64_t pz)
{
    // finest block cell: floor(p * 2^(L-lam) / 16), periodic (power-of-two wrap)
    const int sh = g.lut_level - lam;
    const int64_t m = (1LL << g.lut_level) - 1;
    const int64_t fx = ((px << sh) >> 4) & m, fy = ((py << sh) >> 4) & m, fz = ((pz << sh) >> 4) & m;
    return lut[fx + (m + 1) * (fy + (m + 1) * fz)];
}
__device__ __forceinline__ int entry_slot(int e) { return e & LUT_SLOT_MASK; }
__device__ __forceinline__ int entry_level(int e) { return e >> LUT_LEVEL_SHIFT; }

struct DfsFrame {
    int lam;
    int64_t p[3];
    double w;
    int64_t s[3], e[3], i[3];
};

// Depth-first walk from (lam0, p0, w0).  mode 0: acc += w*v over the leaf
// values of fld; 1: count terms (added to *count); 2: emit terms at src/wout
// (advancing *count).
template <int MODE, int NF = 1>
__device__ void ghost_dfs(const LabArgs &g, const int *__restrict__ lut, const float *const *flds, int lam0,
                          int64_t p0x, int64_t p0y, int64_t p0z, double w0, double *acc, int *count, int *src,
                          float *wout)
{
    const int lo = g.order == 4 ? -3 : -1, hi = g.order == 4 ? 4 : 2;
    DfsFrame st[MAXD];
    int sp = 0;
    int lam = lam0;
    int64_t p[3] = {p0x, p0y, p0z};
    double w = w0;
    while (true) {
        // resolve (lam, p, w) to a value or a synthesis frame
        while (true) {
            const int e = lying_entry(g, lut, lam, p[0], p[1], p[2]);
            const int slot = entry_slot(e), l = entry_level(e);
            if (l == lam) {
                const int64_t msk = ((int64_t)16 << lam) - 1;
                const int cx = (int)((p[0] & msk) - (int64_t)g.index[slot * 3 + 0] * 16);
                const int cy = (int)((p[1] & msk) - (int64_t)g.index[slot * 3 + 1] * 16);
                const int cz = (int)((p[2] & msk) - (int64_t)g.index[slot * 3 + 2] * 16);
                const int cell = cx + 16 * (cy + 16 * cz);
                if (MODE == 0) {
#pragma unroll
                    for (int f = 0; f < NF; f++)
                        acc[f] = __dadd_rn(acc[f], __dmul_rn(w, (double)flds[f][(size_t)slot * 4096 + cell]));
                }
                else if (MODE == 2) {
                    src[*count] = slot * 4096 + cell;
                    wout[*count] = (float)w;
                }
                if (MODE != 0) (*count)++;
                break;
            }
            if (l > lam) {  // finer leaf: analysis
                lam++;
                p[0] *= 2;
                p[1] *= 2;
                p[2] *= 2;
                continue;
            }
            // coarser leaf: synthesis frame (ceil((p+lo)/2) .. floor((p+hi-1)/2))
            if (sp >= MAXD) return;  // unreachable for 2:1 balanced grids
            DfsFrame &f = st[sp++];
            f.lam = lam;
            f.w = w;
            for (int d = 0; d < 3; d++) {
                f.p[d] = p[d];
                f.s[d] = (p[d] + lo + 1) >> 1;       // ceil((p+lo)/2) with arithmetic shift
                f.e[d] = ((p[d] + hi - 1) >> 1) + 1; // 1 + floor((p+hi-1)/2)
                f.i[d] = f.s[d];
            }
            f.i[0] = f.s[0] - 1;  // before the first child
            break;
        }
        // next child with a non-zero weight
        bool found = false;
        while (sp > 0 && !found) {
            DfsFrame &f = st[sp - 1];
            while (true) {
                if (++f.i[0] >= f.e[0]) {
                    f.i[0] = f.s[0];
                    if (++f.i[1] >= f.e[1]) {
                        f.i[1] = f.s[1];
                        if (++f.i[2] >= f.e[2]) break;
                    }
                }
                const double ww = __dmul_rn(__dmul_rn(hs_w(g.order, (int)(2 * f.i[0] - f.p[0])),
                                                      hs_w(g.order, (int)(2 * f.i[1] - f.p[1]))),
                                            hs_w(g.order, (int)(2 * f.i[2] - f.p[2])));
                if (ww == 0.0) continue;
                lam = f.lam - 1;
                p[0] = f.i[0];
                p[1] = f.i[1];
                p[2] = f.i[2];
                w = __dmul_rn(f.w, ww);
                found = true;
                break;
            }
            if (!found) sp--;
        }
        if (!found) break;
    }
}

// Synthesis children of the level-lam point p (in a coarser leaf): the
// children with a non-zero weight w * ((hx*hy)*hz), in increasing (z, y, x)
// order (per axis: p odd -> the 4 (Interp4) or 2 (Interp2) points around
// p/2, p even -> p/2 alone; the other points of the support have weight 0
// and are skipped, as in the general walk).  A batch of children at a time:
// their look-ups, levels and values are loaded together and, when every child
// of the batch resolves directly, consumed in order from registers; otherwise the
// remaining children are walked one by one with the general walk where needed
// (same order, same arithmetic).  A point in a leaf of its own level sits at
// cell (p mod 16) per axis.
template <int MODE, int NF>
__device__ void synth_children(const LabArgs &g, const int *__restrict__ lut, const float *const *flds, int lam,
                               const int64_t p[3], int ep, double w, double *acc, int *count, int *src, float *wout)
{
    const int cl = lam - 1;
    // ep: look-up entry of the coarser leaf C holding p.  When C has level
    // cl, a child inside C's box is a point of C itself (cell = i mod 16 per
    // axis) and needs no look-up; C's block index at level cl is p >> 5 per
    // axis (p / 2 / 16), compared modulo the 2^cl blocks of that level.
    const bool c_own = entry_level(ep) == cl;
    const int64_t mcl = ((int64_t)1 << cl) - 1;
    const int64_t cbx = (p[0] >> 5) & mcl, cby = (p[1] >> 5) & mcl, cbz = (p[2] >> 5) & mcl;
    // per axis: p odd -> 4 (Interp4) or 2 (Interp2) children from i0, p even
    // -> the child p/2 alone; child k of an axis has weight Hs(2(i0 + k) - p)
    // (Interp4: Hs(-3), Hs(-1), Hs(1), Hs(3); Interp2: Hs(-1), Hs(1); even: Hs(0))
    int64_t i0[3];
    int ln[3];  // log2 of the children per axis
    bool odd[3];
    for (int d = 0; d < 3; d++) {
        odd[d] = (p[d] & 1) != 0;
        ln[d] = odd[d] ? (g.order == 4 ? 2 : 1) : 0;
        i0[d] = odd[d] ? (p[d] - (g.order == 4 ? 3 : 1)) >> 1 : p[d] >> 1;
    }
    const int nc = 1 << (ln[0] + ln[1] + ln[2]);
    auto hw = [&](int d, int k) -> double {
        if (!odd[d]) return 1.0;
        if (g.order != 4) return 0.5;
        return (k == 0 || k == 3) ? -1.0 / 16.0 : 9.0 / 16.0;
    };
    // children per batch (register pressure of the loads and stores): a
    // batch resolves with two rounds of loads, look-up entries (slot and
    // level in one word) and then values
    constexpr int NB = MODE == 1 ? 8 : 4;
    for (int c0 = 0; c0 < nc; c0 += NB) {
        // a child i at level cl lies in a leaf of level cl (value at i) or of a
        // finer level (the walk descends once: the point 2i at level cl+1,
        // normally in a leaf of that level); both look-ups are issued at once
        double cw[NB];
        int sa[NB], sb[NB], ca[NB], cb[NB];
#pragma unroll
        for (int k = 0; k < NB; k++) {
            const int c = c0 + k < nc ? c0 + k : c0;
            const int kx = c & ((1 << ln[0]) - 1), ky = (c >> ln[0]) & ((1 << ln[1]) - 1), kz = c >> (ln[0] + ln[1]);
            const int64_t ix = i0[0] + kx, iy = i0[1] + ky, iz = i0[2] + kz;
            const double ww = __dmul_rn(__dmul_rn(hw(0, kx), hw(1, ky)), hw(2, kz));
            cw[k] = __dmul_rn(w, ww);
            const bool in = c0 + k < nc;
            const bool in_c = c_own && ((ix >> 4) & mcl) == cbx && ((iy >> 4) & mcl) == cby && ((iz >> 4) & mcl) == cbz;
            sa[k] = in ? (in_c ? ep : lying_entry(g, lut, cl, ix, iy, iz)) : 0;
            sb[k] = in && !in_c ? lying_entry(g, lut, cl + 1, 2 * ix, 2 * iy, 2 * iz) : 0;
            ca[k] = (int)((ix & 15) + 16 * ((iy & 15) + 16 * (iz & 15)));
            cb[k] = (int)(((2 * ix) & 15) + 16 * (((2 * iy) & 15) + 16 * ((2 * iz) & 15)));
        }
        int sl[NB], cell[NB];
        bool plain = true;  // every child resolved by at most one descent
#pragma unroll
        for (int k = 0; k < NB; k++) {
            const int la = entry_level(sa[k]), lb = entry_level(sb[k]);
            const bool fine = la > cl;
            sl[k] = entry_slot(fine ? sb[k] : sa[k]);
            cell[k] = fine ? cb[k] : ca[k];
            plain = plain && (c0 + k >= nc || la == cl || (fine && lb == cl + 1));
        }
        float val[NB][NF];
#pragma unroll
        for (int k = 0; k < NB; k++)
#pragma unroll
            for (int f = 0; f < NF; f++)
                val[k][f] = MODE == 0 && c0 + k < nc ? flds[f][(size_t)sl[k] * 4096 + cell[k]] : 0.f;
        if (!plain) {
            for (int c = c0; c < nc; c++) {
                const int kx = c & ((1 << ln[0]) - 1), ky = (c >> ln[0]) & ((1 << ln[1]) - 1),
                          kz = c >> (ln[0] + ln[1]);
                const int64_t ix = i0[0] + kx, iy = i0[1] + ky, iz = i0[2] + kz;
                const double ww = __dmul_rn(__dmul_rn(hw(0, kx), hw(1, ky)), hw(2, kz));
                const double cwc = __dmul_rn(w, ww);
                const int e = lying_entry(g, lut, cl, ix, iy, iz);
                const int slot = entry_slot(e);
                if (entry_level(e) == cl) {
                    const int ce = (int)((ix & 15) + 16 * ((iy & 15) + 16 * (iz & 15)));
                    if (MODE == 0) {
#pragma unroll
                        for (int f = 0; f < NF; f++)
                            acc[f] = __dadd_rn(acc[f], __dmul_rn(cwc, (double)flds[f][(size_t)slot * 4096 + ce]));
                    } else {
                        if (MODE == 2) {
                            src[*count] = slot * 4096 + ce;
                            wout[*count] = (float)cwc;
                        }
                        (*count)++;
                    }
                } else {
                    ghost_dfs<MODE, NF>(g, lut, flds, cl, ix, iy, iz, cwc, acc, count, src, wout);
                }
            }
            return;
        }
#pragma unroll
        for (int k = 0; k < NB; k++) {
            if (c0 + k >= nc) break;
            if (MODE == 0) {
#pragma unroll
                for (int f = 0; f < NF; f++) acc[f] = __dadd_rn(acc[f], __dmul_rn(cw[k], (double)val[k][f]));
            } else {
                if (MODE == 2) {
                    src[*count] = sl[k] * 4096 + cell[k];
                    wout[*count] = (float)cw[k];
                }
                (*count)++;
            }
        }
    }
}

// Value of one ghost point: the same walk with a stack-free fast path for
// the common cases (the point lies in a finer leaf, or in a coarser leaf
// whose synthesis children lie in leaves of their own level or one level
// finer); anything else continues with the general walk at the child, in the
// same order.
template <int NF>
__device__ void ghost_value(const LabArgs &g, const int *__restrict__ lut, const float *const *flds, int lam,
                            int64_t px, int64_t py, int64_t pz, double *acc)
{
    int e;
    for (;;) {
        e = lying_entry(g, lut, lam, px, py, pz);
        const int slot = entry_slot(e), l = entry_level(e);
        if (l == lam) {
            const size_t i = (size_t)slot * 4096 + (px & 15) + 16 * ((py & 15) + 16 * (pz & 15));
#pragma unroll
            for (int f = 0; f < NF; f++) acc[f] = __dmul_rn(1.0, (double)flds[f][i]);
            return;
        }
        if (l > lam) {
            lam++;
            px *= 2;
            py *= 2;
            pz *= 2;
            continue;
        }
        break;
    }
#pragma unroll
    for (int f = 0; f < NF; f++) acc[f] = 0.0;
    const int64_t p[3] = {px, py, pz};
    synth_children<0, NF>(g, lut, flds, lam, p, e, 1.0, acc, nullptr, nullptr, nullptr);
}

template <int S0, int S1>
struct LabGeom {
    static constexpr int L = B + S1 - S0 - 1;
    static constexpr int N = L * L * L;
};

// Lab coordinates a consumer reads outside the block: the FWT filters touch
// only {-2, 16, 18} (Interp4; the taps at odd offsets -3 and +1 are zero) or
// {16} (Interp2); the split reads [-1, 17] (Interp4) or [0, 16] (Interp2).
template <int ORDER, bool FOR_SPLIT>
__device__ __forceinline__ bool lab_coord_used(int v)
{
    if (v >= 0 && v < B) return true;
    if (FOR_SPLIT) return ORDER == 4 ? (v >= -1 && v <= B + 1) : (v >= 0 && v <= B);
    return ORDER == 4 ? (v == -2 || v == B || v == B + 2) : (v == B);
}

// Lab coordinate of index i of the coordinates a consumer reads (interior
// and the ghost coordinates of lab_coord_used, increasing).
template <int ORDER, bool FOR_SPLIT>
struct UsedCoords {
    static constexpr int n = ORDER == 4 ? 19 : 17;
    __device__ static int at(int i)
    {
        if (ORDER == 2) return i;
        if (FOR_SPLIT) return i - 1;
        return i == 0 ? -2 : (i <= 16 ? i - 1 : (i == 17 ? B : B + 2));
    }
};

// Lab points an FWT reads (the block and the used ghost coordinates), FP64,
// from the block, its same-level neighbours or the evaluated cross-level
// ghosts.  Each thread gathers K points with all loads issued before any is
// consumed (one memory round trip per K points, not per point); unread lab
// points are left unset.
template <int ORDER>
__device__ void fill_lab_used(double *lab, const float *__restrict__ fld, const double *__restrict__ ghost, int slot,
                              const LabArgs &a)
{
    constexpr int S0 = ORDER == 4 ? -2 : -1;
    constexpr int S1 = ORDER == 4 ? 4 : 2;
    constexpr int L = LabGeom<S0, S1>::L;
    constexpr int NU = UsedCoords<ORDER, false>::n;
    constexpr int NQ = NU * NU * NU;
    constexpr int K = 8;
    __shared__ int s_nb[26];
    if (threadIdx.x < 26) s_nb[threadIdx.x] = a.nbr26[slot * 26 + threadIdx.x];
    __syncthreads();
    const size_t sbase = (size_t)slot * B3;
    for (int q0 = threadIdx.x; q0 < NQ; q0 += K * (int)blockDim.x) {
        size_t fi[K];
        int gi[K], li[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int q = q0 + k * (int)blockDim.x;
            const int qq = q < NQ ? q : 0;
            const int x = UsedCoords<ORDER, false>::at(qq % NU), y = UsedCoords<ORDER, false>::at((qq / NU) % NU),
                      z = UsedCoords<ORDER, false>::at(qq / (NU * NU));
            const int dx = x < 0 ? -1 : (x < B ? 0 : 1), dy = y < 0 ? -1 : (y < B ? 0 : 1),
                      dz = z < 0 ? -1 : (z < B ? 0 : 1);
            const int l = (x - S0) + L * ((y - S0) + L * (z - S0));
            li[k] = q < NQ ? l : -1;
            fi[k] = sbase;
            gi[k] = -1;
            if ((dx | dy | dz) == 0) {
                fi[k] = sbase + x + B * (y + B * z);
            } else {
                const int c = (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1);
                const int nb = s_nb[c < 13 ? c : c - 1];
                if (nb >= 0) fi[k] = (size_t)nb * B3 + (x - dx * B) + B * ((y - dy * B) + B * (z - dz * B));
                else gi[k] = l;
            }
        }
        float fv[K];
        double gv[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            fv[k] = fld[fi[k]];
            gv[k] = ghost[gi[k] >= 0 ? gi[k] : 0];
        }
#pragma unroll
        for (int k = 0; k < K; k++)
            if (li[k] >= 0) lab[li[k]] = gi[k] >= 0 ? gv[k] : (double)fv[k];
    }
}
}  // namespace

// Index range of the used lab coordinates of one side of an axis (d = -1:
// below the block, 0: inside, +1: above) in UsedCoords order.
template <int ORDER>
__device__ __forceinline__ void used_range(int d, int &start, int &count)
{
    if (ORDER == 4) {
        start = d < 0 ? 0 : (d == 0 ? 1 : 17);
        count = d < 0 ? 1 : (d == 0 ? 16 : 2);
    } else {
        start = d < 0 ? 0 : (d == 0 ? 0 : 16);
        count = d < 0 ? 0 : (d == 0 ? 16 : 1);
    }
}

// The lab points a consumer reads in the cross-level regions of the listed
// slots.  One 64-thread workgroup per item = (listed slot j, region r, chunk
// of 64 points): j << 8 | r << 3 | chunk, the 26 regions around the block
// with the centre left out; the host lists only regions whose neighbour is
// not a same-level leaf (fill_lab_used and k_split copy the others directly).
// Each thread evaluates one point (NF fields in one walk).
template <int ORDER, bool FOR_SPLIT, int NF>
__global__ __launch_bounds__(64) void k_lab_ghosts(LabArgs a, FieldPtrs flds, const int *__restrict__ slots,
                                                    const int *__restrict__ items, double *__restrict__ ghost,
                                                    size_t field_stride)
{
    constexpr int S0 = ORDER == 4 ? -2 : -1;
    constexpr int S1 = ORDER == 4 ? 4 : 2;
    constexpr int L = LabGeom<S0, S1>::L;
    const int it = items[blockIdx.x];
    const int j = it >> 8, r = (it >> 3) & 31;
    const int c = r < 13 ? r : r + 1;  // region code (dx+1) + 3(dy+1) + 9(dz+1)
    int sx, nx, sy, ny, sz, nz;
    used_range<ORDER>(c % 3 - 1, sx, nx);
    used_range<ORDER>((c / 3) % 3 - 1, sy, ny);
    used_range<ORDER>(c / 9 - 1, sz, nz);
    const int npts = nx * ny * nz;
    const int slot = slots[j];
    const int64_t ox = (int64_t)a.index[slot * 3 + 0] * B, oy = (int64_t)a.index[slot * 3 + 1] * B,
                  oz = (int64_t)a.index[slot * 3 + 2] * B;
    const int *lut = a.lut + a.lutoff[slot];
    const int lvl = a.level[slot];
    const int q = (it & 7) * 64 + (int)threadIdx.x;
    if (q < npts) {
        // the counts per axis are 1, 2 or 16: shifts instead of divisions
        const int lx = __builtin_ctz(nx), ly = __builtin_ctz(ny);
        const int x = UsedCoords<ORDER, FOR_SPLIT>::at(sx + (q & (nx - 1))),
                  y = UsedCoords<ORDER, FOR_SPLIT>::at(sy + ((q >> lx) & (ny - 1))),
                  z = UsedCoords<ORDER, FOR_SPLIT>::at(sz + (q >> (lx + ly)));
        double v[NF];
        ghost_value<NF>(a, lut, flds.p, lvl, ox + x, oy + y, oz + z, v);
        const size_t o = (size_t)j * LabGeom<S0, S1>::N + (x - S0) + L * ((y - S0) + L * (z - S0));
#pragma unroll
        for (int f = 0; f < NF; f++) ghost[f * field_stride + o] = v[f];
    }
}

namespace {
// Terms of one ghost point (mode 1: count, 2: emit (slot*4096+cell, weight)),
// in the order of the depth-first walk, with the stack-free fast path of
// ghost_value for children that lie in leaves of their own level.
template <int MODE>
__device__ void ghost_terms(const LabArgs &g, const int *__restrict__ lut, int lam, int64_t px, int64_t py,
                            int64_t pz, int *count, int *src, float *wout)
{
    int e;
    for (;;) {
        e = lying_entry(g, lut, lam, px, py, pz);
        const int slot = entry_slot(e), l = entry_level(e);
        if (l == lam) {
            if (MODE == 2) {
                src[*count] = slot * 4096 + (int)((px & 15) + 16 * ((py & 15) + 16 * (pz & 15)));
                wout[*count] = 1.0f;
            }
            (*count)++;
            return;
        }
        if (l > lam) {
            lam++;
            px *= 2;
            py *= 2;
            pz *= 2;
            continue;
        }
        break;
    }
    const int64_t p[3] = {px, py, pz};
    double unused = 0.0;
    synth_children<MODE, 1>(g, lut, nullptr, lam, p, e, 1.0, &unused, count, src, wout);
}
}  // namespace

// Step rows of cross-level faces (count / fill passes).
__device__ __forceinline__ void face_point(int f, int u, int v, int &x, int &y, int &z)
{
    switch (f) {
    case 0: x = -1; y = u; z = v; break;
    case 1: x = B; y = u; z = v; break;
    case 2: x = u; y = -1; z = v; break;
    case 3: x = u; y = B; z = v; break;
    case 4: x = u; y = v; z = -1; break;
    default: x = u; y = v; z = B; break;
    }
}

__global__ __launch_bounds__(256) void k_face_rows_count(LabArgs a, const int *__restrict__ faces, int nfaces,
                                                         int *__restrict__ counts)
{
    const int k = blockIdx.x;
    const int sf = faces[k];
    const int slot = sf / 6, f = sf % 6;
    int x, y, z;
    face_point(f, threadIdx.x & 15, threadIdx.x >> 4, x, y, z);
    const int64_t ox = (int64_t)a.index[slot * 3 + 0] * B, oy = (int64_t)a.index[slot * 3 + 1] * B,
                  oz = (int64_t)a.index[slot * 3 + 2] * B;
    int n = 0;
    ghost_terms<1>(a, a.lut + a.lutoff[slot], a.level[slot], ox + x, oy + y, oz + z, &n, nullptr, nullptr);
    counts[(int64_t)k * 256 + threadIdx.x] = n;
}

__global__ __launch_bounds__(256) void k_face_rows_fill(LabArgs a, const int *__restrict__ faces, int nfaces,
                                                        const int64_t *__restrict__ ptr, int *__restrict__ src,
                                                        float *__restrict__ w)
{
    const int k = blockIdx.x;
    const int sf = faces[k];
    const int slot = sf / 6, f = sf % 6;
    int x, y, z;
    face_point(f, threadIdx.x & 15, threadIdx.x >> 4, x, y, z);
    const int64_t ox = (int64_t)a.index[slot * 3 + 0] * B, oy = (int64_t)a.index[slot * 3 + 1] * B,
                  oz = (int64_t)a.index[slot * 3 + 2] * B;
    const int64_t r = (int64_t)k * 256 + threadIdx.x;
    int n = 0;
    ghost_terms<2>(a, a.lut + a.lutoff[slot], a.level[slot], ox + x, oy + y, oz + z, &n, src + ptr[r], w + ptr[r]);
}

namespace {
// Non-zero taps of the FWT filters in increasing i (the reference loops over
// the full supports; a zero-weight tap adds an exact zero to a finite sum, so
// skipping it leaves every bit of the result unchanged).
template <int ORDER, bool G>
struct Taps;
template <int ORDER>
struct Taps<ORDER, false> {  // Ha = {1} on [0,1)
    static constexpr int n = 1;
    __device__ static constexpr int i(int) { return 0; }
    __device__ static constexpr double w(int) { return 1.0; }
};
template <>
struct Taps<4, true> {  // Ga on [-4,3): {1/16, 0, -9/16, 1, -9/16, 0, 1/16}
    static constexpr int n = 5;
    __device__ static constexpr int i(int k) { return k == 0 ? -4 : (k == 1 ? -2 : (k == 2 ? -1 : (k == 3 ? 0 : 2))); }
    __device__ static constexpr double w(int k)
    {
        return k == 0 ? 1.0 / 16.0 : (k == 1 ? -9.0 / 16.0 : (k == 2 ? 1.0 : (k == 3 ? -9.0 / 16.0 : 1.0 / 16.0)));
    }
};
template <>
struct Taps<2, true> {  // Ga on [-2,1): {-1/2, 1, -1/2}
    static constexpr int n = 3;
    __device__ static constexpr int i(int k) { return k - 2; }
    __device__ static constexpr double w(int k) { return k == 1 ? 1.0 : -0.5; }
};


// All seven detail codes of coarse point d in one pass over the lab: the G
// taps of each axis contain the H tap (i = 0), so every code's taps are a
// subset of the G x G x G taps.  Each lab value is read once and added to
// every code whose filter has that tap, with that code's weight
// ((wZ*wY)*wX); each code still sees its taps in its own z, y, x loop order
// (the reference's: r += lab(2d - i) * ((wZ*wY)*wX), i over the filter's
// support, skipping the zero taps adds exact zeros only).  Codes c = 1..7 in r[c-1]
// (bit 0: G in x, bit 1: G in y, bit 2: G in z).
template <int ORDER, int L, int S0>
__device__ __forceinline__ void fwt_codes(const double *lab, int dx, int dy, int dz, double (&r)[7])
{
    typedef Taps<ORDER, true> G;
    constexpr int KH = ORDER == 4 ? 3 : 2;  // the G tap with i = 0
#pragma unroll
    for (int c = 0; c < 7; c++) r[c] = 0.0;
#pragma unroll
    for (int kz = 0; kz < G::n; kz++) {
        const int z = 2 * dz - G::i(kz) - S0;
#pragma unroll
        for (int ky = 0; ky < G::n; ky++) {
            const int y = 2 * dy - G::i(ky) - S0;
#pragma unroll
            for (int kx = 0; kx < G::n; kx++) {
                const int x = 2 * dx - G::i(kx) - S0;
                const double val = lab[x + L * (y + L * z)];
#pragma unroll
                for (int c = 1; c <= 7; c++) {
                    const bool gx = c & 1, gy = c & 2, gz = c & 4;
                    if ((gx || kx == KH) && (gy || ky == KH) && (gz || kz == KH)) {
                        const double wz = gz ? G::w(kz) : 1.0, wy = gy ? G::w(ky) : 1.0, wx = gx ? G::w(kx) : 1.0;
                        r[c - 1] = __dadd_rn(r[c - 1], __dmul_rn(val, __dmul_rn(__dmul_rn(wz, wy), wx)));
                    }
                }
            }
        }
    }
}
}  // namespace

// BlockFWT::_fwt on phi (MRAGBlockFWT.h:220-298): detail of coarse point d,
// code (bx,by,bz) = sum over the tensor filter of lab(2d - i) * ((wZ*wY)*wX),
// z, y, x loop order, FP64 without contraction; block detail = max over
// codes 1..7 of |value|.
template <int ORDER>
__global__ __launch_bounds__(512) void k_fwt_detail(LabArgs a, const double *__restrict__ ghost,
                                                    const float *__restrict__ phi, const int *__restrict__ cand,
                                                    double *__restrict__ out)
{
    constexpr int S0 = ORDER == 4 ? -2 : -1;
    constexpr int S1 = ORDER == 4 ? 4 : 2;
    constexpr int L = LabGeom<S0, S1>::L;
    __shared__ double lab[LabGeom<S0, S1>::N];
    __shared__ double red[8];
    const int slot = cand[blockIdx.x];
    fill_lab_used<ORDER>(lab, phi, ghost + (size_t)blockIdx.x * LabGeom<S0, S1>::N, slot, a);
    __syncthreads();
    double m = 0.0;
    {
        const int q = threadIdx.x;  // one coarse point per thread
        const int dx = q & 7, dy = (q >> 3) & 7, dz = q >> 6;
        double v[7];
        fwt_codes<ORDER, L, S0>(lab, dx, dy, dz, v);
#pragma unroll
        for (int c = 0; c < 7; c++) {
            const double av = fabs(v[c]);
            m = m > av ? m : av;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const double u = __shfl_down(m, o, 64);
        m = m > u ? m : u;
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = red[0];
        for (int k = 1; k < 8; k++) r = r > red[k] ? r : red[k];
        out[blockIdx.x] = r;
    }
}

// BlockSplitter::_splitBlocks (MRAGBlockSplitter.h:24-99) as a gather: child
// point d = sum over parent lab points s (z, y, x increasing) of
// lab(s) * ((wZ*wY)*wX), w = Hs(2s - d), for phi, pw and pg.  One workgroup
// per (parent, child, field): the child's 16^3 points read parent points
// m-1..m+2 (Interp4) or m..m+1 (Interp2) around m = d/2, an 11^3 (9^3)
// window of the parent lab staged in LDS.  Four workgroups per child (one
// per quarter of its z range); each wave takes the points of one (x, y)
// parity class, so the tap loops of its lanes have the same trip counts.
template <int ORDER>
__global__ __launch_bounds__(256) void k_split(LabArgs a, const double *__restrict__ ghost3, float *phi, float *pw,
                                               float *pg, const int *__restrict__ parents,
                                               const int *__restrict__ children, int nparents)
{
    constexpr int S0 = ORDER == 4 ? -2 : -1;
    constexpr int S1 = ORDER == 4 ? 4 : 2;
    constexpr int L = LabGeom<S0, S1>::L;
    constexpr int W0 = ORDER == 4 ? -1 : 0;  // window start relative to 8*child
    constexpr int WL = ORDER == 4 ? 11 : 9;  // window width
    __shared__ double win[WL * WL * WL];
    const int pidx = blockIdx.x;
    const int child = blockIdx.y;
    const int f = blockIdx.z >> 2, zq = blockIdx.z & 3;
    const int pslot = parents[pidx];
    const int cx = child & 1, cy = (child >> 1) & 1, cz = child >> 2;
    float *fld = f == 0 ? phi : (f == 1 ? pw : pg);
    const double *ghost = ghost3 + ((size_t)f * nparents + pidx) * LabGeom<S0, S1>::N;
    __shared__ int s_nb[26];
    if (threadIdx.x < 26) s_nb[threadIdx.x] = a.nbr26[pslot * 26 + threadIdx.x];
    __syncthreads();
    {
        // every thread gathers its window points with all loads in flight at once
        constexpr int NW = WL * WL * WL, K = (NW + 255) / 256;
        size_t fi[K];
        int gi[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int q = threadIdx.x + 256 * k;
            const int qq = q < NW ? q : 0;
            const int x = 8 * cx + W0 + qq % WL, y = 8 * cy + W0 + (qq / WL) % WL, z = 8 * cz + W0 + qq / (WL * WL);
            const int dx = x < 0 ? -1 : (x < B ? 0 : 1), dy = y < 0 ? -1 : (y < B ? 0 : 1),
                      dz = z < 0 ? -1 : (z < B ? 0 : 1);
            fi[k] = (size_t)pslot * B3;
            gi[k] = -1;
            if ((dx | dy | dz) == 0) {
                fi[k] = (size_t)pslot * B3 + x + B * (y + B * z);
            } else {
                const int c = (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1);
                const int nb = s_nb[c < 13 ? c : c - 1];
                if (nb >= 0) fi[k] = (size_t)nb * B3 + (x - dx * B) + B * ((y - dy * B) + B * (z - dz * B));
                else gi[k] = (x - S0) + L * ((y - S0) + L * (z - S0));
            }
        }
        float fv[K];
        double gv[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            fv[k] = fld[fi[k]];
            gv[k] = ghost[gi[k] >= 0 ? gi[k] : 0];
        }
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int q = threadIdx.x + 256 * k;
            if (q < NW) win[q] = gi[k] >= 0 ? gv[k] : (double)fv[k];
        }
    }
    __syncthreads();
    const int cs = children[pidx * 8 + child];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int lx = 2 * (lane & 7) + (wave & 1), ly = 2 * (lane >> 3) + (wave >> 1);
    for (int it = 0; it < 4; it++) {
        const int lz = 4 * zq + it;
        const int q = lx + B * (ly + B * lz);
        // child point in parent-lab coordinates: m = (16*c + l) / 2, parity of l
        const int mx = 8 * cx + (lx >> 1), my = 8 * cy + (ly >> 1), mz = 8 * cz + (lz >> 1);
        const int px = lx & 1, py = ly & 1, pz = lz & 1;
        const int nz = pz ? (ORDER == 4 ? 4 : 2) : 1;
        const int ny = py ? (ORDER == 4 ? 4 : 2) : 1;
        const int nx = px ? (ORDER == 4 ? 4 : 2) : 1;
        double acc = 0.0;
        for (int kz = 0; kz < nz; kz++) {
            const int oz = pz ? (ORDER == 4 ? kz - 1 : kz) : 0;
            const double wz = pz ? (ORDER == 4 ? ((kz == 0 || kz == 3) ? -1.0 / 16.0 : 9.0 / 16.0) : 0.5) : 1.0;
            for (int ky = 0; ky < ny; ky++) {
                const int oy = py ? (ORDER == 4 ? ky - 1 : ky) : 0;
                const double wy = py ? (ORDER == 4 ? ((ky == 0 || ky == 3) ? -1.0 / 16.0 : 9.0 / 16.0) : 0.5) : 1.0;
                const double wzy = __dmul_rn(wz, wy);
                for (int kx = 0; kx < nx; kx++) {
                    const int ox = px ? (ORDER == 4 ? kx - 1 : kx) : 0;
                    const double wx = px ? (ORDER == 4 ? ((kx == 0 || kx == 3) ? -1.0 / 16.0 : 9.0 / 16.0) : 0.5) : 1.0;
                    const int sidx = (mx + ox - 8 * cx - W0) + WL * ((my + oy - 8 * cy - W0) + WL * (mz + oz - 8 * cz - W0));
                    acc = __dadd_rn(acc, __dmul_rn(win[sidx], __dmul_rn(wzy, wx)));
                }
            }
        }
        fld[(size_t)cs * B3 + q] = (float)acc;
    }
}

// BlockCollapser::_collapseBlocks (MRAGBlockCollapser.h:62-85): Ha = {1},
// the parent takes every other point of its 8 children.
__global__ __launch_bounds__(256) void k_collapse(float *phi, float *pw, float *pg, const int *__restrict__ parents,
                                                  const int *__restrict__ children)
{
    // blockIdx.y = child c: its 8^3 even points fill octant c of the parent,
    // two cells per thread, every load issued before the first store (the
    // parent's slot is a fresh slot: no child reads it)
    const int c = blockIdx.y;
    const int pslot = parents[blockIdx.x];
    const int cs = children[blockIdx.x * 8 + c];
    float v[2][3];
    size_t dst[2];
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const int q = threadIdx.x + 256 * k;  // 0..511 over the 8^3 octant
        const int x = q & 7, y = (q >> 3) & 7, z = q >> 6;
        const size_t src = (size_t)cs * B3 + 2 * x + B * (2 * y + B * 2 * z);
        dst[k] = (size_t)pslot * B3 + (x + 8 * (c & 1)) + B * ((y + 8 * ((c >> 1) & 1)) + B * (z + 8 * (c >> 2)));
        v[k][0] = phi[src];
        v[k][1] = pw[src];
        v[k][2] = pg[src];
    }
#pragma unroll
    for (int k = 0; k < 2; k++) {
        phi[dst[k]] = v[k][0];
        pw[dst[k]] = v[k][1];
        pg[dst[k]] = v[k][2];
    }
}

// ---------------------------------------------------------------------------
// Uniform view (Glioma_HG_UQ::_dumpUQoutput, Glioma_HG_UQ.cpp:309-372) and
// likelihood reductions.  A finest voxel v maps to the leaf covering it; a
// leaf r = gpd/(16*2^level) times coarser writes only the first min(r,4)
// voxels of each cell per axis (the reference's `else` branch), the rest
// are holes (value `hole`).
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool voxel_value(const UniformArgs &u, int64_t v, double &val)
{
    if (u.vol) {
        val = u.vol[v];
        return true;
    }
    const int gpd = u.gpd;
    const int vx = (int)(v % gpd), vy = (int)((v / gpd) % gpd), vz = (int)(v / ((int64_t)gpd * gpd));
    const int bpd = gpd >> 4;
    const int slot = u.lut_slot[(size_t)blockIdx.y * bpd * bpd * bpd + (vx >> 4) + bpd * ((vy >> 4) + bpd * (vz >> 4))];
    const int lvl = u.level[slot];
    const int r = bpd >> lvl;  // voxels per leaf cell per axis (>= 1)
    const int ox = u.index[slot * 3 + 0] * B * r, oy = u.index[slot * 3 + 1] * B * r, oz = u.index[slot * 3 + 2] * B * r;
    const int lx = vx - ox, ly = vy - oy, lz = vz - oz;
    const int ix = lx / r, iy = ly / r, iz = lz / r;
    const int rep = r < 4 ? r : 4;
    if (lx - ix * r >= rep || ly - iy * r >= rep || lz - iz * r >= rep) return false;
    val = (double)u.phi[(size_t)slot * B3 + ix + B * (iy + B * iz)];
    return true;
}

__global__ __launch_bounds__(256) void k_dump_uniform(UniformArgs u, double *out)
{
    const int64_t n = (int64_t)u.gpd * u.gpd * u.gpd;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        double val;
        if (voxel_value(u, v, val)) out[v] = val;
    }
}

namespace {
__device__ __forceinline__ double wave_sum(double x)
{
    for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
    return x;
}
// block sum of NV doubles into partial[blockIdx.x*NV + k]
template <int NV>
__device__ void block_sum_store(double (&v)[NV], double *partial)
{
    __shared__ double sh[NV][4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; k++) {
        const double s = wave_sum(v[k]);
        if (lane == 0) sh[k][wv] = s;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        const int k = threadIdx.x;
        partial[(size_t)blockIdx.x * NV + k] = ((sh[k][0] + sh[k][1]) + sh[k][2]) + sh[k][3];
    }
}
}  // namespace

// Flip-rate Bernoulli + Gaussian PET log-likelihood, per-block partials.
__global__ __launch_bounds__(256) void k_lik_flip(UniformArgs u, const uint8_t *__restrict__ t1,
                                                  const uint8_t *__restrict__ t2, const float *__restrict__ pet,
                                                  double th1, double th2, double lp1, double lp0, double c0,
                                                  double is2, double *partial)
{
    const int s = u.obs_stride;
    const int go = u.gpd / s;
    const int64_t n = (int64_t)go * go * go;
    double acc[3] = {0.0, 0.0, 0.0};
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (int64_t)gridDim.x * blockDim.x) {
        double m = 0.0;
        const int64_t vm = s == 1 ? v
                                  : (int64_t)(v % go) * s +
                                        (int64_t)u.gpd * ((int64_t)((v / go) % go) * s +
                                                          (int64_t)u.gpd * ((v / ((int64_t)go * go)) * s));
        if (!voxel_value(u, vm, m)) m = u.hole;
        const int m1 = m > th1, m2 = m > th2;
        acc[0] += (m1 == (int)t1[v]) ? lp1 : lp0;
        acc[1] += (m2 == (int)t2[v]) ? lp1 : lp0;
        const float p = pet[v];
        if (!__builtin_isnan(p)) {
            const double r = (double)p - m;
            acc[2] += c0 - r * r * is2;
        }
    }
    block_sum_store<3>(acc, partial + (size_t)blockIdx.y * gridDim.x * 3);
}

// Flip-rate likelihood over quads of 4 consecutive observation voxels of a
// row, for observation strides 2^ls <= 4: a quad's four model voxels lie in
// one base-level block, so the leaf look-up (slot, level, index) and the y/z
// replication tests are done once for the four.  U grid-stride iterations run
// at once in three rounds of loads (observations and look-up slots, leaf
// level and index, model rows), and the terms are added in the order of the
// one-at-a-time loop.  VOL: the model is a uniform FP64 volume (u.vol).
template <bool VOL>
__global__ __launch_bounds__(256) void k_lik_flip4(UniformArgs u, int lgpd, int ls,
                                                   const uint8_t *__restrict__ t1, const uint8_t *__restrict__ t2,
                                                   const float *__restrict__ pet, double th1, double th2, double lp1,
                                                   double lp0, double c0, double is2, double *partial)
{
    const int lg = lgpd - ls;  // log2 of the observation grid per dimension
    const int lb = lgpd - 4;   // log2 of the base blocks per dimension
    const int64_t mask = ((int64_t)1 << lg) - 1;
    const int64_t nq = ((int64_t)1 << (3 * lg)) >> 2;
    const int64_t S = (int64_t)gridDim.x * blockDim.x;
    const int *lut = u.lut_slot + ((size_t)blockIdx.y << (3 * lb));
    constexpr int U = 4;
    double acc[3] = {0.0, 0.0, 0.0};
    for (int64_t q0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q0 < nq; q0 += U * S) {
        uint32_t m1[U], m2[U];
        float4 pv[U];
        int mx[U], my[U], mz[U], slot[U];
#pragma unroll
        for (int k = 0; k < U; k++) {
            const int64_t q = q0 + k * S < nq ? q0 + k * S : q0;
            const int64_t v = q << 2;
            m1[k] = *reinterpret_cast<const uint32_t *>(t1 + v);
            m2[k] = *reinterpret_cast<const uint32_t *>(t2 + v);
            pv[k] = *reinterpret_cast<const float4 *>(pet + v);
            mx[k] = (int)((v & mask) << ls);
            my[k] = (int)(((v >> lg) & mask) << ls);
            mz[k] = (int)((v >> (2 * lg)) << ls);
            if (!VOL) slot[k] = lut[(mx[k] >> 4) + ((my[k] >> 4) << lb) + ((mz[k] >> 4) << (2 * lb))];
        }
        double m[U][4];
        if (VOL) {
#pragma unroll
            for (int k = 0; k < U; k++) {
                const double *p = u.vol + mx[k] + ((int64_t)my[k] << lgpd) + ((int64_t)mz[k] << (2 * lgpd));
#pragma unroll
                for (int j = 0; j < 4; j++) m[k][j] = p[j << ls];
            }
        } else {
            int lv[U], i0[U], i1[U], i2[U];
#pragma unroll
            for (int k = 0; k < U; k++) {
                lv[k] = u.level[slot[k]];
                i0[k] = u.index[slot[k] * 3 + 0];
                i1[k] = u.index[slot[k] * 3 + 1];
                i2[k] = u.index[slot[k] * 3 + 2];
            }
            float f[U][4];
            int hx[U][4];
            bool yz[U];
            int rep[U];
#pragma unroll
            for (int k = 0; k < U; k++) {
                const int lr = lb - lv[k];  // log2 of the voxels per leaf cell per axis
                rep[k] = lr < 2 ? (1 << lr) : 4;
                const int lx = mx[k] - (i0[k] << (4 + lr));
                const int ly = my[k] - (i1[k] << (4 + lr));
                const int lz = mz[k] - (i2[k] << (4 + lr));
                const int iy = ly >> lr, iz = lz >> lr;
                yz[k] = ly - (iy << lr) < rep[k] && lz - (iz << lr) < rep[k];
                const float *row = u.phi + (size_t)slot[k] * B3 + B * (iy + B * iz);
                // every voxel's cell lies in the block (a hole reads its cell
                // too): loads without branches, the hole chosen afterwards
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int x = lx + (j << ls);
                    const int ix = x >> lr;
                    f[k][j] = row[ix];
                    hx[k][j] = x - (ix << lr);
                }
            }
#pragma unroll
            for (int k = 0; k < U; k++)
#pragma unroll
                for (int j = 0; j < 4; j++) m[k][j] = (yz[k] && hx[k][j] < rep[k]) ? (double)f[k][j] : u.hole;
        }
#pragma unroll
        for (int k = 0; k < U; k++) {
            if (q0 + k * S >= nq) break;
            const float pp[4] = {pv[k].x, pv[k].y, pv[k].z, pv[k].w};
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int o1 = (int)((m1[k] >> (8 * j)) & 0xff), o2 = (int)((m2[k] >> (8 * j)) & 0xff);
                acc[0] += ((int)(m[k][j] > th1) == o1) ? lp1 : lp0;
                acc[1] += ((int)(m[k][j] > th2) == o2) ? lp1 : lp0;
                if (!__builtin_isnan(pp[j])) {
                    const double r = (double)pp[j] - m[k][j];
                    acc[2] += c0 - r * r * is2;
                }
            }
        }
    }
    block_sum_store<3>(acc, partial + (size_t)blockIdx.y * gridDim.x * 3);
}

// Reference model, PET part (HGG_Likelihood.cpp:40-75): lattice
// ix,iy,iz % step == 0, PET > 0: sum (u - scale*PET)^2 and count.
__global__ __launch_bounds__(256) void k_lik_ref_pet(UniformArgs u, const double *__restrict__ pet, int step,
                                                     double scale, double *partial)
{
    const int gpd = u.gpd;
    const int nl = (gpd + step - 1) / step;
    const int64_t n = (int64_t)nl * nl * nl;
    double acc[2] = {0.0, 0.0};
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        const int64_t ix = (q % nl) * step, iy = ((q / nl) % nl) * step, iz = (q / ((int64_t)nl * nl)) * step;
        const int64_t v = ix + gpd * (iy + (int64_t)gpd * iz);
        const double pv = pet[v];
        if (pv > 0.) {
            double m = 0.0;
            if (!voxel_value(u, v, m)) m = u.hole;
            const double d = m - scale * pv;
            acc[0] += d * d;
            acc[1] += 1.0;
        }
    }
    block_sum_store<2>(acc, partial);
}

namespace {
__device__ __forceinline__ double log_bernoulli(double m, double y, double uc, double slope)
{
    // HGG_Likelihood::_computeLogBernoulli (HGG_Likelihood.cpp:136-160)
    const double is2 = 1. / slope;
    const double diff = m - uc;
    const int sg = diff < -0.0 ? -1 : (diff > 0.0 ? 1 : 0);
    const double alpha = 0.5 + 0.5 * sg * (1. - exp(-diff * diff * is2));
    return (y == 1) ? log(alpha) : log(1. - alpha);
}
}  // namespace

// Reference model, T1/T2 parts (HGG_Likelihood.cpp:77-128): selective points
// (filtered by % step) or every voxel.
__global__ __launch_bounds__(256) void k_lik_ref_ti(UniformArgs u, const double *__restrict__ t1,
                                                    const double *__restrict__ t2, const int *__restrict__ pts,
                                                    int npts, int selective, int step, double t1uc, double t2uc,
                                                    double slope, double *partial)
{
    const int gpd = u.gpd;
    const int64_t n = selective ? (int64_t)npts : (int64_t)gpd * gpd * gpd;
    double acc[2] = {0.0, 0.0};
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        int64_t v;
        if (selective) {
            const int ix = pts[3 * q], iy = pts[3 * q + 1], iz = pts[3 * q + 2];
            if ((ix % step) || (iy % step) || (iz % step)) continue;
            v = ix + gpd * (iy + (int64_t)gpd * iz);
        } else {
            v = q;
        }
        double m = 0.0;
        if (!voxel_value(u, v, m)) m = u.hole;
        acc[0] += log_bernoulli(m, t1[v], t1uc, slope);
        acc[1] += log_bernoulli(m, t2[v], t2uc, slope);
    }
    block_sum_store<2>(acc, partial);
}

// Sum per-block partials: workgroup (s, k) sums value k of sample s over its
// nblocks partials (layout [s][block][nv]); each thread a fixed strided
// subset in increasing order, then a fixed reduction tree (deterministic).
__global__ __launch_bounds__(256) void k_sum_partials(const double *partial, int nblocks, int nv, double *out)
{
    __shared__ double sh[4];
    const int k = blockIdx.x, smp = blockIdx.y;
    const double *p = partial + (size_t)smp * nblocks * nv;
    double s = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += 256) s += p[(size_t)b * nv + k];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[(size_t)smp * nv + k] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
hipError_t launch_fk_step(const StepArgs &a, int zt, hipStream_t st, hipEvent_t e0, hipEvent_t e1)
{
    if (a.nwork <= 0) return hipSuccess;
    // with events: timestamps taken by the dispatch itself (no separate event packets)
    if (zt == 1) {
        const int grid = (a.nwork + 7) & ~7;
        hipExtLaunchKernelGGL(k_fk_step_col, dim3(grid), dim3(256), 0, st, e0, e1, 0, a);
        return hipGetLastError();
    }
    const int items = a.nwork * (16 / zt);
    const int grid = (items + 7) & ~7;
    switch (zt) {
    case 16: hipExtLaunchKernelGGL(k_fk_step<16>, dim3(grid), dim3(256), 0, st, e0, e1, 0, a); break;
    case 8: hipExtLaunchKernelGGL(k_fk_step<8>, dim3(grid), dim3(256), 0, st, e0, e1, 0, a); break;
    case 4: hipExtLaunchKernelGGL(k_fk_step<4>, dim3(grid), dim3(256), 0, st, e0, e1, 0, a); break;
    case 2: hipExtLaunchKernelGGL(k_fk_step<2>, dim3(grid), dim3(256), 0, st, e0, e1, 0, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int ZT>
static hipError_t persistent_zt(const PersistArgs &p, hipStream_t st, int max_items)
{
    constexpr int MAXDEV = 64;
    static int slots[MAXDEV] = {0};  // resident workgroups per CU of this slab depth, by device
    static int ncus[MAXDEV] = {0};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= MAXDEV) return hipErrorNotSupported;
    int &ncu = ncus[dev];
    if (!slots[dev]) {
        int coop = 0;
        e = hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev);
        if (e != hipSuccess) return e;
        if (!coop) return hipErrorNotSupported;
        e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
        if (e != hipSuccess) return e;
        int n = 0;
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, reinterpret_cast<const void *>(k_fk_persistent<ZT>), 256,
                                                         0);
        if (e != hipSuccess) return e;
        if (n <= 0) return hipErrorNotSupported;
        slots[dev] = n;
    }
    int grid = slots[dev] * ncu;
    grid = std::min(grid, std::max(8, (max_items + 7) & ~7));
    grid &= ~7;
    if (grid < 8) return hipErrorNotSupported;
    PersistArgs q = p;
    void *args[] = {&q};
    return hipLaunchCooperativeKernel(reinterpret_cast<const void *>(k_fk_persistent<ZT>), dim3(grid), dim3(256),
                                      args, 0, st);
}

hipError_t launch_fk_persistent(const PersistArgs &p, int zt, hipStream_t st)
{
    if (p.m <= 0) return hipSuccess;
    if (p.m > PMAX || !p.a.wrec) return hipErrorInvalidValue;
    int max_items = 0;
    for (int j = 0; j < p.m; j++) max_items = std::max(max_items, p.nwork[j] * (16 / std::max(zt, 1)));
    switch (zt) {
    case 16: return persistent_zt<16>(p, st, max_items);
    case 8: return persistent_zt<8>(p, st, max_items);
    case 4: return persistent_zt<4>(p, st, max_items);
    case 2: return persistent_zt<2>(p, st, max_items);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_fk_step2(const StepArgs &a, hipStream_t st, hipEvent_t e0, hipEvent_t e1)
{
    if (a.nwork <= 0) return hipSuccess;
    if (!a.ikc || !a.wrec || !a.nbr26 || a.phi_in == a.phi_out) return hipErrorInvalidValue;
    const int grid = (a.nwork + 7) & ~7;
    hipExtLaunchKernelGGL(k_fk_step2, dim3(grid), dim3(1024), 0, st, e0, e1, 0, a);
    return hipGetLastError();
}

hipError_t launch_face_ghosts(const float *phi, const int64_t *ptr, const int *src, const float *w, float *out,
                              int nrows, hipStream_t st)
{
    if (nrows <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_face_ghosts, dim3((4 * nrows + 255) / 256), dim3(256), 0, st, phi, ptr, src, w, out, nrows);
    return hipGetLastError();
}

hipError_t launch_face_prep(const float *pw, const float *pg, const int *work, int nwork, const int *fnbr,
                            const int *fbase, const int *sfbase, const int *slot_sample, const int64_t *ptr,
                            const int *src, const float *w, float *fik, float *ikc, float Dw, float Dg,
                            const float *sprm, hipStream_t st)
{
    if (nwork <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_face_prep, dim3(nwork, ikc ? 7 : 6), dim3(256), 0, st, pw, pg, work, nwork, fnbr, fbase, sfbase,
                       slot_sample, ptr, src, w, fik, ikc, Dw, Dg, sprm);
    return hipGetLastError();
}

int lab_points(int order) { return order == 4 ? 21 * 21 * 21 : 18 * 18 * 18; }

void lab_items(const std::vector<int> &slots, const std::vector<int> &nbr26, int order, std::vector<int> &items)
{
    items.clear();
    for (size_t j = 0; j < slots.size(); j++) {
        const int *nb = nbr26.data() + (size_t)slots[j] * 26;
        for (int r = 0; r < 26; r++) {
            if (nb[r] >= 0) continue;  // same-level neighbour: copied, not evaluated
            const int c = r < 13 ? r : r + 1;
            // Interp2 reads no lab coordinate below the block
            if (order == 2 && (c % 3 == 0 || (c / 3) % 3 == 0 || c / 9 == 0)) continue;
            // used points per axis side: below 1 (Interp4), inside 16, above 2 (Interp4) or 1 (Interp2)
            int npts = 1;
            for (int d : {c % 3, (c / 3) % 3, c / 9}) npts *= d == 0 ? 1 : (d == 1 ? 16 : (order == 4 ? 2 : 1));
            for (int k = 0; k * 64 < npts; k++) items.push_back((int)j << 8 | r << 3 | k);
        }
    }
}

hipError_t launch_lab_ghosts(const LabArgs &a, const float *const *flds, int nf, const int *slots,
                             const int *items, int nitems, double *ghost, size_t field_stride, bool for_split,
                             hipStream_t st)
{
    if (nitems <= 0) return hipSuccess;
    if (nf != 1 && nf != 3) return hipErrorInvalidValue;
    FieldPtrs fp;
    for (int f = 0; f < 3; f++) fp.p[f] = flds[f < nf ? f : 0];
    const dim3 grid(nitems);
#define GQ_LAB(O, S, N) \
    hipLaunchKernelGGL((k_lab_ghosts<O, S, N>), grid, dim3(64), 0, st, a, fp, slots, items, ghost, field_stride)
    if (a.order == 4) {
        if (for_split) {
            if (nf == 3) GQ_LAB(4, true, 3);
            else GQ_LAB(4, true, 1);
        } else {
            if (nf == 3) GQ_LAB(4, false, 3);
            else GQ_LAB(4, false, 1);
        }
    } else {
        if (for_split) {
            if (nf == 3) GQ_LAB(2, true, 3);
            else GQ_LAB(2, true, 1);
        } else {
            if (nf == 3) GQ_LAB(2, false, 3);
            else GQ_LAB(2, false, 1);
        }
    }
#undef GQ_LAB
    return hipGetLastError();
}

hipError_t launch_fwt_detail(int order, const LabArgs &a, const double *ghost, const float *phi, const int *cand,
                             int ncand, double *out, hipStream_t st)
{
    if (ncand <= 0) return hipSuccess;
    if (order == 4) hipLaunchKernelGGL(k_fwt_detail<4>, dim3(ncand), dim3(512), 0, st, a, ghost, phi, cand, out);
    else hipLaunchKernelGGL(k_fwt_detail<2>, dim3(ncand), dim3(512), 0, st, a, ghost, phi, cand, out);
    return hipGetLastError();
}

hipError_t launch_split(int order, const LabArgs &a, const double *ghost3, float *phi, float *pw, float *pg,
                        const int *parents, const int *children, int nparents, hipStream_t st)
{
    if (nparents <= 0) return hipSuccess;
    const dim3 grid(nparents, 8, 3 * 4);  // (parent, child, field x z quarter)
    if (order == 4)
        hipLaunchKernelGGL(k_split<4>, grid, dim3(256), 0, st, a, ghost3, phi, pw, pg, parents, children, nparents);
    else
        hipLaunchKernelGGL(k_split<2>, grid, dim3(256), 0, st, a, ghost3, phi, pw, pg, parents, children, nparents);
    return hipGetLastError();
}

hipError_t launch_face_rows_count(const LabArgs &a, const int *faces, int nfaces, int *counts, hipStream_t st)
{
    if (nfaces <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_face_rows_count, dim3(nfaces), dim3(256), 0, st, a, faces, nfaces, counts);
    return hipGetLastError();
}

hipError_t launch_face_rows_fill(const LabArgs &a, const int *faces, int nfaces, const int64_t *ptr, int *src,
                                 float *w, hipStream_t st)
{
    if (nfaces <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_face_rows_fill, dim3(nfaces), dim3(256), 0, st, a, faces, nfaces, ptr, src, w);
    return hipGetLastError();
}

hipError_t launch_collapse(float *phi, float *pw, float *pg, const int *parents, const int *children, int nparents,
                           hipStream_t st)
{
    if (nparents <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_collapse, dim3(nparents, 8), dim3(256), 0, st, phi, pw, pg, parents, children);
    return hipGetLastError();
}

hipError_t launch_dump_uniform(const UniformArgs &u, double *out, hipStream_t st)
{
    const int64_t n = (int64_t)u.gpd * u.gpd * u.gpd;
    int nb = (int)((n + 255) / 256);
    if (nb > 65536) nb = 65536;
    hipLaunchKernelGGL(k_dump_uniform, dim3(nb), dim3(256), 0, st, u, out);
    return hipGetLastError();
}

int flip_blocks(int64_t n_obs)
{
    // at least 16 observation voxels per thread: one pass of k_lik_flip4's
    // four quads, so no iteration loads a clamped duplicate
    int64_t nb = n_obs / 4096;
    if (nb > 2048) nb = 2048;
    if (nb < 1) nb = 1;
    return (int)nb;
}

int reduce_blocks(int64_t n)
{
    int64_t nb = (n + 255) / 256;
    if (nb > 2048) nb = 2048;
    if (nb < 1) nb = 1;
    return (int)nb;
}

hipError_t launch_lik_flip(const UniformArgs &u, int nsamples, const uint8_t *t1, const uint8_t *t2, const float *pet,
                           double th1, double th2, double lp1, double lp0, double c0, double is2, double *partial,
                           int nblocks, double *out, hipStream_t st)
{
    auto lg2 = [](int x) {
        int l = 0;
        while ((1 << l) < x) l++;
        return (1 << l) == x ? l : -1;
    };
    const int lgpd = lg2(u.gpd), ls = lg2(u.obs_stride);
    const bool pow2 = lgpd >= 4 && ls >= 0 && ls <= 2 &&  // a quad's model voxels in one base block
                      ((uintptr_t)t1 & 3) == 0 && ((uintptr_t)t2 & 3) == 0 && ((uintptr_t)pet & 15) == 0;
    if (pow2 && u.vol)
        hipLaunchKernelGGL(k_lik_flip4<true>, dim3(nblocks, nsamples), dim3(256), 0, st, u, lgpd, ls, t1, t2, pet, th1,
                           th2, lp1, lp0, c0, is2, partial);
    else if (pow2)
        hipLaunchKernelGGL(k_lik_flip4<false>, dim3(nblocks, nsamples), dim3(256), 0, st, u, lgpd, ls, t1, t2, pet,
                           th1, th2, lp1, lp0, c0, is2, partial);
    else
        hipLaunchKernelGGL(k_lik_flip, dim3(nblocks, nsamples), dim3(256), 0, st, u, t1, t2, pet, th1, th2, lp1, lp0,
                           c0, is2, partial);
    hipLaunchKernelGGL(k_sum_partials, dim3(3, nsamples), dim3(256), 0, st, partial, nblocks, 3, out);
    return hipGetLastError();
}

hipError_t launch_lik_ref(const UniformArgs &u, const double *pet, const double *t1, const double *t2,
                          const int *pts, int npts, int selective, int step_pet, int step_ti, double scale,
                          double t1uc, double t2uc, double slope, double *partial, int nblocks, double *out,
                          hipStream_t st)
{
    hipLaunchKernelGGL(k_lik_ref_pet, dim3(nblocks), dim3(256), 0, st, u, pet, step_pet, scale, partial);
    hipLaunchKernelGGL(k_sum_partials, dim3(2, 1), dim3(256), 0, st, partial, nblocks, 2, out);
    hipLaunchKernelGGL(k_lik_ref_ti, dim3(nblocks), dim3(256), 0, st, u, t1, t2, pts, npts, selective, step_ti, t1uc,
                       t2uc, slope, partial + (size_t)nblocks * 2);
    hipLaunchKernelGGL(k_sum_partials, dim3(2, 1), dim3(256), 0, st, partial + (size_t)nblocks * 2, nblocks, 2,
                       out + 2);
    return hipGetLastError();
}

}  // namespace gq

namespace gq {
// Block transfer between a dense host-order staging array and pool slots.
__global__ __launch_bounds__(256) void k_scatter_blocks(float *pool, const float *__restrict__ dense,
                                                        const int *__restrict__ slots)
{
    const int k = blockIdx.x;
    const size_t d = (size_t)slots[k] * 4096, s = (size_t)k * 4096;
    for (int i = threadIdx.x; i < 4096; i += 256) pool[d + i] = dense[s + i];
}
__global__ __launch_bounds__(256) void k_gather_blocks(const float *__restrict__ pool, float *dense,
                                                       const int *__restrict__ slots)
{
    const int k = blockIdx.x;
    const size_t s = (size_t)slots[k] * 4096, d = (size_t)k * 4096;
    for (int i = threadIdx.x; i < 4096; i += 256) dense[d + i] = pool[s + i];
}
hipError_t launch_scatter_blocks(float *pool, const float *dense, const int *slots, int n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_scatter_blocks, dim3(n), dim3(256), 0, st, pool, dense, slots);
    return hipGetLastError();
}
hipError_t launch_gather_blocks(const float *pool, float *dense, const int *slots, int n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_blocks, dim3(n), dim3(256), 0, st, pool, dense, slots);
    return hipGetLastError();
}
}  // namespace gq

namespace gq {
__global__ __launch_bounds__(256) void k_copy_blocks(float *dst, const float *__restrict__ src,
                                                     const int *__restrict__ slots)
{
    const size_t b = (size_t)slots[blockIdx.x] * 4096;
    for (int i = threadIdx.x; i < 4096; i += 256) dst[b + i] = src[b + i];
}
hipError_t launch_copy_blocks(float *dst, const float *src, const int *slots, int n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_copy_blocks, dim3(n), dim3(256), 0, st, dst, src, slots);
    return hipGetLastError();
}
}  // namespace gq

namespace gq {
// Point-sample a uniform gpd^3 volume (x fastest) onto the listed slots: a
// leaf at level l takes every 2^(base-l)-th point of its region.
__global__ __launch_bounds__(256) void k_sample_volume(float *__restrict__ pool, const float *__restrict__ vol,
                                                       int gpd, int base_level, const int *__restrict__ slots,
                                                       const int *__restrict__ level, const int *__restrict__ index)
{
    const int slot = slots[blockIdx.x];
    const int r = 1 << (base_level - level[slot]);
    const int64_t ox = (int64_t)index[slot * 3 + 0] * B * r, oy = (int64_t)index[slot * 3 + 1] * B * r,
                  oz = (int64_t)index[slot * 3 + 2] * B * r;
    const size_t b = (size_t)slot * B3;
    for (int i = threadIdx.x; i < B3; i += 256) {
        const int x = i & 15, y = (i >> 4) & 15, z = i >> 8;
        pool[b + i] = vol[(ox + x * r) + gpd * ((oy + y * r) + (int64_t)gpd * (oz + z * r))];
    }
}

hipError_t launch_sample_volume(float *pool, const float *vol, int gpd, int base_level, const int *slots, int n,
                                const int *level, const int *index, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sample_volume, dim3(n), dim3(256), 0, st, pool, vol, gpd, base_level, slots, level, index);
    return hipGetLastError();
}

hipError_t scan_counts(const int *counts, int64_t *ptr, int n, void *temp, size_t &temp_bytes, hipStream_t st)
{
    if (!temp) return hipcub::DeviceScan::InclusiveSum(nullptr, temp_bytes, counts, ptr + 1, n, st);
    hipError_t e = hipMemsetAsync(ptr, 0, sizeof(int64_t), st);
    if (e != hipSuccess) return e;
    if (n <= 0) return hipSuccess;
    return hipcub::DeviceScan::InclusiveSum(temp, temp_bytes, counts, ptr + 1, n, st);
}
}  // namespace gq

namespace gq {
__global__ __launch_bounds__(64) void k_scatter_slot_rows(SlotTables t, const int *__restrict__ packed, int nrows)
{
    const int r = blockIdx.x;
    if (r >= nrows) return;
    const int *q = packed + (size_t)r * 45;
    const int sl = q[0];
    const int k = threadIdx.x;
    if (k == 0) {
        t.level[sl] = q[1];
        t.slot_sample[sl] = q[5];
        t.lutoff[sl] = q[6];
    }
    if (k < 3) t.index[sl * 3 + k] = q[2 + k];
    if (k < 6) {
        t.fnbr[sl * 6 + k] = q[7 + k];
        t.fbase[sl * 6 + k] = q[13 + k];
    }
    if (k < 26) t.nbr26[sl * 26 + k] = q[19 + k];
}

hipError_t launch_scatter_slot_rows(const SlotTables &t, const int *packed, int nrows, hipStream_t st)
{
    if (nrows <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_scatter_slot_rows, dim3(nrows), dim3(64), 0, st, t, packed, nrows);
    return hipGetLastError();
}
}  // namespace gq
