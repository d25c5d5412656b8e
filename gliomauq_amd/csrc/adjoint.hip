// adjoint.hip -- discrete adjoint of the fused FK step on uniform grids.
//
// Forward (one step, per cell c; the map k_fk_step computes):
//   v_c  = u_c + dt * (ih2 * sum_k df_ck (u_k - u_c) + rho u_c (1 - u_c))
//   u'_c = min(1, max(0, v_c))             (NaN -> 0)
// with df_ck = 2 / (1/K_c + 1/K_k) when tissue_c = pw+pg+u_c > 0, else 0.
// The reference has no adjoint; this is the exact reverse-mode derivative of
// that discrete map (df is piecewise constant in u: the tissue switch and the
// clamp have zero derivative almost everywhere).  With
//   mask_c = [0 < v_c < 1] = [0 < u'_c < 1]   and   mu_c = mask_c * lambda'_c:
//   lambda_c = mu_c (1 + dt (rho (1 - 2 u_c) - ih2 sum_k df_ck)) + dt ih2 sum_k df_kc mu_k
//   dL/dlog Dw += mu_c dt ih2 sum_k df_ck (u_k - u_c)     (df ~ Dw when Dg = Dw/10)
//   dL/dlog rho += mu_c dt rho u_c (1 - u_c)
// Coefficients can be NaN next to CSF (pw = pg = 0), where the forward clamps
// NaN to 0; the adjoint factors of those terms are zero, and the kernel's flux
// form (k_adj_step) never evaluates them.  As for the step kernels, pw and pg
// are non-negative (tissue probabilities).
//
// One backward step is one kernel: it stages u_n, 1/K (precomputed once per
// solve) and mu_{n+1} with a one-cell halo in LDS, computes lambda_n and
// writes mu_n = [0 < u_n < 1] lambda_n -- the mask of the next backward step,
// taken from the clamped state u_n this step already reads, equal to the
// forward's branch bit for bit -- or lambda_n itself at the last step; it adds
// the two parameter-gradient contributions of its slab to per-slab FP64
// accumulators (each owned by one workgroup: deterministic).  Per cell it
// moves 16 bytes: u_n, 1/K, mu_{n+1} in, mu_n out.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "kernels.hpp"

namespace gq {

namespace {

constexpr int B = 16;
constexpr int B3 = 4096;
// LDS tile of a slab: ZT + 2 planes of 18 rows (y = -1 .. 16) x 16 columns,
// then the x halo columns (x = -1, x = 16) of the ZT interior planes; rows of
// 16 floats keep the 64 lanes of a wave (4 rows of 16) on distinct banks.
constexpr int RW = 16;
constexpr int PL = 18 * RW;
template <int ZT>
struct Tile {
    static constexpr int XW = (ZT + 2) * PL, XE = XW + 16 * ZT, SIZE = XE + 16 * ZT;
    __device__ static int cell(int x, int y, int zl) { return x + RW * (y + 1) + PL * zl; }
    __device__ static int west(int y, int k) { return XW + y + 16 * k; }
    __device__ static int east(int y, int k) { return XE + y + 16 * k; }
};

__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }

// the step kernels' 1/K, rounded the same way (contraction off, explicit FMA)
__device__ __forceinline__ float inv_k(float w, float g, float Dw, float Dg)
{
#pragma clang fp contract(off)
    const float b = w + g;
    const float D = __builtin_fmaf(g, Dg, w * Dw);
    return b * frcp(D);
}

__device__ __forceinline__ int nbr_face_cell(int f, int u, int v)
{
    switch (f) {
    case 0: return (B - 1) + B * (u + B * v);
    case 1: return 0 + B * (u + B * v);
    case 2: return u + B * ((B - 1) + B * v);
    case 3: return u + B * (0 + B * v);
    case 4: return u + B * (v + B * (B - 1));
    default: return u + B * (v + B * 0);
    }
}

__device__ __forceinline__ float mu_of(float lam, float u) { return (u > 0.f && u < 1.f) ? lam : 0.f; }

// Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a
// contiguous chunk of the Morton-ordered work list so the halo reads of
// neighbouring blocks hit the same L2 (nwg is a multiple of 8).
__device__ __forceinline__ int xcd_remap(int wg, int nwg)
{
    const int per = nwg >> 3;
    return (wg & 7) * per + (wg >> 3);
}

__device__ __forceinline__ double wave_sum(double x)
{
    for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
    return x;
}

}  // namespace

// ZT planes of one block per workgroup of 256 threads (thread = one (x,y) column).
//
// The update in flux form.  With fc_ck = 2 r_ck, r_ck = 1 / (1/K_c + 1/K_k):
//   lambda_c = mu_c (1 + dt rho (1 - 2 u_c)) + 2 dt ih2 sum_k r_ck (mu_k - mu_c)
// which equals the per-cell form above because mu_c != 0 implies tissue_c and
// finite coefficients of c (else the forward's v_c is NaN or 0 and the mask
// is 0), and likewise for every neighbour k with mu_k != 0.  The 1/K array
// holds +inf where the forward's 1/K is NaN (pw + pg = 0, k_adj_prep), so a
// face touching such a cell has r = 0 -- and there mu_k = mu_c = 0 anyway:
// no term needs a guard.  The z faces of a column are shared between the
// two cells they separate (a + b == b + a: the same bits either way).
template <int ZT>
__global__ __launch_bounds__(256) void k_adj_step(AdjArgs a)
{
#pragma clang fp contract(off)
    constexpr int NS = B / ZT;
    typedef Tile<ZT> T;
    __shared__ float s_u[T::SIZE];
    __shared__ float s_ik[T::SIZE];
    __shared__ float s_mu[T::SIZE];
    __shared__ double s_red[2][4];
    const int item = xcd_remap(blockIdx.x, gridDim.x);
    if (item >= a.nwork * NS) return;  // padding workgroups (whole workgroup exits)
    const int w = item / NS;
    const int slab = item - w * NS;
    const int z0 = slab * ZT;
    const int slot = a.work[w];
    const int t = threadIdx.x;
    const int tx = t & 15, ty = t >> 4;
    const size_t base = (size_t)slot * B3;
    const float *q = a.sprm + 4 * a.slot_sample[slot];
    const float rho = q[2], dt = q[3];

    // every cell this thread stages -- its ZT interior cells, one cell of each
    // z-halo plane and ZT/4 cells of the four side faces -- with all loads in
    // flight before the first LDS store
    static_assert(ZT % 4 == 0, "side faces: ZT/4 cells per thread");
    constexpr int NSIDE = ZT / 4;
    constexpr int NL = ZT + 2 + NSIDE;
    const int *fnb = a.fnbr + slot * 6;
    const int fn0 = fnb[0], fn1 = fnb[1], fn2 = fnb[2], fn3 = fnb[3], fn4 = fnb[4], fn5 = fnb[5];
    size_t gi[NL];
    int li[NL];
#pragma unroll
    for (int k = 0; k < ZT; k++) {
        gi[k] = base + (z0 + k) * 256 + t;
        li[k] = T::cell(tx, ty, k + 1);
    }
    li[ZT] = T::cell(tx, ty, 0);
    gi[ZT] = z0 > 0 ? base + (z0 - 1) * 256 + t : (size_t)fn4 * B3 + nbr_face_cell(4, tx, ty);
    li[ZT + 1] = T::cell(tx, ty, ZT + 1);
    gi[ZT + 1] = z0 + ZT < B ? base + (z0 + ZT) * 256 + t : (size_t)fn5 * B3 + nbr_face_cell(5, tx, ty);
#pragma unroll
    for (int j = 0; j < NSIDE; j++) {
        const int r4 = t + 256 * j;
        const int f = r4 / (16 * ZT);
        const int r = r4 - f * 16 * ZT;
        const int u = r & 15, v = z0 + (r >> 4);
        const int nb = f == 0 ? fn0 : f == 1 ? fn1 : f == 2 ? fn2 : fn3;
        gi[ZT + 2 + j] = (size_t)nb * B3 + nbr_face_cell(f, u, v);
        int l;
        switch (f) {
        case 0: l = T::west(u, v - z0); break;
        case 1: l = T::east(u, v - z0); break;
        case 2: l = T::cell(u, -1, v - z0 + 1); break;
        default: l = T::cell(u, B, v - z0 + 1); break;
        }
        li[ZT + 2 + j] = l;
    }
    float vu[NL], vk[NL], vm[NL];
#pragma unroll
    for (int j = 0; j < NL; j++) {
        vu[j] = a.u[gi[j]];
        vk[j] = a.ik[gi[j]];
        vm[j] = a.mu_in[gi[j]];
    }
#pragma unroll
    for (int j = 0; j < NL; j++) {
        s_u[li[j]] = vu[j];
        s_ik[li[j]] = vk[j];
        s_mu[li[j]] = vm[j];
    }
    __syncthreads();

    const float ih2 = a.ih2_lvl[a.level[slot]];
    const float dti = dt * ih2;
    double gD = 0.0, gR = 0.0;
    float r_up = 0.f;  // r of the +z face of the previous cell of this column, when computed
    bool have_up = false;
#pragma unroll
    for (int k = 0; k < ZT; k++) {
        const int c = T::cell(tx, ty, k + 1);
        const int cw = tx > 0 ? c - 1 : T::west(ty, k);
        const int ce = tx < B - 1 ? c + 1 : T::east(ty, k);
        const float muc = s_mu[c];
        const float m0 = s_mu[cw], m1 = s_mu[ce], m2 = s_mu[c - RW], m3 = s_mu[c + RW], m4 = s_mu[c - PL],
                    m5 = s_mu[c + PL];
        if ((muc == 0.f) & (m0 == 0.f) & (m1 == 0.f) & (m2 == 0.f) & (m3 == 0.f) & (m4 == 0.f) & (m5 == 0.f)) {
            // no adjoint flows through this cell: lambda = 0 and no gradient terms
            a.out[base + (z0 + k) * 256 + t] = 0.f;
            have_up = false;
            continue;
        }
        const float uc = s_u[c];
        const float kc = s_ik[c];
        const float r0 = frcp(kc + s_ik[cw]), r1 = frcp(kc + s_ik[ce]);
        const float r2 = frcp(kc + s_ik[c - RW]), r3 = frcp(kc + s_ik[c + RW]);
        const float r4 = have_up ? r_up : frcp(kc + s_ik[c - PL]);
        const float r5 = frcp(kc + s_ik[c + PL]);
        r_up = r5;
        have_up = true;
        float fl = r0 * (m0 - muc);
        fl = __builtin_fmaf(r1, m1 - muc, fl);
        fl = __builtin_fmaf(r2, m2 - muc, fl);
        fl = __builtin_fmaf(r3, m3 - muc, fl);
        fl = __builtin_fmaf(r4, m4 - muc, fl);
        fl = __builtin_fmaf(r5, m5 - muc, fl);
        const float lam = __builtin_fmaf(2.f * dti, fl, muc * (1.f + dt * (rho * (1.f - 2.f * uc))));
        if (muc != 0.f) {
            float df = r0 * (s_u[cw] - uc);
            df = __builtin_fmaf(r1, s_u[ce] - uc, df);
            df = __builtin_fmaf(r2, s_u[c - RW] - uc, df);
            df = __builtin_fmaf(r3, s_u[c + RW] - uc, df);
            df = __builtin_fmaf(r4, s_u[c - PL] - uc, df);
            df = __builtin_fmaf(r5, s_u[c + PL] - uc, df);
            // the per-block factors 2 dt ih2 and dt rho are applied once to the sums
            gD = __builtin_fma((double)muc, (double)df, gD);
            gR = __builtin_fma((double)muc, (double)uc * (1.0 - (double)uc), gR);
        }
        a.out[base + (z0 + k) * 256 + t] = a.out_lambda ? lam : mu_of(lam, uc);
    }
    // block sums of the two parameter gradients -> this slab's accumulators
    gD = wave_sum(gD) * (2.0 * (double)dt * (double)ih2);
    gR = wave_sum(gR) * ((double)dt * (double)rho);
    const int lane = t & 63, wv = t >> 6;
    if (lane == 0) {
        s_red[0][wv] = gD;
        s_red[1][wv] = gR;
    }
    __syncthreads();
    if (t < 2) {
        const double s = ((s_red[t][0] + s_red[t][1]) + s_red[t][2]) + s_red[t][3];
        a.gacc[((size_t)slot * NS + slab) * 2 + t] += s;
    }
}

// 1/K of every cell of the listed slots (the adjoint's coefficients are fixed
// during a solve: pw, pg and the sample's Dw, Dg do not change), +inf where
// the step's 1/K is NaN (pw + pg = 0: no tissue), so that every face
// coefficient touching such a cell is 0 in the adjoint step.
__global__ __launch_bounds__(256) void k_adj_prep(const float *__restrict__ pw, const float *__restrict__ pg,
                                                  const int *__restrict__ work, const int *__restrict__ slot_sample,
                                                  const float *__restrict__ sprm, float *__restrict__ ik)
{
    const int slot = work[blockIdx.x];
    const float *q = sprm + 4 * slot_sample[slot];
    const float Dw = q[0], Dg = q[1];
    const size_t base = (size_t)slot * B3;
    for (int i = threadIdx.x; i < B3; i += 256) {
        const float v = inv_k(pw[base + i], pg[base + i], Dw, Dg);
        ik[base + i] = v == v ? v : INFINITY;
    }
}

// mu_N = [0 < u_N < 1] dL/du of the flip-rate model at the final state
// (uniform grids): only the Gaussian PET term depends smoothly on u; the
// Bernoulli terms are piecewise constant in u (derivative zero almost
// everywhere).
__global__ __launch_bounds__(256) void k_lik_flip_grad(const float *__restrict__ u, const float *__restrict__ pet,
                                                       const int *__restrict__ work, const int *__restrict__ index,
                                                       int gpd, int stride, double inv_s2, float *__restrict__ mu)
{
    // observations on the grid coarser by `stride`: only model voxels at multiples of it are observed
    const int slot = work[blockIdx.x];
    const size_t base = (size_t)slot * B3;
    const int64_t ox = (int64_t)index[slot * 3] * B, oy = (int64_t)index[slot * 3 + 1] * B,
                  oz = (int64_t)index[slot * 3 + 2] * B;
    const int64_t go = gpd / stride;
    for (int i = threadIdx.x; i < B3; i += 256) {
        const int64_t gx = ox + (i & 15), gy = oy + ((i >> 4) & 15), gz = oz + (i >> 8);
        float g = 0.f;
        if (gx % stride == 0 && gy % stride == 0 && gz % stride == 0) {
            const float p = pet[gx / stride + go * (gy / stride + go * (gz / stride))];
            if (!__builtin_isnan(p)) g = (float)(((double)p - (double)u[base + i]) * inv_s2);
        }
        mu[base + i] = mu_of(g, u[base + i]);
    }
}

namespace {
// signed periodic displacement x - x0 in [-n/2, n/2)
__device__ __forceinline__ double pdisp(double x, double x0, double n)
{
    double d = x - x0;
    d -= n * floor(d / n + 0.5);
    return d;
}
}  // namespace

// Gaussian initial condition u0 = A exp(-|x - x0|^2 / (2 s^2)) (voxel units,
// periodic distance); icp[sample*5] = x0, y0, z0, A, s.
__global__ __launch_bounds__(256) void k_ic_gaussian(float *__restrict__ phi, const int *__restrict__ work,
                                                     const int *__restrict__ index,
                                                     const int *__restrict__ slot_sample,
                                                     const double *__restrict__ icp, int gpd)
{
    const int slot = work[blockIdx.x];
    const double *p = icp + 5 * slot_sample[slot];
    const double n = gpd;
    const double is2 = 1.0 / (2.0 * p[4] * p[4]);
    for (int i = threadIdx.x; i < B3; i += 256) {
        const int x = i & 15, y = (i >> 4) & 15, z = i >> 8;
        const double dx = pdisp(index[slot * 3] * B + x, p[0], n), dy = pdisp(index[slot * 3 + 1] * B + y, p[1], n),
                     dz = pdisp(index[slot * 3 + 2] * B + z, p[2], n);
        phi[(size_t)slot * B3 + i] = (float)(p[3] * exp(-(dx * dx + dy * dy + dz * dz) * is2));
    }
}

// dL/d(x0, y0, z0, A) of the Gaussian initial condition from lambda_0, per slot.
__global__ __launch_bounds__(256) void k_ic_grad(const float *__restrict__ lam0, const int *__restrict__ work,
                                                 const int *__restrict__ index, const int *__restrict__ slot_sample,
                                                 const double *__restrict__ icp, int gpd, double *__restrict__ out)
{
    __shared__ double s_red[4][4];
    const int slot = work[blockIdx.x];
    const double *p = icp + 5 * slot_sample[slot];
    const double n = gpd;
    const double s2 = p[4] * p[4];
    const double is2 = 1.0 / (2.0 * s2);
    double g[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < B3; i += 256) {
        const int x = i & 15, y = (i >> 4) & 15, z = i >> 8;
        const double dx = pdisp(index[slot * 3] * B + x, p[0], n), dy = pdisp(index[slot * 3 + 1] * B + y, p[1], n),
                     dz = pdisp(index[slot * 3 + 2] * B + z, p[2], n);
        const double e = exp(-(dx * dx + dy * dy + dz * dz) * is2);
        const double l = lam0[(size_t)slot * B3 + i];
        const double lu = l * p[3] * e;
        g[0] += lu * dx / s2;
        g[1] += lu * dy / s2;
        g[2] += lu * dz / s2;
        g[3] += l * e;
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const double s = wave_sum(g[k]);
        if (lane == 0) s_red[k][wv] = s;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        const int k = threadIdx.x;
        out[(size_t)blockIdx.x * 4 + k] = ((s_red[k][0] + s_red[k][1]) + s_red[k][2]) + s_red[k][3];
    }
}

constexpr int ADJ_ZT = 4;  // slab depth: 23 KB of LDS per workgroup, six workgroups per CU

int adj_slabs() { return B / ADJ_ZT; }

hipError_t launch_adj_step(const AdjArgs &a, hipStream_t st)
{
    if (a.nwork <= 0) return hipSuccess;
    const int grid = (a.nwork * (B / ADJ_ZT) + 7) & ~7;
    hipLaunchKernelGGL(k_adj_step<ADJ_ZT>, dim3(grid), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_adj_prep(const float *pw, const float *pg, const int *work, int nwork, const int *slot_sample,
                           const float *sprm, float *ik, hipStream_t st)
{
    if (nwork <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_adj_prep, dim3(nwork), dim3(256), 0, st, pw, pg, work, slot_sample, sprm, ik);
    return hipGetLastError();
}

hipError_t launch_lik_flip_grad(const float *u, const float *pet, const int *work, int nwork, const int *index,
                                int gpd, int stride, double inv_s2, float *lam, hipStream_t st)
{
    if (nwork <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_lik_flip_grad, dim3(nwork), dim3(256), 0, st, u, pet, work, index, gpd, stride, inv_s2,
                       lam);
    return hipGetLastError();
}

hipError_t launch_ic_gaussian(float *phi, const int *work, int nwork, const int *index, const int *slot_sample,
                              const double *icp, int gpd, hipStream_t st)
{
    if (nwork <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_ic_gaussian, dim3(nwork), dim3(256), 0, st, phi, work, index, slot_sample, icp, gpd);
    return hipGetLastError();
}

hipError_t launch_ic_grad(const float *lam0, const int *work, int nwork, const int *index, const int *slot_sample,
                          const double *icp, int gpd, double *out, hipStream_t st)
{
    if (nwork <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_ic_grad, dim3(nwork), dim3(256), 0, st, lam0, work, index, slot_sample, icp, gpd, out);
    return hipGetLastError();
}

}  // namespace gq
