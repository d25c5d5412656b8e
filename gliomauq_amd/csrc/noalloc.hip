// noalloc.hip -- the flip-rate log-likelihood of a model volume that already
// lies in device memory, with no allocation of any kind: for callers that
// cannot allocate (code under HIP graph capture, tight sampler loops,
// harnesses that forbid hipMalloc).
//
// The kernels are those of gq_loglik_flip (k_lik_flip + k_sum_partials,
// kernels.hip); their per-workgroup FP64 partial sums go to `scratch`, a
// file-scope device array of SCRATCH_DOUBLES doubles that every GPU the
// library runs on holds until the process ends.  Its device address is looked
// up once per device (hipGetSymbolAddress, on the first call with that device
// current) and only that address is used on the host side.  Every call on one
// GPU shares the array: calls that overlap on one GPU (two streams, two host
// threads) race on it.  No part of the array needs clearing: every workgroup
// writes all of its partials before the reduction reads them.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>

#include "../../include/gq_api.h"
#include "kernels.hpp"

namespace {

// 3 partial sums per workgroup of the likelihood kernel, at most 2048
// workgroups (gq::flip_blocks)
constexpr int SCRATCH_DOUBLES = 3 * 2048;
constexpr int MAX_DEVICES = 64;

}  // namespace

static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

// out[0] = L_T1 + L_FLAIR + L_PET in the order gq_loglik_flip adds them
__global__ void k_lik_total(double *out) { out[0] = out[1] + out[2] + out[3]; }

namespace {

std::mutex g_mu;
double *g_scratch[MAX_DEVICES] = {};

// device address of `scratch` on the current device (looked up once per device)
hipError_t scratch_ptr(double **out)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch[dev] = (double *)p;
    }
    *out = g_scratch[dev];
    return hipSuccess;
}

}  // namespace

extern "C" gq_status gq_loglik_flip_volume_noalloc(const double *model, int gpd, const uint8_t *t1, const uint8_t *t2,
                                                   const float *pet, int obs_gpd, const double params[4], double *out,
                                                   void *stream)
{
    if (!model || !t1 || !t2 || !pet || !params || !out) return GQ_ERR_ARG;
    if (gpd <= 0 || gpd > GQ_NOALLOC_MAX_GPD || obs_gpd <= 0 || gpd % obs_gpd != 0) return GQ_ERR_ARG;
    const double th1 = params[0], th2 = params[1], eps = params[2], sigma = params[3];
    if (!(eps > 0 && eps < 1) || !(sigma > 0)) return GQ_ERR_NUMERIC;
    const int64_t og = obs_gpd;
    const int nb = gq::flip_blocks(og * og * og);
    if (3 * nb > SCRATCH_DOUBLES) return GQ_ERR_UNSUPPORTED;
    double *partial = nullptr;
    if (scratch_ptr(&partial) != hipSuccess) return GQ_ERR_HIP;
    gq::UniformArgs u;
    u.vol = model;
    u.phi = nullptr;
    u.lut_slot = nullptr;
    u.level = nullptr;
    u.index = nullptr;
    u.gpd = gpd;
    u.hole = 0.0;
    u.obs_stride = gpd / obs_gpd;
    const double lp1 = std::log(1. - eps), lp0 = std::log(eps);
    const double c0 = -0.5 * std::log(2. * M_PI * sigma * sigma);
    const double is2 = 1. / (2. * sigma * sigma);
    const hipStream_t st = (hipStream_t)stream;
    hipError_t e = gq::launch_lik_flip(u, 1, t1, t2, pet, th1, th2, lp1, lp0, c0, is2, partial, nb, out + 1, st);
    if (e != hipSuccess) return GQ_ERR_HIP;
    hipLaunchKernelGGL(k_lik_total, dim3(1), dim3(1), 0, st, out);
    return hipGetLastError() == hipSuccess ? GQ_OK : GQ_ERR_HIP;
}
