// surrogate.hpp -- argument blocks and launchers of the surrogate kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace gq {
namespace sur {

// channels-last tensor view: element (voxel v, channel c) at p[v*ld + off + c]
struct View {
    __bf16 *p;
    int D;
    int64_t ld;
    int64_t off;
};
struct FView {
    float *p;
    int D;
    int64_t ld;
    int64_t off;
};

// A layer input of cin channels: channels [0, cup) are the half-resolution
// tensor `up` upsampled 2x (nearest; the upsampling is folded into the
// read), channels [cup, cin) come from `in` at the layer's resolution.  cup =
// 0: every channel from `in`.
struct ConvArgs {
    View in;          // input (D^3 per batch item), channels cup .. cin-1 at in.off
    View up;          // channels 0 .. cup-1 at half resolution (cup > 0)
    int cup;
    int cin, cout, batch;
    const __bf16 *w;  // [cout][27*cin]
    const float *bias;
    int relu;
    __bf16 *out_bf;   // bf16 output (activations) or null
    float *out_f;     // FP32 output (gradients) when out_bf is null
    int accumulate;   // FP32 output: += instead of =
    int64_t out_ld, out_off;
};

struct WgradArgs {
    View x;   // layer input, channels cup .. cin-1 (as ConvArgs::in)
    View up;  // channels 0 .. cup-1 at half resolution (cup > 0)
    int cup;
    View dz;  // gradient at the pre-activation, cout channels
    int cin, cout, batch;
    float *partial;  // [groups][cout][27*cin]
};

struct HeadArgs {
    View h;               // 16 channels
    const float *w;       // [16] (host-updated copy on the device)
    float b;
    int64_t nvox;
    float *pred;          // [nvox]
    const float *target;  // [nvox] or null (prediction only)
    float th, kappa, lam;
    float *dh;            // [nvox][16] dL/dh or null
    float *partial;       // [blocks][18]: dL/dw (16), loss, dL/db; or null
};

hipError_t launch_conv3(const ConvArgs &a, hipStream_t st);
// Brick version (Cout multiple of 16, Cin multiple of 8, D multiple of 8): partials
// [ngroups][cout][27*cin] and bias partials [ngroups][cout]; ngroups from wgrad_groups.
hipError_t launch_conv3_wgrad_brick(const WgradArgs &a, int ngroups, float *bias_partial, hipStream_t st);
int wgrad_groups(int cin, int cout, int D, int batch);
hipError_t launch_reduce_chunks(const float *partial, int nchunks, int64_t n, float *out, hipStream_t st);
hipError_t launch_avgpool(const View &in, const View &out, int C, int batch, hipStream_t st);
hipError_t launch_avgpool_bwd(const float *dout, int64_t dout_ld, const FView &din, int C, int batch, hipStream_t st);
hipError_t launch_upsample_bwd(const FView &dout, float *din, int64_t din_ld, int C, int batch, hipStream_t st);
hipError_t launch_relu_bwd(const FView &dy, const View &y, __bf16 *dz, int C, int64_t nvox, hipStream_t st);
hipError_t launch_head(const HeadArgs &a, hipStream_t st);
int head_blocks(int64_t nvox);
hipError_t launch_assemble_input(const float *params, const float *u0, const float *pw, const float *pg, int64_t D3,
                                 int batch, __bf16 *x, hipStream_t st);
hipError_t launch_adam(float *w, const float *g, float *m, float *v, int64_t n, float lr, float b1, float b2,
                       float eps, float c1, float c2, hipStream_t st);
// out[b*3 + {0,1,2}] = T1, FLAIR, PET log-likelihood terms of item b (partial: batch*nb*3)
hipError_t launch_sur_lik_flip(const float *pred, int64_t n, int batch, const uint8_t *t1, const uint8_t *t2,
                               const float *pet, double th1, double th2, double lp1, double lp0, double c0,
                               double is2, double *partial, int nb, double *out, hipStream_t st);
hipError_t launch_pack_weights(const float *w, int cin, int cout, __bf16 *wf, __bf16 *wt, hipStream_t st);

}  // namespace sur
}  // namespace gq
