// surrogate_api.cpp -- C-ABI of the 3D convolutional surrogate (gq_sur_*).
//
// Network (64^3, channels-last bf16 activations, FP32 master weights, Adam):
//   input 8 ch = {4 parameters broadcast, u0, pw, pg, 1}
//   E1 conv 8->16 @64   E2 conv 16->32 @32   E3 conv 32->64 @16   E4 conv 64->128 @8
//   (2x average pooling between levels; leaky ReLU (0.05) after every conv)
//   D3 conv [up(E4) | E3] 192->64 @16   D2 conv [up(D3) | E2] 96->32 @32
//   D1 conv [up(D2) | E1] 48->16 @64    head 1x1x1 16->1, sigmoid
// Loss: mean (p - t)^2 + lam * mean BCE(sigmoid(kappa (p - th)), [t > th]).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/gq_api.h"
#include "surrogate.hpp"

namespace {

struct sur_error : std::runtime_error {
    gq_status code;
    sur_error(gq_status c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define SCHK(x)                                                                                     \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess)                                                                       \
            throw sur_error(e_ == hipErrorOutOfMemory ? GQ_ERR_NOMEM : GQ_ERR_HIP,                  \
                            std::string(#x) + ": " + hipGetErrorString(e_));                        \
    } while (0)

template <typename T>
T *dalloc(size_t n)
{
    void *p = nullptr;
    SCHK(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
    return (T *)p;
}

struct Conv {
    int cin, cout, D;
    int64_t nw() const { return (int64_t)cout * 27 * cin; }
    float *w = nullptr, *b = nullptr, *gw = nullptr, *gb = nullptr, *mw = nullptr, *vw = nullptr, *mb = nullptr,
          *vb = nullptr;
    __bf16 *wf = nullptr, *wt = nullptr;
};

constexpr int NCONV = 7;

}  // namespace

struct gq_sur {
    int device = 0;
    hipStream_t st = nullptr;
    std::string err;
    int D = 64, bmax = 0;
    Conv cv[NCONV];  // E1 E2 E3 E4 D3 D2 D1
    float hw[16], hb = 0.f, hmw[17], hvw[17];
    float *d_hw = nullptr;
    int64_t step = 0;
    // activations (bf16)
    __bf16 *x0 = nullptr, *e1 = nullptr, *p1 = nullptr, *e2 = nullptr, *p2 = nullptr, *e3 = nullptr,
           *p3 = nullptr, *e4 = nullptr, *d3 = nullptr, *d2 = nullptr, *d1 = nullptr, *dz = nullptr;
    // gradients (FP32)
    float *dh = nullptr, *dcat1 = nullptr, *dcat2 = nullptr, *dcat3 = nullptr, *dd2 = nullptr, *dd3 = nullptr,
          *de4 = nullptr, *dp3 = nullptr, *dp2 = nullptr, *dp1 = nullptr, *partial = nullptr, *bpart = nullptr,
          *hpart = nullptr, *hred = nullptr;
    int64_t partial_n = 0;
    // inputs
    float *pw = nullptr, *pg = nullptr, *prm = nullptr, *u0 = nullptr, *tgt = nullptr, *pred = nullptr;
    // flip-rate observations on the 64^3 grid
    uint8_t *o_t1 = nullptr, *o_t2 = nullptr;
    float *o_pet = nullptr;
    double *lpart = nullptr, *lres = nullptr;
    bool have_obs = false;
    double last_fwd_ms = 0, last_bwd_ms = 0;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    std::vector<void *> allocs;

    template <typename T>
    T *A(size_t n)
    {
        T *p = dalloc<T>(n);
        allocs.push_back(p);
        return p;
    }
    ~gq_sur()
    {
        if (st) (void)hipStreamSynchronize(st);
        for (void *p : allocs) (void)hipFree(p);
        for (auto e : ev)
            if (e) (void)hipEventDestroy(e);
        if (st) (void)hipStreamDestroy(st);
    }

    int64_t nparams() const
    {
        int64_t n = 17;
        for (const auto &c : cv) n += c.nw() + c.cout;
        return n;
    }

    void init(int batch, uint64_t seed)
    {
        bmax = batch;
        const int spec[NCONV][3] = {{8, 16, 64},  {16, 32, 32}, {32, 64, 16}, {64, 128, 8},
                                    {192, 64, 16}, {96, 32, 32}, {48, 16, 64}};
        std::mt19937_64 rng(seed);
        std::normal_distribution<float> nd(0.f, 1.f);
        int64_t maxpart = 0, maxbpart = 0;
        for (int l = 0; l < NCONV; l++) {
            Conv &c = cv[l];
            c.cin = spec[l][0];
            c.cout = spec[l][1];
            c.D = spec[l][2];
            const int64_t n = c.nw();
            std::vector<float> w((size_t)n);
            const float sd = std::sqrt(2.f / (27.f * c.cin));  // He initialisation
            for (auto &x : w) x = sd * nd(rng);
            c.w = A<float>(n);
            c.gw = A<float>(n);
            c.mw = A<float>(n);
            c.vw = A<float>(n);
            c.b = A<float>(c.cout);
            c.gb = A<float>(c.cout);
            c.mb = A<float>(c.cout);
            c.vb = A<float>(c.cout);
            c.wf = A<__bf16>(n);
            c.wt = A<__bf16>(n);
            SCHK(hipMemcpy(c.w, w.data(), n * 4, hipMemcpyHostToDevice));
            for (float *p : {c.mw, c.vw}) SCHK(hipMemset(p, 0, n * 4));
            for (float *p : {c.b, c.mb, c.vb}) SCHK(hipMemset(p, 0, c.cout * 4));
            const int64_t ng = gq::sur::wgrad_groups(c.cin, c.cout, c.D, batch);
            maxpart = std::max(maxpart, ng * n);
            maxbpart = std::max(maxbpart, ng * c.cout);
        }
        for (int c = 0; c < 16; c++) hw[c] = 0.25f * nd(rng);
        hb = -3.f;  // start near the background value (most voxels hold no tumour)
        std::fill(hmw, hmw + 17, 0.f);
        std::fill(hvw, hvw + 17, 0.f);
        d_hw = A<float>(16);
        const int64_t V64 = (int64_t)64 * 64 * 64 * batch, V32 = V64 / 8, V16 = V32 / 8, V8 = V16 / 8;
        x0 = A<__bf16>(V64 * 8);
        e1 = A<__bf16>(V64 * 16);
        p1 = A<__bf16>(V32 * 16);
        e2 = A<__bf16>(V32 * 32);
        p2 = A<__bf16>(V16 * 32);
        e3 = A<__bf16>(V16 * 64);
        p3 = A<__bf16>(V8 * 64);
        e4 = A<__bf16>(V8 * 128);
        d3 = A<__bf16>(V16 * 64);
        d2 = A<__bf16>(V32 * 32);
        d1 = A<__bf16>(V64 * 16);
        dz = A<__bf16>(V64 * 16);  // largest dZ: 64^3 x 16 (also 32^3 x 32, 16^3 x 64, 8^3 x 128)
        dh = A<float>(V64 * 16);
        dcat1 = A<float>(V64 * 48);
        dcat2 = A<float>(V32 * 96);
        dcat3 = A<float>(V16 * 192);
        dd2 = A<float>(V32 * 32);
        dd3 = A<float>(V16 * 64);
        de4 = A<float>(V8 * 128);
        dp3 = A<float>(V8 * 64);
        dp2 = A<float>(V16 * 32);
        dp1 = A<float>(V32 * 16);
        partial_n = std::max<int64_t>(maxpart, 1);
        partial = A<float>(partial_n);
        bpart = A<float>(std::max<int64_t>(maxbpart, 1));
        hpart = A<float>((size_t)gq::sur::head_blocks(V64) * 18);
        hred = A<float>(18);
        pw = A<float>(64 * 64 * 64);
        pg = A<float>(64 * 64 * 64);
        prm = A<float>((size_t)batch * 4);
        u0 = A<float>(V64);
        tgt = A<float>(V64);
        pred = A<float>(V64);
        o_t1 = A<uint8_t>(64 * 64 * 64);
        o_t2 = A<uint8_t>(64 * 64 * 64);
        o_pet = A<float>(64 * 64 * 64);
        lpart = A<double>((size_t)batch * 256 * 3);
        lres = A<double>((size_t)batch * 3);
        for (auto &e : ev) SCHK(hipEventCreate(&e));
        SCHK(hipDeviceSynchronize());  // the synchronous copies and fills above precede any work on st
        pack();
    }

    void pack()
    {
        for (auto &c : cv) SCHK(gq::sur::launch_pack_weights(c.w, c.cin, c.cout, c.wf, c.wt, st));
        SCHK(hipMemcpyAsync(d_hw, hw, sizeof(hw), hipMemcpyHostToDevice, st));
        SCHK(hipStreamSynchronize(st));  // hw is read by the copy
    }

    static gq::sur::View V(__bf16 *p, int D, int64_t ld, int64_t off) { return gq::sur::View{p, D, ld, off}; }
    static gq::sur::FView F(float *p, int D, int64_t ld, int64_t off) { return gq::sur::FView{p, D, ld, off}; }

    // conv l on input channels [up (upsampled, cup channels), in]; bf16 out at (ld, off)
    void conv_fwd(int l, gq::sur::View in, __bf16 *out, int64_t ld, int64_t off, int B,
                  gq::sur::View up = gq::sur::View{}, int cup = 0)
    {
        const Conv &c = cv[l];
        gq::sur::ConvArgs a{};
        a.in = in;
        a.up = up;
        a.cup = cup;
        a.cin = c.cin;
        a.cout = c.cout;
        a.batch = B;
        a.w = c.wf;
        a.bias = c.b;
        a.relu = 1;
        a.out_bf = out;
        a.out_f = nullptr;
        a.accumulate = 0;
        a.out_ld = ld;
        a.out_off = off;
        SCHK(gq::sur::launch_conv3(a, st));
    }

    // dX of conv l from dZ (bf16, cout channels): FP32 into out (ld, off)
    void conv_bwd_data(int l, float *out, int64_t ld, int64_t off, int B)
    {
        const Conv &c = cv[l];
        gq::sur::ConvArgs a{};
        a.in = V(dz, c.D, c.cout, 0);
        a.cin = c.cout;
        a.cout = c.cin;
        a.batch = B;
        a.w = c.wt;
        a.bias = nullptr;
        a.relu = 0;
        a.out_bf = nullptr;
        a.out_f = out;
        a.accumulate = 0;
        a.out_ld = ld;
        a.out_off = off;
        SCHK(gq::sur::launch_conv3(a, st));
    }

    void conv_wgrad(int l, gq::sur::View x, int B, gq::sur::View up = gq::sur::View{}, int cup = 0)
    {
        const Conv &c = cv[l];
        const int ng = gq::sur::wgrad_groups(c.cin, c.cout, c.D, B);
        gq::sur::WgradArgs a{};
        a.x = x;
        a.up = up;
        a.cup = cup;
        a.dz = V(dz, c.D, c.cout, 0);
        a.cin = c.cin;
        a.cout = c.cout;
        a.batch = B;
        a.partial = partial;
        SCHK(gq::sur::launch_conv3_wgrad_brick(a, ng, bpart, st));
        SCHK(gq::sur::launch_reduce_chunks(partial, ng, c.nw(), c.gw, st));
        SCHK(gq::sur::launch_reduce_chunks(bpart, ng, c.cout, c.gb, st));
    }

    void upload_batch(const float *params, const float *u0h, const float *th, int B)
    {
        if (B < 1 || B > bmax) throw sur_error(GQ_ERR_ARG, "batch outside [1, batch capacity]");
        const size_t V = (size_t)64 * 64 * 64 * B;
        SCHK(hipMemcpyAsync(prm, params, (size_t)B * 16, hipMemcpyHostToDevice, st));
        SCHK(hipMemcpyAsync(u0, u0h, V * 4, hipMemcpyHostToDevice, st));
        if (th) SCHK(hipMemcpyAsync(tgt, th, V * 4, hipMemcpyHostToDevice, st));
    }

    // forward; with a target also the loss, head gradients and dL/dh
    void forward(int B, bool with_target, bool want_dh)
    {
        const int64_t D3 = 64 * 64 * 64;
        SCHK(gq::sur::launch_assemble_input(prm, u0, pw, pg, D3, B, x0, st));
        // decoder inputs are [upsampled coarser output, skip]: the upsampling and the
        // concatenation happen in the convolutions' reads (ConvArgs::up / cup)
        conv_fwd(0, V(x0, 64, 8, 0), e1, 16, 0, B);                                   // E1
        SCHK(gq::sur::launch_avgpool(V(e1, 64, 16, 0), V(p1, 32, 16, 0), 16, B, st));
        conv_fwd(1, V(p1, 32, 16, 0), e2, 32, 0, B);                                  // E2
        SCHK(gq::sur::launch_avgpool(V(e2, 32, 32, 0), V(p2, 16, 32, 0), 32, B, st));
        conv_fwd(2, V(p2, 16, 32, 0), e3, 64, 0, B);                                  // E3
        SCHK(gq::sur::launch_avgpool(V(e3, 16, 64, 0), V(p3, 8, 64, 0), 64, B, st));
        conv_fwd(3, V(p3, 8, 64, 0), e4, 128, 0, B);                                  // E4
        conv_fwd(4, V(e3, 16, 64, 0), d3, 64, 0, B, V(e4, 8, 128, 0), 128);          // D3 on [up(E4), E3]
        conv_fwd(5, V(e2, 32, 32, 0), d2, 32, 0, B, V(d3, 16, 64, 0), 64);           // D2 on [up(D3), E2]
        conv_fwd(6, V(e1, 64, 16, 0), d1, 16, 0, B, V(d2, 32, 32, 0), 32);           // D1 on [up(D2), E1]
        gq::sur::HeadArgs h{};
        h.h = V(d1, 64, 16, 0);
        h.w = d_hw;
        h.b = hb;
        h.nvox = D3 * B;
        h.pred = pred;
        h.target = with_target ? tgt : nullptr;
        h.th = 0.25f;
        h.kappa = 20.f;
        h.lam = 0.1f;
        h.dh = want_dh ? dh : nullptr;
        h.partial = with_target ? hpart : nullptr;
        SCHK(gq::sur::launch_head(h, st));
        if (with_target) SCHK(gq::sur::launch_reduce_chunks(hpart, gq::sur::head_blocks(h.nvox), 18, hred, st));
    }

    void backward(int B)
    {
        const int64_t V64 = (int64_t)64 * 64 * 64 * B, V32 = V64 / 8, V16 = V32 / 8, V8 = V16 / 8;
        // D1
        SCHK(gq::sur::launch_relu_bwd(F(dh, 64, 16, 0), V(d1, 64, 16, 0), dz, 16, V64, st));
        conv_wgrad(6, V(e1, 64, 16, 0), B, V(d2, 32, 32, 0), 32);
        conv_bwd_data(6, dcat1, 48, 0, B);
        // D2
        SCHK(gq::sur::launch_upsample_bwd(F(dcat1, 64, 48, 0), dd2, 32, 32, B, st));
        SCHK(gq::sur::launch_relu_bwd(F(dd2, 32, 32, 0), V(d2, 32, 32, 0), dz, 32, V32, st));
        conv_wgrad(5, V(e2, 32, 32, 0), B, V(d3, 16, 64, 0), 64);
        conv_bwd_data(5, dcat2, 96, 0, B);
        // D3
        SCHK(gq::sur::launch_upsample_bwd(F(dcat2, 32, 96, 0), dd3, 64, 64, B, st));
        SCHK(gq::sur::launch_relu_bwd(F(dd3, 16, 64, 0), V(d3, 16, 64, 0), dz, 64, V16, st));
        conv_wgrad(4, V(e3, 16, 64, 0), B, V(e4, 8, 128, 0), 128);
        conv_bwd_data(4, dcat3, 192, 0, B);
        // E4
        SCHK(gq::sur::launch_upsample_bwd(F(dcat3, 16, 192, 0), de4, 128, 128, B, st));
        SCHK(gq::sur::launch_relu_bwd(F(de4, 8, 128, 0), V(e4, 8, 128, 0), dz, 128, V8, st));
        conv_wgrad(3, V(p3, 8, 64, 0), B);
        conv_bwd_data(3, dp3, 64, 0, B);
        // E3: skip gradient in dcat3[128..191] plus the pooled path
        SCHK(gq::sur::launch_avgpool_bwd(dp3, 64, F(dcat3, 16, 192, 128), 64, B, st));
        SCHK(gq::sur::launch_relu_bwd(F(dcat3, 16, 192, 128), V(e3, 16, 64, 0), dz, 64, V16, st));
        conv_wgrad(2, V(p2, 16, 32, 0), B);
        conv_bwd_data(2, dp2, 32, 0, B);
        // E2
        SCHK(gq::sur::launch_avgpool_bwd(dp2, 32, F(dcat2, 32, 96, 64), 32, B, st));
        SCHK(gq::sur::launch_relu_bwd(F(dcat2, 32, 96, 64), V(e2, 32, 32, 0), dz, 32, V32, st));
        conv_wgrad(1, V(p1, 32, 16, 0), B);
        conv_bwd_data(1, dp1, 16, 0, B);
        // E1
        SCHK(gq::sur::launch_avgpool_bwd(dp1, 16, F(dcat1, 64, 48, 32), 16, B, st));
        SCHK(gq::sur::launch_relu_bwd(F(dcat1, 64, 48, 32), V(e1, 64, 16, 0), dz, 16, V64, st));
        conv_wgrad(0, V(x0, 64, 8, 0), B);
    }

    // loss and the head gradients (17) from the reduced head partials
    double read_head(float ghead[17])
    {
        float h[18];
        SCHK(hipMemcpyAsync(h, hred, sizeof(h), hipMemcpyDeviceToHost, st));
        SCHK(hipStreamSynchronize(st));
        for (int c = 0; c < 16; c++) ghead[c] = h[c];
        ghead[16] = h[17];
        return h[16];
    }

    void adam(const float ghead[17], float lr)
    {
        step++;
        const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
        const float c1 = 1.f - std::pow(b1, (float)step), c2 = 1.f - std::pow(b2, (float)step);
        for (auto &c : cv) {
            SCHK(gq::sur::launch_adam(c.w, c.gw, c.mw, c.vw, c.nw(), lr, b1, b2, eps, c1, c2, st));
            SCHK(gq::sur::launch_adam(c.b, c.gb, c.mb, c.vb, c.cout, lr, b1, b2, eps, c1, c2, st));
        }
        for (int k = 0; k < 17; k++) {
            const float g = ghead[k];
            hmw[k] = b1 * hmw[k] + (1.f - b1) * g;
            hvw[k] = b2 * hvw[k] + (1.f - b2) * g * g;
            const float d = lr * (hmw[k] / c1) / (std::sqrt(hvw[k] / c2) + eps);
            if (k < 16) hw[k] -= d;
            else hb -= d;
        }
        pack();
    }
};

namespace {
gq_status sur_guard(gq_sur *s, const std::function<void()> &f)
{
    if (!s) return GQ_ERR_ARG;
    try {
        s->err.clear();
        SCHK(hipSetDevice(s->device));
        f();
        return GQ_OK;
    } catch (const sur_error &e) {
        s->err = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        s->err = "host allocation failed";
        return GQ_ERR_NOMEM;
    } catch (const std::exception &e) {
        s->err = e.what();
        return GQ_ERR_HIP;
    }
}
}  // namespace

extern "C" {

gq_status gq_sur_create(gq_sur **out, int device, int batch, uint64_t seed)
{
    if (!out || batch < 1 || batch > 64) return GQ_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return GQ_ERR_HIP;
    if (device < 0 || device >= ndev) return GQ_ERR_ARG;
    gq_sur *s = new (std::nothrow) gq_sur();
    if (!s) return GQ_ERR_NOMEM;
    s->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking) != hipSuccess) {
        delete s;
        return GQ_ERR_HIP;
    }
    try {
        s->init(batch, seed);
    } catch (const sur_error &e) {
        delete s;
        return e.code;
    } catch (...) {
        delete s;
        return GQ_ERR_NOMEM;
    }
    *out = s;
    return GQ_OK;
}

gq_status gq_sur_destroy(gq_sur *s)
{
    if (!s) return GQ_ERR_ARG;
    (void)hipSetDevice(s->device);
    delete s;
    return GQ_OK;
}

const char *gq_sur_last_error(const gq_sur *s) { return s ? s->err.c_str() : ""; }

gq_status gq_sur_param_count(gq_sur *s, int64_t *n)
{
    return sur_guard(s, [&] {
        if (!n) throw sur_error(GQ_ERR_ARG, "null output");
        *n = s->nparams();
    });
}

gq_status gq_sur_get_params(gq_sur *s, float *out)
{
    return sur_guard(s, [&] {
        if (!out) throw sur_error(GQ_ERR_ARG, "null output");
        float *o = out;
        for (auto &c : s->cv) {
            SCHK(hipMemcpy(o, c.w, c.nw() * 4, hipMemcpyDeviceToHost));
            o += c.nw();
            SCHK(hipMemcpy(o, c.b, c.cout * 4, hipMemcpyDeviceToHost));
            o += c.cout;
        }
        std::memcpy(o, s->hw, 64);
        o[16] = s->hb;
    });
}

gq_status gq_sur_set_params(gq_sur *s, const float *in)
{
    return sur_guard(s, [&] {
        if (!in) throw sur_error(GQ_ERR_ARG, "null input");
        const float *p = in;
        for (auto &c : s->cv) {
            SCHK(hipMemcpy(c.w, p, c.nw() * 4, hipMemcpyHostToDevice));
            p += c.nw();
            SCHK(hipMemcpy(c.b, p, c.cout * 4, hipMemcpyHostToDevice));
            p += c.cout;
        }
        std::memcpy(s->hw, p, 64);
        s->hb = p[16];
        s->pack();
    });
}

gq_status gq_sur_set_anatomy(gq_sur *s, const float *pw, const float *pg)
{
    return sur_guard(s, [&] {
        if (!pw || !pg) throw sur_error(GQ_ERR_ARG, "null anatomy");
        SCHK(hipMemcpy(s->pw, pw, 64 * 64 * 64 * 4, hipMemcpyHostToDevice));
        SCHK(hipMemcpy(s->pg, pg, 64 * 64 * 64 * 4, hipMemcpyHostToDevice));
    });
}

gq_status gq_sur_predict(gq_sur *s, const float *params, const float *u0, int batch, float *out)
{
    return sur_guard(s, [&] {
        if (!params || !u0 || !out) throw sur_error(GQ_ERR_ARG, "null argument");
        s->upload_batch(params, u0, nullptr, batch);
        SCHK(hipEventRecord(s->ev[0], s->st));
        s->forward(batch, false, false);
        SCHK(hipEventRecord(s->ev[1], s->st));
        SCHK(hipMemcpyAsync(out, s->pred, (size_t)64 * 64 * 64 * batch * 4, hipMemcpyDeviceToHost, s->st));
        SCHK(hipStreamSynchronize(s->st));
        float ms = 0;
        SCHK(hipEventElapsedTime(&ms, s->ev[0], s->ev[1]));
        s->last_fwd_ms = ms;
    });
}

// loss and gradients without an update (grads: gq_sur_param_count floats, same layout as gq_sur_get_params)
gq_status gq_sur_loss_grad(gq_sur *s, const float *params, const float *u0, const float *target, int batch,
                           double *loss, float *grads)
{
    return sur_guard(s, [&] {
        if (!params || !u0 || !target) throw sur_error(GQ_ERR_ARG, "null argument");
        s->upload_batch(params, u0, target, batch);
        SCHK(hipEventRecord(s->ev[0], s->st));
        s->forward(batch, true, true);
        SCHK(hipEventRecord(s->ev[1], s->st));
        s->backward(batch);
        SCHK(hipEventRecord(s->ev[2], s->st));
        float gh[17];
        const double l = s->read_head(gh);
        if (loss) *loss = l;
        if (grads) {
            float *o = grads;
            for (auto &c : s->cv) {
                SCHK(hipMemcpy(o, c.gw, c.nw() * 4, hipMemcpyDeviceToHost));
                o += c.nw();
                SCHK(hipMemcpy(o, c.gb, c.cout * 4, hipMemcpyDeviceToHost));
                o += c.cout;
            }
            std::memcpy(o, gh, 17 * 4);
        }
        float f = 0, b = 0;
        SCHK(hipEventElapsedTime(&f, s->ev[0], s->ev[1]));
        SCHK(hipEventElapsedTime(&b, s->ev[1], s->ev[2]));
        s->last_fwd_ms = f;
        s->last_bwd_ms = b;
    });
}

gq_status gq_sur_train_step(gq_sur *s, const float *params, const float *u0, const float *target, int batch, float lr,
                            double *loss)
{
    return sur_guard(s, [&] {
        if (!params || !u0 || !target || !(lr > 0)) throw sur_error(GQ_ERR_ARG, "bad arguments");
        s->upload_batch(params, u0, target, batch);
        SCHK(hipEventRecord(s->ev[0], s->st));
        s->forward(batch, true, true);
        SCHK(hipEventRecord(s->ev[1], s->st));
        s->backward(batch);
        SCHK(hipEventRecord(s->ev[2], s->st));
        float gh[17];
        const double l = s->read_head(gh);
        if (loss) *loss = l;
        s->adam(gh, lr);
        float f = 0, b = 0;
        SCHK(hipEventElapsedTime(&f, s->ev[0], s->ev[1]));
        SCHK(hipEventElapsedTime(&b, s->ev[1], s->ev[2]));
        s->last_fwd_ms = f;
        s->last_bwd_ms = b;
    });
}

gq_status gq_sur_obs_flip_upload(gq_sur *s, const uint8_t *t1, const uint8_t *t2, const float *pet)
{
    return sur_guard(s, [&] {
        if (!t1 || !t2 || !pet) throw sur_error(GQ_ERR_ARG, "null observations");
        const size_t n = (size_t)64 * 64 * 64;
        SCHK(hipMemcpy(s->o_t1, t1, n, hipMemcpyHostToDevice));
        SCHK(hipMemcpy(s->o_t2, t2, n, hipMemcpyHostToDevice));
        SCHK(hipMemcpy(s->o_pet, pet, n * 4, hipMemcpyHostToDevice));
        s->have_obs = true;
    });
}

gq_status gq_sur_loglik_flip(gq_sur *s, const float *params, const float *u0, int batch, const double lik[4],
                             double *out)
{
    return sur_guard(s, [&] {
        if (!params || !u0 || !lik || !out) throw sur_error(GQ_ERR_ARG, "null argument");
        if (!s->have_obs) throw sur_error(GQ_ERR_STATE, "observations missing");
        const double eps = lik[2], sigma = lik[3];
        if (!(eps > 0 && eps < 1) || !(sigma > 0)) throw sur_error(GQ_ERR_NUMERIC, "flip rate / sigma out of range");
        s->upload_batch(params, u0, nullptr, batch);
        SCHK(hipEventRecord(s->ev[0], s->st));
        s->forward(batch, false, false);
        SCHK(hipEventRecord(s->ev[1], s->st));
        const double lp1 = std::log(1. - eps), lp0 = std::log(eps);
        const double c0 = -0.5 * std::log(2. * M_PI * sigma * sigma);
        const double is2 = 1. / (2. * sigma * sigma);
        SCHK(gq::sur::launch_sur_lik_flip(s->pred, (int64_t)64 * 64 * 64, batch, s->o_t1, s->o_t2, s->o_pet, lik[0],
                                          lik[1], lp1, lp0, c0, is2, s->lpart, 256, s->lres, s->st));
        std::vector<double> h((size_t)batch * 3);
        SCHK(hipMemcpyAsync(h.data(), s->lres, h.size() * 8, hipMemcpyDeviceToHost, s->st));
        SCHK(hipStreamSynchronize(s->st));
        for (int b = 0; b < batch; b++) {
            out[4 * b + 1] = h[3 * b];
            out[4 * b + 2] = h[3 * b + 1];
            out[4 * b + 3] = h[3 * b + 2];
            out[4 * b] = h[3 * b] + h[3 * b + 1] + h[3 * b + 2];
        }
        float f = 0;
        SCHK(hipEventElapsedTime(&f, s->ev[0], s->ev[1]));
        s->last_fwd_ms = f;
    });
}

gq_status gq_sur_timing(gq_sur *s, double out[2])
{
    return sur_guard(s, [&] {
        if (!out) throw sur_error(GQ_ERR_ARG, "null output");
        out[0] = s->last_fwd_ms;
        out[1] = s->last_bwd_ms;
    });
}

}  // extern "C"
