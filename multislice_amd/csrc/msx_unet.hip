// msx_unet.hip -- the frozen-phonon surrogate: a U-Net that maps the mean ADF
// image of a few frozen-phonon configurations to the converged many-config
// mean.  Host-side orchestration of the kernels in msx_unet_kernels.hip and the
// msx_unet_* entry points of include/msx.h.
//
// Architecture (levels L, base channels B, C_l = B << l):
//   encoder l = 0..L-1 : conv3x3(C_{l-1} -> C_l)+ReLU, conv3x3(C_l -> C_l)+ReLU,
//                        2x2 average pool between levels (the input has
//                        C_{-1} = in_channels channels: one image per
//                        frozen-phonon configuration, or their mean)
//   decoder l = L-2..0 : nearest x2 upsample, concat [up, skip_l],
//                        conv3x3(C_{l+1}+C_l -> C_l)+ReLU, conv3x3(C_l -> C_l)+ReLU
//   head               : conv1x1(C_0 -> 1), output = channel mean of the input
//                        + head (residual)
// Loss: mean squared error over all pixels; optimiser: Adam.
// Parameters are one packed FP32 vector, conv by conv in the order above, each
// conv as W[Cout][Cin][k*k] then bias[Cout].
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "../../include/msx.h"
#include "msx_internal.h"
#include "msx_unet.h"

using namespace msx;
using namespace msx::unet;

#define UHIP(x)                                                                                      \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess)                                                                        \
            return fail(e_ == hipErrorOutOfMemory ? MSX_ENOMEM : MSX_EHIP, "%s: %s (%s:%d)", #x,     \
                        hipGetErrorString(e_), __FILE__, __LINE__);                                  \
    } while (0)
#define UCHK(x)                      \
    do {                             \
        int r_ = (x);                \
        if (r_ != MSX_OK) return r_; \
    } while (0)

namespace {

struct ConvSpec {
    int cin, cout, ks, relu;
    size_t w_off, b_off;
};

constexpr int kSqBlocks = 256;

}  // namespace

struct msx_unet_s {
    msx_ctx ctx = nullptr;  // must outlive every call except msx_unet_destroy
    int device = 0;
    int levels = 0, base = 0, cin0 = 1;
    std::vector<ConvSpec> convs;
    size_t nparam = 0;
    float* pbuf = nullptr;  // params | grads | adam m | adam v
    long long step = 0;
    // activation workspace, laid out for (wn, wh, ww)
    int wn = 0, wh = 0, ww = 0;
    float* ws = nullptr;
    size_t ws_bytes = 0;
    double* partial = nullptr;
    float* wscr = nullptr;  // weight-gradient slices (conv_wgrad_scratch)
    size_t wscr_n = 0;
    std::vector<float*> ea, eb, pl, cat, da, db, g_ea, g_eb, g_pl, g_cat, g_da, g_db;
    float *xin = nullptr, *tin = nullptr, *z = nullptr, *out = nullptr, *gz = nullptr;
    float* xbase = nullptr;  // channel mean of xin (xin itself with one channel)
    float* grad_begin = nullptr;
    size_t grad_bytes = 0;

    float* params() { return pbuf; }
    float* grads() { return pbuf + nparam; }
    float* adam_m() { return pbuf + 2 * nparam; }
    float* adam_v() { return pbuf + 3 * nparam; }
    int C(int l) const { return base << l; }
    int enc_a(int l) const { return 2 * l; }
    int enc_b(int l) const { return 2 * l + 1; }
    int dec_a(int l) const { return 2 * levels + 2 * (levels - 2 - l); }
    int dec_b(int l) const { return dec_a(l) + 1; }
    int head() const { return (int)convs.size() - 1; }
    hipStream_t stream() const { return msx_ctx_stream(ctx); }

    ~msx_unet_s() {
        if (pbuf) hipFree(pbuf);
        if (ws) hipFree(ws);
        if (partial) hipFree(partial);
        if (wscr) hipFree(wscr);
    }

    // lay out every activation and gradient tensor for batch n of h x w images
    int workspace(int n, int h, int w) {
        if (ws && n == wn && h == wh && w == ww) return MSX_OK;
        const int L = levels;
        auto hw = [&](int l) { return (size_t)(h >> l) * (size_t)(w >> l); };
        size_t total = 0;
        struct Slot {
            float** dst;
            size_t n;
        };
        std::vector<Slot> slots;
        ea.assign(L, nullptr);
        eb.assign(L, nullptr);
        pl.assign(L, nullptr);
        cat.assign(L, nullptr);
        da.assign(L, nullptr);
        db.assign(L, nullptr);
        g_ea.assign(L, nullptr);
        g_eb.assign(L, nullptr);
        g_pl.assign(L, nullptr);
        g_cat.assign(L, nullptr);
        g_da.assign(L, nullptr);
        g_db.assign(L, nullptr);
        const size_t img = (size_t)n * h * w;
        slots.push_back({&xin, img * cin0});
        if (cin0 > 1) slots.push_back({&xbase, img});
        slots.push_back({&tin, img});
        slots.push_back({&z, img});
        slots.push_back({&out, img});
        for (int l = 0; l < L; l++) {
            slots.push_back({&ea[l], (size_t)n * C(l) * hw(l)});
            slots.push_back({&eb[l], (size_t)n * C(l) * hw(l)});
            if (l < L - 1) {
                slots.push_back({&pl[l], (size_t)n * C(l) * hw(l + 1)});
                slots.push_back({&cat[l], (size_t)n * (C(l + 1) + C(l)) * hw(l)});
                slots.push_back({&da[l], (size_t)n * C(l) * hw(l)});
                slots.push_back({&db[l], (size_t)n * C(l) * hw(l)});
            }
        }
        const size_t grad_first = slots.size();
        slots.push_back({&gz, img});
        for (int l = 0; l < L; l++) {
            slots.push_back({&g_ea[l], (size_t)n * C(l) * hw(l)});
            slots.push_back({&g_eb[l], (size_t)n * C(l) * hw(l)});
            if (l < L - 1) {
                slots.push_back({&g_pl[l], (size_t)n * C(l) * hw(l + 1)});
                slots.push_back({&g_cat[l], (size_t)n * (C(l + 1) + C(l)) * hw(l)});
                slots.push_back({&g_da[l], (size_t)n * C(l) * hw(l)});
                slots.push_back({&g_db[l], (size_t)n * C(l) * hw(l)});
            }
        }
        std::vector<size_t> offs(slots.size());
        for (size_t i = 0; i < slots.size(); i++) {
            offs[i] = total;
            total += (slots[i].n + 63) / 64 * 64;  // 256-byte aligned tensors
        }
        const size_t bytes = total * sizeof(float);
        if (bytes > ws_bytes) {
            UHIP(hipStreamSynchronize(stream()));
            if (ws) hipFree(ws);
            ws = nullptr;
            ws_bytes = 0;
            UHIP(hipMalloc(&ws, bytes));
            ws_bytes = bytes;
        }
        for (size_t i = 0; i < slots.size(); i++) *slots[i].dst = ws + offs[i];
        if (cin0 == 1) xbase = xin;
        grad_begin = ws + offs[grad_first];
        grad_bytes = (total - offs[grad_first]) * sizeof(float);
        wn = n;
        wh = h;
        ww = w;
        return MSX_OK;
    }

    ConvArgs conv_args(int i, int n, int l) {
        const ConvSpec& c = convs[i];
        ConvArgs a{};
        a.w = params() + c.w_off;
        a.bias = params() + c.b_off;
        a.N = n;
        a.Cin = c.cin;
        a.Cout = c.cout;
        a.H = wh >> l;
        a.W = ww >> l;
        a.relu = c.relu;
        return a;
    }
    int conv_fwd(int i, int l, const float* x, float* y) {
        ConvArgs a = conv_args(i, wn, l);
        a.x = x;
        a.out = y;
        a.out_c = a.Cout;
        UHIP(launch_conv(kFwd, convs[i].ks, a, stream()));
        return MSX_OK;
    }
    // weight, bias and (if gx) input gradients of conv i; x = its input,
    // y = its output, g = gradient w.r.t. y (masked by ReLU inside the kernels)
    int conv_bwd(int i, int l, const float* x, const float* y, const float* g, float* gx) {
        const ConvSpec& c = convs[i];
        ConvArgs a = conv_args(i, wn, l);
        a.x = x;
        a.y = y;
        a.g = g;
        a.out = grads() + c.w_off;
        const size_t need = conv_wgrad_scratch(c.ks, a);
        if (need > wscr_n) {
            UHIP(hipStreamSynchronize(stream()));
            if (wscr) hipFree(wscr);
            wscr = nullptr;
            wscr_n = 0;
            UHIP(hipMalloc(&wscr, need * sizeof(float)));
            wscr_n = need;
        }
        a.partial = wscr;
        UHIP(launch_conv(kBwdWeight, c.ks, a, stream()));
        UHIP(launch_bias_grad(g, y, c.relu, wn, c.cout, a.H * a.W, grads() + c.b_off, stream()));
        if (gx) {
            a.out = gx;
            a.out_c = c.cin;
            a.accumulate = 1;
            UHIP(launch_conv(kBwdData, c.ks, a, stream()));
        }
        return MSX_OK;
    }

    // forward from xin; leaves out = xin + head(...)
    int forward() {
        const int L = levels, n = wn;
        for (int l = 0; l < L; l++) {
            UCHK(conv_fwd(enc_a(l), l, l == 0 ? xin : pl[l - 1], ea[l]));
            UCHK(conv_fwd(enc_b(l), l, ea[l], eb[l]));
            if (l < L - 1) UHIP(launch_avgpool_fwd(eb[l], pl[l], (long long)n * C(l), wh >> l, ww >> l, stream()));
        }
        const float* prev = eb[L - 1];
        for (int l = L - 2; l >= 0; l--) {
            const int Ct = C(l + 1) + C(l), hl = wh >> l, wl = ww >> l;
            UHIP(launch_upsample_fwd(prev, cat[l], n, C(l + 1), hl / 2, wl / 2, Ct, 0, stream()));
            UHIP(launch_copy_channels(eb[l], C(l), 0, cat[l], Ct, C(l + 1), n, C(l), hl * wl, 0, stream()));
            UCHK(conv_fwd(dec_a(l), l, cat[l], da[l]));
            UCHK(conv_fwd(dec_b(l), l, da[l], db[l]));
            prev = db[l];
        }
        UCHK(conv_fwd(head(), 0, prev, z));
        return MSX_OK;
    }

    // loss = mean (out - t)^2 (out computed here), and every gradient
    int backward(double* loss) {
        const int L = levels, n = wn;
        const long long cnt = (long long)n * wh * ww;
        // clear the accumulated gradients first: gz is the first tensor of the region
        UHIP(hipMemsetAsync(grads(), 0, nparam * sizeof(float), stream()));
        UHIP(hipMemsetAsync(grad_begin, 0, grad_bytes, stream()));
        UHIP(launch_residual_mse(xbase, z, tin, out, gz, cnt, 2.0f / (float)cnt, stream()));
        UHIP(launch_sq_diff(out, tin, cnt, partial, kSqBlocks, stream()));
        const float* prev = L > 1 ? db[0] : eb[0];
        float* g_prev = L > 1 ? g_db[0] : g_eb[0];
        UCHK(conv_bwd(head(), 0, prev, z, gz, g_prev));
        for (int l = 0; l <= L - 2; l++) {
            const int Ct = C(l + 1) + C(l), hl = wh >> l, wl = ww >> l;
            UCHK(conv_bwd(dec_b(l), l, da[l], db[l], g_db[l], g_da[l]));
            UCHK(conv_bwd(dec_a(l), l, cat[l], da[l], g_da[l], g_cat[l]));
            float* g_up = (l + 1 <= L - 2) ? g_db[l + 1] : g_eb[L - 1];
            UHIP(launch_upsample_bwd(g_cat[l], g_up, n, C(l + 1), hl / 2, wl / 2, Ct, 0, stream()));
            UHIP(launch_copy_channels(g_cat[l], Ct, C(l + 1), g_eb[l], C(l), 0, n, C(l), hl * wl, 1, stream()));
        }
        for (int l = L - 1; l >= 0; l--) {
            UCHK(conv_bwd(enc_b(l), l, ea[l], eb[l], g_eb[l], g_ea[l]));
            UCHK(conv_bwd(enc_a(l), l, l == 0 ? xin : pl[l - 1], ea[l], g_ea[l], l > 0 ? g_pl[l - 1] : nullptr));
            if (l > 0)
                UHIP(launch_avgpool_bwd(g_pl[l - 1], g_eb[l - 1], (long long)n * C(l - 1), wh >> (l - 1),
                                        ww >> (l - 1), stream()));
        }
        if (loss) {
            std::vector<double> h(kSqBlocks);
            UHIP(hipMemcpyAsync(h.data(), partial, sizeof(double) * kSqBlocks, hipMemcpyDeviceToHost, stream()));
            UHIP(hipStreamSynchronize(stream()));
            double s = 0.0;
            for (double v : h) s += v;
            *loss = s / (double)cnt;
        }
        return MSX_OK;
    }
};

namespace {

int check_net(msx_unet net) {
    if (!net) return fail(MSX_ESTATE, "net is NULL");
    return MSX_OK;
}
int check_shape(msx_unet net, int n, int h, int w, int where) {
    if (where != MSX_HOST && where != MSX_DEVICE) return fail(MSX_EINVAL, "where must be MSX_HOST or MSX_DEVICE");
    const int q = 1 << (net->levels - 1);
    if (n < 1 || h < q || w < q || h % q || w % q || h > 4096 || w > 4096)
        return fail(MSX_EINVAL, "image batch %d x %d x %d: h and w must be multiples of %d (levels %d), <= 4096", n,
                    h, w, q, net->levels);
    if ((long long)n * h * w > (1LL << 26)) return fail(MSX_EINVAL, "batch of %lld pixels too large", (long long)n * h * w);
    return MSX_OK;
}
// stage the inputs (x [n][cin0][h][w], optional target [n][h][w]) into the
// workspace, with the channel mean of x for the residual
int stage_inputs(msx_unet net, const float* x, const float* t, int where) {
    const size_t npix = (size_t)net->wn * net->wh * net->ww, bytes = sizeof(float) * npix;
    const hipMemcpyKind k = where == MSX_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
    UHIP(hipMemcpyAsync(net->xin, x, bytes * net->cin0, k, net->stream()));
    if (net->cin0 > 1)
        UHIP(launch_channel_mean(net->xin, net->xbase, net->wn, net->cin0, (long long)net->wh * net->ww,
                                 net->stream()));
    if (t) UHIP(hipMemcpyAsync(net->tin, t, bytes, k, net->stream()));
    return MSX_OK;
}

}  // namespace

// ============================================================== C ABI
extern "C" {

MSX_API int msx_unet_create_in(msx_ctx ctx, int levels, int base_channels, int in_channels, uint64_t seed,
                               msx_unet* out) {
    if (!out) return fail(MSX_EINVAL, "out is NULL");
    *out = nullptr;
    if (!ctx) return fail(MSX_ESTATE, "ctx is NULL");
    if (levels < 1 || levels > 6) return fail(MSX_EINVAL, "levels %d outside 1..6", levels);
    if (in_channels < 1 || in_channels > 64) return fail(MSX_EINVAL, "in_channels %d outside 1..64", in_channels);
    if (base_channels < 1 || (base_channels << (levels - 1)) > 1024)
        return fail(MSX_EINVAL, "base_channels %d: widest level must be <= 1024 channels", base_channels);
    UHIP(hipSetDevice(msx_ctx_device(ctx)));
    msx_unet net = new msx_unet_s();
    net->ctx = ctx;
    net->device = msx_ctx_device(ctx);
    net->levels = levels;
    net->base = base_channels;
    net->cin0 = in_channels;
    size_t off = 0;
    auto add = [&](int cin, int cout, int ks, int relu) {
        ConvSpec c{cin, cout, ks, relu, off, off + (size_t)cout * cin * ks * ks};
        off = c.b_off + cout;
        net->convs.push_back(c);
    };
    for (int l = 0; l < levels; l++) {
        add(l == 0 ? in_channels : net->C(l - 1), net->C(l), 3, 1);
        add(net->C(l), net->C(l), 3, 1);
    }
    for (int l = levels - 2; l >= 0; l--) {
        add(net->C(l + 1) + net->C(l), net->C(l), 3, 1);
        add(net->C(l), net->C(l), 3, 1);
    }
    add(net->C(0), 1, 1, 0);
    net->nparam = off;
    // He-normal weights (std sqrt(2 / fan_in)), zero biases; the head starts
    // 10x smaller so the untrained network is close to the identity
    std::vector<float> p(off, 0.f);
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> nd(0.0, 1.0);
    for (size_t i = 0; i < net->convs.size(); i++) {
        const ConvSpec& c = net->convs[i];
        double sd = std::sqrt(2.0 / (c.cin * c.ks * c.ks));
        if ((int)i == net->head()) sd *= 0.1;
        for (size_t k = c.w_off; k < c.b_off; k++) p[k] = (float)(sd * nd(rng));
    }
    hipError_t e = hipMalloc(&net->pbuf, 4 * off * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&net->partial, kSqBlocks * sizeof(double));
    if (e == hipSuccess) e = hipMemcpy(net->pbuf, p.data(), off * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(net->pbuf + off, 0, 3 * off * sizeof(float));
    if (e != hipSuccess) {
        delete net;
        return fail(e == hipErrorOutOfMemory ? MSX_ENOMEM : MSX_EHIP, "unet allocation: %s", hipGetErrorString(e));
    }
    *out = net;
    return MSX_OK;
}

MSX_API int msx_unet_create(msx_ctx ctx, int levels, int base_channels, uint64_t seed, msx_unet* out) {
    return msx_unet_create_in(ctx, levels, base_channels, 1, seed, out);
}

MSX_API int msx_unet_in_channels(msx_unet net) { return net ? net->cin0 : 0; }

MSX_API int msx_unet_destroy(msx_unet net) {
    if (!net) return MSX_OK;
    // does not touch the context, which may already be gone; hipFree waits
    // for the device
    hipSetDevice(net->device);
    delete net;
    return MSX_OK;
}

MSX_API int msx_unet_conv(msx_ctx ctx, int mode, int ks, const float* x, const float* w, const float* bias,
                          const float* g, const float* y, float* out, int n, int cin, int cout, int h, int wd,
                          int relu) {
    if (!ctx) return fail(MSX_ESTATE, "ctx is NULL");
    if (mode < 0 || mode > 2) return fail(MSX_EINVAL, "mode %d outside 0..2", mode);
    if (ks != 1 && ks != 3) return fail(MSX_EINVAL, "ks must be 1 or 3");
    if (n < 1 || cin < 1 || cout < 1 || h < 1 || wd < 1 || (long long)n * h * wd > (1LL << 26) ||
        (long long)cin * cout > (1LL << 24))
        return fail(MSX_EINVAL, "bad shape n %d cin %d cout %d h %d w %d", n, cin, cout, h, wd);
    if (!out || !w) return fail(MSX_EINVAL, "w and out must be non-NULL");
    if ((mode != 1 && !x) || (mode != 0 && !g) || (mode != 0 && relu && !y))
        return fail(MSX_EINVAL, "missing x, g or y for mode %d", mode);
    UHIP(hipSetDevice(msx_ctx_device(ctx)));
    ConvArgs a{};
    a.x = x;
    a.w = w;
    a.bias = mode == 0 ? bias : nullptr;
    a.g = g;
    a.y = y;
    a.out = out;
    a.N = n;
    a.Cin = cin;
    a.Cout = cout;
    a.H = h;
    a.W = wd;
    a.relu = relu != 0;
    a.out_c = mode == 1 ? cin : cout;
    a.accumulate = 0;
    const size_t need = mode == kBwdWeight ? conv_wgrad_scratch(ks, a) : 0;
    if (need) UHIP(hipMalloc(&a.partial, need * sizeof(float)));
    hipError_t e = launch_conv(mode, ks, a, msx_ctx_stream(ctx));
    if (e == hipSuccess) e = hipStreamSynchronize(msx_ctx_stream(ctx));
    if (a.partial) hipFree(a.partial);
    UHIP(e);
    return MSX_OK;
}

MSX_API int64_t msx_unet_param_count(msx_unet net) { return net ? (int64_t)net->nparam : 0; }

MSX_API int msx_unet_params(msx_unet net, float* dst, int where) {
    UCHK(check_net(net));
    if (!dst) return fail(MSX_EINVAL, "dst is NULL");
    UHIP(hipSetDevice(msx_ctx_device(net->ctx)));
    UHIP(hipMemcpyAsync(dst, net->params(), net->nparam * sizeof(float),
                        where == MSX_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice, net->stream()));
    UHIP(hipStreamSynchronize(net->stream()));
    return MSX_OK;
}

MSX_API int msx_unet_set_params(msx_unet net, const float* src, int where) {
    UCHK(check_net(net));
    if (!src) return fail(MSX_EINVAL, "src is NULL");
    UHIP(hipSetDevice(msx_ctx_device(net->ctx)));
    UHIP(hipMemcpyAsync(net->params(), src, net->nparam * sizeof(float),
                        where == MSX_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, net->stream()));
    UHIP(hipMemsetAsync(net->adam_m(), 0, 2 * net->nparam * sizeof(float), net->stream()));
    net->step = 0;
    UHIP(hipStreamSynchronize(net->stream()));
    return MSX_OK;
}

MSX_API int msx_unet_forward(msx_unet net, const float* x, float* y, int n, int h, int w, int where) {
    UCHK(check_net(net));
    UCHK(check_shape(net, n, h, w, where));
    if (!x || !y) return fail(MSX_EINVAL, "x and y must be non-NULL");
    UHIP(hipSetDevice(msx_ctx_device(net->ctx)));
    UCHK(net->workspace(n, h, w));
    UCHK(stage_inputs(net, x, nullptr, where));
    UCHK(net->forward());
    const long long cnt = (long long)n * h * w;
    UHIP(launch_residual_mse(net->xbase, net->z, nullptr, net->out, nullptr, cnt, 0.f, net->stream()));
    UHIP(hipMemcpyAsync(y, net->out, sizeof(float) * cnt,
                        where == MSX_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice, net->stream()));
    UHIP(hipStreamSynchronize(net->stream()));
    return MSX_OK;
}

MSX_API int msx_unet_gradient(msx_unet net, const float* x, const float* target, int n, int h, int w, double* loss,
                              float* grad, int where) {
    UCHK(check_net(net));
    UCHK(check_shape(net, n, h, w, where));
    if (!x || !target) return fail(MSX_EINVAL, "x and target must be non-NULL");
    UHIP(hipSetDevice(msx_ctx_device(net->ctx)));
    UCHK(net->workspace(n, h, w));
    UCHK(stage_inputs(net, x, target, where));
    UCHK(net->forward());
    UCHK(net->backward(loss));
    if (grad)
        UHIP(hipMemcpyAsync(grad, net->grads(), net->nparam * sizeof(float),
                            where == MSX_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice, net->stream()));
    UHIP(hipStreamSynchronize(net->stream()));
    return MSX_OK;
}

MSX_API int msx_unet_train_step(msx_unet net, const float* x, const float* target, int n, int h, int w, float lr,
                                double* loss, int where) {
    UCHK(check_net(net));
    UCHK(check_shape(net, n, h, w, where));
    if (!x || !target) return fail(MSX_EINVAL, "x and target must be non-NULL");
    if (!(lr > 0.f)) return fail(MSX_EINVAL, "lr must be > 0");
    UHIP(hipSetDevice(msx_ctx_device(net->ctx)));
    UCHK(net->workspace(n, h, w));
    UCHK(stage_inputs(net, x, target, where));
    UCHK(net->forward());
    UCHK(net->backward(loss));
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    net->step++;
    const float c1 = (float)(1.0 - std::pow((double)b1, (double)net->step));
    const float c2 = (float)(1.0 - std::pow((double)b2, (double)net->step));
    UHIP(launch_adam(net->params(), net->grads(), net->adam_m(), net->adam_v(), (long long)net->nparam, lr, b1, b2,
                     eps, c1, c2, net->stream()));
    UHIP(hipStreamSynchronize(net->stream()));
    return MSX_OK;
}

}  // extern "C"
