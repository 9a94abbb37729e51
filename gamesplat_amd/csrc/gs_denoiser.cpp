// gs_denoiser.cpp - the learned denoiser of the stochastic path: U-Net
// parameters, activations, forward / backward / Adam orchestration, the
// training-crop pool and inference on a render context's GPS accumulation.
//
// No counterpart in the reference: gamesplat shows the accumulated GPS
// samples directly.  The network maps the running mean (RGBA) and the GPS
// depth of 1-4 accumulated frames to the sorted-alpha image; sorted alpha
// stays the fallback and the image its output is scored against.
//
// Network (4 levels, 32/48/64/96 channels):
//   e0a 8->32, e0b 32->32 -> skip0 | pool | e1a 32->48, e1b 48->48 -> skip1
//   | pool | e2a 48->64, e2b 64->64 -> skip2 | pool | ba 64->96, bb 96->96
//   | up, [.., skip2] d2a 160->64, d2b | up, [.., skip1] d1a 112->48, d1b
//   | up, [.., skip0] d0a 80->32, d0b | 1x1 head 32->4, + input mean RGBA.
// Every 3x3 convolution is followed by ReLU; activations and their
// gradients are bf16, accumulation fp32, parameters and Adam state fp32.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <random>
#include <string>
#include <vector>

#include "gs_internal.h"

using namespace gs;

namespace {

struct ConvSpec {
    const char* name;
    int cin, cout;
};
constexpr ConvSpec kConvs[14] = {
    {"e0a", 8, 32},    {"e0b", 32, 32},  {"e1a", 32, 48}, {"e1b", 48, 48}, {"e2a", 48, 64},
    {"e2b", 64, 64},   {"ba", 64, 96},   {"bb", 96, 96},  {"d2a", 160, 64}, {"d2b", 64, 64},
    {"d1a", 112, 48},  {"d1b", 48, 48},  {"d0a", 80, 32}, {"d0b", 32, 32},
};
enum { E0A, E0B, E1A, E1B, E2A, E2B, BA, BB, D2A, D2B, D1A, D1B, D0A, D0B };

constexpr int kWgChunk = 2048;  // pixels per weight-gradient block
constexpr uint32_t kCrop = 256; // training crop size

struct Layer {
    int cin, cout, kpad, kpad_t;
    size_t w_off, b_off;
    uint16_t* wf;  // [cout][kpad]
    uint16_t* wb;  // [cin][kpad_t]
};

int round_up(int v, int m) { return (v + m - 1) / m * m; }

#define DN_HIP(expr)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) {                                                                   \
            std::string m_ = std::string(#expr) + " failed: " + hipGetErrorString(e_);            \
            return api_error(GS_ERR_HIP, m_.c_str());                                             \
        }                                                                                         \
    } while (0)
#define DN_CHECK(expr)              \
    do {                            \
        int r_ = (expr);            \
        if (r_ != GS_OK) return r_; \
    } while (0)

}  // namespace

struct gs_denoiser {
    gs_ctx* ctx = nullptr;
    hipStream_t s = nullptr;
    int device = 0;
    Layer layers[14];
    size_t head_off = 0;
    uint32_t P = 0;
    float *params = nullptr, *grads = nullptr, *m = nullptr, *v = nullptr;
    uint16_t* wbuf = nullptr;
    int step = 0;

    // activations and gradients for an (an, ah, aw) batch
    int an = 0, ah = 0, aw = 0;
    bool grads_alloc = false;
    std::vector<void*> act_allocs;
    uint16_t *inp = nullptr, *t0, *cat0, *t0d, *u0, *p1, *t1, *cat1, *t1d, *u1, *p2, *t2, *cat2, *t2d, *u2, *p3, *t3,
             *u3;
    uint16_t *dt0, *dcat0, *dt0d, *du0, *dp1, *dt1, *dcat1, *dt1d, *du1, *dp2, *dt2, *dcat2, *dt2d, *du2, *dp3, *dt3,
             *du3;
    float *out = nullptr, *tgt = nullptr;
    float *partial = nullptr, *bias_partial = nullptr, *head_partial = nullptr, *loss = nullptr;
    uint16_t* image = nullptr;  // head output (RGBA16F) at an x ah x aw

    // training pool of kCrop x kCrop crops
    uint16_t* pool_in = nullptr;
    float* pool_tgt = nullptr;
    uint32_t pool_cap = 0, pool_n = 0, pool_next = 0;
    uint32_t* idx = nullptr;
    float* losses = nullptr;
    uint32_t losses_cap = 0;

    // last denoised frame
    uint16_t* den = nullptr;
    size_t den_bytes = 0;
    float2* judge = nullptr;  // per-block (squared-error sum, max abs) of gs_denoiser_judge
    int den_w = 0, den_h = 0;
    float den_flip = 1.0f;
};

namespace {

template <typename T>
int dmalloc(T** p, size_t bytes) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, bytes ? bytes : 16);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        std::string m = "hipMalloc(" + std::to_string(bytes) + ") for the denoiser failed";
        return api_error(GS_ERR_OOM, m.c_str());
    }
    *p = static_cast<T*>(q);
    return GS_OK;
}

void free_acts(gs_denoiser* d) {
    for (void* p : d->act_allocs) (void)hipFree(p);
    d->act_allocs.clear();
    d->an = d->ah = d->aw = 0;
    d->grads_alloc = false;
}

// Activation (and, for training, gradient) buffers for n x h x w.
int ensure_acts(gs_denoiser* d, int n, int h, int w, bool train) {
    if (d->an == n && d->ah == h && d->aw == w && (d->grads_alloc || !train)) return GS_OK;
    free_acts(d);
    const size_t p0 = (size_t)n * h * w, p1 = p0 / 4, p2 = p0 / 16, p3 = p0 / 64;
    auto bf = [&](uint16_t** ptr, size_t px, int c) {
        int r = dmalloc(ptr, px * c * 2);
        if (r == GS_OK) d->act_allocs.push_back(*ptr);
        return r;
    };
    auto f32 = [&](float** ptr, size_t count) {
        int r = dmalloc(ptr, count * 4);
        if (r == GS_OK) d->act_allocs.push_back(*ptr);
        return r;
    };
    DN_CHECK(bf(&d->inp, p0, 8));
    DN_CHECK(bf(&d->t0, p0, 32));
    DN_CHECK(bf(&d->cat0, p0, 80));
    DN_CHECK(bf(&d->t0d, p0, 32));
    DN_CHECK(bf(&d->u0, p0, 32));
    DN_CHECK(bf(&d->p1, p1, 32));
    DN_CHECK(bf(&d->t1, p1, 48));
    DN_CHECK(bf(&d->cat1, p1, 112));
    DN_CHECK(bf(&d->t1d, p1, 48));
    DN_CHECK(bf(&d->u1, p1, 48));
    DN_CHECK(bf(&d->p2, p2, 48));
    DN_CHECK(bf(&d->t2, p2, 64));
    DN_CHECK(bf(&d->cat2, p2, 160));
    DN_CHECK(bf(&d->t2d, p2, 64));
    DN_CHECK(bf(&d->u2, p2, 64));
    DN_CHECK(bf(&d->p3, p3, 64));
    DN_CHECK(bf(&d->t3, p3, 96));
    DN_CHECK(bf(&d->u3, p3, 96));
    DN_CHECK(bf(&d->image, p0, 4));
    DN_CHECK(f32(&d->loss, 4));
    if (train) {
        DN_CHECK(bf(&d->dt0, p0, 32));
        DN_CHECK(bf(&d->dcat0, p0, 80));
        DN_CHECK(bf(&d->dt0d, p0, 32));
        DN_CHECK(bf(&d->du0, p0, 32));
        DN_CHECK(bf(&d->dp1, p1, 32));
        DN_CHECK(bf(&d->dt1, p1, 48));
        DN_CHECK(bf(&d->dcat1, p1, 112));
        DN_CHECK(bf(&d->dt1d, p1, 48));
        DN_CHECK(bf(&d->du1, p1, 48));
        DN_CHECK(bf(&d->dp2, p2, 48));
        DN_CHECK(bf(&d->dt2, p2, 64));
        DN_CHECK(bf(&d->dcat2, p2, 160));
        DN_CHECK(bf(&d->dt2d, p2, 64));
        DN_CHECK(bf(&d->du2, p2, 64));
        DN_CHECK(bf(&d->dp3, p3, 64));
        DN_CHECK(bf(&d->dt3, p3, 96));
        DN_CHECK(bf(&d->du3, p3, 96));
        DN_CHECK(f32(&d->out, p0 * 4));
        DN_CHECK(f32(&d->tgt, p0 * 4));
        size_t chunks = (p0 + kWgChunk - 1) / kWgChunk;
        size_t most = 0;
        for (const ConvSpec& c : kConvs) most = std::max(most, (size_t)9 * c.cin * c.cout);
        DN_CHECK(f32(&d->partial, chunks * most));
        DN_CHECK(f32(&d->bias_partial, chunks * 96));
        DN_CHECK(f32(&d->head_partial, head_partial_floats((long)p0)));
        d->grads_alloc = true;
    }
    d->an = n;
    d->ah = h;
    d->aw = w;
    return GS_OK;
}

int pack_all(gs_denoiser* d) {
    for (Layer& L : d->layers)
        DN_HIP(launch_pack_weights(d->params + L.w_off, L.cin, L.cout, L.kpad, L.kpad_t, L.wf, L.wb, d->s));
    return GS_OK;
}

// Conv li with bias and ReLU; optionally its output 2x2-pooled into `pool`
// and its first `split` input channels upsampled from the half-size x2.
struct ConvExtra {
    uint16_t* pool = nullptr;
    int ps = 0;
    const uint16_t* x2 = nullptr;
    int x2s = 0, split = 0;
    const float* head = nullptr;  // fused 1x1 head (see ConvLaunch)
    const uint16_t* head_inp = nullptr;
    uint16_t* head_out = nullptr;
    int head_w = 0, head_h = 0;
};

int conv(gs_denoiser* d, int li, const uint16_t* x, int xs, int xo, uint16_t* y, int ys, int yo, int n, int h,
         int w, const ConvExtra& ex = ConvExtra()) {
    const Layer& L = d->layers[li];
    ConvLaunch a{};
    a.pool = ex.pool;
    a.ps = ex.ps;
    a.x2 = ex.x2;
    a.x2s = ex.x2s;
    a.split = ex.split;
    a.head = ex.head;
    a.head_inp = ex.head_inp;
    a.head_out = ex.head_out;
    a.head_w = ex.head_w;
    a.head_h = ex.head_h;
    a.x = x; a.xs = xs; a.xo = xo;
    a.y = y; a.ys = ys; a.yo = yo;
    a.wt = L.wf;
    a.bias = d->params + L.b_off;
    a.n = n; a.h = h; a.w = w;
    a.cin = L.cin; a.cout = L.cout; a.kpad = L.kpad; a.relu = 1;
    DN_HIP(launch_conv3x3(a, L.cout / 16, d->s));
    return GS_OK;
}

// Backward of conv li: dy (gradient of its stored ReLU output y) is masked in
// place, the weight and bias gradients are written, and when dx is given the
// input gradient (bf16) is computed.
int conv_bwd(gs_denoiser* d, int li, const uint16_t* x, int xs, int xo, TensorView y, TensorView dy, uint16_t* dx,
             int dxs, int dxo, int n, int h, int w) {
    const Layer& L = d->layers[li];
    const long pixels = (long)n * h * w;
    DN_HIP(launch_relu_mask(dy, y, pixels, L.cout, d->s));
    WGradLaunch g{};
    g.x = x; g.xs = xs; g.xo = xo;
    g.dz = dy.p; g.dzs = dy.cs; g.dzo = dy.co;
    g.partial = d->partial;
    g.bias_partial = d->bias_partial;
    g.n = n; g.h = h; g.w = w; g.cin = L.cin; g.cout = L.cout;
    g.chunk_pixels = kWgChunk;
    const int chunks = (int)((pixels + kWgChunk - 1) / kWgChunk);
    DN_HIP(launch_conv3x3_wgrad(g, chunks, d->grads + L.w_off, d->grads + L.b_off, d->s));
    if (dx) {
        ConvLaunch a{};
        a.x = dy.p; a.xs = dy.cs; a.xo = dy.co;
        a.y = dx; a.ys = dxs; a.yo = dxo;
        a.wt = L.wb;
        a.bias = nullptr;
        a.n = n; a.h = h; a.w = w;
        a.cin = L.cout; a.cout = L.cin; a.kpad = L.kpad_t; a.relu = 0;
        DN_HIP(launch_conv3x3(a, L.cin / 16, d->s));
    }
    return GS_OK;
}

TensorView tv(uint16_t* p, int cs, int co) { return TensorView{p, cs, co}; }

// Forward pass over d->inp (an x ah x aw); the head writes fp32 out, the
// RGBA16F image, and with a target the loss, head gradients and du0.  With
// `den` (inference on one frame) the last convolution runs the head itself
// and writes its den_w x den_h corner to den as RGBA16F; u0, out and the
// padded image are then not written.
int forward(gs_denoiser* d, bool with_target, uint16_t* den = nullptr, int den_w = 0, int den_h = 0) {
    const int n = d->an, h0 = d->ah, w0 = d->aw;
    const int h1 = h0 / 2, w1 = w0 / 2, h2 = h0 / 4, w2 = w0 / 4, h3 = h0 / 8, w3 = w0 / 8;
    hipStream_t s = d->s;
    // the encoder's second convolutions also write the pooled tensor; for
    // inference the decoder's first convolutions read the upsampled half of
    // their concatenation straight from the half-size tensor (training keeps
    // the concatenation whole: the weight gradient reads it)
    auto pooled = [](uint16_t* p, int ps) {
        ConvExtra e;
        e.pool = p;
        e.ps = ps;
        return e;
    };
    auto upsampled = [&](const uint16_t* x2, int c) {
        ConvExtra e;
        if (!with_target) {
            e.x2 = x2;
            e.x2s = c;
            e.split = c;
        }
        return e;
    };
    DN_CHECK(conv(d, E0A, d->inp, 8, 0, d->t0, 32, 0, n, h0, w0));
    DN_CHECK(conv(d, E0B, d->t0, 32, 0, d->cat0, 80, 48, n, h0, w0, pooled(d->p1, 32)));
    DN_CHECK(conv(d, E1A, d->p1, 32, 0, d->t1, 48, 0, n, h1, w1));
    DN_CHECK(conv(d, E1B, d->t1, 48, 0, d->cat1, 112, 64, n, h1, w1, pooled(d->p2, 48)));
    DN_CHECK(conv(d, E2A, d->p2, 48, 0, d->t2, 64, 0, n, h2, w2));
    DN_CHECK(conv(d, E2B, d->t2, 64, 0, d->cat2, 160, 96, n, h2, w2, pooled(d->p3, 64)));
    DN_CHECK(conv(d, BA, d->p3, 64, 0, d->t3, 96, 0, n, h3, w3));
    DN_CHECK(conv(d, BB, d->t3, 96, 0, d->u3, 96, 0, n, h3, w3));
    if (with_target) DN_HIP(launch_up2(tv(d->u3, 96, 0), tv(d->cat2, 160, 0), n, h3, w3, 96, s));
    DN_CHECK(conv(d, D2A, d->cat2, 160, 0, d->t2d, 64, 0, n, h2, w2, upsampled(d->u3, 96)));
    DN_CHECK(conv(d, D2B, d->t2d, 64, 0, d->u2, 64, 0, n, h2, w2));
    if (with_target) DN_HIP(launch_up2(tv(d->u2, 64, 0), tv(d->cat1, 112, 0), n, h2, w2, 64, s));
    DN_CHECK(conv(d, D1A, d->cat1, 112, 0, d->t1d, 48, 0, n, h1, w1, upsampled(d->u2, 64)));
    DN_CHECK(conv(d, D1B, d->t1d, 48, 0, d->u1, 48, 0, n, h1, w1));
    if (with_target) DN_HIP(launch_up2(tv(d->u1, 48, 0), tv(d->cat0, 80, 0), n, h1, w1, 48, s));
    DN_CHECK(conv(d, D0A, d->cat0, 80, 0, d->t0d, 32, 0, n, h0, w0, upsampled(d->u1, 48)));
    if (den && !with_target) {
        ConvExtra e;
        e.head = d->params + d->head_off;
        e.head_inp = d->inp;
        e.head_out = den;
        e.head_w = den_w;
        e.head_h = den_h;
        return conv(d, D0B, d->t0d, 32, 0, d->u0, 32, 0, n, h0, w0, e);
    }
    DN_CHECK(conv(d, D0B, d->t0d, 32, 0, d->u0, 32, 0, n, h0, w0));
    HeadLaunch hl{};
    hl.u = d->u0;
    hl.inp = d->inp;
    hl.params = d->params + d->head_off;
    hl.pixels = (long)n * h0 * w0;
    hl.out = with_target ? d->out : nullptr;
    hl.image = d->image;
    hl.target = with_target ? d->tgt : nullptr;
    hl.inv_count = 1.0f / (float)(hl.pixels * 4);
    hl.du = d->du0;
    hl.partial = d->head_partial;
    DN_HIP(launch_head(hl, d->loss, d->grads + d->head_off, s));
    return GS_OK;
}

int backward(gs_denoiser* d) {
    const int n = d->an, h0 = d->ah, w0 = d->aw;
    const int h1 = h0 / 2, w1 = w0 / 2, h2 = h0 / 4, w2 = w0 / 4, h3 = h0 / 8, w3 = w0 / 8;
    hipStream_t s = d->s;
    DN_CHECK(conv_bwd(d, D0B, d->t0d, 32, 0, tv(d->u0, 32, 0), tv(d->du0, 32, 0), d->dt0d, 32, 0, n, h0, w0));
    DN_CHECK(conv_bwd(d, D0A, d->cat0, 80, 0, tv(d->t0d, 32, 0), tv(d->dt0d, 32, 0), d->dcat0, 80, 0, n, h0, w0));
    DN_HIP(launch_up2_bwd(tv(d->dcat0, 80, 0), tv(d->du1, 48, 0), n, h1, w1, 48, s));
    DN_CHECK(conv_bwd(d, D1B, d->t1d, 48, 0, tv(d->u1, 48, 0), tv(d->du1, 48, 0), d->dt1d, 48, 0, n, h1, w1));
    DN_CHECK(conv_bwd(d, D1A, d->cat1, 112, 0, tv(d->t1d, 48, 0), tv(d->dt1d, 48, 0), d->dcat1, 112, 0, n, h1, w1));
    DN_HIP(launch_up2_bwd(tv(d->dcat1, 112, 0), tv(d->du2, 64, 0), n, h2, w2, 64, s));
    DN_CHECK(conv_bwd(d, D2B, d->t2d, 64, 0, tv(d->u2, 64, 0), tv(d->du2, 64, 0), d->dt2d, 64, 0, n, h2, w2));
    DN_CHECK(conv_bwd(d, D2A, d->cat2, 160, 0, tv(d->t2d, 64, 0), tv(d->dt2d, 64, 0), d->dcat2, 160, 0, n, h2, w2));
    DN_HIP(launch_up2_bwd(tv(d->dcat2, 160, 0), tv(d->du3, 96, 0), n, h3, w3, 96, s));
    DN_CHECK(conv_bwd(d, BB, d->t3, 96, 0, tv(d->u3, 96, 0), tv(d->du3, 96, 0), d->dt3, 96, 0, n, h3, w3));
    DN_CHECK(conv_bwd(d, BA, d->p3, 64, 0, tv(d->t3, 96, 0), tv(d->dt3, 96, 0), d->dp3, 64, 0, n, h3, w3));
    DN_HIP(launch_pool2_bwd(tv(d->dp3, 64, 0), tv(d->dcat2, 160, 96), n, h2, w2, 64, s));
    DN_CHECK(conv_bwd(d, E2B, d->t2, 64, 0, tv(d->cat2, 160, 96), tv(d->dcat2, 160, 96), d->dt2, 64, 0, n, h2, w2));
    DN_CHECK(conv_bwd(d, E2A, d->p2, 48, 0, tv(d->t2, 64, 0), tv(d->dt2, 64, 0), d->dp2, 48, 0, n, h2, w2));
    DN_HIP(launch_pool2_bwd(tv(d->dp2, 48, 0), tv(d->dcat1, 112, 64), n, h1, w1, 48, s));
    DN_CHECK(conv_bwd(d, E1B, d->t1, 48, 0, tv(d->cat1, 112, 64), tv(d->dcat1, 112, 64), d->dt1, 48, 0, n, h1, w1));
    DN_CHECK(conv_bwd(d, E1A, d->p1, 32, 0, tv(d->t1, 48, 0), tv(d->dt1, 48, 0), d->dp1, 32, 0, n, h1, w1));
    DN_HIP(launch_pool2_bwd(tv(d->dp1, 32, 0), tv(d->dcat0, 80, 48), n, h0, w0, 32, s));
    DN_CHECK(conv_bwd(d, E0B, d->t0, 32, 0, tv(d->cat0, 80, 48), tv(d->dcat0, 80, 48), d->dt0, 32, 0, n, h0, w0));
    DN_CHECK(conv_bwd(d, E0A, d->inp, 8, 0, tv(d->t0, 32, 0), tv(d->dt0, 32, 0), nullptr, 0, 0, n, h0, w0));
    return GS_OK;
}

int adam(gs_denoiser* d, float lr) {
    d->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - powf(b1, (float)d->step), c2 = 1.0f - powf(b2, (float)d->step);
    DN_HIP(launch_adam(d->params, d->grads, d->m, d->v, d->P, lr, b1, b2, 1e-8f, c1, c2, d->s));
    return pack_all(d);
}

uint16_t host_bf16(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

}  // namespace

extern "C" {

uint32_t gs_denoiser_param_count(void) {
    uint32_t n = 0;
    for (const ConvSpec& c : kConvs) n += (uint32_t)(c.cout * 9 * c.cin + c.cout);
    return n + 4 * 32 + 4;
}

void gs_denoiser_destroy(gs_denoiser* d) {
    if (!d) return;
    (void)hipSetDevice(d->device);
    if (d->s) (void)hipStreamSynchronize(d->s);
    free_acts(d);
    void* bufs[] = {d->params, d->grads, d->m, d->v, d->wbuf, d->pool_in, d->pool_tgt, d->idx, d->losses, d->den,
                    d->judge};
    for (void* p : bufs)
        if (p) (void)hipFree(p);
    delete d;
}

int gs_denoiser_create(gs_ctx* ctx, uint64_t seed, gs_denoiser** out) {
    if (!ctx || !out) return api_error(GS_ERR_INVALID, "bad arguments");
    *out = nullptr;
    CtxFrames f;
    DN_CHECK(ctx_frames(ctx, &f));
    gs_denoiser* d = new gs_denoiser();
    d->ctx = ctx;
    d->s = f.stream;
    d->device = f.device;
    size_t off = 0, wwords = 0;
    for (int i = 0; i < 14; ++i) {
        Layer& L = d->layers[i];
        L.cin = kConvs[i].cin;
        L.cout = kConvs[i].cout;
        L.kpad = round_up(9 * L.cin, 32);
        L.kpad_t = round_up(9 * L.cout, 32);
        L.w_off = off;
        off += (size_t)L.cout * 9 * L.cin;
        L.b_off = off;
        off += L.cout;
        wwords += (size_t)L.cout * L.kpad + (size_t)L.cin * L.kpad_t;
    }
    d->head_off = off;
    off += 4 * 32 + 4;
    d->P = (uint32_t)off;
    int r = GS_OK;
    if (r == GS_OK) r = dmalloc(&d->params, off * 4);
    if (r == GS_OK) r = dmalloc(&d->grads, off * 4);
    if (r == GS_OK) r = dmalloc(&d->m, off * 4);
    if (r == GS_OK) r = dmalloc(&d->v, off * 4);
    if (r == GS_OK) r = dmalloc(&d->wbuf, wwords * 2);
    if (r != GS_OK) {
        gs_denoiser_destroy(d);
        return r;
    }
    size_t wo = 0;
    for (Layer& L : d->layers) {
        L.wf = d->wbuf + wo;
        wo += (size_t)L.cout * L.kpad;
        L.wb = d->wbuf + wo;
        wo += (size_t)L.cin * L.kpad_t;
    }
    // He-normal weights, zero biases; the three zero input channels of e0a
    // get zero weights; the head starts small so the output starts near the
    // input mean.
    std::vector<float> h(off, 0.0f);
    std::mt19937_64 rng(seed);
    std::normal_distribution<float> nd(0.0f, 1.0f);
    for (int i = 0; i < 14; ++i) {
        const Layer& L = d->layers[i];
        const float sd = sqrtf(2.0f / (9.0f * L.cin));
        for (int co = 0; co < L.cout; ++co)
            for (int t = 0; t < 9; ++t)
                for (int ci = 0; ci < L.cin; ++ci) {
                    float v = nd(rng) * sd;
                    if (i == E0A && ci >= 5) v = 0.0f;
                    h[L.w_off + ((size_t)co * 9 + t) * L.cin + ci] = v;
                }
    }
    for (int k = 0; k < 128; ++k) h[d->head_off + k] = nd(rng) * sqrtf(2.0f / 32.0f) * 0.1f;
    hipError_t e = hipMemcpyAsync(d->params, h.data(), off * 4, hipMemcpyHostToDevice, d->s);
    if (e == hipSuccess) e = hipMemsetAsync(d->m, 0, off * 4, d->s);
    if (e == hipSuccess) e = hipMemsetAsync(d->v, 0, off * 4, d->s);
    if (e == hipSuccess) e = hipMemsetAsync(d->grads, 0, off * 4, d->s);
    if (e == hipSuccess) e = hipStreamSynchronize(d->s);
    if (e != hipSuccess) {
        gs_denoiser_destroy(d);
        return api_error(GS_ERR_HIP, "denoiser parameter upload failed");
    }
    r = pack_all(d);
    if (r != GS_OK) {
        gs_denoiser_destroy(d);
        return r;
    }
    *out = d;
    return api_ok();
}

int gs_denoiser_get_params(gs_denoiser* d, float* host, size_t count) {
    if (!d || !host || count != d->P) return api_error(GS_ERR_INVALID, "parameter buffer must hold gs_denoiser_param_count() floats");
    DN_HIP(hipSetDevice(d->device));
    DN_HIP(hipMemcpyAsync(host, d->params, (size_t)d->P * 4, hipMemcpyDeviceToHost, d->s));
    DN_HIP(hipStreamSynchronize(d->s));
    return api_ok();
}

int gs_denoiser_set_params(gs_denoiser* d, const float* host, size_t count) {
    if (!d || !host || count != d->P) return api_error(GS_ERR_INVALID, "parameter buffer must hold gs_denoiser_param_count() floats");
    DN_HIP(hipSetDevice(d->device));
    DN_HIP(hipMemcpyAsync(d->params, host, (size_t)d->P * 4, hipMemcpyHostToDevice, d->s));
    DN_HIP(hipMemsetAsync(d->m, 0, (size_t)d->P * 4, d->s));
    DN_HIP(hipMemsetAsync(d->v, 0, (size_t)d->P * 4, d->s));
    DN_HIP(hipStreamSynchronize(d->s));
    d->step = 0;
    DN_CHECK(pack_all(d));
    return api_ok();
}

int gs_denoiser_forward_backward(gs_denoiser* d, const float* inp, const float* target, uint32_t n, uint32_t h,
                                 uint32_t w, float* loss, float* out, float* grad) {
    if (!d || !inp || !target) return api_error(GS_ERR_INVALID, "bad arguments");
    if (!n || !h || !w || h % 8 || w % 8 || (size_t)n * h * w > (1u << 26))
        return api_error(GS_ERR_INVALID, "batch dims must be positive, h and w multiples of 8");
    DN_HIP(hipSetDevice(d->device));
    DN_CHECK(ensure_acts(d, (int)n, (int)h, (int)w, true));
    const size_t px = (size_t)n * h * w;
    std::vector<uint16_t> hb(px * 8);
    for (size_t i = 0; i < px * 8; ++i) hb[i] = host_bf16(inp[i]);
    DN_HIP(hipMemcpyAsync(d->inp, hb.data(), px * 16, hipMemcpyHostToDevice, d->s));
    DN_HIP(hipMemcpyAsync(d->tgt, target, px * 16, hipMemcpyHostToDevice, d->s));
    DN_CHECK(forward(d, true));
    DN_CHECK(backward(d));
    float l = 0.0f;
    DN_HIP(hipMemcpyAsync(&l, d->loss, 4, hipMemcpyDeviceToHost, d->s));
    if (out) DN_HIP(hipMemcpyAsync(out, d->out, px * 16, hipMemcpyDeviceToHost, d->s));
    if (grad) DN_HIP(hipMemcpyAsync(grad, d->grads, (size_t)d->P * 4, hipMemcpyDeviceToHost, d->s));
    DN_HIP(hipStreamSynchronize(d->s));
    if (loss) *loss = l / (float)(px * 4);
    return api_ok();
}

int gs_denoiser_adam_step(gs_denoiser* d, float lr) {
    if (!d) return api_error(GS_ERR_INVALID, "bad arguments");
    DN_HIP(hipSetDevice(d->device));
    DN_CHECK(adam(d, lr));
    return api_ok();
}

int gs_denoiser_capture(gs_denoiser* d, uint32_t crops, uint64_t seed) {
    if (!d || crops == 0 || crops > 4096) return api_error(GS_ERR_INVALID, "crops must be in [1, 4096]");
    CtxFrames f;
    DN_CHECK(ctx_frames(d->ctx, &f));
    if (!f.accum || !f.alpha) return api_error(GS_ERR_STATE, "capture needs a GPS accumulation and a sorted-alpha frame");
    if (f.gps_w != f.alpha_w || f.gps_h != f.alpha_h)
        return api_error(GS_ERR_STATE, "GPS accumulation and sorted-alpha frame differ in size");
    if (f.gps_w < (int)kCrop || f.gps_h < (int)kCrop) return api_error(GS_ERR_INVALID, "frame smaller than a 256x256 crop");
    const size_t per = (size_t)kCrop * kCrop;
    if (!d->pool_in) {
        d->pool_cap = 1024;
        DN_CHECK(dmalloc(&d->pool_in, d->pool_cap * per * 16));
        DN_CHECK(dmalloc(&d->pool_tgt, d->pool_cap * per * 16));
    }
    std::mt19937_64 rng(seed);
    for (uint32_t k = 0; k < crops; ++k) {
        const uint32_t slot = d->pool_next;
        d->pool_next = (d->pool_next + 1) % d->pool_cap;
        if (d->pool_n < d->pool_cap) d->pool_n++;
        PackLaunch p{};
        p.accum = f.accum;
        p.depth = f.depth;
        p.alpha = f.alpha;
        p.width = f.gps_w;
        p.height = f.gps_h;
        p.x0 = (int)(rng() % (uint64_t)(f.gps_w - kCrop + 1));
        p.y0 = (int)(rng() % (uint64_t)(f.gps_h - kCrop + 1));
        p.hp = kCrop;
        p.wp = kCrop;
        p.inp = d->pool_in + slot * per * 8;
        p.target = d->pool_tgt + slot * per * 4;
        DN_HIP(launch_pack_input(p, d->s));
    }
    return api_ok();
}

int gs_denoiser_pool_size(gs_denoiser* d, uint32_t* count) {
    if (!d || !count) return api_error(GS_ERR_INVALID, "bad arguments");
    *count = d->pool_n;
    return api_ok();
}

int gs_denoiser_train(gs_denoiser* d, uint32_t steps, uint32_t batch, float lr, uint64_t seed, float* losses) {
    if (!d || steps == 0 || batch == 0 || batch > 64) return api_error(GS_ERR_INVALID, "steps >= 1, batch in [1, 64]");
    if (!d->pool_n) return api_error(GS_ERR_STATE, "no captured training crops");
    DN_HIP(hipSetDevice(d->device));
    DN_CHECK(ensure_acts(d, (int)batch, (int)kCrop, (int)kCrop, true));
    if (!d->idx) DN_CHECK(dmalloc(&d->idx, 64 * 4));
    if (d->losses_cap < steps) {
        if (d->losses) (void)hipFree(d->losses);
        d->losses = nullptr;
        DN_CHECK(dmalloc(&d->losses, (size_t)steps * 4));
        d->losses_cap = steps;
    }
    std::mt19937_64 rng(seed);
    std::vector<uint32_t> pick(batch);
    const uint32_t per = kCrop * kCrop;
    // one pinned page of indices per step keeps the host copies ordered
    uint32_t* hidx = nullptr;
    DN_HIP(hipHostMalloc((void**)&hidx, (size_t)steps * batch * 4, hipHostMallocDefault));
    int r = GS_OK;
    for (uint32_t st = 0; st < steps && r == GS_OK; ++st) {
        for (uint32_t b = 0; b < batch; ++b) hidx[st * batch + b] = (uint32_t)(rng() % d->pool_n);
        hipError_t e = hipMemcpyAsync(d->idx, hidx + st * batch, batch * 4, hipMemcpyHostToDevice, d->s);
        if (e == hipSuccess) e = launch_gather(d->pool_in, d->pool_tgt, d->idx, per, batch, d->inp, d->tgt, d->s);
        if (e != hipSuccess) r = api_error(GS_ERR_HIP, "training batch gather failed");
        if (r == GS_OK) r = forward(d, true);
        if (r == GS_OK) r = backward(d);
        if (r == GS_OK) r = adam(d, lr);
        if (r == GS_OK && hipMemcpyAsync(d->losses + st, d->loss, 4, hipMemcpyDeviceToDevice, d->s) != hipSuccess)
            r = api_error(GS_ERR_HIP, "loss copy failed");
    }
    hipError_t e = hipStreamSynchronize(d->s);
    (void)hipHostFree(hidx);
    if (r != GS_OK) return r;
    DN_HIP(e);
    if (losses) {
        DN_HIP(hipMemcpy(losses, d->losses, (size_t)steps * 4, hipMemcpyDeviceToHost));
        const float inv = 1.0f / (float)((size_t)batch * per * 4);
        for (uint32_t st = 0; st < steps; ++st) losses[st] *= inv;
    }
    return api_ok();
}

int gs_denoise(gs_denoiser* d) {
    if (!d) return api_error(GS_ERR_INVALID, "bad arguments");
    CtxFrames f;
    DN_CHECK(ctx_frames(d->ctx, &f));
    if (!f.accum) return api_error(GS_ERR_STATE, "gs_denoise needs a GPS accumulation");
    const int hp = round_up(f.gps_h, 8), wp = round_up(f.gps_w, 8);
    DN_CHECK(ensure_acts(d, 1, hp, wp, false));
    const size_t bytes = (size_t)f.gps_w * f.gps_h * 8;
    if (d->den_bytes < bytes) {
        if (d->den) (void)hipFree(d->den);
        d->den = nullptr;
        d->den_bytes = 0;
        DN_CHECK(dmalloc(&d->den, bytes));
        d->den_bytes = bytes;
    }
    PackLaunch p{};
    p.accum = f.accum;
    p.depth = f.depth;
    p.width = f.gps_w;
    p.height = f.gps_h;
    p.hp = hp;
    p.wp = wp;
    p.inp = d->inp;
    DN_HIP(launch_pack_input(p, d->s));
    DN_CHECK(forward(d, false, d->den, f.gps_w, f.gps_h));
    d->den_w = f.gps_w;
    d->den_h = f.gps_h;
    d->den_flip = f.clip_y_sign;
    return api_ok();
}

int gs_denoiser_judge(gs_denoiser* d, double* psnr_db, float* max_abs) {
    if (!d || !psnr_db || !max_abs) return api_error(GS_ERR_INVALID, "bad arguments");
    if (!d->den_w) return api_error(GS_ERR_STATE, "no denoised frame yet");
    CtxFrames f;
    DN_CHECK(ctx_frames(d->ctx, &f));
    if (!f.alpha) return api_error(GS_ERR_STATE, "gs_denoiser_judge needs a sorted-alpha frame of the same cull");
    if (f.alpha_w != d->den_w || f.alpha_h != d->den_h)
        return api_error(GS_ERR_INVALID, "sorted-alpha and denoised images differ in size");
    constexpr uint32_t kBlocks = 1024;
    if (!d->judge) DN_CHECK(dmalloc(&d->judge, sizeof(float2) * kBlocks));
    const uint32_t count = (uint32_t)d->den_w * (uint32_t)d->den_h * 4u;
    DN_HIP(launch_image_error(d->den, f.alpha, count, d->judge, kBlocks, d->s));
    std::vector<float2> part(kBlocks);
    DN_HIP(hipMemcpyAsync(part.data(), d->judge, sizeof(float2) * kBlocks, hipMemcpyDeviceToHost, d->s));
    DN_HIP(hipStreamSynchronize(d->s));
    double sum = 0.0;
    float mx = 0.0f;
    for (const float2& p : part) {
        sum += (double)p.x;
        mx = p.y > mx ? p.y : mx;
    }
    const double mse = sum / (double)count;
    *psnr_db = mse > 0.0 ? 10.0 * log10(1.0 / mse) : 1.0 / 0.0;
    *max_abs = mx;
    return api_ok();
}

int gs_denoiser_read_image(gs_denoiser* d, void* host, size_t bytes) {
    if (!d || !host) return api_error(GS_ERR_INVALID, "bad arguments");
    if (!d->den_w) return api_error(GS_ERR_STATE, "no denoised frame yet");
    const size_t need = (size_t)d->den_w * d->den_h * 8;
    if (bytes != need) return api_error(GS_ERR_INVALID, "denoised image needs width * height * 8 bytes");
    DN_HIP(hipSetDevice(d->device));
    std::vector<uint8_t> tmp(need);
    DN_HIP(hipMemcpyAsync(tmp.data(), d->den, need, hipMemcpyDeviceToHost, d->s));
    DN_HIP(hipStreamSynchronize(d->s));
    const size_t row = (size_t)d->den_w * 8;
    for (int y = 0; y < d->den_h; ++y) {
        const int src = d->den_flip > 0.0f ? d->den_h - 1 - y : y;
        memcpy(static_cast<uint8_t*>(host) + (size_t)y * row, tmp.data() + (size_t)src * row, row);
    }
    return api_ok();
}

}  // extern "C"
