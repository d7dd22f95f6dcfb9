// mbl_net.hip -- the warm-start network: a 4-level convolutional
// encoder-decoder (U-Net shape) over an H x W raster, trained and evaluated
// by the kernels of mbl_net_kernels.hip (BF16 MFMA, FP32 accumulation, FP32
// master weights, Adam). C ABI in include/mbl.h (mbl_net_*).
//
// Layers (3x3 convolutions with ReLU unless noted, c = base channels):
//   encoder level l = 0..3 at H/2^l x W/2^l: conv a, conv b (channels c, 2c, 4c, 4c);
//   2x2 max pool between levels; level 3 is the bottleneck.
//   decoder level l = 2, 1, 0: nearest x2 upsample of the level below,
//   concatenated with the encoder output of level l, conv a, conv b.
//   head: 1x1 convolution to the outputs, linear, FP32.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/mbl.h"
#include "mbl_net.h"

using namespace mbl::net;

namespace {

struct Layer {
    int cin, cin_real, cout, cout_real, ksize;
    size_t woff, boff; // offsets into the flat FP32 parameter vector
    size_t nw() const { return (size_t)cout * ksize * ksize * cin; }
};

struct Tensor {
    bf16 *p = nullptr;
    int C = 0, level = 0;
};

} // namespace

struct mbl_net {
    hipStream_t s = nullptr;
    int H = 0, W = 0, in_ch = 0, out_ch = 0, base = 0, max_batch = 0;
    int ch[4] = {0, 0, 0, 0};
    std::vector<Layer> L;
    size_t nparam = 0;
    float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr; // FP32 master, grad, Adam moments
    bf16 *wb = nullptr, *wf = nullptr;                           // BF16 weights, flipped BF16 weights
    int adam_t = 0;
    // activations and their gradients
    Tensor in0, e[4][2], pool[3], cat[3], d[3][2];
    Tensor ge[4][2], gpool[3], gcat[3], gd[3][2], tmp[3];
    uint8_t *arg[3] = {nullptr, nullptr, nullptr};
    float *out = nullptr;  // [P0][32] FP32 head output
    bf16 *gout = nullptr;  // [P0][32]
    float *partial = nullptr;
    size_t partial_n = 0;
    float *lossp = nullptr;
    float *din = nullptr, *dtarget = nullptr; // FP32 staging of one batch
    float *dout = nullptr;                    // the real output channels of one batch
    std::vector<void *> allocs;
};

static int P_of(const mbl_net *n, int N, int level) { return N * (n->H >> level) * (n->W >> level); }

static void *dev(mbl_net *n, size_t bytes)
{
    void *p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 16) != hipSuccess) return nullptr;
    n->allocs.push_back(p);
    return p;
}

// counter-based normal (SplitMix64 finaliser + Box-Muller) for weight initialisation
static double nrand(unsigned long long seed, unsigned long long i)
{
    auto mix = [](unsigned long long z) {
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    };
    unsigned long long a = mix(seed + (2 * i + 1) * 0x9E3779B97F4A7C15ULL);
    unsigned long long b = mix(seed + (2 * i + 2) * 0x9E3779B97F4A7C15ULL);
    double u1 = (double)((a >> 11) + 1) * (1.0 / 9007199254740992.0);
    double u2 = (double)(b >> 11) * (1.0 / 9007199254740992.0);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

static Tensor tens(mbl_net *n, int C, int level)
{
    Tensor t;
    t.C = C;
    t.level = level;
    t.p = (bf16 *)dev(n, sizeof(bf16) * (size_t)P_of(n, n->max_batch, level) * C);
    return t;
}

extern "C" int mbl_net_create(mbl_handle_t h, int H, int W, int in_ch, int out_ch, int base, int max_batch,
                              unsigned long long seed, mbl_net_t *outp)
{
    if (!outp || H % 8 || W % 8 || H <= 0 || W <= 0 || in_ch < 1 || in_ch > 32 || out_ch < 1 || out_ch > 32 ||
        base % 32 || base <= 0 || max_batch <= 0)
        return MBL_E_INVALID;
    (void)h;
    mbl_net *n = new mbl_net();
    *outp = nullptr;
    if (hipStreamCreateWithFlags(&n->s, hipStreamNonBlocking) != hipSuccess) {
        delete n;
        return MBL_E_HIP;
    }
    n->H = H;
    n->W = W;
    n->in_ch = in_ch;
    n->out_ch = out_ch;
    n->base = base;
    n->max_batch = max_batch;
    const int c[4] = {base, 2 * base, 4 * base, 4 * base};
    for (int i = 0; i < 4; i++) n->ch[i] = c[i];
    auto add = [&](int cin, int cin_real, int cout, int cout_real, int k) {
        Layer l{cin, cin_real, cout, cout_real, k, 0, 0};
        l.woff = n->nparam;
        n->nparam += l.nw();
        l.boff = n->nparam;
        n->nparam += cout;
        n->L.push_back(l);
    };
    // order: e0a e0b e1a e1b e2a e2b e3a e3b d2a d2b d1a d1b d0a d0b head
    add(32, in_ch, c[0], c[0], 3);
    add(c[0], c[0], c[0], c[0], 3);
    for (int l = 1; l < 4; l++) {
        add(c[l - 1], c[l - 1], c[l], c[l], 3);
        add(c[l], c[l], c[l], c[l], 3);
    }
    for (int l = 2; l >= 0; l--) {
        add(c[l + 1] + c[l], c[l + 1] + c[l], c[l], c[l], 3);
        add(c[l], c[l], c[l], c[l], 3);
    }
    add(c[0], c[0], 32, out_ch, 1);
    // parameters: He-normal weights on the real channels, zero elsewhere, zero biases
    std::vector<float> hw(n->nparam, 0.f);
    unsigned long long idx = 0;
    for (const Layer &l : n->L) {
        const int taps = l.ksize * l.ksize;
        const double sd = std::sqrt(2.0 / (l.cin_real * taps));
        for (int co = 0; co < l.cout; co++)
            for (int t = 0; t < taps; t++)
                for (int ci = 0; ci < l.cin; ci++) {
                    double v = nrand(seed, idx++) * sd;
                    if (co < l.cout_real && ci < l.cin_real) hw[l.woff + ((size_t)co * taps + t) * l.cin + ci] = (float)v;
                }
    }
    bool ok = true;
    n->w = (float *)dev(n, sizeof(float) * n->nparam);
    n->g = (float *)dev(n, sizeof(float) * n->nparam);
    n->m = (float *)dev(n, sizeof(float) * n->nparam);
    n->v = (float *)dev(n, sizeof(float) * n->nparam);
    n->wb = (bf16 *)dev(n, sizeof(bf16) * n->nparam);
    n->wf = (bf16 *)dev(n, sizeof(bf16) * n->nparam);
    ok = n->w && n->g && n->m && n->v && n->wb && n->wf;
    if (ok) ok = hipMemcpy(n->w, hw.data(), sizeof(float) * n->nparam, hipMemcpyHostToDevice) == hipSuccess;
    if (ok) ok = hipMemset(n->m, 0, sizeof(float) * n->nparam) == hipSuccess;
    if (ok) ok = hipMemset(n->v, 0, sizeof(float) * n->nparam) == hipSuccess;
    // activations
    n->in0 = tens(n, 32, 0);
    for (int l = 0; l < 4; l++)
        for (int k = 0; k < 2; k++) {
            n->e[l][k] = tens(n, c[l], l);
            n->ge[l][k] = tens(n, c[l], l);
        }
    for (int l = 0; l < 3; l++) {
        n->pool[l] = tens(n, c[l], l + 1);
        n->gpool[l] = tens(n, c[l], l + 1);
        n->arg[l] = (uint8_t *)dev(n, (size_t)P_of(n, max_batch, l + 1) * c[l]);
        n->cat[l] = tens(n, c[l + 1] + c[l], l);
        n->gcat[l] = tens(n, c[l + 1] + c[l], l);
        n->tmp[l] = tens(n, c[l], l);
        for (int k = 0; k < 2; k++) {
            n->d[l][k] = tens(n, c[l], l);
            n->gd[l][k] = tens(n, c[l], l);
        }
    }
    const int P0 = P_of(n, max_batch, 0);
    n->out = (float *)dev(n, sizeof(float) * (size_t)P0 * 32);
    n->gout = (bf16 *)dev(n, sizeof(bf16) * (size_t)P0 * 32);
    n->lossp = (float *)dev(n, sizeof(float) * ((size_t)P0 * 32 / 256 + 1));
    n->din = (float *)dev(n, sizeof(float) * (size_t)P0 * in_ch);
    n->dtarget = (float *)dev(n, sizeof(float) * (size_t)P0 * out_ch);
    n->dout = (float *)dev(n, sizeof(float) * (size_t)P0 * out_ch);
    // split-K partial buffer: the largest layer x its split count
    for (size_t i = 0; i < n->L.size(); i++) {
        const Layer &l = n->L[i];
        const int level = i < 8 ? (int)(i / 2) : (i < 14 ? 2 - (int)((i - 8) / 2) : 0);
        const int nsplit = wgrad_splits(max_batch, H >> level, W >> level, l.cin, l.cout, l.ksize);
        n->partial_n = std::max(n->partial_n, (size_t)nsplit * (l.nw() + l.cout));
    }
    n->partial = (float *)dev(n, sizeof(float) * n->partial_n);
    for (void *p : n->allocs) ok = ok && p;
    if (!ok) {
        mbl_net_destroy(n);
        return MBL_E_NOMEM;
    }
    *outp = n;
    return 0;
}

extern "C" void mbl_net_destroy(mbl_net_t n)
{
    if (!n) return;
    // best-effort teardown: nothing to report an error to
    if (n->s) (void)hipStreamSynchronize(n->s);
    for (void *p : n->allocs)
        if (p) (void)hipFree(p);
    if (n->s) (void)hipStreamDestroy(n->s);
    delete n;
}

extern "C" long long mbl_net_num_params(mbl_net_t n) { return n ? (long long)n->nparam : 0; }

extern "C" int mbl_net_params(mbl_net_t n, float *host, int get)
{
    if (!n || !host) return MBL_E_INVALID;
    hipError_t e = get ? hipMemcpy(host, n->w, sizeof(float) * n->nparam, hipMemcpyDeviceToHost)
                       : hipMemcpy(n->w, host, sizeof(float) * n->nparam, hipMemcpyHostToDevice);
    return e == hipSuccess ? 0 : MBL_E_HIP;
}

// level of layer i (for its pixel count)
static int layer_level(int i) { return i < 8 ? i / 2 : (i < 14 ? 2 - (i - 8) / 2 : 0); }

static void refresh_weights(mbl_net *n)
{
    to_bf16(n->s, n->w, n->nparam, n->wb);
    FlipArgs fa;
    memset(&fa, 0, sizeof(fa));
    for (const Layer &l : n->L) { // mbl_net_create always builds the same 15 layers (FLIP_MAX = 16)
        fa.L[fa.n] = FlipLayer{l.woff, fa.total, l.cout, l.ksize * l.ksize, l.cin};
        fa.total += l.nw();
        fa.n++;
    }
    flip_weights_all(n->s, n->w, fa, n->wf);
}

static void conv_fwd(mbl_net *n, int li, int N, const bf16 *X, int ldx, int xoff, bf16 *Y, int ldy, int yoff, bool relu,
                     float *Yf = nullptr, int ldyf = 0)
{
    const Layer &l = n->L[li];
    const int lev = layer_level(li);
    ConvArgs a;
    memset(&a, 0, sizeof(a));
    a.X = X;
    a.ldx = ldx;
    a.xoff = xoff;
    a.N = N;
    a.H = n->H >> lev;
    a.W = n->W >> lev;
    a.Cin = l.cin;
    a.Wt = n->wb + l.woff;
    a.bias = n->w + l.boff;
    a.Cout = l.cout;
    a.ksize = l.ksize;
    a.Y = Y;
    a.ldy = ldy;
    a.yoff = yoff;
    a.Yf = Yf;
    a.ldyf = ldyf;
    a.relu = relu ? 1 : 0;
    conv(n->s, a);
}

static void forward(mbl_net *n, int N)
{
    const int *c = n->ch;
    pack_input(n->s, n->din, n->in_ch, P_of(n, N, 0), n->in0.p, 32);
    const bf16 *x = n->in0.p;
    int ldx = 32;
    int li = 0;
    for (int l = 0; l < 4; l++) {
        conv_fwd(n, li++, N, x, ldx, 0, n->e[l][0].p, c[l], 0, true);
        conv_fwd(n, li++, N, n->e[l][0].p, c[l], 0, n->e[l][1].p, c[l], 0, true);
        if (l < 3) {
            maxpool(n->s, n->e[l][1].p, N, n->H >> l, n->W >> l, c[l], n->pool[l].p, n->arg[l]);
            x = n->pool[l].p;
            ldx = c[l];
        }
    }
    const bf16 *below = n->e[3][1].p;
    for (int l = 2; l >= 0; l--) {
        const int cc = c[l + 1] + c[l];
        upsample(n->s, below, N, n->H >> (l + 1), n->W >> (l + 1), c[l + 1], n->cat[l].p, cc, 0);
        copy_ch(n->s, n->e[l][1].p, c[l], 0, n->cat[l].p, cc, c[l + 1], c[l], P_of(n, N, l), 0);
        conv_fwd(n, li++, N, n->cat[l].p, cc, 0, n->d[l][0].p, c[l], 0, true);
        conv_fwd(n, li++, N, n->d[l][0].p, c[l], 0, n->d[l][1].p, c[l], 0, true);
        below = n->d[l][1].p;
    }
    conv_fwd(n, li, N, n->d[0][1].p, c[0], 0, nullptr, 0, 0, false, n->out, 32);
}

// Gradient of layer li from dZ (its pre-activation gradient): parameter
// gradients from the layer input X, and (if dX) the input gradient masked by
// (X > 0), i.e. the pre-activation gradient of the ReLU layer that produced X.
static void conv_bwd(mbl_net *n, int li, int N, const bf16 *dZ, int ldd, const bf16 *X, int ldx, bf16 *dX)
{
    const Layer &l = n->L[li];
    const int lev = layer_level(li);
    const int nsplit = wgrad_splits(N, n->H >> lev, n->W >> lev, l.cin, l.cout, l.ksize);
    WgradArgs wa;
    wa.dY = dZ;
    wa.ldd = ldd;
    wa.doff = 0;
    wa.X = X;
    wa.ldx = ldx;
    wa.xoff = 0;
    wa.N = N;
    wa.H = n->H >> lev;
    wa.W = n->W >> lev;
    wa.Cin = l.cin;
    wa.Cout = l.cout;
    wa.ksize = l.ksize;
    wa.partial = n->partial;
    conv_wgrad(n->s, wa, nsplit); // weight and bias partials (the bias follows the weights)
    reduce_partials(n->s, n->partial, nsplit, l.nw() + l.cout, n->g + l.woff, 0);
    if (!dX) return;
    ConvArgs a;
    memset(&a, 0, sizeof(a));
    a.X = dZ;
    a.ldx = ldd;
    a.N = N;
    a.H = n->H >> lev;
    a.W = n->W >> lev;
    a.Cin = l.cout;
    a.Wt = n->wf + l.woff;
    a.Cout = l.cin;
    a.ksize = l.ksize;
    a.Y = dX;
    a.ldy = l.cin;
    a.M = X;
    a.ldm = ldx;
    conv(n->s, a);
}

static void backward(mbl_net *n, int N)
{
    const int *c = n->ch;
    int li = 14;
    // head (1x1) and decoder
    conv_bwd(n, li--, N, n->gout, 32, n->d[0][1].p, c[0], n->gd[0][1].p);
    for (int l = 0; l <= 2; l++) {
        const int cc = c[l + 1] + c[l];
        conv_bwd(n, li--, N, n->gd[l][1].p, c[l], n->d[l][0].p, c[l], n->gd[l][0].p);
        conv_bwd(n, li--, N, n->gd[l][0].p, c[l], n->cat[l].p, cc, n->gcat[l].p);
        // split the concatenation: upsampled part -> level l+1, skip part -> encoder level l
        bf16 *gbelow = (l < 2) ? n->gd[l + 1][1].p : n->ge[3][1].p;
        upsample_bwd(n->s, n->gcat[l].p, cc, 0, N, n->H >> (l + 1), n->W >> (l + 1), c[l + 1], gbelow);
        copy_ch(n->s, n->gcat[l].p, cc, c[l + 1], n->ge[l][1].p, c[l], 0, c[l], P_of(n, N, l), 0);
    }
    // encoder, bottleneck first
    for (int l = 3; l >= 0; l--) {
        const bf16 *xin = (l == 0) ? n->in0.p : n->pool[l - 1].p;
        const int ldx = (l == 0) ? 32 : c[l - 1];
        conv_bwd(n, 2 * l + 1, N, n->ge[l][1].p, c[l], n->e[l][0].p, c[l], n->ge[l][0].p);
        conv_bwd(n, 2 * l, N, n->ge[l][0].p, c[l], xin, ldx, l == 0 ? nullptr : n->gpool[l - 1].p);
        if (l > 0) {
            maxpool_bwd(n->s, n->gpool[l - 1].p, n->arg[l - 1], N, n->H >> (l - 1), n->W >> (l - 1), c[l - 1],
                        n->tmp[l - 1].p);
            copy_ch(n->s, n->tmp[l - 1].p, c[l - 1], 0, n->ge[l - 1][1].p, c[l - 1], 0, c[l - 1], P_of(n, N, l - 1), 1);
        }
    }
    (void)li;
}

static int stage(mbl_net *n, const float *in, const float *target, int N, bool device_src)
{
    const size_t P0 = (size_t)P_of(n, N, 0);
    const hipMemcpyKind k = device_src ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    if (hipMemcpyAsync(n->din, in, sizeof(float) * P0 * n->in_ch, k, n->s) != hipSuccess) return MBL_E_HIP;
    if (target && hipMemcpyAsync(n->dtarget, target, sizeof(float) * P0 * n->out_ch, k, n->s) != hipSuccess)
        return MBL_E_HIP;
    return 0;
}

static int loss_and_grad(mbl_net *n, int N, double *loss)
{
    const int P0 = P_of(n, N, 0);
    int nb = mse(n->s, n->out, 32, n->dtarget, n->out_ch, P0, n->gout, 32, n->lossp);
    backward(n, N);
    std::vector<float> lp(nb);
    if (hipMemcpyAsync(lp.data(), n->lossp, sizeof(float) * nb, hipMemcpyDeviceToHost, n->s) != hipSuccess ||
        hipStreamSynchronize(n->s) != hipSuccess)
        return MBL_E_HIP;
    double s = 0.0;
    for (float x : lp) s += x;
    *loss = s / ((double)P0 * n->out_ch);
    return 0;
}

extern "C" int mbl_net_forward(mbl_net_t n, const float *in, int N, float *out)
{
    if (!n || !in || !out || N <= 0 || N > n->max_batch) return MBL_E_INVALID;
    refresh_weights(n);
    if (stage(n, in, nullptr, N, false)) return MBL_E_HIP;
    forward(n, N);
    const size_t P0 = (size_t)P_of(n, N, 0);
    // only the real output channels of the 32-wide head cross to the host
    take_channels(n->s, n->out, 32, n->out_ch, (int)P0, n->dout);
    if (hipMemcpyAsync(out, n->dout, sizeof(float) * P0 * n->out_ch, hipMemcpyDeviceToHost, n->s) != hipSuccess ||
        hipStreamSynchronize(n->s) != hipSuccess)
        return MBL_E_HIP;
    return 0;
}

extern "C" int mbl_net_loss_grad(mbl_net_t n, const float *in, const float *target, int N, double *loss, float *grad)
{
    if (!n || !in || !target || N <= 0 || N > n->max_batch) return MBL_E_INVALID;
    refresh_weights(n);
    if (stage(n, in, target, N, false)) return MBL_E_HIP;
    forward(n, N);
    double l = 0.0;
    if (loss_and_grad(n, N, &l)) return MBL_E_HIP;
    if (loss) *loss = l;
    if (grad && hipMemcpy(grad, n->g, sizeof(float) * n->nparam, hipMemcpyDeviceToHost) != hipSuccess)
        return MBL_E_HIP;
    return 0;
}

extern "C" int mbl_net_train(mbl_net_t n, const float *in, const float *target, int nsamples, int batch, int epochs,
                             float lr, unsigned long long seed, double *epoch_loss)
{
    if (!n || !in || !target || nsamples <= 0 || batch <= 0 || batch > n->max_batch || epochs <= 0)
        return MBL_E_INVALID;
    const size_t P1 = (size_t)n->H * n->W;
    // the training set stays resident in HBM for all epochs
    float *din = nullptr, *dt = nullptr;
    if (hipMalloc(&din, sizeof(float) * P1 * n->in_ch * nsamples) != hipSuccess) return MBL_E_NOMEM;
    if (hipMalloc(&dt, sizeof(float) * P1 * n->out_ch * nsamples) != hipSuccess) {
        (void)hipFree(din);
        return MBL_E_NOMEM;
    }
    int *dord = nullptr; // the epoch's sample order, read by the batch gather
    if (hipMalloc(&dord, sizeof(int) * nsamples) != hipSuccess) {
        (void)hipFree(din);
        (void)hipFree(dt);
        return MBL_E_NOMEM;
    }
    int rc = 0;
    if (hipMemcpy(din, in, sizeof(float) * P1 * n->in_ch * nsamples, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dt, target, sizeof(float) * P1 * n->out_ch * nsamples, hipMemcpyHostToDevice) != hipSuccess)
        rc = MBL_E_HIP;
    std::vector<int> order(nsamples);
    for (int i = 0; i < nsamples; i++) order[i] = i;
    unsigned long long st = seed ? seed : 1;
    for (int ep = 0; ep < epochs && !rc; ep++) {
        for (int i = nsamples - 1; i > 0; i--) { // Fisher-Yates with a xorshift64 stream
            st ^= st << 13;
            st ^= st >> 7;
            st ^= st << 17;
            int j = (int)(st % (unsigned long long)(i + 1));
            std::swap(order[i], order[j]);
        }
        double sum = 0.0;
        int nb = 0;
        // the previous epoch's gathers have read the old order before it is replaced
        if (hipStreamSynchronize(n->s) != hipSuccess ||
            hipMemcpy(dord, order.data(), sizeof(int) * nsamples, hipMemcpyHostToDevice) != hipSuccess) {
            rc = MBL_E_HIP;
            break;
        }
        for (int b0 = 0; b0 + batch <= nsamples && !rc; b0 += batch) {
            gather_samples(n->s, din, dt, dord, b0, batch, P1 * n->in_ch, P1 * n->out_ch, n->din, n->dtarget);
            if (hipGetLastError() != hipSuccess) { // a batch that was not staged is not trained on
                rc = MBL_E_HIP;
                break;
            }
            refresh_weights(n);
            forward(n, batch);
            double l = 0.0;
            if ((rc = loss_and_grad(n, batch, &l)) != 0) break;
            sum += l;
            nb++;
            adam(n->s, n->w, n->g, n->m, n->v, n->nparam, lr, ++n->adam_t);
        }
        if (epoch_loss) epoch_loss[ep] = nb ? sum / nb : 0.0;
    }
    if (hipStreamSynchronize(n->s) != hipSuccess) rc = MBL_E_HIP;
    (void)hipFree(din);
    (void)hipFree(dt);
    (void)hipFree(dord);
    return rc;
}
