// mbl_net.h -- internal declarations of the warm-start network kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace mbl {
namespace net {

typedef uint16_t bf16; // bfloat16 bits

struct ConvArgs {
    const bf16 *X;
    int ldx, xoff;      // input leading dimension (channels per pixel) and channel offset
    int N, H, W, Cin;   // Cin % 32 == 0
    const bf16 *Wt;     // [Cout][ksize*ksize][Cin]
    const float *bias;  // [Cout] or null
    int Cout;           // % 32 == 0
    int ksize;          // 3 or 1
    bf16 *Y;            // BF16 output or null
    int ldy, yoff;
    float *Yf;          // FP32 output or null
    int ldyf;
    const bf16 *M;      // multiply by (M > 0) or null
    int ldm, moff;
    int relu;
};

struct WgradArgs {
    const bf16 *dY;
    int ldd, doff;
    const bf16 *X;
    int ldx, xoff;
    int N, H, W, Cin, Cout, ksize;
    float *partial;     // [nsplit][Cout*ksize*ksize*Cin + Cout]: weights, then the bias
};

void conv(hipStream_t s, const ConvArgs &a);
// split count of a layer's weight gradient (partial sets reduced by reduce_partials)
int wgrad_splits(int N, int H, int W, int Cin, int Cout, int ksize);
void conv_wgrad(hipStream_t s, const WgradArgs &a, int nsplit);
void reduce_partials(hipStream_t s, const float *partial, int nsplit, size_t n, float *out, int accumulate);
void flip_weights(hipStream_t s, const float *W, int Cout, int taps, int Cin, bf16 *Wb);

// the flipped BF16 weights of every layer in one launch
constexpr int FLIP_MAX = 16;
struct FlipLayer {
    size_t woff;  // offset of the layer's weights in W and Wb
    size_t start; // first index of the layer in the concatenation of all layers' weights
    int cout, taps, cin;
};
struct FlipArgs {
    FlipLayer L[FLIP_MAX];
    int n;
    size_t total;
};
void flip_weights_all(hipStream_t s, const float *W, const FlipArgs &fa, bf16 *Wb);
// dst[k][e] = src[order[b0 + k]][e], for the inputs and the targets of one batch
void gather_samples(hipStream_t s, const float *src_in, const float *src_t, const int *order, int b0, int batch,
                    size_t per_in, size_t per_t, float *dst_in, float *dst_t);
void to_bf16(hipStream_t s, const float *src, size_t n, bf16 *dst);
void maxpool(hipStream_t s, const bf16 *X, int N, int H, int W, int C, bf16 *Y, uint8_t *arg);
void maxpool_bwd(hipStream_t s, const bf16 *dY, const uint8_t *arg, int N, int H, int W, int C, bf16 *dX);
void upsample(hipStream_t s, const bf16 *X, int N, int H, int W, int C, bf16 *Y, int ld, int off);
void upsample_bwd(hipStream_t s, const bf16 *dY, int ld, int off, int N, int H, int W, int C, bf16 *dX);
void copy_ch(hipStream_t s, const bf16 *src, int lds, int soff, bf16 *dst, int ldd, int doff, int C, int P, int add);
int mse(hipStream_t s, const float *Y, int ldy, const float *T, int C, int P, bf16 *dY, int ldg, float *partial);
void adam(hipStream_t s, float *w, const float *g, float *m, float *v, size_t n, float lr, int t);
void pack_input(hipStream_t s, const float *src, int Creal, int P, bf16 *dst, int Cpad);
void take_channels(hipStream_t s, const float *src, int lds, int C, int P, float *dst);

} // namespace net
} // namespace mbl
