// mbl_net_kernels.hip -- kernels of the warm-start network (a small conv
// encoder-decoder) for gfx950: BF16 activations in NHWC layout, BF16 MFMA
// (v_mfma_f32_16x16x32_bf16) with FP32 accumulation, FP32 master weights.
//
// MFMA operand layout used here (16x16 output tile, K = 32 per instruction):
// lane l supplies row (l % 16) of A and column (l % 16) of B, eight K values
// each from group g = l / 16; the accumulator lane l holds rows 4g..4g+3 of
// column l % 16. Both operands take their K values from the same (g, element)
// positions, so the K order inside a group does not affect the product.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mbl_net.h"

namespace mbl {
namespace net {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;

__device__ inline float bf2f(bf16 v) { return __uint_as_float(((uint32_t)v) << 16); }
__device__ inline bf16 f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u); // round to nearest even (finite values)
    return (bf16)(u >> 16);
}

__device__ inline bf16x8 load8(const bf16 *p)
{
    uint4 v = *reinterpret_cast<const uint4 *>(p);
    return *reinterpret_cast<bf16x8 *>(&v);
}

__device__ inline bf16x8 zero8()
{
    uint4 v = {0u, 0u, 0u, 0u};
    return *reinterpret_cast<bf16x8 *>(&v);
}

// ---------------------------------------------------------------------------
// 3x3 (or 1x1) convolution, zero padding, as an implicit GEMM:
//   out[p][co] = bias[co] + sum_{tap, ci} X[p + tap][ci] * Wt[co][tap][ci]
// Workgroup: 4 waves = 128 output pixels x COT output channels (COT = 64, or
// 32 when Cout is not a multiple of 64); wave w owns pixels 32w..32w+31 as two
// 16-pixel MFMA rows and all COT/16 channel tiles, so each activation load
// feeds COT/16 MFMAs. Per k-step (one tap, 32 input channels) the COT x 32
// weight tile is loaded once per workgroup into LDS (one 16-byte load per
// thread) and read from there by the four waves. Cin % 32 == 0, Cout % 32 == 0.
// Epilogue: optional bias, optional multiply by (mask > 0) (ReLU derivative of
// the layer input in the backward-data pass), optional ReLU, BF16 store into a
// channel slice of Y (leading dimension ldy, offset yoff) and/or FP32 store.
// ---------------------------------------------------------------------------
constexpr int CONV_PX = 128;  // output pixels per workgroup
constexpr int CONV_WLD = 40;  // LDS row of the weight tile (bf16): 32 + 8, rows stay 16-byte aligned

template <int COT> __global__ __launch_bounds__(256) void k_conv(ConvArgs a)
{
    constexpr int NC = COT / 16;
    __shared__ __attribute__((aligned(16))) bf16 sW[COT * CONV_WLD];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int g = lane >> 4, r16 = lane & 15;
    const int HW = a.H * a.W, P = a.N * HW;
    const int p0 = blockIdx.x * CONV_PX + 32 * w;
    const int c0 = blockIdx.y * COT;
    const int taps = a.ksize * a.ksize, half = a.ksize / 2;
    // the two pixels this lane supplies (row r16 of each 16-pixel MFMA row)
    int pn[2], py[2], px[2];
    bool pin[2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const int p = p0 + 16 * h + r16;
        pin[h] = p < P;
        pn[h] = p / HW;
        const int rem = p - pn[h] * HW;
        py[h] = rem / a.W;
        px[h] = rem - py[h] * a.W;
    }
    f32x4 acc[2][NC];
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
        for (int c = 0; c < NC; c++) acc[h][c] = f32x4{0.f, 0.f, 0.f, 0.f};
    // weight staging role: row t >> 2 of the tile, input channels (t & 3) * 8 .. + 7
    const int wrow = t >> 2, wcc = (t & 3) * 8;
    for (int tap = 0; tap < taps; tap++) {
        const int ty = tap / a.ksize - half, tx = tap % a.ksize - half;
        const bf16 *xrow[2];
        bool valid[2];
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int yy = py[h] + ty, xx = px[h] + tx;
            valid[h] = pin[h] && yy >= 0 && yy < a.H && xx >= 0 && xx < a.W;
            xrow[h] = valid[h] ? a.X + ((size_t)(pn[h] * a.H + yy) * a.W + xx) * a.ldx + a.xoff : a.X;
        }
        const bf16 *wsrc = a.Wt + ((size_t)(c0 + wrow) * taps + tap) * a.Cin + wcc;
        for (int ci = 0; ci < a.Cin; ci += 32) {
            __syncthreads(); // the previous k-step's reads of sW are done
            if (wrow < COT) *reinterpret_cast<uint4 *>(sW + wrow * CONV_WLD + wcc) = *reinterpret_cast<const uint4 *>(wsrc + ci);
            bf16x8 av[2];
#pragma unroll
            for (int h = 0; h < 2; h++) av[h] = valid[h] ? load8(xrow[h] + ci + 8 * g) : zero8();
            __syncthreads();
#pragma unroll
            for (int c = 0; c < NC; c++) {
                const bf16x8 bv = load8(sW + (16 * c + r16) * CONV_WLD + 8 * g);
#pragma unroll
                for (int h = 0; h < 2; h++) acc[h][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[h], bv, acc[h][c], 0, 0, 0);
            }
        }
    }
    // acc[h][c][r] = out[pixel p0 + 16h + 4g + r][channel c0 + 16c + r16]
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
        for (int c = 0; c < NC; c++) {
            const int co = c0 + 16 * c + r16;
            const float bias = a.bias ? a.bias[co] : 0.f;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int pix = p0 + 16 * h + 4 * g + r;
                if (pix >= P) continue;
                float v = acc[h][c][r] + bias;
                if (a.M && !(bf2f(a.M[(size_t)pix * a.ldm + a.moff + co]) > 0.f)) v = 0.f;
                if (a.relu) v = fmaxf(v, 0.f);
                if (a.Yf) a.Yf[(size_t)pix * a.ldyf + co] = v;
                if (a.Y) a.Y[(size_t)pix * a.ldy + a.yoff + co] = f2bf(v);
            }
        }
}

// floor(v / d) for 0 <= v < 2^24 through a float reciprocal, corrected to the exact quotient
__device__ inline int fdiv(int v, int d, float inv)
{
    int q = (int)((float)v * inv);
    q -= (q * d > v);
    q += ((q + 1) * d <= v);
    return q;
}

// ---------------------------------------------------------------------------
// Weight gradient: dW[co][tap][ci] = sum_p dY[p][co] * X[p + tap][ci] over one
// pixel chunk per blockIdx.z, written as FP32 partials (reduced in order by
// k_reduce_partials). A workgroup (4 waves) computes a 32 co x 32 ci tile of
// one tap. Per stage it loads 64 pixels of dY[.][co0..co0+31] and of the
// tap-shifted X[.][ci0..ci0+31] with one 16-byte load per thread and operand,
// stores both transposed in LDS ([channel][pixel], so each lane's 8 K values
// are one 16-byte LDS read), and each wave runs two MFMA k-steps (K = 32
// pixels) on its 16 x 16 sub-tile while the next stage's loads are in flight. The workgroups of the first tap and input
// channel tile also sum the dY values they load: the bias gradient
// db[co] = sum_p dY[p][co], written after the weights' partials (the bias
// follows the weights in the parameter vector), so one ordered reduction
// yields both.
// ---------------------------------------------------------------------------
constexpr int WG_PX = 64;      // pixels per stage
constexpr int WG_LD = WG_PX + 8; // LDS row length (bf16): rows stay 16-byte aligned, banks spread

__global__ __launch_bounds__(256) void k_conv_wgrad(WgradArgs a)
{
    // two LDS buffers: stage s fills one while the waves may still read stage s-1's from the
    // other, so one barrier per stage; the next stage's global loads are in flight during this
    // stage's MFMAs
    __shared__ __attribute__((aligned(16))) bf16 sA[2][32 * WG_LD]; // dY^T
    __shared__ __attribute__((aligned(16))) bf16 sB[2][32 * WG_LD]; // X^T (tap-shifted)
    __shared__ float sBias[64][33];                                 // per pixel lane: its 32 co sums
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int g = lane >> 4, r16 = lane & 15;
    const int taps = a.ksize * a.ksize, half = a.ksize / 2;
    const int co0 = blockIdx.x * 32;
    const int tap = blockIdx.y % taps;
    const int ci0 = (blockIdx.y / taps) * 32;
    const int P = a.N * a.H * a.W;
    const int chunk = (P + gridDim.z - 1) / gridDim.z;
    const int pb = blockIdx.z * chunk, pe = min(P, pb + chunk);
    const int dy = tap / a.ksize - half, dx = tap % a.ksize - half;
    const int HW = a.H * a.W;
    const float invHW = 1.0f / (float)HW, invW = 1.0f / (float)a.W;
    // loader role: pixel lp of the stage, channels lc..lc+7 of both operands
    const int lp = t >> 2, lc = (t & 3) * 8;
    // compute role: wave w owns co rows 16*(w&1).. and ci columns 16*(w>>1)..
    const int wco = (w & 1) * 16, wci = (w >> 1) * 16;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    const bool bias = (blockIdx.y == 0); // tap 0, input channels 0..31
    float bs[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    auto fetch = [&](int s0, uint4 &va, uint4 &vb) {
        const int p = s0 + lp;
        va = uint4{0u, 0u, 0u, 0u};
        vb = uint4{0u, 0u, 0u, 0u};
        if (p < pe) {
            va = *reinterpret_cast<const uint4 *>(a.dY + (size_t)p * a.ldd + a.doff + co0 + lc);
            const int n = fdiv(p, HW, invHW), rem = p - n * HW;
            const int yq = fdiv(rem, a.W, invW);
            const int yy = yq + dy, xx = rem - yq * a.W + dx;
            if (yy >= 0 && yy < a.H && xx >= 0 && xx < a.W)
                vb = *reinterpret_cast<const uint4 *>(a.X + ((size_t)(n * a.H + yy) * a.W + xx) * a.ldx + a.xoff +
                                                      ci0 + lc);
        }
    };
    uint4 va, vb;
    fetch(pb, va, vb);
    int buf = 0;
    for (int s0 = pb; s0 < pe; s0 += WG_PX) {
        const bf16 *ea = reinterpret_cast<const bf16 *>(&va);
        const bf16 *eb = reinterpret_cast<const bf16 *>(&vb);
        if (bias) {
#pragma unroll
            for (int i = 0; i < 8; i++) bs[i] += bf2f(ea[i]);
        }
#pragma unroll
        for (int i = 0; i < 8; i++) {
            sA[buf][(lc + i) * WG_LD + lp] = ea[i];
            sB[buf][(lc + i) * WG_LD + lp] = eb[i];
        }
        __syncthreads();
        if (s0 + WG_PX < pe) fetch(s0 + WG_PX, va, vb);
#pragma unroll
        for (int ks = 0; ks < WG_PX / 32; ks++) {
            const bf16x8 A = load8(sA[buf] + (wco + r16) * WG_LD + ks * 32 + 8 * g);
            const bf16x8 B = load8(sB[buf] + (wci + r16) * WG_LD + ks * 32 + 8 * g);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, acc, 0, 0, 0);
        }
        buf ^= 1;
    }
    // acc[r] = D[co = co0 + wco + 4g + r][ci = ci0 + wci + r16]
    const size_t nw = (size_t)a.Cout * taps * a.Cin;
    float *out = a.partial + (size_t)blockIdx.z * (nw + a.Cout);
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int co = co0 + wco + 4 * g + r;
        out[((size_t)co * taps + tap) * a.Cin + ci0 + wci + r16] = acc[r];
    }
    if (bias) { // the 64 pixel lanes' sums of each co, added in lane order
#pragma unroll
        for (int i = 0; i < 8; i++) sBias[lp][lc + i] = bs[i];
        __syncthreads();
        if (t < 32) {
            float sum = 0.f;
            for (int q = 0; q < 64; q++) sum += sBias[q][t];
            out[nw + co0 + t] = sum;
        }
    }
}

// out[i] (+)= sum_s partial[s][i], in split order (deterministic)
__global__ void k_reduce_partials(const float *partial, int nsplit, size_t n, float *out, int accumulate)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.f;
    for (int k = 0; k < nsplit; k++) s += partial[(size_t)k * n + i];
    out[i] = accumulate ? out[i] + s : s;
}

// Backward-data weights: Wb[ci][tap][co] = W[co][taps-1-tap][ci] (BF16)
__global__ void k_flip_weights(const float *W, int Cout, int taps, int Cin, bf16 *Wb)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)Cout * taps * Cin;
    if (i >= n) return;
    const int ci = (int)(i % Cin);
    const int tap = (int)((i / Cin) % taps);
    const int co = (int)(i / ((size_t)Cin * taps));
    Wb[((size_t)ci * taps + (taps - 1 - tap)) * Cout + co] = f2bf(W[i]);
}

__global__ void k_to_bf16(const float *src, size_t n, bf16 *dst)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = f2bf(src[i]);
}

// 2x2 max pool (NHWC); arg holds the winning child 0..3 per output element
__global__ void k_maxpool(const bf16 *X, int N, int H, int W, int C, bf16 *Y, uint8_t *arg)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int Ho = H / 2, Wo = W / 2;
    const size_t n_out = (size_t)N * Ho * Wo * C;
    if (i >= n_out) return;
    const int c = (int)(i % C);
    const size_t q = i / C;
    const int xo = (int)(q % Wo), yo = (int)((q / Wo) % Ho), n = (int)(q / ((size_t)Wo * Ho));
    float best = -3.4e38f;
    int bi = 0;
    for (int k = 0; k < 4; k++) {
        const int yy = 2 * yo + k / 2, xx = 2 * xo + k % 2;
        const float v = bf2f(X[((size_t)(n * H + yy) * W + xx) * C + c]);
        if (v > best) {
            best = v;
            bi = k;
        }
    }
    Y[i] = f2bf(best);
    arg[i] = (uint8_t)bi;
}

__global__ void k_maxpool_bwd(const bf16 *dY, const uint8_t *arg, int N, int H, int W, int C, bf16 *dX)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int Ho = H / 2, Wo = W / 2;
    const size_t n_out = (size_t)N * Ho * Wo * C;
    if (i >= n_out) return;
    const int c = (int)(i % C);
    const size_t q = i / C;
    const int xo = (int)(q % Wo), yo = (int)((q / Wo) % Ho), n = (int)(q / ((size_t)Wo * Ho));
    for (int k = 0; k < 4; k++) {
        const int yy = 2 * yo + k / 2, xx = 2 * xo + k % 2;
        dX[((size_t)(n * H + yy) * W + xx) * C + c] = (k == arg[i]) ? dY[i] : (bf16)0;
    }
}

// nearest x2 upsample of X[N][H][W][C] into channels [off, off+C) of Y (ld)
__global__ void k_upsample(const bf16 *X, int N, int H, int W, int C, bf16 *Y, int ld, int off)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n_out = (size_t)N * 2 * H * 2 * W * C;
    if (i >= n_out) return;
    const int c = (int)(i % C);
    const size_t q = i / C;
    const int x = (int)(q % (2 * W)), y = (int)((q / (2 * W)) % (2 * H)), n = (int)(q / ((size_t)4 * W * H));
    Y[q * ld + off + c] = X[((size_t)(n * H + y / 2) * W + x / 2) * C + c];
}

// dX[n][y][x][c] = sum of the four children of channels [off, off+C) of dY
__global__ void k_upsample_bwd(const bf16 *dY, int ld, int off, int N, int H, int W, int C, bf16 *dX)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n_in = (size_t)N * H * W * C;
    if (i >= n_in) return;
    const int c = (int)(i % C);
    const size_t q = i / C;
    const int x = (int)(q % W), y = (int)((q / W) % H), n = (int)(q / ((size_t)W * H));
    float s = 0.f;
    for (int k = 0; k < 4; k++) {
        const size_t pc = (size_t)(n * 2 * H + 2 * y + k / 2) * (2 * W) + 2 * x + k % 2;
        s += bf2f(dY[pc * ld + off + c]);
    }
    dX[i] = f2bf(s);
}

// channel-slice copy / accumulate: dst[p][doff+c] (+)= src[p][soff+c]
__global__ void k_copy_ch(const bf16 *src, int lds, int soff, bf16 *dst, int ldd, int doff, int C, int P, int add)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)P * C) return;
    const int c = (int)(i % C);
    const size_t p = i / C;
    float v = bf2f(src[p * lds + soff + c]);
    if (add) v += bf2f(dst[p * ldd + doff + c]);
    dst[p * ldd + doff + c] = f2bf(v);
}

// mean squared error over P pixels x C channels of an FP32 output; dY in BF16
// padded to ldg channels (the head's MFMA width); loss partials per workgroup
__global__ __launch_bounds__(256) void k_mse(const float *Y, int ldy, const float *T, int C, int P, bf16 *dY, int ldg,
                                             float *partial)
{
    __shared__ float sh[256];
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (i < (size_t)P * ldg) {
        const int c = (int)(i % ldg);
        const size_t p = i / ldg;
        float d = 0.f;
        if (c < C) {
            d = Y[p * ldy + c] - T[p * C + c];
            l = d * d;
        }
        dY[i] = f2bf(2.f * d / (float)((size_t)P * C));
    }
    sh[threadIdx.x] = l;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// Adam (Kingma & Ba 2015) on FP32 master parameters
__global__ void k_adam(float *w, const float *g, float *m, float *v, size_t n, float lr, float b1, float b2, float eps,
                       float c1, float c2)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// FP32 host-layout input [P][Cin_real] -> BF16 padded [P][Cpad]
__global__ void k_pack_input(const float *src, int Creal, int P, bf16 *dst, int Cpad)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)P * Cpad) return;
    const int c = (int)(i % Cpad);
    const size_t p = i / Cpad;
    dst[i] = c < Creal ? f2bf(src[p * Creal + c]) : (bf16)0;
}

// dst[p][c] = src[p][c] for the first C of lds channels
__global__ void k_take_channels(const float *src, int lds, int C, int P, float *dst)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)P * C) return;
    const size_t p = i / C;
    dst[i] = src[p * lds + (i - p * C)];
}

// ---- launchers ----
static inline dim3 g1(size_t n) { return dim3((unsigned)((n + 255) / 256)); }

void conv(hipStream_t s, const ConvArgs &a)
{
    const int P = a.N * a.H * a.W;
    if (a.Cout % 64 == 0)
        hipLaunchKernelGGL((k_conv<64>), dim3((P + CONV_PX - 1) / CONV_PX, a.Cout / 64), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((k_conv<32>), dim3((P + CONV_PX - 1) / CONV_PX, a.Cout / 32), dim3(256), 0, s, a);
}
void conv_wgrad(hipStream_t s, const WgradArgs &a, int nsplit)
{
    const int taps = a.ksize * a.ksize;
    hipLaunchKernelGGL(k_conv_wgrad, dim3(a.Cout / 32, taps * (a.Cin / 32), nsplit), dim3(256), 0, s, a);
}
void reduce_partials(hipStream_t s, const float *partial, int nsplit, size_t n, float *out, int acc)
{
    hipLaunchKernelGGL(k_reduce_partials, g1(n), dim3(256), 0, s, partial, nsplit, n, out, acc);
}
void flip_weights(hipStream_t s, const float *W, int Cout, int taps, int Cin, bf16 *Wb)
{
    hipLaunchKernelGGL(k_flip_weights, g1((size_t)Cout * taps * Cin), dim3(256), 0, s, W, Cout, taps, Cin, Wb);
}
void to_bf16(hipStream_t s, const float *src, size_t n, bf16 *dst)
{
    hipLaunchKernelGGL(k_to_bf16, g1(n), dim3(256), 0, s, src, n, dst);
}
void maxpool(hipStream_t s, const bf16 *X, int N, int H, int W, int C, bf16 *Y, uint8_t *arg)
{
    hipLaunchKernelGGL(k_maxpool, g1((size_t)N * (H / 2) * (W / 2) * C), dim3(256), 0, s, X, N, H, W, C, Y, arg);
}
void maxpool_bwd(hipStream_t s, const bf16 *dY, const uint8_t *arg, int N, int H, int W, int C, bf16 *dX)
{
    hipLaunchKernelGGL(k_maxpool_bwd, g1((size_t)N * (H / 2) * (W / 2) * C), dim3(256), 0, s, dY, arg, N, H, W, C,
                       dX);
}
void upsample(hipStream_t s, const bf16 *X, int N, int H, int W, int C, bf16 *Y, int ld, int off)
{
    hipLaunchKernelGGL(k_upsample, g1((size_t)N * 4 * H * W * C), dim3(256), 0, s, X, N, H, W, C, Y, ld, off);
}
void upsample_bwd(hipStream_t s, const bf16 *dY, int ld, int off, int N, int H, int W, int C, bf16 *dX)
{
    hipLaunchKernelGGL(k_upsample_bwd, g1((size_t)N * H * W * C), dim3(256), 0, s, dY, ld, off, N, H, W, C, dX);
}
void copy_ch(hipStream_t s, const bf16 *src, int lds, int soff, bf16 *dst, int ldd, int doff, int C, int P, int add)
{
    hipLaunchKernelGGL(k_copy_ch, g1((size_t)P * C), dim3(256), 0, s, src, lds, soff, dst, ldd, doff, C, P, add);
}
int mse(hipStream_t s, const float *Y, int ldy, const float *T, int C, int P, bf16 *dY, int ldg, float *partial)
{
    const size_t n = (size_t)P * ldg;
    hipLaunchKernelGGL(k_mse, g1(n), dim3(256), 0, s, Y, ldy, T, C, P, dY, ldg, partial);
    return (int)((n + 255) / 256);
}
void adam(hipStream_t s, float *w, const float *g, float *m, float *v, size_t n, float lr, int t)
{
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
    hipLaunchKernelGGL(k_adam, g1(n), dim3(256), 0, s, w, g, m, v, n, lr, b1, b2, 1e-8f, c1, c2);
}
void take_channels(hipStream_t s, const float *src, int lds, int C, int P, float *dst)
{
    hipLaunchKernelGGL(k_take_channels, g1((size_t)P * C), dim3(256), 0, s, src, lds, C, P, dst);
}
void pack_input(hipStream_t s, const float *src, int Creal, int P, bf16 *dst, int Cpad)
{
    hipLaunchKernelGGL(k_pack_input, g1((size_t)P * Cpad), dim3(256), 0, s, src, Creal, P, dst, Cpad);
}

} // namespace net
} // namespace mbl
