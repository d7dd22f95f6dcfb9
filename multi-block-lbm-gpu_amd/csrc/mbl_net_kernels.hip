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

#include <algorithm>
#include <type_traits>
#include <utility>

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

// f(std::integral_constant<int, 0>{}) ... f(std::integral_constant<int, N - 1>{}): loop bodies whose
// index is a compile-time constant from the start
template <class F, int... I> __device__ __forceinline__ void static_for_impl(F &f, std::integer_sequence<int, I...>)
{
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F> __device__ __forceinline__ void static_for(F &&f)
{
    static_for_impl(f, std::make_integer_sequence<int, N>{});
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
// Launched for the 1x1 head; 3x3 layers run k_conv3_tile.
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

// ---------------------------------------------------------------------------
// 3x3 convolution on 2D pixel tiles: a workgroup computes an 8 x TW pixel tile
// of one image for COT output channels (TW = 16 on 4 waves, or 32 on 8 waves).
// Per block of 32 input channels it stages the tile's 10 x (TW + 2) halo of X
// and the COT x 9 x 32 weights in LDS, and each wave runs 9 taps x 2 groups of
// 16 pixels x COT/16 MFMAs from there, so an
// activation is read from memory once per channel block instead of once per
// tap, and two barriers cover 9 taps. The next block's global loads are in
// flight during this block's MFMAs (held in registers). Same epilogue as
// k_conv. Sums run channel block by channel block, the taps inside.
// ---------------------------------------------------------------------------
constexpr int CT_H = 8, CT_W = 16;                 // output tile (rows x columns)
constexpr int CT_HH = CT_H + 2, CT_HW = CT_W + 2;  // halo tile
constexpr int CT_LD = 40;                          // LDS row: 32 channels + 8 (rows 16-byte aligned)

__host__ __device__ inline int conv_tiles_x(int W) { return (W + CT_W - 1) / CT_W; }
__host__ __device__ inline int conv_tiles_y(int H) { return (H + CT_H - 1) / CT_H; }

template <int COT, int TW> __global__ __launch_bounds__(256 * TW / 16, 2) void k_conv3_tile(ConvArgs a)
{
    constexpr int NT = 256 * TW / 16;       // threads: 4 waves per 16 tile columns
    constexpr int HW = TW + 2;              // halo width
    constexpr int NC = COT / 16;
    constexpr int NXV = CT_HH * HW * 4;     // 16-byte vectors of one halo block
    constexpr int NXL = (NXV + NT - 1) / NT; // per thread
    constexpr int NWV = COT * 9 * 4;        // 16-byte vectors of one weight block
    constexpr int NWL = (NWV + NT - 1) / NT;
    __shared__ __attribute__((aligned(16))) bf16 sX[CT_HH * HW * CT_LD];
    __shared__ __attribute__((aligned(16))) bf16 sW[COT * 9 * CT_LD];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int g = lane >> 4, r16 = lane & 15;
    const int tx_n = (a.W + TW - 1) / TW, ty_n = conv_tiles_y(a.H);
    int b = blockIdx.x;
    const int n = b / (tx_n * ty_n);
    b -= n * tx_n * ty_n;
    const int y0 = (b / tx_n) * CT_H, x0 = (b % tx_n) * TW;
    const int c0 = blockIdx.y * COT;
    // MFMA row group q = 2w + h: tile row q / (TW / 16), columns 16 (q % (TW / 16)) .. + 15
    constexpr int CG = TW / 16;
    uint4 xv[NXL], wv[NWL];
    // register arrays indexed by compile-time constants only (static_for), so they stay in registers
    auto fetch = [&](int ci) __attribute__((always_inline)) {
        static_for<NXL>([&](auto m) __attribute__((always_inline)) {
            const int i = t + NT * m;
            xv[m] = uint4{0u, 0u, 0u, 0u};
            const int hp = i >> 2, part = i & 3;
            const int yy = y0 - 1 + hp / HW, xx = x0 - 1 + hp % HW;
            if (i < NXV && yy >= 0 && yy < a.H && xx >= 0 && xx < a.W)
                xv[m] = *reinterpret_cast<const uint4 *>(a.X + ((size_t)(n * a.H + yy) * a.W + xx) * a.ldx + a.xoff + ci +
                                                         8 * part);
        });
        static_for<NWL>([&](auto m) __attribute__((always_inline)) {
            const int i = t + NT * m;
            wv[m] = uint4{0u, 0u, 0u, 0u};
            if (i < NWV) // row = local output channel * 9 + tap
                wv[m] = *reinterpret_cast<const uint4 *>(a.Wt + (size_t)(c0 * 9 + (i >> 2)) * a.Cin + ci + 8 * (i & 3));
        });
    };
    f32x4 acc[2][NC];
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
        for (int c = 0; c < NC; c++) acc[h][c] = f32x4{0.f, 0.f, 0.f, 0.f};
    fetch(0);
    for (int ci = 0; ci < a.Cin; ci += 32) {
        __syncthreads(); // the previous block's reads of sX and sW are done
        static_for<NXL>([&](auto m) __attribute__((always_inline)) {
            const int i = t + NT * m;
            if (i < NXV) *reinterpret_cast<uint4 *>(sX + (i >> 2) * CT_LD + 8 * (i & 3)) = xv[m];
        });
        static_for<NWL>([&](auto m) __attribute__((always_inline)) {
            const int i = t + NT * m;
            if (i < NWV) *reinterpret_cast<uint4 *>(sW + (i >> 2) * CT_LD + 8 * (i & 3)) = wv[m];
        });
        __syncthreads();
        if (ci + 32 < a.Cin) fetch(ci + 32);
#pragma unroll
        for (int tap = 0; tap < 9; tap++) {
            const int ty = tap / 3, tx = tap % 3; // halo offsets of the tap
            bf16x8 av[2];
#pragma unroll
            for (int h = 0; h < 2; h++)
                av[h] = load8(sX + (((2 * w + h) / CG + ty) * HW + 16 * ((2 * w + h) % CG) + r16 + tx) * CT_LD + 8 * g);
#pragma unroll
            for (int c = 0; c < NC; c++) {
                const bf16x8 bv = load8(sW + ((16 * c + r16) * 9 + tap) * CT_LD + 8 * g);
#pragma unroll
                for (int h = 0; h < 2; h++) acc[h][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[h], bv, acc[h][c], 0, 0, 0);
            }
        }
    }
    // acc[h][c][r] = out[row y0 + (2w + h) / CG, column x0 + 16 ((2w + h) % CG) + 4g + r][channel c0 + 16c + r16]:
    // rounded to BF16
    // with the bias into LDS, then written (mask and ReLU applied) as 16-byte vectors of 8 channels;
    // the mask and ReLU commute with the rounding. The FP32 output, when asked for, is stored directly.
    constexpr int OLD = COT + 8; // LDS row of the output tile: one pixel
    static_assert(CT_H * TW * OLD <= COT * 9 * CT_LD, "output tile fits in the weight buffer");
    bf16 *sO = sW;
    __syncthreads(); // every wave is done with sW
#pragma unroll
    for (int h = 0; h < 2; h++)
#pragma unroll
        for (int c = 0; c < NC; c++) {
            const int co = c0 + 16 * c + r16;
            const float bias = a.bias ? a.bias[co] : 0.f;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int p = ((2 * w + h) / CG) * TW + 16 * ((2 * w + h) % CG) + 4 * g + r;
                const float v = acc[h][c][r] + bias;
                sO[p * OLD + 16 * c + r16] = f2bf(v);
                if (a.Yf) {
                    const int yy = y0 + p / TW, xx = x0 + p % TW;
                    if (yy < a.H && xx < a.W) {
                        float u = v;
                        const size_t pix = (size_t)(n * a.H + yy) * a.W + xx;
                        if (a.M && !(bf2f(a.M[pix * a.ldm + a.moff + co]) > 0.f)) u = 0.f;
                        if (a.relu) u = fmaxf(u, 0.f);
                        a.Yf[pix * a.ldyf + co] = u;
                    }
                }
            }
        }
    if (!a.Y) return;
    __syncthreads();
    constexpr int NV = CT_H * TW * COT / 8 / NT; // output vectors per thread
#pragma unroll
    for (int m = 0; m < NV; m++) {
        const int i = t + NT * m, p = i / (COT / 8), c8 = 8 * (i % (COT / 8));
        const int yy = y0 + p / TW, xx = x0 + p % TW;
        if (yy >= a.H || xx >= a.W) continue;
        const size_t pix = (size_t)(n * a.H + yy) * a.W + xx;
        uint4 v = *reinterpret_cast<const uint4 *>(sO + p * OLD + c8);
        uint32_t u[4] = {v.x, v.y, v.z, v.w};
        // BF16 bit tests: v > 0 <=> 1 <= bits <= 0x7F80 (+0 and NaN excluded); max(v, 0) keeps bits <= 0x7F80
        uint32_t mw[4] = {0x3F803F80u, 0x3F803F80u, 0x3F803F80u, 0x3F803F80u}; // 1.0: no mask
        if (a.M) {
            const uint4 mv = *reinterpret_cast<const uint4 *>(a.M + pix * a.ldm + a.moff + c0 + c8);
            mw[0] = mv.x, mw[1] = mv.y, mw[2] = mv.z, mw[3] = mv.w;
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            uint32_t lo = u[k] & 0xFFFFu, hi = u[k] >> 16;
            if (((mw[k] & 0xFFFFu) - 1u) >= 0x7F80u) lo = 0;
            if (((mw[k] >> 16) - 1u) >= 0x7F80u) hi = 0;
            if (a.relu) {
                if (lo > 0x7F80u) lo = 0;
                if (hi > 0x7F80u) hi = 0;
            }
            u[k] = lo | (hi << 16);
        }
        *reinterpret_cast<uint4 *>(a.Y + pix * a.ldy + a.yoff + c0 + c8) = uint4{u[0], u[1], u[2], u[3]};
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
    auto fetch = [&](int s0, uint4 &va, uint4 &vb) __attribute__((always_inline)) {
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

// ---------------------------------------------------------------------------
// 3x3 weight gradient on 2D pixel tiles: a workgroup computes the 32 co x 32 ci
// block of all 9 taps, dW[co][tap][ci] = sum_p dY[p][co] * X[p + tap][ci], over
// a contiguous range of 8 x 16 pixel tiles (blockIdx.z of the split), written
// as FP32 partials for the ordered reduction. Per tile it stages dY^T (32 co x
// 128 pixels) and the tile's 10 x 18 halo of X^T (32 ci x 10 rows, each row 32
// columns wide with column c at x = x0 - 8 + c) in LDS. The MFMA K dimension is
// pixels: a lane's 8 K values are 8 pixels of a tile row, one 16-byte LDS read;
// the taps shifted by one column are assembled from the aligned reads of the
// same row with alignbyte (no misaligned LDS access, no shifted copies). Wave w
// owns co rows 16(w & 1).. and ci columns 16(w >> 1).. of all 9 taps; the dY
// operand of a k-step serves the 9 taps. The next tile's loads are in flight
// during this tile's MFMAs. Workgroups of input-channel block 0 also sum dY for
// the bias gradient (written after the weights' partials).
// ---------------------------------------------------------------------------
constexpr int WT_LDA = CT_H * CT_W + 8;  // LDS row of dY^T: 128 pixels + 8
constexpr int WT_ROW = 32;               // LDS row of X^T (one halo row)
constexpr int WT_LDB = CT_HH * WT_ROW + 8; // per input channel: 10 halo rows + 8

__global__ __launch_bounds__(256, 2) void k_wgrad3_tile(WgradArgs a)
{
    constexpr int NAV = CT_H * CT_W * 4 / 256;          // dY vectors per thread (2)
    constexpr int NXV = CT_HH * CT_HW * 4;              // X halo vectors (720)
    constexpr int NXL = (NXV + 255) / 256;              // per thread (3)
    __shared__ __attribute__((aligned(16))) bf16 sA[32 * WT_LDA];
    __shared__ __attribute__((aligned(16))) bf16 sB[32 * WT_LDB];
    __shared__ float sBias[64][33];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int g = lane >> 4, r16 = lane & 15;
    const int co0 = blockIdx.x * 32, ci0 = blockIdx.y * 32;
    const int tx_n = conv_tiles_x(a.W), ty_n = conv_tiles_y(a.H);
    const int T = a.N * tx_n * ty_n;
    const int tb = (int)((long long)T * blockIdx.z / gridDim.z), te = (int)((long long)T * (blockIdx.z + 1) / gridDim.z);
    const bool bias = (blockIdx.y == 0);
    const int wco = 16 * (w & 1), wci = 16 * (w >> 1);
    uint4 av[NAV], xv[NXL];
    // inlined: the register arrays it fills must not go to the stack
    auto fetch = [&](int tile) __attribute__((always_inline)) {
        const int n = tile / (tx_n * ty_n), b = tile - n * tx_n * ty_n;
        const int y0 = (b / tx_n) * CT_H, x0 = (b % tx_n) * CT_W;
#pragma unroll
        for (int m = 0; m < NAV; m++) {
            const int i = t + 256 * m, p = i >> 2;
            const int yy = y0 + p / CT_W, xx = x0 + p % CT_W;
            av[m] = uint4{0u, 0u, 0u, 0u};
            if (yy < a.H && xx < a.W)
                av[m] = *reinterpret_cast<const uint4 *>(a.dY + ((size_t)(n * a.H + yy) * a.W + xx) * a.ldd + a.doff + co0 +
                                                         8 * (i & 3));
        }
#pragma unroll
        for (int m = 0; m < NXL; m++) {
            const int i = t + 256 * m;
            xv[m] = uint4{0u, 0u, 0u, 0u};
            if (i < NXV) {
                const int hp = i >> 2;
                const int yy = y0 - 1 + hp / CT_HW, xx = x0 - 1 + hp % CT_HW;
                if (yy >= 0 && yy < a.H && xx >= 0 && xx < a.W)
                    xv[m] = *reinterpret_cast<const uint4 *>(a.X + ((size_t)(n * a.H + yy) * a.W + xx) * a.ldx + a.xoff +
                                                             ci0 + 8 * (i & 3));
            }
        }
    };
    f32x4 acc[9];
#pragma unroll
    for (int k = 0; k < 9; k++) acc[k] = f32x4{0.f, 0.f, 0.f, 0.f};
    float bs[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (tb < te) fetch(tb);
    for (int tile = tb; tile < te; tile++) {
        __syncthreads(); // the previous tile's reads of sA and sB are done
#pragma unroll
        for (int m = 0; m < NAV; m++) {
            const int i = t + 256 * m, p = i >> 2, c8 = 8 * (i & 3);
            const bf16 *e = reinterpret_cast<const bf16 *>(&av[m]);
#pragma unroll
            for (int k = 0; k < 8; k++) {
                sA[(c8 + k) * WT_LDA + p] = e[k];
                if (bias) bs[k] += bf2f(e[k]);
            }
        }
#pragma unroll
        for (int m = 0; m < NXL; m++) {
            const int i = t + 256 * m;
            if (i < NXV) {
                const int hp = i >> 2, c8 = 8 * (i & 3);
                const int col = (hp / CT_HW) * WT_ROW + hp % CT_HW + 7; // halo column hx sits at x0 - 1 + hx = x0 - 8 + col
                const bf16 *e = reinterpret_cast<const bf16 *>(&xv[m]);
#pragma unroll
                for (int k = 0; k < 8; k++) sB[(c8 + k) * WT_LDB + col] = e[k];
            }
        }
        __syncthreads();
        if (tile + 1 < te) fetch(tile + 1);
#pragma unroll
        for (int ks = 0; ks < CT_H / 2; ks++) {
            // K step: tile rows 2ks, 2ks + 1; lane group g holds row 2ks + (g >> 1), columns 8(g & 1) .. + 7
            const bf16x8 A = load8(sA + (wco + r16) * WT_LDA + 32 * ks + 8 * g);
            const int row = 2 * ks + (g >> 1), cb = 8 * (g & 1);
#pragma unroll
            for (int ty = 0; ty < 3; ty++) {
                const bf16 *rp = sB + (wci + r16) * WT_LDB + (row + ty) * WT_ROW + cb;
                const uint4 mid = *reinterpret_cast<const uint4 *>(rp + 8);   // columns cb + 8 .. cb + 15: tap column 0
                const uint32_t lo = *reinterpret_cast<const uint32_t *>(rp + 6);  // columns cb + 6, cb + 7
                const uint32_t hi = *reinterpret_cast<const uint32_t *>(rp + 16); // columns cb + 16, cb + 17
                // tap column -1: columns cb + 7 .. cb + 14; +1: cb + 9 .. cb + 16
                const uint4 left = {__builtin_amdgcn_alignbyte(mid.x, lo, 2), __builtin_amdgcn_alignbyte(mid.y, mid.x, 2),
                                    __builtin_amdgcn_alignbyte(mid.z, mid.y, 2), __builtin_amdgcn_alignbyte(mid.w, mid.z, 2)};
                const uint4 right = {__builtin_amdgcn_alignbyte(mid.y, mid.x, 2), __builtin_amdgcn_alignbyte(mid.z, mid.y, 2),
                                     __builtin_amdgcn_alignbyte(mid.w, mid.z, 2), __builtin_amdgcn_alignbyte(hi, mid.w, 2)};
                acc[3 * ty + 0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, *reinterpret_cast<const bf16x8 *>(&left),
                                                                          acc[3 * ty + 0], 0, 0, 0);
                acc[3 * ty + 1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, *reinterpret_cast<const bf16x8 *>(&mid),
                                                                          acc[3 * ty + 1], 0, 0, 0);
                acc[3 * ty + 2] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, *reinterpret_cast<const bf16x8 *>(&right),
                                                                          acc[3 * ty + 2], 0, 0, 0);
            }
        }
    }
    // acc[tap][r] = dW[co = co0 + wco + 4g + r][tap][ci = ci0 + wci + r16]
    const size_t nw = (size_t)a.Cout * 9 * a.Cin;
    float *out = a.partial + (size_t)blockIdx.z * (nw + a.Cout);
#pragma unroll
    for (int tap = 0; tap < 9; tap++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int co = co0 + wco + 4 * g + r;
            out[((size_t)co * 9 + tap) * a.Cin + ci0 + wci + r16] = acc[tap][r];
        }
    if (bias) { // lane q of the loaders held pixels q and 64 + q of every tile; their sums added in lane order
        const int q = t >> 2, c8 = 8 * (t & 3);
#pragma unroll
        for (int k = 0; k < 8; k++) sBias[q][c8 + k] = bs[k];
        __syncthreads();
        if (t < 32) {
            float sum = 0.f;
            for (int j = 0; j < 64; j++) sum += sBias[j][t];
            out[nw + co0 + t] = sum;
        }
    }
}

// out[i] (+)= sum_s partial[s][i] in a fixed order (deterministic): a workgroup takes 32 outputs;
// thread group q = 0..7 sums splits q, q + 8, q + 16, ... of each, then the 8 group sums are added
// in group order
__global__ __launch_bounds__(256) void k_reduce_partials(const float *partial, int nsplit, size_t n, float *out,
                                                         int accumulate)
{
    __shared__ float sh[8][32];
    const int j = threadIdx.x & 31, q = threadIdx.x >> 5;
    const size_t i = (size_t)blockIdx.x * 32 + j;
    // four running sums per thread (splits q + 8(4m + r) in sum r), so four loads are in flight
    float s[4] = {0.f, 0.f, 0.f, 0.f};
    if (i < n) {
        int k = q;
        for (; k + 24 < nsplit; k += 32)
#pragma unroll
            for (int r = 0; r < 4; r++) s[r] += partial[(size_t)(k + 8 * r) * n + i];
        for (int r = 0; k < nsplit; k += 8, r++) s[r] += partial[(size_t)k * n + i];
    }
    sh[q][j] = (s[0] + s[1]) + (s[2] + s[3]);
    __syncthreads();
    if (q == 0 && i < n) {
        float v = sh[0][j];
        for (int r = 1; r < 8; r++) v += sh[r][j];
        out[i] = accumulate ? out[i] + v : v;
    }
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

__global__ void k_flip_weights_all(const float *W, FlipArgs fa, bf16 *Wb)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= fa.total) return;
    int l = 0;
    while (l + 1 < fa.n && i >= fa.L[l + 1].start) l++;
    const FlipLayer &L = fa.L[l];
    const size_t j = i - L.start;
    const int ci = (int)(j % L.cin);
    const int tap = (int)((j / L.cin) % L.taps);
    const int co = (int)(j / ((size_t)L.cin * L.taps));
    Wb[L.woff + ((size_t)ci * L.taps + (L.taps - 1 - tap)) * L.cout + co] = f2bf(W[L.woff + j]);
}

__global__ void k_gather_samples(const float *src_in, const float *src_t, const int *order, int b0, size_t per_in,
                                 size_t per_t, float *dst_in, float *dst_t)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    const size_t sidx = (size_t)order[b0 + k];
    if (e < per_in) dst_in[k * per_in + e] = src_in[sidx * per_in + e];
    if (e < per_t) dst_t[k * per_t + e] = src_t[sidx * per_t + e];
}

__global__ void k_to_bf16(const float *src, size_t n, bf16 *dst)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = f2bf(src[i]);
}

// The elementwise kernels below move 8 channels (one 16-byte vector) per thread: every channel
// count, leading dimension and channel offset they are given is a multiple of 8.
__device__ inline void unpack8(const uint4 &v, float f[8])
{
    const uint32_t u[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        f[2 * k] = __uint_as_float(u[k] << 16);
        f[2 * k + 1] = __uint_as_float(u[k] & 0xFFFF0000u);
    }
}
__device__ inline uint4 pack8(const float f[8])
{
    uint32_t u[4];
#pragma unroll
    for (int k = 0; k < 4; k++) u[k] = (uint32_t)f2bf(f[2 * k]) | ((uint32_t)f2bf(f[2 * k + 1]) << 16);
    return uint4{u[0], u[1], u[2], u[3]};
}
__device__ inline uint4 ld16(const bf16 *p) { return *reinterpret_cast<const uint4 *>(p); }
__device__ inline void st16(bf16 *p, const uint4 &v) { *reinterpret_cast<uint4 *>(p) = v; }

// 2x2 max pool (NHWC); arg holds the winning child 0..3 per output element (the first of equal ones)
__global__ void k_maxpool(const bf16 *X, int N, int H, int W, int C, bf16 *Y, uint8_t *arg)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; // one vector of 8 channels
    const int Ho = H / 2, Wo = W / 2, C8 = C / 8;
    const size_t n_out = (size_t)N * Ho * Wo * C8;
    if (i >= n_out) return;
    const int c = (int)(i % C8) * 8;
    const size_t q = i / C8;
    const int xo = (int)(q % Wo), yo = (int)((q / Wo) % Ho), n = (int)(q / ((size_t)Wo * Ho));
    float best[8], v[8];
    uint8_t bi[8];
    for (int k = 0; k < 4; k++) {
        const int yy = 2 * yo + k / 2, xx = 2 * xo + k % 2;
        unpack8(ld16(X + ((size_t)(n * H + yy) * W + xx) * C + c), v);
#pragma unroll
        for (int e = 0; e < 8; e++)
            if (k == 0 || v[e] > best[e]) {
                best[e] = v[e];
                bi[e] = (uint8_t)k;
            }
    }
    st16(Y + q * C + c, pack8(best));
    uint2 packed;
    packed.x = bi[0] | (bi[1] << 8) | (bi[2] << 16) | ((uint32_t)bi[3] << 24);
    packed.y = bi[4] | (bi[5] << 8) | (bi[6] << 16) | ((uint32_t)bi[7] << 24);
    *reinterpret_cast<uint2 *>(arg + q * C + c) = packed;
}

__global__ void k_maxpool_bwd(const bf16 *dY, const uint8_t *arg, int N, int H, int W, int C, bf16 *dX)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int Ho = H / 2, Wo = W / 2, C8 = C / 8;
    const size_t n_out = (size_t)N * Ho * Wo * C8;
    if (i >= n_out) return;
    const int c = (int)(i % C8) * 8;
    const size_t q = i / C8;
    const int xo = (int)(q % Wo), yo = (int)((q / Wo) % Ho), n = (int)(q / ((size_t)Wo * Ho));
    const uint4 d = ld16(dY + q * C + c);
    const uint32_t dw[4] = {d.x, d.y, d.z, d.w};
    const uint2 am = *reinterpret_cast<const uint2 *>(arg + q * C + c);
    for (int k = 0; k < 4; k++) {
        const int yy = 2 * yo + k / 2, xx = 2 * xo + k % 2;
        uint32_t o[4];
#pragma unroll
        for (int j = 0; j < 4; j++) { // channels 2j, 2j + 1: keep each half whose winner is child k
            const uint32_t word = j < 2 ? am.x : am.y;
            const uint32_t lo = (word >> (16 * (j & 1))) & 0xFFu, hi = (word >> (16 * (j & 1) + 8)) & 0xFFu;
            o[j] = (lo == (uint32_t)k ? (dw[j] & 0xFFFFu) : 0u) | (hi == (uint32_t)k ? (dw[j] & 0xFFFF0000u) : 0u);
        }
        st16(dX + ((size_t)(n * H + yy) * W + xx) * C + c, uint4{o[0], o[1], o[2], o[3]});
    }
}

// nearest x2 upsample of X[N][H][W][C] into channels [off, off+C) of Y (ld)
__global__ void k_upsample(const bf16 *X, int N, int H, int W, int C, bf16 *Y, int ld, int off)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int C8 = C / 8;
    const size_t n_out = (size_t)N * 2 * H * 2 * W * C8;
    if (i >= n_out) return;
    const int c = (int)(i % C8) * 8;
    const size_t q = i / C8;
    const int x = (int)(q % (2 * W)), y = (int)((q / (2 * W)) % (2 * H)), n = (int)(q / ((size_t)4 * W * H));
    st16(Y + q * ld + off + c, ld16(X + ((size_t)(n * H + y / 2) * W + x / 2) * C + c));
}

// dX[n][y][x][c] = sum of the four children of channels [off, off+C) of dY
__global__ void k_upsample_bwd(const bf16 *dY, int ld, int off, int N, int H, int W, int C, bf16 *dX)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int C8 = C / 8;
    const size_t n_in = (size_t)N * H * W * C8;
    if (i >= n_in) return;
    const int c = (int)(i % C8) * 8;
    const size_t q = i / C8;
    const int x = (int)(q % W), y = (int)((q / W) % H), n = (int)(q / ((size_t)W * H));
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, v[8];
    for (int k = 0; k < 4; k++) {
        const size_t pc = (size_t)(n * 2 * H + 2 * y + k / 2) * (2 * W) + 2 * x + k % 2;
        unpack8(ld16(dY + pc * ld + off + c), v);
#pragma unroll
        for (int e = 0; e < 8; e++) s[e] += v[e];
    }
    st16(dX + q * C + c, pack8(s));
}

// channel-slice copy / accumulate: dst[p][doff+c] (+)= src[p][soff+c]
__global__ void k_copy_ch(const bf16 *src, int lds, int soff, bf16 *dst, int ldd, int doff, int C, int P, int add)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int C8 = C / 8;
    if (i >= (size_t)P * C8) return;
    const int c = (int)(i % C8) * 8;
    const size_t p = i / C8;
    const uint4 v = ld16(src + p * lds + soff + c);
    if (!add) {
        st16(dst + p * ldd + doff + c, v);
        return;
    }
    float a[8], b[8];
    unpack8(v, a);
    unpack8(ld16(dst + p * ldd + doff + c), b);
#pragma unroll
    for (int e = 0; e < 8; e++) a[e] += b[e];
    st16(dst + p * ldd + doff + c, pack8(a));
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
    if (a.ksize == 3) {
        // 8 x 32 tiles on 8 waves where the width is a multiple of 32 (twice the pixels per staged
        // weight block, twice the waves per CU), else 8 x 16 tiles on 4 waves; the same sums in
        // the same order either way
        const bool wide = a.W % 32 == 0;
        const unsigned tiles = (unsigned)(a.N * conv_tiles_y(a.H) * ((a.W + (wide ? 31 : 15)) / (wide ? 32 : 16)));
        const dim3 blk(wide ? 512 : 256);
        if (a.Cout % 64 == 0) {
            if (wide)
                hipLaunchKernelGGL((k_conv3_tile<64, 32>), dim3(tiles, a.Cout / 64), blk, 0, s, a);
            else
                hipLaunchKernelGGL((k_conv3_tile<64, 16>), dim3(tiles, a.Cout / 64), blk, 0, s, a);
        } else {
            if (wide)
                hipLaunchKernelGGL((k_conv3_tile<32, 32>), dim3(tiles, a.Cout / 32), blk, 0, s, a);
            else
                hipLaunchKernelGGL((k_conv3_tile<32, 16>), dim3(tiles, a.Cout / 32), blk, 0, s, a);
        }
        return;
    }
    if (a.Cout % 64 == 0)
        hipLaunchKernelGGL((k_conv<64>), dim3((P + CONV_PX - 1) / CONV_PX, a.Cout / 64), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((k_conv<32>), dim3((P + CONV_PX - 1) / CONV_PX, a.Cout / 32), dim3(256), 0, s, a);
}
int wgrad_splits(int N, int H, int W, int Cin, int Cout, int ksize)
{
    if (ksize != 3) return std::max(1, N * H * W / 4096);
    // about 1024 workgroups, at least two tiles each, at most 512 partial sets
    const int tiles = N * ((H + CT_H - 1) / CT_H) * ((W + CT_W - 1) / CT_W);
    const int blocks = (Cout / 32) * (Cin / 32);
    return std::max(1, std::min(std::min(tiles / 2, 1024 / blocks), 512));
}
void conv_wgrad(hipStream_t s, const WgradArgs &a, int nsplit)
{
    if (a.ksize == 3) {
        hipLaunchKernelGGL(k_wgrad3_tile, dim3(a.Cout / 32, a.Cin / 32, nsplit), dim3(256), 0, s, a);
        return;
    }
    const int taps = a.ksize * a.ksize;
    hipLaunchKernelGGL(k_conv_wgrad, dim3(a.Cout / 32, taps * (a.Cin / 32), nsplit), dim3(256), 0, s, a);
}
void reduce_partials(hipStream_t s, const float *partial, int nsplit, size_t n, float *out, int acc)
{
    hipLaunchKernelGGL(k_reduce_partials, dim3((unsigned)((n + 31) / 32)), dim3(256), 0, s, partial, nsplit, n, out, acc);
}
void flip_weights(hipStream_t s, const float *W, int Cout, int taps, int Cin, bf16 *Wb)
{
    hipLaunchKernelGGL(k_flip_weights, g1((size_t)Cout * taps * Cin), dim3(256), 0, s, W, Cout, taps, Cin, Wb);
}
void flip_weights_all(hipStream_t s, const float *W, const FlipArgs &fa, bf16 *Wb)
{
    hipLaunchKernelGGL(k_flip_weights_all, g1(fa.total), dim3(256), 0, s, W, fa, Wb);
}
void gather_samples(hipStream_t s, const float *src_in, const float *src_t, const int *order, int b0, int batch,
                    size_t per_in, size_t per_t, float *dst_in, float *dst_t)
{
    const size_t per = std::max(per_in, per_t);
    hipLaunchKernelGGL(k_gather_samples, dim3((unsigned)((per + 255) / 256), (unsigned)batch), dim3(256), 0, s, src_in,
                       src_t, order, b0, per_in, per_t, dst_in, dst_t);
}
void to_bf16(hipStream_t s, const float *src, size_t n, bf16 *dst)
{
    hipLaunchKernelGGL(k_to_bf16, g1(n), dim3(256), 0, s, src, n, dst);
}
void maxpool(hipStream_t s, const bf16 *X, int N, int H, int W, int C, bf16 *Y, uint8_t *arg)
{
    hipLaunchKernelGGL(k_maxpool, g1((size_t)N * (H / 2) * (W / 2) * (C / 8)), dim3(256), 0, s, X, N, H, W, C, Y, arg);
}
void maxpool_bwd(hipStream_t s, const bf16 *dY, const uint8_t *arg, int N, int H, int W, int C, bf16 *dX)
{
    hipLaunchKernelGGL(k_maxpool_bwd, g1((size_t)N * (H / 2) * (W / 2) * (C / 8)), dim3(256), 0, s, dY, arg, N, H, W, C,
                       dX);
}
void upsample(hipStream_t s, const bf16 *X, int N, int H, int W, int C, bf16 *Y, int ld, int off)
{
    hipLaunchKernelGGL(k_upsample, g1((size_t)N * 4 * H * W * (C / 8)), dim3(256), 0, s, X, N, H, W, C, Y, ld, off);
}
void upsample_bwd(hipStream_t s, const bf16 *dY, int ld, int off, int N, int H, int W, int C, bf16 *dX)
{
    hipLaunchKernelGGL(k_upsample_bwd, g1((size_t)N * H * W * (C / 8)), dim3(256), 0, s, dY, ld, off, N, H, W, C, dX);
}
void copy_ch(hipStream_t s, const bf16 *src, int lds, int soff, bf16 *dst, int ldd, int doff, int C, int P, int add)
{
    hipLaunchKernelGGL(k_copy_ch, g1((size_t)P * (C / 8)), dim3(256), 0, s, src, lds, soff, dst, ldd, doff, C, P, add);
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
