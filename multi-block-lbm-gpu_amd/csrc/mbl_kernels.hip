// mbl_kernels.hip -- HIP kernels of the multi-block D2Q9 BGK time-stepping
// path for gfx950 (MI355X), and their asynchronous launchers.
//
// Kernels (reference behaviour restated, file:line):
//   k_stream_collide  fused pull-stream + macros + BGK + bounce-back + ghost
//                     carry-over of multi/FH/evolution_block.h:3-126 (stb1/2/3)
//   k_init_general    lbm/init_block.h:2-51
//   k_line_ops        temporal_interpolation (multi/FH/assit_multi.h:213-247),
//                     c_c_boundary_transfer (:2-46), c_f_boundary_transfer
//                     (:249-271), f_c_boundary_transfer (:273-314)
//   k_spline_solve    init_Spline3 gather+RHS (:48-116) and the tridiagonal
//                     solve that cusparseDgtsv2 performs (:193-205), as an
//                     LDS-resident parallel cyclic reduction in FP64
//   k_spline_eval     Spline3_S + calculate_Spline3_gpu (:118-185)
//   k_ring_fill       init_Fstore (multi/init_multi.h:3-28)
//   k_tri_A           init_tri_A (multi/init_multi.h:30-51)
//   k_drag            kernel_sum_Force (lbm/force_block.h:129-178), FP64
//   k_residual_*      error_L2_area + host sum (lbm/error_L2.h:3-91), FP64
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>

#include "mbl_internal.h"

namespace mbl {

// D2Q9 lattice, lbm/lbm.h:43-45 ordering.
__device__ constexpr int EX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
__device__ constexpr int EY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
__device__ constexpr int OPP[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
__device__ constexpr double W64[9] = {4. / 9, 1. / 9, 1. / 9, 1. / 9, 1. / 9, 1. / 36, 1. / 36, 1. / 36, 1. / 36};

// ---------------------------------------------------------------------------
// Plane storage codecs. `ld` returns the stored quantity q in the codec's
// arithmetic type R: q = f for ST_NATIVE, q = f - w_k for the shifted modes.
// Planes of binary32 (and binary64) are updated in binary64 arithmetic: the
// fused kernel is bound by HBM traffic there, and MI355X's binary64 vector rate
// leaves its time per launch unchanged (16384^2 periodic block, measured), while
// binary32 arithmetic biased the shedding of the three-level Re = 100 channel by
// 1e-4 (Strouhal number, mean drag) against a binary64 run, and binary64
// arithmetic over the same binary32 planes removes it (1e-7). The 16-bit
// formats, whose kernels are bound by vector instructions, compute in binary32.
// ---------------------------------------------------------------------------
template <int MODE> struct Codec;

// Pointers into HBM are cast to the global address space so the backend emits
// global_load/global_store (SGPR base + VGPR offset) rather than flat accesses.
#define GLOBAL_AS __attribute__((address_space(1)))

// Element i of a plane through a 32-bit byte offset from the plane's (uniform)
// base, so loads and stores take the scalar-base + vector-offset form instead
// of 64-bit per-lane address arithmetic. Planes are at most 4 GiB (checked on
// the host when a layout is built).
template <class T> __device__ inline const GLOBAL_AS T *off32(const GLOBAL_AS T *b, uint32_t i)
{
    return (const GLOBAL_AS T *)((const GLOBAL_AS char *)b + i * (uint32_t)sizeof(T));
}
template <class T> __device__ inline GLOBAL_AS T *off32(GLOBAL_AS T *b, uint32_t i)
{
    return (GLOBAL_AS T *)((GLOBAL_AS char *)b + i * (uint32_t)sizeof(T));
}

template <> struct Codec<ST_NATIVE> {
    typedef Lbd T;
    typedef GLOBAL_AS Lbd GT;
    typedef double R; // arithmetic type of the node update (see Codec's comment)
    static constexpr bool SHIFT = false;
    __device__ static inline R ld(T v) { return (R)v; }
    __device__ static inline T st(R v) { return (T)v; }
};

#ifndef MBL_FP64
template <> struct Codec<ST_F32S> {
    typedef float T;
    typedef GLOBAL_AS float GT;
    typedef double R;
    static constexpr bool SHIFT = true;
    __device__ static inline R ld(T v) { return (R)v; }
    __device__ static inline T st(R v) { return (T)v; }
};

template <> struct Codec<ST_F16S> {
    typedef _Float16 T;
    typedef GLOBAL_AS _Float16 GT;
    typedef float R;
    static constexpr bool SHIFT = true;
    __device__ static inline float ld(T v) { return (float)v; }
    __device__ static inline T st(float v) { return (_Float16)v; } // round to nearest even
};

template <> struct Codec<ST_BF16S> {
    typedef uint16_t T;
    typedef GLOBAL_AS uint16_t GT;
    typedef float R;
    static constexpr bool SHIFT = true;
    __device__ static inline float ld(T v) { return __uint_as_float(((uint32_t)v) << 16); }
    __device__ static inline T st(float v)
    {
        // round to nearest even on the upper 16 bits (finite inputs)
        uint32_t u = __float_as_uint(v);
        u += 0x7FFFu + ((u >> 16) & 1u);
        return (T)(u >> 16);
    }
};

// Custom 16-bit floats for shifted populations: sign, EB exponent bits with
// bias 2^EB - 1 (every code finite, the top binade is [1, 2)), 15 - EB
// fraction bits, subnormals below 2^(2 - 2^EB). |f - w_k| stays well inside
// (2^-14, 2), so the exponent bits IEEE binary16 spends on range are better
// spent on the significand. Code 0xFFFF is NaN so a diverging run stays
// visible; overflow saturates at magnitude code 0x7FFE (2 - 2^(2-FB)). Decoding places the 15 magnitude bits under
// the binary32 exponent field and rescales by a power of two (subnormal
// codes come out as binary32 subnormals and scale exactly); encoding rounds
// to nearest even on the dropped bits, and in the subnormal range rounds the
// exactly scaled value to an integer with one addition. Both are branch-free.
template <int EB> struct Half16 {
    static constexpr int FB = 15 - EB;                       // fraction bits
    static constexpr int BIAS = (1 << EB) - 1;
    static constexpr uint32_t OFF = (uint32_t)(127 - BIAS);  // binary32 exponent offset
    __device__ static inline float ld(uint16_t v)
    {
        const uint32_t b = v;
        const float r = __uint_as_float(((b & 0x7FFFu) << (23 - FB)) | ((b & 0x8000u) << 16)) *
                        __uint_as_float((OFF + 127u) << 23);
        return b == 0xFFFFu ? __uint_as_float(0x7FC00000u) : r;
    }
    __device__ static inline uint16_t st(float v)
    {
        const uint32_t u = __float_as_uint(v);
        const uint32_t s = (u >> 16) & 0x8000u, a = u & 0x7FFFFFFFu;
        constexpr uint32_t SH = 23 - FB, HALF = (1u << (SH - 1)) - 1u, MAGIC = 0x4B400000u; // bits of 1.5 * 2^23
        // Two candidate codes, both offset by MAGIC (whose low 16 bits are zero):
        // sub: |v| 2^(BIAS - 1 + FB) rounded to the nearest even integer by adding 1.5 * 2^23 in the
        //   fused multiply-add: the code below 2^(1 - BIAS), and at least the normal code above it
        //   (equal in the first normal binade, larger beyond).
        // nrm: the normal code, rounded to nearest even on the SH dropped bits with the exponent
        //   offset of the first normal binade removed before the shift (one 3-input add, one
        //   shift, one add); below that binade the sum wraps and nrm exceeds MAGIC + 2^19.
        // So min(nrm, sub) is the code everywhere, and the third operand saturates overflow and
        // infinity one code below 0x7FFF, so no sign gives 0xFFFF: one 3-input minimum.
        const uint32_t nrm = ((a + ((a >> SH) & 1u) + (HALF - ((OFF + 1u) << 23))) >> SH) + (MAGIC + (1u << FB));
        const uint32_t sub = __float_as_uint(
            __builtin_fmaf(__uint_as_float(a), __uint_as_float((uint32_t)(127 + BIAS - 1 + FB) << 23), 12582912.0f));
        const uint32_t code = min(min(nrm, sub), MAGIC + 0x7FFEu) | s;
        return (uint16_t)(a > 0x7F800000u ? 0xFFFFu : code);
    }
};

// 1-4-11: normal range 2^-14 .. 2 - 2^-10, subnormal step 2^-25.
template <> struct Codec<ST_F16C> {
    typedef uint16_t T;
    typedef GLOBAL_AS uint16_t GT;
    typedef float R;
    static constexpr bool SHIFT = true;
    __device__ static inline float ld(T v) { return Half16<4>::ld(v); }
    __device__ static inline T st(float v) { return Half16<4>::st(v); }
};

// 1-3-12: normal range 2^-6 .. 2 - 2^-11, subnormal step 2^-18.
template <> struct Codec<ST_F16D> {
    typedef uint16_t T;
    typedef GLOBAL_AS uint16_t GT;
    typedef float R;
    static constexpr bool SHIFT = true;
    __device__ static inline float ld(T v) { return Half16<3>::ld(v); }
    __device__ static inline T st(float v) { return Half16<3>::st(v); }
};
#endif

// f value in double of a stored entry of direction k
template <int MODE> __device__ inline double dec64(typename Codec<MODE>::T v, int k)
{
    double q = (double)Codec<MODE>::ld(v);
    return Codec<MODE>::SHIFT ? q + W64[k] : q;
}

template <int MODE> __device__ inline typename Codec<MODE>::T enc64(double f, int k)
{
    return Codec<MODE>::st((typename Codec<MODE>::R)(Codec<MODE>::SHIFT ? f - W64[k] : f));
}

// lbm/assit_func.h:2-9 in double
__device__ inline double feq64(int k, double rh, double ux, double uy)
{
    double eu = EX[k] * ux + EY[k] * uy;
    double uv = ux * ux + uy * uy;
    return W64[k] * rh * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * uv);
}

// 1/x: the hardware reciprocal estimate refined by one Newton-Raphson step,
// r' = r + r (1 - x r), which squares the estimate's relative error (within an
// ulp of the correctly rounded quotient, in 3 instructions instead of the ~10
// of the IEEE division). FP64 keeps the IEEE division.
__device__ inline float recip(float x)
{
    const float r = __builtin_amdgcn_rcpf(x);
    return __builtin_fmaf(r, __builtin_fmaf(-x, r, 1.0f), r);
}
__device__ inline double recip(double x) { return 1.0 / x; }

// Fused multiply-add in compute precision. The collision below spells out
// every fused operation and turns the compiler's own contraction off, so its
// rounding is fixed by the source: the vector and the per-node paths, and
// every launch shape, produce the same bits.
__device__ inline float fmar(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ inline double fmar(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Equilibrium in the stored domain, compute precision. drho = rho - 1 is
// passed separately so the shifted form keeps its small magnitude.
// feq = w rho (1 + 3 eu + 4.5 eu^2 - 1.5 u.u)  (lbm/assit_func.h:2-9)
// 3 eu + 4.5 eu^2 - 1.5 u.u for direction k
template <class R> __device__ inline R poly_k(int k, R ux, R uy, R uu)
{
#pragma clang fp contract(off)
    R eu = (R)0.0; // e_k . u with e_k in {-1, 0, 1}^2: one rounding at most
    if (EX[k] == 1) eu = ux;
    if (EX[k] == -1) eu = -ux;
    if (EY[k] == 1) eu = (EX[k] == 0) ? uy : eu + uy;
    if (EY[k] == -1) eu = (EX[k] == 0) ? -uy : eu - uy;
    return fmar(eu, fmar((R)4.5, eu, (R)3.0), (R)-1.5 * uu);
}

template <bool SHIFT, class R> __device__ inline R qeq(int k, R rho, R drho, R ux, R uy, R uu)
{
#pragma clang fp contract(off)
    const R w = (R)W64[k], poly = poly_k<R>(k, ux, uy, uu);
    if (SHIFT) return w * fmar(rho, poly, drho);
    return w * fmar(rho, poly, rho);
}

// BGK on one node: q[9] in, Q[9] out, macros out (evolution_block.h:60-70, 100-111).
// Relaxation in the difference form Q = qa - omega (q - feq), qa the value q stands
// for in the output's domain (q itself, or q + w_k when q is the shifted copy of an
// unshifted input): q - feq is formed (in one fused multiply-add) before the scaling,
// so the result carries one rounding of qa's size.
template <bool SHIFT, class R>
__device__ inline void collide(const R (&q)[9], R omega, R (&Q)[9], R &rho, R &ux, R &uy, const R (&qa)[9])
{
#pragma clang fp contract(off)
    // moments from shared partial sums (14 additions): east/west and north/south movers
    const R e3 = q[1] + q[5] + q[8], w3 = q[3] + q[6] + q[7];
    const R n3 = q[2] + q[5] + q[6], s3 = q[4] + q[7] + q[8];
    const R jx = e3 - w3, jy = n3 - s3;
    const R s = ((q[0] + q[2]) + q[4]) + (e3 + w3);
    R drho = SHIFT ? s : s - (R)1.0;
    rho = SHIFT ? (R)1.0 + s : s;
    const R inv = recip(rho);
    ux = jx * inv;
    uy = jy * inv;
    const R uu = fmar(ux, ux, uy * uy);
#pragma unroll
    for (int k = 0; k < 9; k++) {
        // feq = w_k t (qeq's form); q - feq is one fused multiply-add, one rounding
        const R t = fmar(rho, poly_k<R>(k, ux, uy, uu), SHIFT ? drho : rho);
        Q[k] = fmar(-omega, fmar(-(R)W64[k], t, q[k]), qa[k]);
    }
}

// The collision of one node for codec C. Binary32 planes holding f itself are
// collided on the deviations f - w_k (exact while f lies within [w_k / 2, 2 w_k],
// Sterbenz, which the flow speeds of the method keep; outside it, one rounding of
// the small result), so the moments and q - feq carry no rounding of f's size; the
// relaxation lands on the unshifted f with one fused multiply-add. On the two-level
// cylinder channel at Re = 20 (20,000 coarse steps) this takes the drag coefficient's
// difference from an FP64 run of the same case from 1.1e-5 to 5.5e-6 (relative).
// Shifted planes and the FP64 build collide their loaded values as they are.
template <class C, class R = typename C::R>
__device__ inline void collide_c(const R (&q)[9], R omega, R (&Q)[9], R &rho, R &ux, R &uy)
{
    if constexpr (C::SHIFT || sizeof(R) != 4) {
        collide<C::SHIFT, R>(q, omega, Q, rho, ux, uy, q);
    } else {
        R qs[9];
#pragma unroll
        for (int k = 0; k < 9; k++) qs[k] = q[k] - (R)W64[k];
        collide<true, R>(qs, omega, Q, rho, ux, uy, q);
    }
}

template <bool SHIFT, class R> __device__ inline void equilibrium(R rho, R ux, R uy, R (&Q)[9])
{
#pragma clang fp contract(off)
    const R uu = fmar(ux, ux, uy * uy);
#pragma unroll
    for (int k = 0; k < 9; k++) Q[k] = qeq<SHIFT, R>(k, rho, rho - (R)1.0, ux, uy, uu);
}

// Bijective XCD-aware tile order: hardware workgroup b runs on XCD b % 8;
// give each XCD one contiguous range of tiles so neighbouring tiles (which
// share cache lines at their x seams) meet in the same L2.
__device__ inline int xcd_tile(int b, int G)
{
    if (G < 8) return b;
    int q = G >> 3, r = G & 7;
    int x = b & 7, local = b >> 3;
    return x * q + (x < r ? x : r) + local;
}

template <class T, int VX> struct Vec {
    typedef T type __attribute__((ext_vector_type(VX)));
};

template <class T, int VX, class P> __device__ inline void vload(P p, T (&o)[VX])
{
    if constexpr (VX == 1) {
        o[0] = p[0];
    } else {
        typedef typename Vec<T, VX>::type V;
        typedef const GLOBAL_AS V GV;
        V v = *reinterpret_cast<GV *>(p);
#pragma unroll
        for (int i = 0; i < VX; i++) o[i] = v[i];
    }
}

// VX consecutive elements from an address aligned to one element only (a row
// above or below when the row length is not a multiple of VX); the hardware
// serves it with the same wide load.
template <class T, int VX> struct VecU {
    typedef T type __attribute__((ext_vector_type(VX), aligned(sizeof(T))));
};

template <class T, int VX, class P> __device__ inline void vload_u(P p, T (&o)[VX])
{
    if constexpr (VX == 1) {
        o[0] = p[0];
    } else {
        typedef typename VecU<T, VX>::type V;
        typedef const GLOBAL_AS V GV;
        V v = *reinterpret_cast<GV *>(p);
#pragma unroll
        for (int i = 0; i < VX; i++) o[i] = v[i];
    }
}

template <class T, int VX, class P> __device__ inline void vstore(P p, const T (&o)[VX])
{
    if constexpr (VX == 1) {
        __builtin_nontemporal_store(o[0], p);
    } else {
        typedef typename Vec<T, VX>::type V;
        V v;
#pragma unroll
        for (int i = 0; i < VX; i++) v[i] = o[i];
        typedef GLOBAL_AS V GV;
        __builtin_nontemporal_store(v, reinterpret_cast<GV *>(p));
    }
}

// ---------------------------------------------------------------------------
// General per-node update (boundary flags, block edges, outlet); scalar.
// Returns false when the node keeps its F (nothing written by the caller's
// rules besides the copy done here).
// ---------------------------------------------------------------------------
template <int MODE, bool PERIODIC>
__device__ inline void fluidlike(const BlockDesc &d, const typename Codec<MODE>::GT *const (&S)[9], int i, int j,
                                 const Consts &c, typename Codec<MODE>::R omega, typename Codec<MODE>::R (&Q)[9], typename Codec<MODE>::R &rho, typename Codec<MODE>::R &ux, typename Codec<MODE>::R &uy)
{
    typedef typename Codec<MODE>::R R;
    typedef Codec<MODE> C;
    const int nx = d.nx, ny = d.ny;
    bool edge = !PERIODIC && (i == 0 || j == 0 || i == nx - 1 || j == ny - 1);
    if (!edge) {
        R q[9];
#pragma unroll
        for (int k = 0; k < 9; k++) {
            int id = i - EX[k], jd = j - EY[k];
            if (PERIODIC) {
                id = id < 0 ? nx - 1 : (id >= nx ? 0 : id);
                jd = jd < 0 ? ny - 1 : (jd >= ny ? 0 : jd);
            }
            q[k] = C::ld(S[k][(size_t)jd * nx + id]);
        }
        collide_c<C>(q, omega, Q, rho, ux, uy);
    } else {
        // outer-edge Dirichlet equilibrium (evolution_block.h:47-57)
        rho = (R)c.rho0;
        ux = (R)c.U0x;
        uy = (R)c.U0y;
        equilibrium<C::SHIFT, R>(rho, ux, uy, Q);
    }
}

// error_L2's terms (error_L2.h:3-55) of one node from the macros just stored, and u_prev <- u:
// the partial sums of |u - u_prev|^2 and |u|^2 and the smallest density, in FP64.
struct ResAcc {
    double s1 = 0.0, s2 = 0.0, mn = INFINITY;
};

__device__ inline void resid_node(const BlockDesc &d, uint32_t o, Lbd r, Lbd ux, Lbd uy, Lbd px, Lbd py, ResAcc &a)
{
#pragma clang fp contract(off)
    const double x = (double)ux, y = (double)uy, x0 = (double)px, y0 = (double)py;
    a.mn = fmin(a.mn, (double)r);
    a.s1 += (x - x0) * (x - x0) + (y - y0) * (y - y0);
    a.s2 += x * x + y * y;
    *off32((GLOBAL_AS Lbd *)d.upx, o) = ux;
    *off32((GLOBAL_AS Lbd *)d.upy, o) = uy;
}

// mout (may be null): receives the macros the update wrote (rho, ux, uy); returns whether it wrote them.
template <int MODE, bool PERIODIC, bool FLAGS, bool MACROS>
__device__ inline bool general_node(const BlockDesc &d, const typename Codec<MODE>::GT *const (&S)[9],
                                    typename Codec<MODE>::GT *const (&D)[9], int i, int j, const Consts &c,
                                    typename Codec<MODE>::R omega, Lbd *mout = nullptr)
{
    typedef typename Codec<MODE>::R R;
    typedef Codec<MODE> C;
    typedef typename C::T T;
    const int nx = d.nx, ny = d.ny;
    const size_t ixy = (size_t)j * nx + i;
    const int flag = FLAGS ? d.flag[ixy] : FLUID;
    if (flag == SOLID_SBB || flag == MBL_SOLID_WALL) {
        // full-way bounce-back (evolution_block.h:113-117); pulls leaving the
        // block take the node's own value (they never reach a fluid node).
        T v[9];
#pragma unroll
        for (int k = 0; k < 9; k++) {
            int id = i - EX[k], jd = j - EY[k];
            if (PERIODIC) {
                id = id < 0 ? nx - 1 : (id >= nx ? 0 : id);
                jd = jd < 0 ? ny - 1 : (jd >= ny ? 0 : jd);
            } else if (id < 0 || id >= nx || jd < 0 || jd >= ny) {
                id = i;
                jd = j;
            }
            v[k] = S[k][(size_t)jd * nx + id];
        }
#pragma unroll
        for (int k = 0; k < 9; k++) D[k][ixy] = v[OPP[k]];
        return false;
    }
    if (flag == FLUID) {
        R Q[9], rho, ux, uy;
        fluidlike<MODE, PERIODIC>(d, S, i, j, c, omega, Q, rho, ux, uy);
#pragma unroll
        for (int k = 0; k < 9; k++) D[k][ixy] = C::st(Q[k]);
        if (MACROS) {
            d.rho[ixy] = (Lbd)rho;
            d.ux[ixy] = (Lbd)ux;
            d.uy[ixy] = (Lbd)uy;
            if (mout) {
                mout[0] = (Lbd)rho;
                mout[1] = (Lbd)ux;
                mout[2] = (Lbd)uy;
            }
        }
        return MACROS;
    }
    if (flag == BOUNDARY_INLET || flag == BOUNDARY_OUTLET) {
        // channel ends by non-equilibrium extrapolation (Guo, Zheng & Shi
        // 2002) from the post-collision state of the x+1 (inlet) or x-1
        // (outlet) neighbour, recomputed here: inlet imposes u_in(y) with the
        // neighbour's density, outlet imposes rho0 with the neighbour's velocity.
        const bool inlet = (flag == BOUNDARY_INLET);
        const size_t n = inlet ? ixy + 1 : ixy - 1;
        R QN[9], rN, uN, vN;
        if (d.flag[n] == C_BOUNDARY) {
#pragma unroll
            for (int k = 0; k < 9; k++) QN[k] = C::ld(S[k][n]);
            rN = (R)d.rho[n];
            uN = (R)d.ux[n];
            vN = (R)d.uy[n];
        } else {
            fluidlike<MODE, PERIODIC>(d, S, inlet ? i + 1 : i - 1, j, c, omega, QN, rN, uN, vN);
        }
        R rs, us, vs;
        if (inlet) {
            const double y = j * d.dx + d.base_y;
            rs = rN;
            us = (R)(4.0 * c.umax * (c.LY - 1.5 - y) * (y - 0.5) / (c.LY - 2.0) / (c.LY - 2.0));
            vs = (R)0.0;
        } else {
            rs = (R)c.rho0;
            us = uN;
            vs = vN;
        }
        const R uuN = uN * uN + vN * vN, uus = us * us + vs * vs;
#pragma unroll
        for (int k = 0; k < 9; k++) {
            R v = qeq<C::SHIFT, R>(k, rs, rs - (R)1.0, us, vs, uus) +
                       (QN[k] - qeq<C::SHIFT, R>(k, rN, rN - (R)1.0, uN, vN, uuN));
            D[k][ixy] = C::st(v);
        }
        if (MACROS) {
            d.rho[ixy] = (Lbd)rs;
            d.ux[ixy] = (Lbd)us;
            d.uy[ixy] = (Lbd)vs;
            if (mout) {
                mout[0] = (Lbd)rs;
                mout[1] = (Lbd)us;
                mout[2] = (Lbd)vs;
            }
        }
        return MACROS;
    }
    // C_BOUNDARY (and any other flag): F carried over unchanged
#pragma unroll
    for (int k = 0; k < 9; k++) D[k][ixy] = S[k][ixy];
    return false;
}

// One node of a flagged block in one instruction stream for every flag a block
// edge or body can carry, so narrow blocks (whose waves all hold edge lanes) do
// not run a second path: a lane pulls the 9 sources of a node, collides, and
// keeps the collided, the bounced-back or the copied values by select.
//  * ghost (C_BOUNDARY) nodes pull their own values (carried over);
//  * an inlet (outlet) node pulls for its x+1 (x-1) neighbour, whose
//    post-collision state its non-equilibrium extrapolation starts from (the
//    neighbour's own values when it is a ghost node);
//  * fluid nodes on the block edge, and an inlet/outlet whose neighbour is on
//    the edge, take the equilibrium of (rho0, U0) (evolution_block.h:47-57)
//    in place of the collision.
// Only waves holding an inlet/outlet or edge node run those parts, and no part
// holds more registers than the collision itself. Contraction is off in the
// whole update (the collision spells out its fused operations), so every
// instantiation rounds alike. RESID: with `res`, the nodes
// off the block's outer edge whose macros the update writes (fluid, inlet,
// outlet) also add their error_L2 terms to `acc` and set u_prev <- u (their
// previous velocity is loaded with the flags, off the dependent path); an
// interface op may rewrite the macros of an edge node later in the cycle, so
// edge nodes are summed after the cycle.
template <int MODE, bool MACROS, bool RESID = false>
__device__ inline void flagged_node(const BlockDesc &d, const typename Codec<MODE>::GT *const (&S)[9],
                                    typename Codec<MODE>::GT *const (&D)[9], int i, int j, uint32_t r0,
                                    const Consts &c, typename Codec<MODE>::R omega, bool res = false,
                                    ResAcc *acc = nullptr)
{
#pragma clang fp contract(off)
    typedef typename Codec<MODE>::R R;
    typedef Codec<MODE> C;
    typedef typename C::T T;
    const int nx = d.nx, ny = d.ny;
    const uint32_t unx = (uint32_t)nx;
    const uint32_t o = r0 + (uint32_t)i;
    // the node's flag and its x-neighbours' (same cache lines), in one round trip
    const int flag = d.flag[o];
    const int flm = d.flag[i > 0 ? o - 1 : o];
    const int flp = d.flag[i < nx - 1 ? o + 1 : o];
    const bool rnode = RESID && res && !(i == 0 || j == 0 || i == nx - 1 || j == ny - 1);
    Lbd px = (Lbd)0, py = (Lbd)0;
    if (rnode) {
        px = *off32((const GLOBAL_AS Lbd *)d.upx, o);
        py = *off32((const GLOBAL_AS Lbd *)d.upy, o);
    }
    const bool inlet = flag == BOUNDARY_INLET, outlet = flag == BOUNDARY_OUTLET, nee = inlet || outlet;
    const bool solid = flag == SOLID_SBB || flag == MBL_SOLID_WALL;
    // the node the pulls are for, and its flag
    const int xs = inlet ? i + 1 : (outlet ? i - 1 : i);
    const uint32_t os = r0 + (uint32_t)xs;
    const int sflag = inlet ? flp : (outlet ? flm : flag);
    const bool copy = nee ? (sflag == C_BOUNDARY) : !(flag == FLUID || solid);
    const bool edge_s = (xs == 0 || j == 0 || xs == nx - 1 || j == ny - 1);
    const bool eq = !copy && (nee || flag == FLUID) && edge_s;
    T raw[9];
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const int id = xs - EX[k], jd = j - EY[k];
        const bool own = copy || id < 0 || id >= nx || jd < 0 || jd >= ny;
        // row jd's offset from row j's (wraps only where own holds)
        const uint32_t rk = EY[k] == 1 ? r0 - unx : (EY[k] == -1 ? r0 + unx : r0);
        raw[k] = *off32(S[k], own ? os : rk + (uint32_t)id);
    }
    R q[9];
#pragma unroll
    for (int k = 0; k < 9; k++) q[k] = C::ld(raw[k]);
    R Q[9], rho, ux, uy;
    collide_c<C>(q, omega, Q, rho, ux, uy);
    if (__builtin_amdgcn_ballot_w64(eq)) {
        if (eq) {
            rho = (R)c.rho0;
            ux = (R)c.U0x;
            uy = (R)c.U0y;
            equilibrium<C::SHIFT, R>(rho, ux, uy, Q);
        }
    }
    if (__builtin_amdgcn_ballot_w64(nee)) {
        if (nee) {
            // channel ends by non-equilibrium extrapolation (Guo, Zheng & Shi 2002) from the
            // post-collision state of the neighbour: the inlet imposes u_in(y) with the
            // neighbour's density, the outlet rho0 with the neighbour's velocity
            R rN = rho, uN = ux, vN = uy;
            if (copy) {
                rN = (R)d.rho[os];
                uN = (R)d.ux[os];
                vN = (R)d.uy[os];
#pragma unroll
                for (int k = 0; k < 9; k++) Q[k] = q[k];
            }
            R rs, us, vs;
            if (inlet) {
                const double y = j * d.dx + d.base_y;
                rs = rN;
                us = (R)(4.0 * c.umax * (c.LY - 1.5 - y) * (y - 0.5) / (c.LY - 2.0) / (c.LY - 2.0));
                vs = (R)0.0;
            } else {
                rs = (R)c.rho0;
                us = uN;
                vs = vN;
            }
            const R uuN = uN * uN + vN * vN, uus = us * us + vs * vs;
#pragma unroll
            for (int k = 0; k < 9; k++)
                Q[k] = qeq<C::SHIFT, R>(k, rs, rs - (R)1.0, us, vs, uus) +
                       (Q[k] - qeq<C::SHIFT, R>(k, rN, rN - (R)1.0, uN, vN, uuN));
            rho = rs;
            ux = us;
            uy = vs;
        }
    }
    const bool upd = flag == FLUID || nee;
    // plain stores: the planes of a multi-block layout stay in the L2 and the last-level cache
    // for the launches after this one (non-temporal stores measured 18 % slower on the
    // three-level case, 11 % on the 256-case sweep)
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const T v = upd ? C::st(Q[k]) : (solid ? raw[OPP[k]] : raw[k]);
        *off32(D[k], o) = v;
    }
    if (MACROS && upd) {
        *off32((GLOBAL_AS Lbd *)d.rho, o) = (Lbd)rho;
        *off32((GLOBAL_AS Lbd *)d.ux, o) = (Lbd)ux;
        *off32((GLOBAL_AS Lbd *)d.uy, o) = (Lbd)uy;
        if (rnode) resid_node(d, o, (Lbd)rho, (Lbd)ux, (Lbd)uy, px, py, *acc);
    }
}

// ---------------------------------------------------------------------------
// Fused stream-collide. One wave covers 64*VX consecutive nodes of one row
// (or two rows of 32*VX); a workgroup covers tile_rows rows. Interior
// all-fluid runs use VX-wide loads of each source row plus one scalar for the
// x-shifted planes; every other node takes the flagged per-node update.
// ---------------------------------------------------------------------------
// One row segment of a tile: 64*VX consecutive nodes of row j (element offset
// r0 = j * nx in the planes) starting at the wave's first node.
template <int MODE, int VX, bool PERIODIC, bool FLAGS, bool MACROS, bool RESID = false>
__device__ __forceinline__ void sc_row(const BlockDesc &d, uint32_t parity_mask, int j, uint32_t r0, int x0,
                                       bool res = false, ResAcc *acc = nullptr)
{
    static_assert(!RESID || (FLAGS && MACROS && !PERIODIC && VX == 1), "residual terms: flagged path only");
    typedef typename Codec<MODE>::R R;
    typedef Codec<MODE> C;
    typedef typename C::T T;
    const Consts &c = d.bc;
    const int nx = d.nx, ny = d.ny;
    if (j >= ny || x0 >= nx) return;
    const int par = (parity_mask >> d.level) & 1;
    typedef typename C::GT GT;
    const GT *S[9];
    GT *D[9];
#pragma unroll
    for (int k = 0; k < 9; k++) {
        S[k] = (const GT *)d.P[par][k];
        D[k] = (GT *)d.P[par ^ 1][k];
    }
    const R omega = (R)d.omega;

    if constexpr (FLAGS && !PERIODIC && VX == 1) {
        flagged_node<MODE, MACROS, RESID>(d, S, D, x0, j, r0, c, omega, res, acc);
        return;
    }

    bool fast;
    if (PERIODIC)
        fast = true;
    else
        fast = (x0 >= 1) && (x0 + VX <= nx - 1) && (j >= 1) && (j <= ny - 2);
    if (FLAGS && fast) {
        int fl[VX];
        vload<int, VX>(off32((const GLOBAL_AS int *)d.flag, r0 + (uint32_t)x0), fl);
#pragma unroll
        for (int v = 0; v < VX; v++) fast = fast && (fl[v] == FLUID);
    }
    if (!fast) {
        // block edges, non-fluid nodes and the partial segments at the row ends, node by node
#pragma unroll 1
        for (int v = 0; v < VX; v++) {
            const int i = x0 + v;
            if (i < 0 || i >= nx) continue;
            if constexpr (FLAGS && !PERIODIC)
                flagged_node<MODE, MACROS>(d, S, D, i, j, r0, c, omega);
            else
                general_node<MODE, PERIODIC, FLAGS, MACROS>(d, S, D, i, j, c, omega);
        }
        return;
    }

    // rows j -1 and j + 1 (periodic blocks wrap)
    const uint32_t unx = (uint32_t)nx;
    uint32_t rm = r0 - unx, rp = r0 + unx;
    int xm = x0 - 1, xp = x0 + VX;
    if (PERIODIC) {
        rm = (j == 0) ? (uint32_t)(ny - 1) * unx : rm;
        rp = (j == ny - 1) ? 0u : rp;
        xm = (x0 == 0) ? nx - 1 : xm;
        xp = (xp >= nx) ? 0 : xp;
    }
    const uint32_t rc = r0;
    R q[9][VX];
#pragma unroll
    for (int k = 0; k < 9; k++) {
        // pull: q_k(x) = F_k(x - e_k); EY = +1 reads row j-1, EY = -1 row j+1. The
        // segment is aligned in its own row; in the rows above and below it is
        // aligned only when the row length is a multiple of VX (always so in
        // periodic blocks), else it is read with element alignment.
        const uint32_t rb = (EY[k] == 1) ? rm : ((EY[k] == -1) ? rp : rc);
        const bool al = PERIODIC || EY[k] == 0;
        T tv[VX];
        if (EX[k] == 0 || VX > 1) {
            if (al)
                vload<T, VX>(off32(S[k], rb + x0), tv);
            else
                vload_u<T, VX>(off32(S[k], rb + x0), tv);
        }
        if (EX[k] == 0) {
#pragma unroll
            for (int v = 0; v < VX; v++) q[k][v] = C::ld(tv[v]);
        } else if (EX[k] == 1) {
            q[k][0] = C::ld(*off32(S[k], rb + xm));
#pragma unroll
            for (int v = 1; v < VX; v++) q[k][v] = C::ld(tv[v - 1]);
        } else {
            q[k][VX - 1] = C::ld(*off32(S[k], rb + xp));
#pragma unroll
            for (int v = 0; v < VX - 1; v++) q[k][v] = C::ld(tv[v + 1]);
        }
    }
    T out[9][VX];
    Lbd mr[VX], mx[VX], my[VX];
#pragma unroll
    for (int v = 0; v < VX; v++) {
        R qq[9], Q[9], rho, ux, uy;
#pragma unroll
        for (int k = 0; k < 9; k++) qq[k] = q[k][v];
        collide_c<C>(qq, omega, Q, rho, ux, uy);
#pragma unroll
        for (int k = 0; k < 9; k++) out[k][v] = C::st(Q[k]);
        mr[v] = (Lbd)rho;
        mx[v] = (Lbd)ux;
        my[v] = (Lbd)uy;
    }
#pragma unroll
    for (int k = 0; k < 9; k++) vstore<T, VX>(off32(D[k], rc + x0), out[k]);
    if (MACROS) {
        vstore<Lbd, VX>(off32((GLOBAL_AS Lbd *)d.rho, rc + x0), mr);
        vstore<Lbd, VX>(off32((GLOBAL_AS Lbd *)d.ux, rc + x0), mx);
        vstore<Lbd, VX>(off32((GLOBAL_AS Lbd *)d.uy, rc + x0), my);
    }
}

// One tile: tw*VX columns x rows rows of one block; each wave takes 64/tw
// rows at a time, the workgroup's waves the next ones in turn. The block comes from the per-tile table when
// there is one (one scalar load), else from a search of tile_begin.
// wv: the wave's index among the nwv waves that take the tile.
// The tile's residual sums (weighted by the block's 1/m, as k_residual_tiles weights its
// tiles) into out[0..2]: sum |u - u_prev|^2 / m, sum |u|^2 / m, smallest density. Every thread
// of the workgroup calls it.
__device__ inline double wave_sum(double v);

__device__ inline double wave_min(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ inline void tile_resid_store(const ResAcc &a, double inv_m, double *out)
{
    __shared__ double rsh[3 * 16];
    const double s1 = wave_sum(a.s1), s2 = wave_sum(a.s2), mn = wave_min(a.mn);
    const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads(); // the previous tile's reads of rsh are done
    if ((threadIdx.x & 63) == 0) {
        rsh[3 * w] = s1;
        rsh[3 * w + 1] = s2;
        rsh[3 * w + 2] = mn;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t1 = 0.0, t2 = 0.0, tm = INFINITY;
        for (int q = 0; q < nw; q++) {
            t1 += rsh[3 * q];
            t2 += rsh[3 * q + 1];
            tm = fmin(tm, rsh[3 * q + 2]);
        }
        out[0] = t1 * inv_m;
        out[1] = t2 * inv_m;
        out[2] = tm;
    }
}

template <int MODE, int VX, bool PERIODIC, bool FLAGS, bool MACROS, bool RESID = false>
__device__ __forceinline__ void sc_tile(const BlockDesc *__restrict__ descs, const int *__restrict__ tile_blk, int ndesc,
                               int tile_rows, uint32_t parity_mask, int t, int wv, int nwv, uint32_t rmask = 0,
                               double *rparts = nullptr)
{
    int lo = 0;
    if (tile_blk) {
        lo = tile_blk[t];
    } else {
        int hi = ndesc - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (descs[mid].tile_begin <= t)
                lo = mid;
            else
                hi = mid - 1;
        }
    }
    const BlockDesc &d = descs[lo];
    // the tile's column and row in its block: lt / tiles_x is the multiply-high by the host's
    // 0xFFFFFFFF / tiles_x (the quotient or one less for lt < 2^31) and one correction, in
    // scalar registers
    const uint32_t lt = (uint32_t)(t - d.tile_begin), ntx = (uint32_t)d.tiles_x;
    uint32_t tyi = __umulhi(lt, d.tx_rcp), txi = lt - tyi * ntx;
    if (txi >= ntx) {
        tyi++;
        txi -= ntx;
    }
    // a wave covers one row segment of 64 lanes or two of 32 (tw is 64 or 32; always 64 in
    // periodic blocks, so there a wave's row and row offsets are scalar)
    const int tw = PERIODIC ? 64 : d.tw, lane = (int)(threadIdx.x & 63);
    const bool two = !PERIODIC && (tw == 32);
    const int per_wave = two ? 2 : 1, sub = two ? (lane >> 5) : 0, col = lane & (tw - 1);
    const int base = ((int)txi * tw + col) * VX;
    const int rows = d.trows > 0 ? d.trows : tile_rows;
    const int step = nwv * per_wave;
    const uint32_t unx = (uint32_t)d.nx;
    // residual terms of this tile's nodes when its level's last collide of the cycle is this one
    const bool res = RESID && ((rmask >> d.level) & 1u);
    ResAcc acc;
    // the wave's first row rw and its row offset are scalar; a lane adds its segment's row (sub)
    for (int rw = __builtin_amdgcn_readfirstlane(wv) * per_wave; rw < rows; rw += step) {
        if (rw + sub >= rows) break;
        const int jw = (int)tyi * rows + rw;
        const uint32_t r0 = (uint32_t)jw * unx + (sub ? unx : 0u);
        // Segments start at multiples of VX of the plane index, so a row of a
        // length that is not a multiple of VX begins (j*nx mod VX) nodes into
        // its first segment; the host sizes tiles_x for that shift.
        const int x0 = PERIODIC ? base : base - (int)(r0 & (uint32_t)(VX - 1));
        sc_row<MODE, VX, PERIODIC, FLAGS, MACROS, RESID>(d, parity_mask, jw + sub, r0, x0, res, &acc);
    }
    if constexpr (RESID) {
        if (res) tile_resid_store(acc, d.inv_m, rparts + 3 * (size_t)t);
    }
}

// One tile per workgroup when the grid covers every tile; otherwise a
// persistent grid in which workgroup b (after the XCD-aware reordering) walks
// one contiguous chunk of tiles, so x-neighbouring tiles stay on one XCD.
// RESID: the tiles of the levels in rmask also write their residual sums to rparts[3 t ..]
// (t: the tile's number across all levels).
template <int MODE, int VX, bool PERIODIC, bool FLAGS, bool MACROS, bool RESID = false>
__global__ __launch_bounds__(256) void k_stream_collide(const BlockDesc *__restrict__ descs,
                                                        const int *__restrict__ tile_blk, int ndesc, int tile_rows,
                                                        uint32_t parity_mask, int total_tiles, int tile_base,
                                                        uint32_t rmask, double *rparts)
{
    const int G = (int)gridDim.x;
    const int b = xcd_tile((int)blockIdx.x, G);
    int t0 = b, t1 = min(b + 1, total_tiles);
    if (G < total_tiles) {
        const int per = (total_tiles + G - 1) / G;
        t0 = b * per;
        t1 = min(total_tiles, t0 + per);
    }
    // the launch covers tiles [tile_base, tile_base + total_tiles) of the descriptors' numbering
    t0 += tile_base;
    t1 += tile_base;
    // one call site, inlined: the descriptor and plane bases stay in scalar registers
    for (int t = t0; t < t1; t++)
        sc_tile<MODE, VX, PERIODIC, FLAGS, MACROS, RESID>(descs, tile_blk, ndesc, tile_rows, parity_mask, t,
                                                          (int)(threadIdx.x >> 6), (int)(blockDim.x >> 6), rmask,
                                                          rparts);
}

// ---------------------------------------------------------------------------
// Two time steps of a periodic all-fluid block in one pass (temporal blocking). A workgroup
// owns an output tile of P2_TX x TY nodes: it collides step n+1 on the tile grown by one node
// on every side (pulling step n from the planes, periodic wrap), keeps those post-collision
// values in LDS as the planes would hold them (C::st), then collides step n+2 on the tile
// pulling from LDS and writes it to the other plane set. Every node update is the single-step
// kernel's (the same loads, collide_c and stores, in the same arithmetic), so two single steps
// and one paired launch give the same bits in the plane set that holds the state; the
// intermediate step n+1 is not written (the other plane set keeps step n). Per two node
// updates it moves 36 B read + 36 B written (FP32) instead of twice that; the step-n+1 ring is
// the overhead: (P2_TX + 2)(TY + 2) collisions per P2_TX TY outputs.
// ---------------------------------------------------------------------------
constexpr int P2_TX = 64;

// Waves per SIMD the kernel asks for: three workgroups per CU where the LDS holds three (the
// registers are capped to fit: 512 threads, FP32 planes: 64 VGPRs, no spill; 16384^2 132k-137k
// MLUPS against 121k-123k at the 110 VGPRs the compiler takes unbounded, two workgroups per CU).
constexpr int p2_waves(int threads, int ty, int elem_bytes)
{
    return 3 * 9 * (ty + 2) * (P2_TX + 2) * elem_bytes <= 160 * 1024 ? 3 * threads / 256 : 1;
}

template <int MODE, int TY, int P2_THREADS>
__global__ __launch_bounds__(P2_THREADS, p2_waves(P2_THREADS, TY, (int)sizeof(typename Codec<MODE>::T)))
void k_periodic2(Periodic2Args a)
{
    typedef Codec<MODE> C;
    typedef typename C::T T;
    typedef typename C::R R;
    typedef typename C::GT GT;
    constexpr int AX = P2_TX + 2, AY = TY + 2;
    __shared__ T L[9][AY][AX];
    const int nx = a.nx, ny = a.ny;
    const uint32_t unx = (uint32_t)nx;
    const int b = xcd_tile((int)blockIdx.x, (int)gridDim.x);
    const int ty = b / a.tiles_x, tx = b - ty * a.tiles_x;
    const int x0 = tx * P2_TX, y0 = ty * TY;
    const R omega = (R)a.omega;
    // step n+1 on the grown tile. 16-bit planes: every load of the thread's nodes first (one
    // memory round trip for all of them), then their collisions (the FP32-arithmetic update is
    // short); 4- and 8-byte planes: node by node (their FP64 collision holds the registers;
    // loads-first measured 97k against 117k-131k MLUPS on 16384^2 FP32, 172k against 163k on
    // 1-3-12)
    constexpr int NA = (AX * AY + P2_THREADS - 1) / P2_THREADS;
    if constexpr (sizeof(T) > 2) {
        for (int idx = (int)threadIdx.x; idx < AX * AY; idx += P2_THREADS) {
            const int ly = idx / AX, lx = idx - ly * AX;
            int x = x0 - 1 + lx, y = y0 - 1 + ly;
            x = x < 0 ? x + nx : (x >= nx ? x - nx : x);
            y = y < 0 ? y + ny : (y >= ny ? y - ny : y);
            R q[9];
#pragma unroll
            for (int k = 0; k < 9; k++) {
                int xs = x - EX[k], ys = y - EY[k];
                xs = xs < 0 ? xs + nx : (xs >= nx ? xs - nx : xs);
                ys = ys < 0 ? ys + ny : (ys >= ny ? ys - ny : ys);
                q[k] = C::ld(*off32((const GT *)a.S[k], (uint32_t)ys * unx + (uint32_t)xs));
            }
            R Q[9], rho, ux, uy;
            collide_c<C>(q, omega, Q, rho, ux, uy);
#pragma unroll
            for (int k = 0; k < 9; k++) L[k][ly][lx] = C::st(Q[k]);
        }
    } else {
        T raw[NA][9];
#pragma unroll
        for (int it = 0; it < NA; it++) {
            const int idx = (int)threadIdx.x + it * P2_THREADS;
            if (idx >= AX * AY) continue;
            const int ly = idx / AX, lx = idx - ly * AX;
            int x = x0 - 1 + lx, y = y0 - 1 + ly;
            x = x < 0 ? x + nx : (x >= nx ? x - nx : x);
            y = y < 0 ? y + ny : (y >= ny ? y - ny : y);
#pragma unroll
            for (int k = 0; k < 9; k++) {
                int xs = x - EX[k], ys = y - EY[k];
                xs = xs < 0 ? xs + nx : (xs >= nx ? xs - nx : xs);
                ys = ys < 0 ? ys + ny : (ys >= ny ? ys - ny : ys);
                raw[it][k] = *off32((const GT *)a.S[k], (uint32_t)ys * unx + (uint32_t)xs);
            }
        }
#pragma unroll
        for (int it = 0; it < NA; it++) {
            const int idx = (int)threadIdx.x + it * P2_THREADS;
            if (idx >= AX * AY) continue;
            const int ly = idx / AX, lx = idx - ly * AX;
            R q[9];
#pragma unroll
            for (int k = 0; k < 9; k++) q[k] = C::ld(raw[it][k]);
            R Q[9], rho, ux, uy;
            collide_c<C>(q, omega, Q, rho, ux, uy);
#pragma unroll
            for (int k = 0; k < 9; k++) L[k][ly][lx] = C::st(Q[k]);
        }
    }
    __syncthreads();
    // step n+2 on the tile, pulling step n+1 from LDS
    for (int idx = (int)threadIdx.x; idx < P2_TX * TY; idx += P2_THREADS) {
        const int ly = idx / P2_TX, lx = idx - ly * P2_TX;
        R q[9];
#pragma unroll
        for (int k = 0; k < 9; k++) q[k] = C::ld(L[k][ly + 1 - EY[k]][lx + 1 - EX[k]]);
        R Q[9], rho, ux, uy;
        collide_c<C>(q, omega, Q, rho, ux, uy);
        const uint32_t o = (uint32_t)(y0 + ly) * unx + (uint32_t)(x0 + lx);
#pragma unroll
        for (int k = 0; k < 9; k++) __builtin_nontemporal_store(C::st(Q[k]), off32((GT *)a.D[k], o));
    }
}

// ---------------------------------------------------------------------------
// Three time steps per launch on a periodic all-fluid block (temporal blocking
// two deep): a workgroup owns a P2_TX x TY output tile. It collides step n+1
// on the tile and a two-node ring ((P2_TX + 4) x (TY + 4) nodes, pulling step
// n from the planes with the periodic wrap) into LDS, then step n+2 on the
// tile and a one-node ring from there, then step n+3 on the tile, which it
// writes to the other plane set. Step n+2 is computed in chunks of rows, each
// held in registers until every thread has read the rows it replaces: it lands
// in the rows of step n+1 that no later chunk reads, so one LDS buffer serves
// both steps (49 KB for FP32 planes: three workgroups per CU). Every node
// update is the single-step kernel's (same loads, collide_c, stores), so the
// state equals three single steps bit for bit; the plane set handed back as f
// holds step n. Per launch and output node: 9 loads of step n and 9 stores of
// step n+3 (72 B FP32 for three updates); the rings are recomputed:
// (P2_TX + 4)(TY + 4) + (P2_TX + 2)(TY + 2) extra collisions per P2_TX TY outputs.
// ---------------------------------------------------------------------------
constexpr int P3_THREADS = 512;

// waves per SIMD asked for: as many workgroups per CU as the LDS holds, at most three
constexpr int p3_waves(int ty, int elem_bytes)
{
    return (9 * (ty + 4) * (P2_TX + 4) * elem_bytes * 3 <= 160 * 1024   ? 3
            : 9 * (ty + 4) * (P2_TX + 4) * elem_bytes * 2 <= 160 * 1024 ? 2
                                                                        : 1) *
           P3_THREADS / 256;
}

template <int MODE, int TY>
__global__ __launch_bounds__(P3_THREADS, p3_waves(TY, (int)sizeof(typename Codec<MODE>::T)))
void k_periodic3(Periodic2Args a)
{
    typedef Codec<MODE> C;
    typedef typename C::T T;
    typedef typename C::R R;
    typedef typename C::GT GT;
    constexpr int AX = P2_TX + 4, AY = TY + 4; // step n+1 region: the tile and a two-node ring
    constexpr int BX = P2_TX + 2, BY = TY + 2; // step n+2 region: the tile and a one-node ring
    constexpr int BR = P3_THREADS / BX;        // step n+2 rows per chunk (one node per thread)
    static_assert(BR >= 1, "a chunk holds at least one row");
    // L[k][r][c]: step n+1 at node (x0 - 2 + c, y0 - 2 + r); step n+2 of region row j (node row
    // y0 - 1 + j), region column i (node column x0 - 1 + i) replaces it at L[k][j][i + 1]
    __shared__ T L[9][AY][AX];
    const int nx = a.nx, ny = a.ny;
    const uint32_t unx = (uint32_t)nx;
    const int b = xcd_tile((int)blockIdx.x, (int)gridDim.x);
    const int ty = b / a.tiles_x, tx = b - ty * a.tiles_x;
    const int x0 = tx * P2_TX, y0 = ty * TY;
    const R omega = (R)a.omega;
    // step n+1 on the tile and its two-node ring, node by node
    for (int idx = (int)threadIdx.x; idx < AX * AY; idx += P3_THREADS) {
        const int ly = idx / AX, lx = idx - ly * AX;
        int x = x0 - 2 + lx, y = y0 - 2 + ly;
        x = x < 0 ? x + nx : (x >= nx ? x - nx : x);
        y = y < 0 ? y + ny : (y >= ny ? y - ny : y);
        R q[9];
#pragma unroll
        for (int k = 0; k < 9; k++) {
            int xs = x - EX[k], ys = y - EY[k];
            xs = xs < 0 ? xs + nx : (xs >= nx ? xs - nx : xs);
            ys = ys < 0 ? ys + ny : (ys >= ny ? ys - ny : ys);
            q[k] = C::ld(*off32((const GT *)a.S[k], (uint32_t)ys * unx + (uint32_t)xs));
        }
        R Q[9], rho, ux, uy;
        collide_c<C>(q, omega, Q, rho, ux, uy);
#pragma unroll
        for (int k = 0; k < 9; k++) L[k][ly][lx] = C::st(Q[k]);
    }
    // step n+2 in chunks of BR region rows: chunk rows j0 .. j1-1 read step n+1 rows j0 .. j1+1
    // and replace rows j0 .. j1-1, which later chunks (rows from j1 on) do not read
    for (int j0 = 0; j0 < BY; j0 += BR) {
        const int j1 = j0 + BR < BY ? j0 + BR : BY;
        __syncthreads(); // step n+1 rows complete; the previous chunk's rows written
        const int idx = (int)threadIdx.x;
        const bool mine = idx < (j1 - j0) * BX;
        const int j = j0 + idx / BX, i = idx - (idx / BX) * BX;
        T Qs[9];
        if (mine) {
            R q[9];
#pragma unroll
            for (int k = 0; k < 9; k++) q[k] = C::ld(L[k][j + 1 - EY[k]][i + 1 - EX[k]]);
            R Q[9], rho, ux, uy;
            collide_c<C>(q, omega, Q, rho, ux, uy);
#pragma unroll
            for (int k = 0; k < 9; k++) Qs[k] = C::st(Q[k]);
        }
        __syncthreads(); // every read of rows j0 .. j1+1 is done
        if (mine) {
#pragma unroll
            for (int k = 0; k < 9; k++) L[k][j][i + 1] = Qs[k];
        }
    }
    __syncthreads();
    // step n+3 on the tile, pulling step n+2 (region row ly + 1 - ey, column lx + 1 - ex)
    for (int idx = (int)threadIdx.x; idx < P2_TX * TY; idx += P3_THREADS) {
        const int ly = idx / P2_TX, lx = idx - ly * P2_TX;
        R q[9];
#pragma unroll
        for (int k = 0; k < 9; k++) q[k] = C::ld(L[k][ly + 1 - EY[k]][lx + 2 - EX[k]]);
        R Q[9], rho, ux, uy;
        collide_c<C>(q, omega, Q, rho, ux, uy);
        const uint32_t o = (uint32_t)(y0 + ly) * unx + (uint32_t)(x0 + lx);
#pragma unroll
        for (int k = 0; k < 9; k++) __builtin_nontemporal_store(C::st(Q[k]), off32((GT *)a.D[k], o));
    }
}

// Two time steps per launch, strip form: a workgroup of P2S_W waves walks down a strip of
// 64 P2S_W - 2 output columns and P2S_H rows. Lane c (64 P2S_W of them) collides step n+1 at
// column x0 - 1 + c of one row at a time (pulling step n from the planes) into a 3-row window
// in LDS, and step n+2 of the row above from that window; the step-n loads of the next row are
// issued before the window is read, so they are in flight meanwhile. The ring of recomputed
// step-n+1 nodes is two columns per strip and two rows per strip height (against a one-node
// ring around every 64 x 16 tile), and the window is 3 rows of LDS, so more workgroups share a CU.
// Same node updates as the single-step kernel, so the same bits. Measured on 16384^2 (same box,
// two rounds): FP32 planes 107k MLUPS against 123k-133k for the tiles (two barriers per row
// serialize the FP64 collisions), 1-3-12 planes 174k-175k against 170k; the tiles are the default.
constexpr int P2S_W = 4;   // waves per workgroup
constexpr int P2S_H = 64;  // output rows per strip

template <int MODE>
__global__ __launch_bounds__(64 * P2S_W) void k_periodic2s(Periodic2Args a)
{
    typedef Codec<MODE> C;
    typedef typename C::T T;
    typedef typename C::R R;
    typedef typename C::GT GT;
    constexpr int NC = 64 * P2S_W, OW = NC - 2;
    __shared__ T Wn[3][9][NC];
    const int nx = a.nx, ny = a.ny;
    const uint32_t unx = (uint32_t)nx;
    const int b = xcd_tile((int)blockIdx.x, (int)gridDim.x);
    const int ty = b / a.tiles_x, tx = b - ty * a.tiles_x;
    const int x0 = tx * OW, y0 = ty * P2S_H;
    const int c = (int)threadIdx.x;
    int x = x0 - 1 + c;
    const bool out = c >= 1 && c <= OW && x < nx; // this lane's column is an output of the strip
    x = x < 0 ? x + nx : (x >= nx ? x - nx : x);
    int xs[9];
#pragma unroll
    for (int k = 0; k < 9; k++) {
        int v = x - EX[k];
        xs[k] = v < 0 ? v + nx : (v >= nx ? v - nx : v);
    }
    const R omega = (R)a.omega;
    // step-n values this lane pulls for step n+1 at row y (periodic rows)
    auto pull = [&](int y, T (&v)[9]) {
#pragma unroll
        for (int k = 0; k < 9; k++) {
            int ys = y - EY[k];
            ys = ys < 0 ? ys + ny : (ys >= ny ? ys - ny : ys);
            v[k] = *off32((const GT *)a.S[k], (uint32_t)ys * unx + (uint32_t)xs[k]);
        }
    };
    auto step1 = [&](const T (&v)[9], int slot) {
        R q[9], Q[9], rho, ux, uy;
#pragma unroll
        for (int k = 0; k < 9; k++) q[k] = C::ld(v[k]);
        collide_c<C>(q, omega, Q, rho, ux, uy);
#pragma unroll
        for (int k = 0; k < 9; k++) Wn[slot][k][c] = C::st(Q[k]);
    };
    // step n+1 at rows y0 - 1 and y0 (window slots 0, 1), the loads of row y0 + 1 in flight
    T v[9], nv[9];
    pull(y0 - 1 < 0 ? y0 - 1 + ny : y0 - 1, v);
    pull(y0, nv);
    step1(v, 0);
#pragma unroll
    for (int k = 0; k < 9; k++) v[k] = nv[k];
    pull(y0 + 1 >= ny ? y0 + 1 - ny : y0 + 1, nv);
    step1(v, 1);
    for (int i = 0; i < P2S_H; i++) {
        // step n+1 at row y0 + i + 1 into slot (i + 2) % 3; the loads of the row after it issued
#pragma unroll
        for (int k = 0; k < 9; k++) v[k] = nv[k];
        if (i + 1 < P2S_H) {
            int yn = y0 + i + 2;
            yn = yn >= ny ? yn - ny : yn;
            pull(yn, nv);
        }
        step1(v, (i + 2) % 3);
        __syncthreads();
        // step n+2 at row y0 + i from the window rows y0 + i - 1, y0 + i, y0 + i + 1
        if (out) {
            R q[9], Q[9], rho, ux, uy;
#pragma unroll
            for (int k = 0; k < 9; k++) q[k] = C::ld(Wn[(i + 1 - EY[k]) % 3][k][c - EX[k]]);
            collide_c<C>(q, omega, Q, rho, ux, uy);
            const uint32_t o = (uint32_t)(y0 + i) * unx + (uint32_t)x;
#pragma unroll
            for (int k = 0; k < 9; k++) __builtin_nontemporal_store(C::st(Q[k]), off32((GT *)a.D[k], o));
        }
        __syncthreads(); // the window's oldest row is overwritten next
    }
}

// ---------------------------------------------------------------------------
// init_General (lbm/init_block.h:2-51)
// ---------------------------------------------------------------------------
__device__ inline uint64_t splitmix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// counter-based normal pair: SplitMix64 finaliser on (seed, i, j), Box-Muller
__device__ inline void random_velocity(uint64_t seed, double sigma, int i, int j, double &ux, double &uy)
{
    uint64_t key = ((uint64_t)(uint32_t)j << 32) | (uint64_t)(uint32_t)i;
    uint64_t a = splitmix64(seed + (2 * key + 1) * 0x9E3779B97F4A7C15ULL);
    uint64_t b = splitmix64(seed + (2 * key + 2) * 0x9E3779B97F4A7C15ULL);
    double u1 = (double)((a >> 11) + 1) * (1.0 / 9007199254740992.0);
    double u2 = (double)(b >> 11) * (1.0 / 9007199254740992.0);
    double r = sigma * sqrt(-2.0 * log(u1));
    double th = 6.283185307179586 * u2;
    ux = r * cos(th);
    uy = r * sin(th);
}

struct InitArgs {
    BlockDesc d;
    Lbd *x, *y, *fx, *fy;
    int bc_mode, init_mode;
    unsigned long long seed;
    double sigma, LX, LY, R, CX, CY, base_x;
    int m;
    Consts c;
};

template <int MODE> __global__ __launch_bounds__(256) void k_init_general(InitArgs a)
{
    typedef Codec<MODE> C;
    typedef typename C::T T;
    const BlockDesc &d = a.d;
    const int i = blockIdx.x * 16 + (threadIdx.x & 15);
    const int j = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (i >= d.nx || j >= d.ny) return;
    const size_t ixy = (size_t)j * d.nx + i;
    const double dx = 1.0 / a.m;
    const double X = i * dx + a.base_x;
    const double Y = j * dx + d.base_y;
    a.x[ixy] = (Lbd)X;
    a.y[ixy] = (Lbd)Y;
    int flag;
    if ((X - a.CX) * (X - a.CX) + (Y - a.CY) * (Y - a.CY) < a.R * a.R)
        flag = SOLID_SBB;
    else if (a.bc_mode == MBL_BC_PERIODIC)
        flag = FLUID;
    else if (a.bc_mode == MBL_BC_CHANNEL) {
        if (Y == 0.0 || Y == a.LY - 1.0)
            flag = MBL_SOLID_WALL;
        else if (X == 0.0)
            flag = BOUNDARY_INLET;
        else if (X == a.LX - 1.0)
            flag = BOUNDARY_OUTLET;
        else if (i == 0 || j == 0 || i == d.nx - 1 || j == d.ny - 1)
            flag = C_BOUNDARY;
        else
            flag = FLUID;
    } else {
        if (Y == 0.0 || Y == a.LY - 1.0 || X == 0.0 || X == a.LX - 1.0)
            flag = FLUID;
        else if (i == 0 || j == 0 || i == d.nx - 1 || j == d.ny - 1)
            flag = C_BOUNDARY;
        else
            flag = FLUID;
    }
    ((int *)d.flag)[ixy] = flag;
    double rho = a.c.rho0, ux = a.c.U0x, uy = a.c.U0y;
    if (a.init_mode == MBL_INIT_REST) {
        ux = 0.0;
        uy = 0.0;
    } else if (a.init_mode == MBL_INIT_RANDOM) {
        random_velocity(a.seed, a.sigma, i, j, ux, uy);
    } else if (a.init_mode == MBL_INIT_PROFILE) {
        const bool solid = (flag == SOLID_SBB || flag == MBL_SOLID_WALL);
        ux = solid ? 0.0 : 4.0 * a.c.umax * (a.LY - 1.5 - Y) * (Y - 0.5) / (a.LY - 2.0) / (a.LY - 2.0);
        uy = 0.0;
    }
    d.rho[ixy] = (Lbd)rho;
    d.ux[ixy] = (Lbd)ux;
    d.uy[ixy] = (Lbd)uy;
    if (d.upx) d.upx[ixy] = (Lbd)ux;
    if (d.upy) d.upy[ixy] = (Lbd)uy;
    a.fx[ixy] = (Lbd)0.0;
    a.fy[ixy] = (Lbd)0.0;
    // F = f = feq(rho, u) in double, stored in the plane's domain
    const double uv = ux * ux + uy * uy;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        double eu = EX[k] * ux + EY[k] * uy;
        double poly = 3.0 * eu + 4.5 * eu * eu - 1.5 * uv;
        double q = C::SHIFT ? W64[k] * ((rho - 1.0) + rho * poly) : W64[k] * rho * (1.0 + poly);
        T v = C::st((typename C::R)q);
        ((T *)d.P[0][k])[ixy] = v;
        ((T *)d.P[1][k])[ixy] = v;
    }
}

// ---------------------------------------------------------------------------
// Interface line ops
// ---------------------------------------------------------------------------
// 3-point Lagrange interpolation in time at xm of the ring slots m0, m1, m2
// (assit_multi.h:232-245); one rounding per operation wherever it is inlined
__device__ __forceinline__ double temporal3(double xm, Lbd f0, Lbd f1, Lbd f2)
{
#pragma clang fp contract(off)
    const double a = 0.5 * (xm - 1.0) * xm;
    const double b = (xm - 1.0) * (xm + 1.0);
    const double cc = 0.5 * xm * (xm + 1.0);
    return a * (double)f0 - b * (double)f1 + cc * (double)f2;
}

// Every load an entry needs is issued before its write mask is tested (all
// are in bounds for e < n), so an entry costs one dependent memory round trip
// after the op table (two for CC/FC: index, then source node), not one more
// for the mask.
template <int MODE>
__device__ __forceinline__ void line_op_entry(const LineOp &op, int e, const BlockDesc *__restrict__ blocks,
                                              uint32_t parity_mask, const MarkerSet &ms, double xm)
{
    typedef Codec<MODE> C;
    typedef typename C::T T;
    const int m0 = ms.m[op.mgroup][0], m1 = ms.m[op.mgroup][1], m2 = ms.m[op.mgroup][2];
    const BlockDesc &db = blocks[op.dst_block];
    const int pd = (parity_mask >> db.level) & 1;
    const uint8_t mask = op.mask[e];
    const int didx = op.dst_idx[e];
    T *D[9];
#pragma unroll
    for (int k = 0; k < 9; k++) D[k] = (T *)db.P[pd][k];
    if (op.type == OP_TEMPORAL) {
        double v[9];
#pragma unroll
        for (int k = 0; k < 9; k++) v[k] = temporal3(xm, op.ring[m0][k][e], op.ring[m1][k][e], op.ring[m2][k][e]);
        if (!mask) return;
#pragma unroll
        for (int k = 0; k < 9; k++) D[k][didx] = enc64<MODE>(v[k], k);
        return;
    }
    if (op.type == OP_CF) {
        Lbd v[9];
#pragma unroll
        for (int k = 0; k < 9; k++) v[k] = op.ring[m2][k][e];
        if (!mask) return;
#pragma unroll
        for (int k = 0; k < 9; k++) D[k][didx] = enc64<MODE>((double)v[k], k);
        return;
    }
    const BlockDesc &sb = blocks[op.src_block];
    const int ps = (parity_mask >> sb.level) & 1;
    const int sidx = op.src_idx[e];
    T sv[9];
    int rt = 0;
    if (op.red_n > 0 && e == op.red_e[0]) rt = 1;
    if (op.red_n > 1 && e == op.red_e[1]) rt = 2;
    if (rt) { // a line end read elsewhere (LineOp::red_*): the same bits the node holds in the reference order
        const bool r1 = (rt == 2);
        const int t = r1 ? op.red_t[1] : op.red_t[0];
        if (t == 1) {
            // both redirects' ring pointers come with the op (uniform), one is selected per lane
            const int mc0 = ms.m[op.red_mg[0]][2], mc1 = ms.m[op.red_mg[1]][2];
            const int ce = r1 ? op.red_ce[1] : op.red_ce[0];
#pragma unroll
            for (int k = 0; k < 9; k++) {
                const Lbd *p = r1 ? op.red_ring[1][mc1][k] : op.red_ring[0][mc0][k];
                sv[k] = enc64<MODE>((double)p[ce], k);
            }
        } else {
#pragma unroll
            for (int k = 0; k < 9; k++) sv[k] = ((const T *)sb.P[ps ^ 1][k])[sidx];
        }
    } else {
#pragma unroll
        for (int k = 0; k < 9; k++) sv[k] = ((const T *)sb.P[ps][k])[sidx];
    }
    const Lbd r = sb.rho[sidx], ux = sb.ux[sidx], uy = sb.uy[sidx];
    if (mask & 1) {
        db.rho[didx] = r;
        db.ux[didx] = ux;
        db.uy[didx] = uy;
    }
    if (!(mask & 2)) return;
    if (op.type == OP_CC) {
#pragma unroll
        for (int k = 0; k < 9; k++) D[k][didx] = sv[k];
    } else { // OP_FC: non-equilibrium rescaling fine -> coarse (assit_multi.h:304-311)
#pragma unroll
        for (int k = 0; k < 9; k++) {
            const double ft = dec64<MODE>(sv[k], k);
            const double fe = feq64(k, (double)r, (double)ux, (double)uy);
            D[k][didx] = enc64<MODE>(fe + op.para * (ft - fe), k);
        }
    }
}

template <int MODE>
__global__ __launch_bounds__(256) void k_line_ops(const LineOp *__restrict__ ops, const BlockDesc *__restrict__ blocks,
                                                  uint32_t parity_mask, MarkerSet ms, double xm)
{
    const LineOp &op = ops[blockIdx.y];
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= op.n) return;
    line_op_entry<MODE>(op, e, blocks, parity_mask, ms, xm);
}

// init_Fstore: every ring slot <- the fine edge's F (init_multi.h:3-28)
template <int MODE>
__global__ __launch_bounds__(256) void k_ring_fill(const LineOp *__restrict__ ops, const BlockDesc *__restrict__ blocks,
                                                   uint32_t parity_mask)
{
    typedef typename Codec<MODE>::T T;
    const LineOp &op = ops[blockIdx.y];
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= op.n) return;
    const BlockDesc &db = blocks[op.dst_block];
    const int pd = (parity_mask >> db.level) & 1;
    const int idx = op.dst_idx[e];
#pragma unroll
    for (int k = 0; k < 9; k++) {
        Lbd v = (Lbd)dec64<MODE>(((const T *)db.P[pd][k])[idx], k);
        op.ring[0][k][e] = v;
        op.ring[1][k][e] = v;
        op.ring[2][k][e] = v;
    }
}

// ---------------------------------------------------------------------------
// Spline: workgroup (column c, line op). Gathers one of the 12 columns of the
// coarse line (rho, ux, uy, F0..F8), builds the as-coded right-hand side
// 6g_i - 3g_{i-1} - 3g_{i+1} with zero end rows (assit_multi.h:82-107), and
// solves tridiag(l, m, u) M = rhs by parallel cyclic reduction (Hockney 1965)
// with the rows in registers and one LDS exchange per level, in FP64.
// ---------------------------------------------------------------------------
constexpr int SPL_TPB = 256;
constexpr int SPL_MAX_NC = 2048;

template <int MODE, int R>
__global__ __launch_bounds__(SPL_TPB) void k_spline_solve(const LineOp *__restrict__ ops,
                                                          const BlockDesc *__restrict__ blocks, uint32_t parity_mask)
{
    typedef typename Codec<MODE>::T T;
    const LineOp &op = ops[blockIdx.y];
    const int col = blockIdx.x;
    const int N = op.nc;
    const BlockDesc &sb = blocks[op.src_block];
    const int ps = (parity_mask >> sb.level) & 1;
    extern __shared__ double lds[]; // 4*N doubles: a | b | c | d
    double *La = lds, *Lb = lds + N, *Lc = lds + 2 * N, *Ld = lds + 3 * N;
    double a[R], b[R], c[R], dd[R], g[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int ip = threadIdx.x + r * SPL_TPB;
        g[r] = 0.0;
        if (ip < N) {
            const int node = op.src_idx[ip];
            double v;
            if (col == 0)
                v = (double)sb.rho[node];
            else if (col == 1)
                v = (double)sb.ux[node];
            else if (col == 2)
                v = (double)sb.uy[node];
            else
                v = dec64<MODE>(((const T *)sb.P[ps][col - 3])[node], col - 3);
            g[r] = v;
            op.gcol[col][ip] = (Lbd)v;
            Ld[ip] = v;
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int ip = threadIdx.x + r * SPL_TPB;
        a[r] = 0.0;
        b[r] = 1.0;
        c[r] = 0.0;
        dd[r] = 0.0;
        if (ip < N) {
            a[r] = (double)op.tl[ip];
            b[r] = (double)op.tm[ip];
            c[r] = (double)op.tu[ip];
            dd[r] = (ip == 0 || ip == N - 1) ? 0.0 : (6.0 * g[r] - 3.0 * Ld[ip - 1] - 3.0 * Ld[ip + 1]);
        }
    }
    for (int s = 1; s < N; s <<= 1) {
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int ip = threadIdx.x + r * SPL_TPB;
            if (ip < N) {
                La[ip] = a[r];
                Lb[ip] = b[r];
                Lc[ip] = c[r];
                Ld[ip] = dd[r];
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int ip = threadIdx.x + r * SPL_TPB;
            if (ip >= N) continue;
            const int im = ip - s, iq = ip + s;
            double na = 0.0, nc = 0.0, nb = b[r], nd = dd[r];
            if (im >= 0) {
                const double alpha = -a[r] / Lb[im];
                na = alpha * La[im];
                nb += alpha * Lc[im];
                nd += alpha * Ld[im];
            }
            if (iq < N) {
                const double gamma = -c[r] / Lb[iq];
                nc = gamma * Lc[iq];
                nb += gamma * La[iq];
                nd += gamma * Ld[iq];
            }
            a[r] = na;
            b[r] = nb;
            c[r] = nc;
            dd[r] = nd;
        }
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int ip = threadIdx.x + r * SPL_TPB;
        if (ip < N) {
            const double M = dd[r] / b[r];
            op.M64[(size_t)col * N + ip] = M;
            op.varibs[(size_t)col * N + ip] = (Lbd)M;
        }
    }
}

// assit_multi.h:118-134, the literal 0.1666667 included
__device__ inline double spline3_s(double xx, double x_1, double h, double fi_1, double fi, double mi_1, double mi)
{
    double x = x_1 + h;
    double a = mi_1 * (x - xx) * (x - xx) * (x - xx) * 0.1666667;
    double b = mi * (xx - x_1) * (xx - x_1) * (xx - x_1) * 0.1666667;
    double c = (fi_1 - mi_1 * h * h * 0.1666667) * (x - xx);
    double d = (fi - mi * h * h * 0.1666667) * (xx - x_1);
    return (a + b + c + d);
}

// calculate_Spline3_gpu (assit_multi.h:136-185): evaluate at fine point ip,
// write the fine edge's rho, u and the rescaled F into ring slot `slot`.
template <int MODE>
__global__ __launch_bounds__(256) void k_spline_eval(const LineOp *__restrict__ ops, const BlockDesc *__restrict__ blocks,
                                                     uint32_t parity_mask, MarkerSet ms)
{
    typedef typename Codec<MODE>::T T;
    const LineOp &op = ops[blockIdx.y];
    const int slot = ms.m[op.mgroup][0]; // the oldest slot of the line's ring receives t_{n+1}
    const int ip = blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= op.n) return;
    const int N = op.nc;
    const BlockDesc &sb = blocks[op.src_block];
    const int ps = (parity_mask >> sb.level) & 1;
    const double xx = (double)ip / (double)op.ratio;
    const int I = (ceil(xx) == 0.0) ? (int)(ceil(xx) + 1.0) : (int)ceil(xx);
    const int n0 = op.src_idx[I - 1], n1 = op.src_idx[I];
    const double *M = op.M64;
    const double rho = spline3_s(xx, I - 1.0, 1.0, (double)sb.rho[n0], (double)sb.rho[n1], M[I - 1], M[I]);
    const double ux = spline3_s(xx, I - 1.0, 1.0, (double)sb.ux[n0], (double)sb.ux[n1], M[N + I - 1], M[N + I]);
    const double uy =
        spline3_s(xx, I - 1.0, 1.0, (double)sb.uy[n0], (double)sb.uy[n1], M[2 * N + I - 1], M[2 * N + I]);
    if (op.mask[ip] & 1) {
        const BlockDesc &db = blocks[op.dst_block];
        const int idx = op.dst_idx[ip];
        db.rho[idx] = (Lbd)rho;
        db.ux[idx] = (Lbd)ux;
        db.uy[idx] = (Lbd)uy;
    }
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const T *Sk = (const T *)sb.P[ps][k];
        const double ft = spline3_s(xx, I - 1.0, 1.0, dec64<MODE>(Sk[n0], k), dec64<MODE>(Sk[n1], k),
                                    M[(3 + k) * N + I - 1], M[(3 + k) * N + I]);
        const double fe = feq64(k, rho, ux, uy);
        op.ring[slot][k][ip] = (Lbd)(fe + op.para * (ft - fe));
    }
}

// The whole spline step of one line in one workgroup (lines up to
// SPL_FUSED_MAX_NC coarse points): gather the columns (init_Spline3,
// assit_multi.h:48-116), solve tridiag(l, m, u) M = rhs for all of them by PCR
// at once, and evaluate the fine points (calculate_Spline3_gpu, :136-185) from
// the columns and M held in LDS. The matrix is the same at every step, so the
// elimination factors of every PCR level are computed once per layout
// (LineOp::pcr): a level is one LDS exchange of the right-hand sides and two
// multiply-adds per column, with no division and no vote. The fine points'
// write masks and node indices are loaded before the gather, off the path of
// the dependent loads. Same operations, in the same order per column, as
// k_spline_solve + k_spline_eval. One coarse row per thread; blockDim.x >= nc.
// NG > 1 spreads a line over NG workgroups (grp = 0 .. NG-1): each solves the
// rho, ux, uy columns (every fine point's equilibrium needs them) and the
// distributions k = grp, grp + NG, ..., and writes those; grp 0 also writes the
// macros. A column's arithmetic does not depend on the others, so every split
// gives the same bits.
constexpr int SPL_FUSED_MAX_NC = 303; // 24 * 8 * nc bytes of LDS within 64 KiB (and the drag's)

// Workgroups per line of a spline launch of nlines lines: 9 while the launch is far from filling
// the GPU (each workgroup solves 4 of the 12 columns), 3 for more lines (6 columns each: the
// rho, ux, uy columns are solved by every workgroup). Measured on the 256-case sweep (1,024
// lines): 27.6k MLUPS with 9, 29.0k with 3, 27.7k with 1.
static inline int spline_groups(int nlines) { return nlines * 9 <= 1024 ? 9 : 3; }

template <int MODE, int NG = 1>
__device__ __forceinline__ void spline_line(const LineOp &op, const BlockDesc *__restrict__ blocks,
                                            uint32_t parity_mask, const MarkerSet &ms, double *lds, int grp = 0,
                                            const LineOp *__restrict__ base = nullptr)
{
    static_assert(NG == 1 || NG == 3 || NG == 9, "the 9 distributions split evenly");
    constexpr int NK = 9 / NG, NCOL = 3 + NK; // distributions and columns of this workgroup
    typedef typename Codec<MODE>::T T;
    const int slot = ms.m[op.mgroup][0]; // the oldest slot of the line's ring receives t_{n+1}
    const int N = op.nc, NF = op.n, NL = op.pcr_levels;
    const BlockDesc &sb = blocks[op.src_block];
    const BlockDesc &db = blocks[op.dst_block];
    const int ps = (parity_mask >> sb.level) & 1;
    // column c of this workgroup is global column gc(c): rho, ux, uy, then F_k for its k
    auto gc = [&](int c) { return c < 3 ? c : 3 + grp + NG * (c - 3); };
    // lds: G[NCOL][N] | d[NCOL][N]
    double *G = lds, *Ld = lds + NCOL * N;
    const int ip = threadIdx.x;
    const bool row = ip < N;
    // the first fine point of this thread: its write mask and node index, loaded now
    const int f0 = ip;
    const uint8_t mk0 = f0 < NF ? op.mask[f0] : (uint8_t)0;
    const int di0 = f0 < NF ? op.dst_idx[f0] : 0;
    // this row's elimination factors of every level, and its reduced diagonal
    double al[PCR_MAX_LEVELS], ga[PCR_MAX_LEVELS], bd = 1.0;
    if (row) {
#pragma unroll
        for (int q = 0; q < PCR_MAX_LEVELS; q++) {
            al[q] = q < NL ? op.pcr[(size_t)(2 * q) * N + ip] : 0.0;
            ga[q] = q < NL ? op.pcr[(size_t)(2 * q + 1) * N + ip] : 0.0;
        }
        bd = op.pcr[(size_t)(2 * NL) * N + ip];
    }
    double d[NCOL];
    if (row) {
        const int node = op.src_idx[ip];
        const int2 rd = op.gred ? op.gred[ip] : make_int2(-1, -1);
        double g[NCOL];
        g[0] = (double)sb.rho[node];
        g[1] = (double)sb.ux[node];
        g[2] = (double)sb.uy[node];
#pragma unroll
        for (int c = 3; c < NCOL; c++) {
            const int k = gc(c) - 3;
            g[c] = dec64<MODE>(((const T *)sb.P[ps][k])[node], k);
        }
        // nodes a same-level copy of this step fills: read at the copy's source (same level,
        // same plane set), in the waves that hold one
        if (__builtin_amdgcn_ballot_w64(rd.x >= 0)) {
            if (rd.x >= 0) {
                const BlockDesc &rb = blocks[rd.x];
                g[0] = (double)rb.rho[rd.y];
                g[1] = (double)rb.ux[rd.y];
                g[2] = (double)rb.uy[rd.y];
#pragma unroll
                for (int c = 3; c < NCOL; c++) {
                    const int k = gc(c) - 3;
                    g[c] = dec64<MODE>(((const T *)rb.P[ps][k])[rd.y], k);
                }
            }
        }
#pragma unroll
        for (int c = 0; c < NCOL; c++) {
            G[c * N + ip] = g[c];
            if (c >= 3 || grp == 0) op.gcol[gc(c)][ip] = (Lbd)g[c];
        }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < NCOL; c++)
        d[c] = (!row || ip == 0 || ip == N - 1)
                   ? 0.0
                   : (6.0 * G[c * N + ip] - 3.0 * G[c * N + ip - 1] - 3.0 * G[c * N + ip + 1]);
#pragma unroll
    for (int lv = 0; lv < PCR_MAX_LEVELS; lv++) {
        if (lv >= NL) continue; // NL is the same for the whole workgroup
        const int s = 1 << lv;
        if (row) {
#pragma unroll
            for (int c = 0; c < NCOL; c++) Ld[c * N + ip] = d[c];
        }
        __syncthreads();
        if (row) {
            const int im = ip - s, iq = ip + s;
            if (im >= 0) {
#pragma unroll
                for (int c = 0; c < NCOL; c++) d[c] += al[lv] * Ld[c * N + im];
            }
            if (iq < N) {
#pragma unroll
                for (int c = 0; c < NCOL; c++) d[c] += ga[lv] * Ld[c * N + iq];
            }
        }
        __syncthreads(); // this level's reads before the next level's writes
    }
    if (row) {
#pragma unroll
        for (int c = 0; c < NCOL; c++) {
            const double M = d[c] / bd;
            Ld[c * N + ip] = M;
            if (c >= 3 || grp == 0) op.varibs[(size_t)gc(c) * N + ip] = (Lbd)M;
        }
    }
    __syncthreads();
    const double *M = Ld;
    for (int f = f0; f < NF; f += blockDim.x) {
        const uint8_t mk = f == f0 ? mk0 : op.mask[f];
        const double xx = (double)f / (double)op.ratio;
        const int I = (ceil(xx) == 0.0) ? (int)(ceil(xx) + 1.0) : (int)ceil(xx);
        const double rho = spline3_s(xx, I - 1.0, 1.0, G[I - 1], G[I], M[I - 1], M[I]);
        const double ux = spline3_s(xx, I - 1.0, 1.0, G[N + I - 1], G[N + I], M[N + I - 1], M[N + I]);
        const double uy = spline3_s(xx, I - 1.0, 1.0, G[2 * N + I - 1], G[2 * N + I], M[2 * N + I - 1], M[2 * N + I]);
        if ((mk & 1) && grp == 0) {
            const int idx = f == f0 ? di0 : op.dst_idx[f];
            db.rho[idx] = (Lbd)rho;
            db.ux[idx] = (Lbd)ux;
            db.uy[idx] = (Lbd)uy;
        }
#pragma unroll
        for (int c = 3; c < NCOL; c++) {
            const int k = gc(c) - 3, o = c * N;
            const double ft = spline3_s(xx, I - 1.0, 1.0, G[o + I - 1], G[o + I], M[o + I - 1], M[o + I]);
            const double fe = feq64(k, rho, ux, uy);
            op.ring[slot][k][f] = (Lbd)(fe + op.para * (ft - fe));
        }
    }
    if (op.fuse_t && base) {
        // the temporal op on this line: its fine points after this thread stored them, for this
        // workgroup's distributions (line_op_entry's OP_TEMPORAL, same arithmetic)
        const LineOp &t = base[op.fuse_t - 1];
        int mt[3];
#pragma unroll
        for (int p = 0; p < 3; p++) mt[p] = (ms.m[t.mgroup][p] + 2 * t.trot) % 3;
        const BlockDesc &tb = blocks[t.dst_block];
        const int pt = ((parity_mask >> tb.level) & 1) ^ t.tflip;
        for (int f = f0; f < t.n; f += blockDim.x) {
            const uint8_t mk = t.mask[f];
            const int didx = t.dst_idx[f];
#pragma unroll
            for (int c = 3; c < NCOL; c++) {
                const int k = gc(c) - 3;
                const double v = temporal3(t.txm, t.ring[mt[0]][k][f], t.ring[mt[1]][k][f], t.ring[mt[2]][k][f]);
                if (mk) ((T *)tb.P[pt][k])[didx] = enc64<MODE>(v, k);
            }
        }
    }
}

template <int MODE, int NG>
__global__ __launch_bounds__(1024) void k_spline_fused(const LineOp *__restrict__ ops,
                                                       const BlockDesc *__restrict__ blocks, uint32_t parity_mask,
                                                       MarkerSet ms)
{
    extern __shared__ double lds[];
    spline_line<MODE, NG>(ops[blockIdx.y], blocks, parity_mask, ms, lds, (int)blockIdx.x, ops);
}

// One-workgroup splines and interface ops of one position of the scheduled cycle in one
// launch: workgroup row y < nspl runs spline y (its first NG workgroups, one share of the
// line's columns each), the rows after it the entries of interface op y (one entry per thread).
template <int MODE, int NG>
__global__ __launch_bounds__(1024) void k_iface_mixed(const LineOp *__restrict__ ops, int nspl,
                                                      const BlockDesc *__restrict__ blocks, uint32_t parity_mask,
                                                      MarkerSet ms, double xm)
{
    extern __shared__ double lds[];
    const int y = (int)blockIdx.y;
    if (y < nspl) {
        if (blockIdx.x < NG) spline_line<MODE, NG>(ops[y], blocks, parity_mask, ms, lds, (int)blockIdx.x, ops);
        return;
    }
    const LineOp &op = ops[y];
    const int e = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (e >= op.n) return;
    line_op_entry<MODE>(op, e, blocks, parity_mask, ms, xm);
}

// A stream-collide and the interface ops of one position of the scheduled cycle in one launch
// (no dependence between them, or the schedule would have separated them): workgroups
// [0, tiles) are the collide's tiles (one node per thread, flagged blocks), the others run
// the entries of the ops, opx workgroups per op.
template <int MODE, bool RESID>
__global__ __launch_bounds__(256) void k_collide_ops(CollideOpsArgs a)
{
    const int b = (int)blockIdx.x;
    if (b < a.tiles) {
        const int t = a.tile_base + xcd_tile(b, a.tiles);
        sc_tile<MODE, 1, false, true, true, RESID>(a.descs, a.tile_blk, a.ndesc, a.rows, a.pm, t,
                                                   (int)(threadIdx.x >> 6), (int)(blockDim.x >> 6), a.rmask, a.rparts);
        return;
    }
    const int q = b - a.tiles, y = q / a.opx;
    const LineOp &op = a.ops[y];
    const int e = (q - y * a.opx) * (int)blockDim.x + (int)threadIdx.x;
    if (e >= op.n) return;
    line_op_entry<MODE>(op, e, a.descs, a.pm, a.ms, a.xm);
}

// init_tri_A (multi/init_multi.h:30-51), final race-free state
__global__ void k_tri_A(Lbd *l, Lbd *m, Lbd *u, int N)
{
    const int ip = blockIdx.x * blockDim.x + threadIdx.x;
    if (ip >= N) return;
    m[ip] = (Lbd)2.0;
    l[ip] = (ip == 0 || ip == 1 || ip == N - 1) ? (Lbd)0.0 : (Lbd)0.5;
    u[ip] = (ip == 0 || ip == N - 2 || ip == N - 1) ? (Lbd)0.0 : (Lbd)0.5;
}

// ---------------------------------------------------------------------------
// FP64 reductions
// ---------------------------------------------------------------------------
__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// sums v over the workgroup (blockDim.x multiple of 64, <= 1024); result in thread 0
__device__ inline double group_sum(double v, double *sh)
{
    v = wave_sum(v);
    const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < nw; i++) s += sh[i];
    return s;
}

// minimum of v over the workgroup (NaN ignored, as fmin does); result in thread 0
__device__ inline double group_min(double v, double *sh)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    double s = v;
    if (threadIdx.x == 0)
        for (int i = 0; i < nw; i++) s = fmin(s, sh[i]);
    return s;
}

// Momentum exchange (force_block.h:129-178) from the previous state planes:
// f_post[k][x_s] = Fold[k][x_f], f_post[opp k][x_f] = Fold[opp k][x_s].
template <int MODE> __global__ __launch_bounds__(1024) void k_drag(DragArgs a)
{
    typedef typename Codec<MODE>::T T;
    __shared__ double sh[16];
    double fx = 0.0, fy = 0.0;
    for (int e = threadIdx.x; e < a.n; e += blockDim.x) {
        const Int3 L = a.links[e];
        const int k = L.k, kd = OPP[k];
        const size_t xs = (size_t)L.j * a.nx + L.i;
        const size_t xf = (size_t)(L.j - EY[k]) * a.nx + (L.i - EX[k]);
        const double fk = dec64<MODE>(((const T *)a.Fold[k])[xf], k);
        const double fkd = dec64<MODE>(((const T *)a.Fold[kd])[xs], kd);
        fx += (fk * EX[k] - fkd * EX[kd]);
        fy += (fk * EY[k] - fkd * EY[kd]);
    }
    const double sx = group_sum(fx, sh);
    const double sy = group_sum(fy, sh);
    if (threadIdx.x == 0) {
        a.out[0] = sx;
        a.out[1] = sy;
    }
}

// Momentum exchange of many cases at once: workgroup c reduces case c's
// links into out[2c], out[2c+1] (same per-link terms as k_drag).
template <int MODE> __device__ __forceinline__ void drag_case(const DragArgs &a, double *out2, double *sh)
{
    typedef typename Codec<MODE>::T T;
    double fx = 0.0, fy = 0.0;
    for (int e = threadIdx.x; e < a.n; e += blockDim.x) {
        const Int3 L = a.links[e];
        const int k = L.k, kd = OPP[k];
        const size_t xs = (size_t)L.j * a.nx + L.i;
        const size_t xf = (size_t)(L.j - EY[k]) * a.nx + (L.i - EX[k]);
        const double fk = dec64<MODE>(((const T *)a.Fold[k])[xf], k);
        const double fkd = dec64<MODE>(((const T *)a.Fold[kd])[xs], kd);
        fx += (fk * EX[k] - fkd * EX[kd]);
        fy += (fk * EY[k] - fkd * EY[kd]);
    }
    const double sx = group_sum(fx, sh);
    const double sy = group_sum(fy, sh);
    if (threadIdx.x == 0) {
        out2[0] = sx;
        out2[1] = sy;
    }
}

template <int MODE> __global__ __launch_bounds__(256) void k_drag_batch(const DragArgs *__restrict__ args, double *out)
{
    __shared__ double sh[16];
    drag_case<MODE>(args[blockIdx.x], out + 2 * blockIdx.x, sh);
}

// error_L2_area (error_L2.h:3-55): workgroup t takes tile t of 64 x 16 nodes
// of block tblk[t]: FP64 sums of |u - u_prev|^2 and |u|^2 over the tile,
// u_prev <- u, weighted by the block's 1/m (exact for the power-of-two ratios
// of nested blocks, so the sum over tiles equals the reference's per-block
// sums divided by m, error_L2.h:64-84, up to summation order).
// mode 0: error_L2; mode 1: L2 distance of (ux, uy) on FLUID nodes from the
// analytic plane Poiseuille profile u_a(y) = 4 umax (LY - 1.5 - y)(y - 0.5)/(LY - 2)^2,
// without touching u_prev.
__global__ __launch_bounds__(256) void k_residual_tiles(const BlockDesc *__restrict__ descs,
                                                        const int *__restrict__ tblk, int ndesc, double *partials,
                                                        int mode, double umax, double LY)
{
    __shared__ double sh[16];
    const int t = blockIdx.x;
    int blk;
    if (tblk) {
        blk = tblk[t];
    } else { // no tile table: the last block whose first residual tile is <= t
        int lo = 0, hi = ndesc - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (descs[mid].rtile_begin <= t)
                lo = mid;
            else
                hi = mid - 1;
        }
        blk = lo;
    }
    const BlockDesc &d = descs[blk];
    const int lt = t - d.rtile_begin;
    const int ty = lt / d.rtiles_x, tx = lt - ty * d.rtiles_x;
    const int i = tx * 64 + (threadIdx.x & 63);
    double s1 = 0.0, s2 = 0.0, rmin = INFINITY;
    if (i < d.nx) {
        for (int r = 0; r < 4; r++) {
            const int j = ty * 16 + r * 4 + (threadIdx.x >> 6);
            if (j >= d.ny) break;
            const size_t ixy = (size_t)j * d.nx + i;
            const double ux = (double)d.ux[ixy], uy = (double)d.uy[ixy];
            if (mode == 0) {
                rmin = fmin(rmin, (double)d.rho[ixy]);
                const double upx = (double)d.upx[ixy], upy = (double)d.upy[ixy];
                s1 += (ux - upx) * (ux - upx) + (uy - upy) * (uy - upy);
                s2 += ux * ux + uy * uy;
                d.upx[ixy] = d.ux[ixy];
                d.upy[ixy] = d.uy[ixy];
            } else if (d.flag[ixy] == FLUID) {
                const double y = j * d.dx + d.base_y;
                const double ua = 4.0 * umax * (LY - 1.5 - y) * (y - 0.5) / (LY - 2.0) / (LY - 2.0);
                s1 += (ux - ua) * (ux - ua) + uy * uy;
                s2 += ua * ua;
            }
        }
    }
    const double a1 = group_sum(s1, sh);
    const double a2 = group_sum(s2, sh);
    const double mn = group_min(rmin, sh);
    if (threadIdx.x == 0) {
        partials[3 * (size_t)t] = a1 * d.inv_m;
        partials[3 * (size_t)t + 1] = a2 * d.inv_m;
        partials[3 * (size_t)t + 2] = mn;
    }
}

// The weighted tile sums in one fixed order: result sqrt(s1)/sqrt(s2), s1, s2,
// and the smallest density of any node (a divergence check: rho <= 0).
__device__ __forceinline__ void residual_record(int total, const double *__restrict__ partials, double *out,
                                                double *sh)
{
    double g1 = 0.0, g2 = 0.0, mn = INFINITY;
    for (int q = threadIdx.x; q < total; q += blockDim.x) {
        g1 += partials[3 * (size_t)q];
        g2 += partials[3 * (size_t)q + 1];
        mn = fmin(mn, partials[3 * (size_t)q + 2]);
    }
    const double all1 = group_sum(g1, sh);
    const double all2 = group_sum(g2, sh);
    const double allmn = group_min(mn, sh);
    if (threadIdx.x == 0) {
        out[0] = sqrt(all1) / sqrt(all2);
        out[1] = all1;
        out[2] = all2;
        out[3] = allmn;
    }
}

__global__ __launch_bounds__(1024) void k_residual_final(int total, const double *__restrict__ partials, double *out)
{
    __shared__ double sh[16];
    residual_record(total, partials, out, sh);
}

// The records of consecutive steps, one workgroup each (the same order as k_residual_final).
__global__ __launch_bounds__(1024) void k_residual_steps(int total, size_t stride, const double *__restrict__ partials,
                                                         double *out)
{
    __shared__ double sh[16];
    residual_record(total, partials + 3 * stride * blockIdx.x, out + 4 * (size_t)blockIdx.x, sh);
}

// Per-case residual of a batch: case c's residual tiles are [case_tiles[c], case_tiles[c+1]),
// summed in order; out[4c..4c+3] = residual, s1, s2, smallest density (k_residual_final's record).
__global__ __launch_bounds__(256) void k_residual_cases(const int *__restrict__ case_tiles,
                                                        const double *__restrict__ partials, double *out)
{
    __shared__ double sh[16];
    const int c = blockIdx.x;
    double g1 = 0.0, g2 = 0.0, mn = INFINITY;
    for (int q = case_tiles[c] + threadIdx.x; q < case_tiles[c + 1]; q += blockDim.x) {
        g1 += partials[3 * (size_t)q];
        g2 += partials[3 * (size_t)q + 1];
        mn = fmin(mn, partials[3 * (size_t)q + 2]);
    }
    const double all1 = group_sum(g1, sh);
    const double all2 = group_sum(g2, sh);
    const double allmn = group_min(mn, sh);
    if (threadIdx.x == 0) {
        out[4 * (size_t)c] = sqrt(all1) / sqrt(all2);
        out[4 * (size_t)c + 1] = all1;
        out[4 * (size_t)c + 2] = all2;
        out[4 * (size_t)c + 3] = allmn;
    }
}

// The diagnostics of one coarse step after a cycle whose last stream-collide of each level
// wrote its tiles' residual sums (error_L2.h:57-91 and force_block.h:180-216).
// Workgroups [0, list_wgs) take error_L2's terms of the nodes the collides do not sum (block
// edges, solid and ghost nodes: their macros are final once the cycle is done), u_prev <- u,
// plus one contiguous share of the tiles' partial sums, into their slots after the tiles';
// workgroup list_wgs (with a drag) sums the momentum exchange. The list_wgs partial sums of
// each step stay in a slot of their own until the diagnostics are read back, when one
// launch sums every step's slots in a fixed order (k_residual_steps: within one launch,
// workgroups on different XCDs would need a write-back of the L2 to see each other's sums).
// The slot comes from a device counter, so the launch is the same at every step (it sits in
// the step's captured graph): read from one of two counters, the next slot written to the
// other, which the next step reads (no workgroup writes what another reads).
template <int MODE> __device__ __forceinline__ void step_diag_wg(const StepDiagArgs &a, int wg, double *sh)
{
    // this step's record slot (the previous step's diagnostics, or the host, set it); the next
    // step reads the other counter
    const int slot = *a.slot_rd;
    if (slot < 0 || slot >= a.cap) return;
    if (wg == 0 && threadIdx.x == 0) *a.slot_wr = slot + 1;
    if (wg < a.list_wgs) {
        ResAcc acc;
        const int e = wg * (int)blockDim.x + (int)threadIdx.x;
        if (e < a.nlist) {
            const int4 q = a.list[e];
            const BlockDesc &d = a.descs[q.x], &m = a.descs[q.z];
            const uint32_t o = (uint32_t)q.y, om = (uint32_t)q.w;
            const Lbd ux = m.ux[om], uy = m.uy[om], r = m.rho[om], px = d.upx[o], py = d.upy[o];
            const double x = (double)ux, y = (double)uy, x0 = (double)px, y0 = (double)py;
            acc.mn = (double)r;
            acc.s1 = ((x - x0) * (x - x0) + (y - y0) * (y - y0)) * d.inv_m;
            acc.s2 = (x * x + y * y) * d.inv_m;
            d.upx[o] = ux;
            d.upy[o] = uy;
        }
        const int chunk = (a.ntiles + a.list_wgs - 1) / a.list_wgs;
        const int t1 = min(a.ntiles, (wg + 1) * chunk);
        for (int q = wg * chunk + (int)threadIdx.x; q < t1; q += (int)blockDim.x) {
            acc.s1 += a.rparts[3 * (size_t)q];
            acc.s2 += a.rparts[3 * (size_t)q + 1];
            acc.mn = fmin(acc.mn, a.rparts[3 * (size_t)q + 2]);
        }
        tile_resid_store(acc, 1.0, a.lbase + 3 * (a.lstride * (size_t)slot + (size_t)wg));
    } else {
        drag_case<MODE>(a.drag, a.drag_base + 2 * (size_t)slot, sh);
    }
}

template <int MODE> __global__ __launch_bounds__(256) void k_step_diag(StepDiagArgs a)
{
    __shared__ double sh[16];
    step_diag_wg<MODE>(a, (int)blockIdx.x, sh);
}

// The cycle's last interface ops (workgroups [0, nops opx)) and the step's diagnostics (the
// workgroups after them) in one launch.
template <int MODE>
__global__ __launch_bounds__(256) void k_ops_diag(const LineOp *__restrict__ ops, int nops, int opx,
                                                  const BlockDesc *__restrict__ blocks, uint32_t parity_mask,
                                                  MarkerSet ms, double xm, StepDiagArgs a)
{
    __shared__ double sh[16];
    const int b = (int)blockIdx.x, nb = nops * opx;
    if (b >= nb) {
        step_diag_wg<MODE>(a, b - nb, sh);
        return;
    }
    const int y = b / opx;
    const LineOp &op = ops[y];
    const int e = (b - y * opx) * (int)blockDim.x + (int)threadIdx.x;
    if (e >= op.n) return;
    line_op_entry<MODE>(op, e, blocks, parity_mask, ms, xm);
}

// Zeroth and first moments of one block's 9 planes (sum of f, f e_x, f e_y over every node)
// and the count of non-finite entries, FP64: workgroup partials in a grid-stride order,
// then one ordered sum (k_moments_final).
template <int MODE>
__global__ __launch_bounds__(256) void k_moments(const void *const *planes, size_t n, double *partials)
{
    typedef typename Codec<MODE>::T T;
    __shared__ double sh[16];
    double m = 0.0, mx = 0.0, my = 0.0, bad = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
#pragma unroll
        for (int k = 0; k < 9; k++) {
            const double f = dec64<MODE>(((const T *)planes[k])[i], k);
            if (!isfinite(f)) {
                bad += 1.0;
                continue;
            }
            m += f;
            mx += f * EX[k];
            my += f * EY[k];
        }
    }
    const double a = group_sum(m, sh), b = group_sum(mx, sh), c = group_sum(my, sh), d = group_sum(bad, sh);
    if (threadIdx.x == 0) {
        partials[4 * (size_t)blockIdx.x] = a;
        partials[4 * (size_t)blockIdx.x + 1] = b;
        partials[4 * (size_t)blockIdx.x + 2] = c;
        partials[4 * (size_t)blockIdx.x + 3] = d;
    }
}

__global__ __launch_bounds__(256) void k_moments_final(int nparts, const double *partials, double *out)
{
    __shared__ double sh[16];
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int q = threadIdx.x; q < nparts; q += blockDim.x)
        for (int c = 0; c < 4; c++) v[c] += partials[4 * (size_t)q + c];
    for (int c = 0; c < 4; c++) {
        const double s = group_sum(v[c], sh);
        if (threadIdx.x == 0) out[c] = s;
    }
}

template <int MODE> __global__ void k_decode(const void *plane, int k, size_t n, double *out)
{
    typedef typename Codec<MODE>::T T;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = dec64<MODE>(((const T *)plane)[i], k);
}

template <int MODE> __global__ void k_encode(void *plane, int k, size_t n, const double *in)
{
    typedef typename Codec<MODE>::T T;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) ((T *)plane)[i] = enc64<MODE>(in[i], k);
}

// ---------------------------------------------------------------------------
// Raster exchange of batched cases (learned warm start; no reference
// counterpart). A raster of H x W pixels spans the channel's coarse nodes
// [0, LX-1] x [0, LY-1]; pixel (py, px) sits at X = px (LX-1)/(W-1),
// Y = py (LY-1)/(H-1).
// ---------------------------------------------------------------------------

// Bilinear (rho, ux, uy) of a block at coarse coordinates (X, Y); false when
// the point lies outside the block.
__device__ inline bool block_sample(const RasterBlock &b, double X, double Y, double v[3])
{
    const double fx = (X - b.bx) * b.m, fy = (Y - b.by) * b.m;
    const double eps = 1e-9;
    if (fx < -eps || fy < -eps || fx > b.nx - 1 + eps || fy > b.ny - 1 + eps) return false;
    int i0 = (int)floor(fx), j0 = (int)floor(fy);
    i0 = min(max(i0, 0), b.nx - 2);
    j0 = min(max(j0, 0), b.ny - 2);
    const double tx = fx - i0, ty = fy - j0;
    const size_t n00 = (size_t)j0 * b.nx + i0, n10 = n00 + 1, n01 = n00 + b.nx, n11 = n01 + 1;
    const Lbd *src[3] = {b.rho, b.ux, b.uy};
#pragma unroll
    for (int q = 0; q < 3; q++) {
        const Lbd *s = src[q];
        v[q] = (1 - ty) * ((1 - tx) * (double)s[n00] + tx * (double)s[n10]) +
               ty * ((1 - tx) * (double)s[n01] + tx * (double)s[n11]);
    }
    return true;
}

__global__ __launch_bounds__(256) void k_raster(const RasterBlock *__restrict__ blocks,
                                                const RasterCase *__restrict__ cases, int H, int W,
                                                float *__restrict__ out)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= H * W) return;
    const RasterCase &c = cases[blockIdx.y];
    const int px = p % W, py = p / W;
    const double X = px * (c.LX - 1.0) / (W - 1), Y = py * (c.LY - 1.0) / (H - 1);
    double v[3] = {0.0, 0.0, 0.0};
    for (int q = 0; q < c.nblk; q++)
        if (block_sample(blocks[c.blk[q]], X, Y, v)) break;
    float *o = out + ((size_t)blockIdx.y * H * W + p) * 3;
    o[0] = (float)v[0];
    o[1] = (float)v[1];
    o[2] = (float)v[2];
}

// Bilinear sample of channel q of a case's raster at raster coordinates (gx, gy).
__device__ inline double raster_at(const float *r, int H, int W, double gx, double gy, int q)
{
    gx = fmin(fmax(gx, 0.0), W - 1.0);
    gy = fmin(fmax(gy, 0.0), H - 1.0);
    int i0 = min((int)gx, W - 2), j0 = min((int)gy, H - 2);
    const double tx = gx - i0, ty = gy - j0;
    const float *r00 = r + ((size_t)j0 * W + i0) * 3;
    const float *r01 = r00 + (size_t)W * 3;
    return (1 - ty) * ((1 - tx) * r00[q] + tx * r00[3 + q]) + ty * ((1 - tx) * r01[q] + tx * r01[3 + q]);
}

// F = f = feq(rho, u) - (tau - 1) w_k rho / cs^2 (e_k e_k - cs^2 I) : grad u on
// every node of every block, from the case's raster (first-order
// Chapman-Enskog state after collision; grad u in the block's lattice units is
// the coarse-unit gradient divided by m). Solid and wall nodes rest at rho = 1.
template <int MODE>
__global__ __launch_bounds__(256) void k_init_raster(const RasterBlock *__restrict__ blocks,
                                                     const RasterCase *__restrict__ cases, int H, int W,
                                                     const float *__restrict__ fields)
{
    typedef Codec<MODE> C;
    typedef typename C::T T;
    const RasterBlock &b = blocks[blockIdx.y];
    const size_t nn = (size_t)b.nx * b.ny;
    const size_t ixy = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ixy >= nn) return;
    const RasterCase &c = cases[b.cas];
    const float *r = fields + (size_t)b.cas * H * W * 3;
    const int i = (int)(ixy % b.nx), j = (int)(ixy / b.nx);
    const double X = b.bx + (double)i / b.m, Y = b.by + (double)j / b.m;
    const int flag = b.flag[ixy];
    double rho = 1.0, ux = 0.0, uy = 0.0, dux_dx = 0.0, dux_dy = 0.0, duy_dx = 0.0, duy_dy = 0.0;
    if (flag != SOLID_SBB && flag != MBL_SOLID_WALL) {
        const double sx = (W - 1) / (c.LX - 1.0), sy = (H - 1) / (c.LY - 1.0); // raster pixels per coarse unit
        const double gx = X * sx, gy = Y * sy;
        rho = raster_at(r, H, W, gx, gy, 0);
        ux = raster_at(r, H, W, gx, gy, 1);
        uy = raster_at(r, H, W, gx, gy, 2);
        const double xp = fmin(gx + 1.0, W - 1.0), xm = fmax(gx - 1.0, 0.0);
        const double yp = fmin(gy + 1.0, H - 1.0), ym = fmax(gy - 1.0, 0.0);
        const double hx = (xp - xm) / sx * b.m, hy = (yp - ym) / sy * b.m; // in the block's lattice units
        dux_dx = (raster_at(r, H, W, xp, gy, 1) - raster_at(r, H, W, xm, gy, 1)) / hx;
        duy_dx = (raster_at(r, H, W, xp, gy, 2) - raster_at(r, H, W, xm, gy, 2)) / hx;
        dux_dy = (raster_at(r, H, W, gx, yp, 1) - raster_at(r, H, W, gx, ym, 1)) / hy;
        duy_dy = (raster_at(r, H, W, gx, yp, 2) - raster_at(r, H, W, gx, ym, 2)) / hy;
    }
    b.rho[ixy] = (Lbd)rho;
    b.ux[ixy] = (Lbd)ux;
    b.uy[ixy] = (Lbd)uy;
    const double uv = ux * ux + uy * uy;
    const double div = dux_dx + duy_dy;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const double eu = EX[k] * ux + EY[k] * uy;
        const double poly = 3.0 * eu + 4.5 * eu * eu - 1.5 * uv;
        // e_a e_b d_a u_b - cs^2 div u
        const double qg = EX[k] * EX[k] * dux_dx + EX[k] * EY[k] * (dux_dy + duy_dx) + EY[k] * EY[k] * duy_dy -
                          div / 3.0;
        const double neq = -(b.tau - 1.0) * W64[k] * rho * 3.0 * qg;
        const double q = (C::SHIFT ? W64[k] * ((rho - 1.0) + rho * poly) : W64[k] * rho * (1.0 + poly)) + neq;
        const T v = C::st((typename C::R)q);
        ((T *)b.P[0][k])[ixy] = v;
        ((T *)b.P[1][k])[ixy] = v;
    }
}

// Streaming copy (the measured bandwidth ceiling the fused kernel is compared
// with): 16-byte loads, non-temporal 16-byte stores, 4 per thread, XCD-aware
// workgroup order as in the fused kernel.
// Streaming-copy probes (no reference counterpart), one 16-byte element per
// thread and stream, non-temporal loads and stores, XCD-aware workgroup order:
// the fastest of the copy shapes measured on MI355X (4 or 8 elements per
// thread, plain stores, a linear order and grid-stride loops were all slower).
// k_stream_copy moves one stream; k_plane_copy moves S planes of m elements
// into S planes (the fused kernel's pattern: 9 planes read, 9 written).
__global__ __launch_bounds__(256) void k_stream_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n16)
{
    typedef __attribute__((ext_vector_type(4))) unsigned int u32x4;
    const size_t i = (size_t)xcd_tile((int)blockIdx.x, (int)gridDim.x) * 256 + threadIdx.x;
    if (i >= n16) return;
    const u32x4 v = __builtin_nontemporal_load((const GLOBAL_AS u32x4 *)src + i);
    __builtin_nontemporal_store(v, (GLOBAL_AS u32x4 *)dst + i);
}

template <int S>
__global__ __launch_bounds__(256) void k_plane_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t m16)
{
    typedef __attribute__((ext_vector_type(4))) unsigned int u32x4;
    const size_t i = (size_t)xcd_tile((int)blockIdx.x, (int)gridDim.x) * 256 + threadIdx.x;
    if (i >= m16) return;
    u32x4 v[S];
#pragma unroll
    for (int s = 0; s < S; s++) v[s] = __builtin_nontemporal_load((const GLOBAL_AS u32x4 *)src + s * m16 + i);
#pragma unroll
    for (int s = 0; s < S; s++) __builtin_nontemporal_store(v[s], (GLOBAL_AS u32x4 *)dst + s * m16 + i);
}

// ---------------------------------------------------------------------------
// Batched cases in one persistent launch, one workgroup per case: the workgroup
// runs its case's coarse steps from a table of items in the reference order
// (stream-collide of a level, a sub-batch of interface ops, the spline lines, the
// drag), with a workgroup barrier between items. Cases are independent, so no
// workgroup waits for another. The device code of each item is the separate
// launches' own: the same per-node and per-entry arithmetic.
// ---------------------------------------------------------------------------
// The items of the case kernel (inlined: the launch's constants stay in scalar registers).
template <int MODE>
__device__ __forceinline__ void case_ops(const CaseItem &p, const LineOp *__restrict__ ops,
                                      const BlockDesc *__restrict__ descs)
{
    const int total = p.b * p.max_n;
    for (int g = (int)threadIdx.x; g < total; g += CASE_TPB) {
        const int o = g / p.max_n, e = g - o * p.max_n;
        const LineOp &op = ops[p.a + o];
        if (e < op.n) line_op_entry<MODE>(op, e, descs, p.pm, p.ms, p.xm);
    }
}

template <int MODE>
__device__ __forceinline__ void case_splines(const CaseItem &p, const LineOp *__restrict__ ops,
                                          const BlockDesc *__restrict__ descs, double *lds)
{
    for (int o = 0; o < p.b; o++) {
        spline_line<MODE>(ops[p.a + o], descs, p.pm, p.ms, lds);
        __syncthreads(); // the next line reuses the LDS
    }
}

template <int MODE, int VX>
__device__ __forceinline__ void case_collide(const CaseItem &p, const CaseArgs &a, int grp, int ng, int wv, int nwv)
{
    for (int t = grp; t < p.b; t += ng)
        sc_tile<MODE, VX, false, true, true>(a.descs, a.tile_blk, a.ndesc, a.rows, p.pm, p.a + t, wv, nwv);
}

template <int MODE> __device__ __forceinline__ void case_drag(const DragArgs &d, double *out, double *sh)
{
    drag_case<MODE>(d, out, sh);
}

template <int MODE, int VX> __global__ __launch_bounds__(CASE_TPB) void k_cases(CaseArgs a)
{
    extern __shared__ double lds[];
    __shared__ double sh[16];
    const int c = a.case_of_wg[blockIdx.x];
    const int tid = (int)threadIdx.x;
    // stream-collide tiles: CASE_TPB / tpb at once, one per group of tpb threads
    const int ng = CASE_TPB / a.tpb, grp = tid / a.tpb;
    const int wv = (tid - grp * a.tpb) >> 6, nwv = a.tpb >> 6;
    for (int s = 0; s < a.nsteps; s++) {
        const CaseItem *it = a.items + ((size_t)((a.step0 + s) % a.period) * a.ncases + c) * a.per_case;
        for (int k = 0; k < a.per_case; k++) {
            const CaseItem &p = it[k];
            if (p.kind == ITEM_COLLIDE) {
                case_collide<MODE, VX>(p, a, grp, ng, wv, nwv);
            } else if (p.kind == ITEM_OPS) {
                case_ops<MODE>(p, a.ops, a.descs);
            } else if (p.kind == ITEM_SPLINE) {
                case_splines<MODE>(p, a.ops, a.descs, lds);
            } else { // ITEM_DRAG
                double *out = a.hist + 2 * ((size_t)a.ncases * (size_t)((a.head0 + s) % a.cap) + (size_t)c);
                case_drag<MODE>(a.drag[p.a][c], out, sh);
            }
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
#define MBL_LAUNCH_CHECK()                                                                                             \
    do {                                                                                                               \
        if (hipGetLastError() != hipSuccess) return MBL_E_HIP;                                                         \
    } while (0)

template <int MODE, int VX>
static int sc_dispatch(hipStream_t s, int tpb, bool periodic, bool flags, bool macros, const BlockDesc *d,
                       const int *tb, int nd, int tiles, int rows, uint32_t pm, int groups, int tbase,
                       uint32_t rmask, double *rparts)
{
    dim3 g(groups > 0 && groups < tiles ? groups : tiles), b(tpb);
    if (periodic) {
        if (flags) {
            if (macros)
                hipLaunchKernelGGL((k_stream_collide<MODE, VX, true, true, true>), g, b, 0, s, d, tb, nd, rows, pm, tiles, tbase, 0u, nullptr);
            else
                hipLaunchKernelGGL((k_stream_collide<MODE, VX, true, true, false>), g, b, 0, s, d, tb, nd, rows, pm, tiles, tbase, 0u, nullptr);
        } else {
            if (macros)
                hipLaunchKernelGGL((k_stream_collide<MODE, VX, true, false, true>), g, b, 0, s, d, tb, nd, rows, pm, tiles, tbase, 0u, nullptr);
            else
                hipLaunchKernelGGL((k_stream_collide<MODE, VX, true, false, false>), g, b, 0, s, d, tb, nd, rows, pm, tiles, tbase, 0u, nullptr);
        }
    } else {
        if (macros) {
            if constexpr (VX == 1) {
                if (rmask)
                    hipLaunchKernelGGL((k_stream_collide<MODE, 1, false, true, true, true>), g, b, 0, s, d, tb, nd, rows, pm, tiles, tbase, rmask, rparts);
                else
                    hipLaunchKernelGGL((k_stream_collide<MODE, 1, false, true, true>), g, b, 0, s, d, tb, nd, rows, pm, tiles, tbase, 0u, nullptr);
            } else {
                if (rmask) return MBL_E_UNSUPPORTED;
                hipLaunchKernelGGL((k_stream_collide<MODE, VX, false, true, true>), g, b, 0, s, d, tb, nd, rows, pm, tiles, tbase, 0u, nullptr);
            }
        } else {
            if (rmask) return MBL_E_UNSUPPORTED;
            hipLaunchKernelGGL((k_stream_collide<MODE, VX, false, true, false>), g, b, 0, s, d, tb, nd, rows, pm, tiles, tbase, 0u, nullptr);
        }
    }
    MBL_LAUNCH_CHECK();
    return 0;
}

template <int MODE>
static int sc_mode(hipStream_t s, int vx, int tpb, bool periodic, bool flags, bool macros, const BlockDesc *d,
                   const int *tb, int nd, int tiles, int rows, uint32_t pm, int groups, int tbase, uint32_t rmask,
                   double *rparts)
{
    if (rmask && (periodic || !flags || !macros || vx != 1 || !rparts)) return MBL_E_UNSUPPORTED;
    switch (vx) {
    case 1: return sc_dispatch<MODE, 1>(s, tpb, periodic, flags, macros, d, tb, nd, tiles, rows, pm, groups, tbase, rmask, rparts);
    case 2: return sc_dispatch<MODE, 2>(s, tpb, periodic, flags, macros, d, tb, nd, tiles, rows, pm, groups, tbase, rmask, rparts);
    case 4: return sc_dispatch<MODE, 4>(s, tpb, periodic, flags, macros, d, tb, nd, tiles, rows, pm, groups, tbase, rmask, rparts);
    case 8: return sc_dispatch<MODE, 8>(s, tpb, periodic, flags, macros, d, tb, nd, tiles, rows, pm, groups, tbase, rmask, rparts);
    default: return MBL_E_INVALID;
    }
}

// Instantiates a templated launcher body once per storage mode of the build.
#ifdef MBL_FP64
#define MBL_MODE_SWITCH(mode, CALL)                                                                                    \
    switch (mode) {                                                                                                    \
    case ST_NATIVE: {                                                                                                  \
        constexpr int M_ = ST_NATIVE;                                                                                  \
        CALL;                                                                                                          \
    } break;                                                                                                           \
    default: return MBL_E_UNSUPPORTED;                                                                                 \
    }
#else
#define MBL_MODE_SWITCH(mode, CALL)                                                                                    \
    switch (mode) {                                                                                                    \
    case ST_NATIVE: {                                                                                                  \
        constexpr int M_ = ST_NATIVE;                                                                                  \
        CALL;                                                                                                          \
    } break;                                                                                                           \
    case ST_F32S: {                                                                                                    \
        constexpr int M_ = ST_F32S;                                                                                    \
        CALL;                                                                                                          \
    } break;                                                                                                           \
    case ST_F16S: {                                                                                                    \
        constexpr int M_ = ST_F16S;                                                                                    \
        CALL;                                                                                                          \
    } break;                                                                                                           \
    case ST_BF16S: {                                                                                                   \
        constexpr int M_ = ST_BF16S;                                                                                   \
        CALL;                                                                                                          \
    } break;                                                                                                           \
    case ST_F16C: {                                                                                                    \
        constexpr int M_ = ST_F16C;                                                                                    \
        CALL;                                                                                                          \
    } break;                                                                                                           \
    case ST_F16D: {                                                                                                    \
        constexpr int M_ = ST_F16D;                                                                                    \
        CALL;                                                                                                          \
    } break;                                                                                                           \
    default: return MBL_E_UNSUPPORTED;                                                                                 \
    }
#endif

int launch_stream_collide(void *stream, int mode, int vx, int tpb, bool periodic, bool flags, bool macros,
                          const BlockDesc *d_descs, const int *d_tile_blk, int ndesc, int total_tiles, int tile_rows,
                          uint32_t parity_mask, int groups, int tile_base, uint32_t rmask, double *rparts)
{
    if (total_tiles <= 0) return 0;
    if (tpb % 64 || tpb < 64 || tpb > 256 || tile_rows < tpb / 64 || tile_base < 0) return MBL_E_INVALID;
    hipStream_t s = (hipStream_t)stream;
    MBL_MODE_SWITCH(mode, return sc_mode<M_>(s, vx, tpb, periodic, flags, macros, d_descs, d_tile_blk, ndesc,
                                             total_tiles, tile_rows, parity_mask, groups, tile_base, rmask, rparts));
    return 0;
}

int launch_init_general(void *stream, int mode, const BlockDesc &d, const AREA &area, Lbd *x, Lbd *y, Lbd *forcex,
                        Lbd *forcey, int bc_mode, int init_mode, unsigned long long seed, double sigma, double LX,
                        double LY, double R, double CX, double CY, const Consts &c)
{
    InitArgs a;
    a.d = d;
    a.x = x;
    a.y = y;
    a.fx = forcex;
    a.fy = forcey;
    a.bc_mode = bc_mode;
    a.init_mode = init_mode;
    a.seed = seed;
    a.sigma = sigma;
    a.LX = LX;
    a.LY = LY;
    a.R = R;
    a.CX = CX;
    a.CY = CY;
    a.base_x = area.base_x;
    a.m = area.m;
    a.c = c;
    dim3 g((area.nx + 15) / 16, (area.ny + 15) / 16), b(256);
    hipStream_t s = (hipStream_t)stream;
    MBL_MODE_SWITCH(mode, hipLaunchKernelGGL((k_init_general<M_>), g, b, 0, s, a));
    MBL_LAUNCH_CHECK();
    return 0;
}

int launch_line_ops(void *stream, int mode, const LineOp *d_ops, int nops, int max_n, const BlockDesc *d_blocks,
                    uint32_t parity_mask, const MarkerSet &ms, double xm)
{
    if (nops <= 0 || max_n <= 0) return 0;
    dim3 g((max_n + 255) / 256, nops), b(256);
    hipStream_t s = (hipStream_t)stream;
    MBL_MODE_SWITCH(mode, hipLaunchKernelGGL((k_line_ops<M_>), g, b, 0, s, d_ops, d_blocks, parity_mask, ms, xm));
    MBL_LAUNCH_CHECK();
    return 0;
}

int launch_collide_ops(void *stream, int mode, int tpb, const CollideOpsArgs &args, int max_n)
{
    if (args.tiles <= 0 || args.nops <= 0 || max_n <= 0) return MBL_E_INVALID;
    if (tpb % 64 || tpb < 64 || tpb > 256 || args.rows < tpb / 64 || args.tile_base < 0) return MBL_E_INVALID;
    CollideOpsArgs a = args;
    a.opx = (max_n + tpb - 1) / tpb;
    dim3 g(a.tiles + a.opx * a.nops), b(tpb);
    hipStream_t s = (hipStream_t)stream;
    if (a.rmask) {
        if (!a.rparts) return MBL_E_INVALID;
        MBL_MODE_SWITCH(mode, hipLaunchKernelGGL((k_collide_ops<M_, true>), g, b, 0, s, a));
    } else {
        MBL_MODE_SWITCH(mode, hipLaunchKernelGGL((k_collide_ops<M_, false>), g, b, 0, s, a));
    }
    MBL_LAUNCH_CHECK();
    return 0;
}

int launch_ring_fill(void *stream, int mode, const LineOp *d_ops, int nops, int max_n, const BlockDesc *d_blocks,
                     uint32_t parity_mask)
{
    if (nops <= 0 || max_n <= 0) return 0;
    dim3 g((max_n + 255) / 256, nops), b(256);
    hipStream_t s = (hipStream_t)stream;
    MBL_MODE_SWITCH(mode, hipLaunchKernelGGL((k_ring_fill<M_>), g, b, 0, s, d_ops, d_blocks, parity_mask));
    MBL_LAUNCH_CHECK();
    return 0;
}

template <int MODE>
static int spline_solve_r(hipStream_t s, int R, const LineOp *ops, int nops, size_t lds, const BlockDesc *blocks,
                          uint32_t pm)
{
    dim3 g(12, nops), b(SPL_TPB);
    switch (R) {
    case 1: hipLaunchKernelGGL((k_spline_solve<MODE, 1>), g, b, lds, s, ops, blocks, pm); break;
    case 2: hipLaunchKernelGGL((k_spline_solve<MODE, 2>), g, b, lds, s, ops, blocks, pm); break;
    case 4: hipLaunchKernelGGL((k_spline_solve<MODE, 4>), g, b, lds, s, ops, blocks, pm); break;
    case 8: hipLaunchKernelGGL((k_spline_solve<MODE, 8>), g, b, lds, s, ops, blocks, pm); break;
    default: return MBL_E_UNSUPPORTED;
    }
    return 0;
}

int launch_spline_path(int max_nc) { return max_nc <= SPL_FUSED_MAX_NC ? 0 : 1; }

size_t cases_lds(int max_nc) { return 24 * sizeof(double) * (size_t)(max_nc > 0 ? max_nc : 0); }

int cases_spline_ok(int max_nc) { return max_nc <= SPL_FUSED_MAX_NC && max_nc <= CASE_TPB; }

template <int MODE> static int cases_vx(hipStream_t s, int vx, const CaseArgs &a, size_t lds, int grid)
{
    switch (vx) {
    case 1: hipLaunchKernelGGL((k_cases<MODE, 1>), dim3(grid), dim3(CASE_TPB), lds, s, a); break;
    case 2: hipLaunchKernelGGL((k_cases<MODE, 2>), dim3(grid), dim3(CASE_TPB), lds, s, a); break;
    case 4: hipLaunchKernelGGL((k_cases<MODE, 4>), dim3(grid), dim3(CASE_TPB), lds, s, a); break;
    case 8: hipLaunchKernelGGL((k_cases<MODE, 8>), dim3(grid), dim3(CASE_TPB), lds, s, a); break;
    default: return MBL_E_UNSUPPORTED;
    }
    return 0;
}

int launch_cases(void *stream, int mode, int vx, const CaseArgs &a, size_t lds)
{
    if (a.ncases <= 0 || a.per_case <= 0 || a.period <= 0 || a.step0 < 0 || a.step0 >= a.period || a.nsteps < 0 ||
        a.cap <= 0 || a.tpb < 64 ||
        a.tpb % 64 || CASE_TPB % a.tpb || a.rows < a.tpb / 64 || lds > 64 * 1024)
        return MBL_E_INVALID;
    if (a.nsteps == 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    int rc = 0;
    MBL_MODE_SWITCH(mode, rc = cases_vx<M_>(s, vx, a, lds, a.ncases));
    if (rc) return rc;
    MBL_LAUNCH_CHECK();
    return 0;
}

int launch_spline(void *stream, int mode, const LineOp *d_ops, int nops, int max_nc, int max_nf,
                  const BlockDesc *d_blocks, uint32_t parity_mask, const MarkerSet &ms)
{
    if (nops <= 0) return 0;
    if (max_nc > SPL_MAX_NC) return MBL_E_UNSUPPORTED;
    if (max_nc <= SPL_FUSED_MAX_NC) {
        // one thread per coarse row, and per fine point where the workgroup allows (up to 1024)
        const int threads = std::min(1024, (std::max(max_nc, max_nf) + 63) / 64 * 64);
        const int ng = spline_groups(nops);
        const size_t lds = 2 * (3 + 9 / ng) * sizeof(double) * (size_t)max_nc;
        hipStream_t s = (hipStream_t)stream;
        const dim3 g(ng, nops), b(threads);
        if (ng == 9) {
            MBL_MODE_SWITCH(mode, hipLaunchKernelGGL((k_spline_fused<M_, 9>), g, b, lds, s, d_ops, d_blocks, parity_mask, ms));
        } else {
            MBL_MODE_SWITCH(mode, hipLaunchKernelGGL((k_spline_fused<M_, 3>), g, b, lds, s, d_ops, d_blocks, parity_mask, ms));
        }
        MBL_LAUNCH_CHECK();
        return 0;
    }
    int R = 1;
    while (R * SPL_TPB < max_nc) R <<= 1;
    const size_t lds = 4 * sizeof(double) * (size_t)max_nc;
    hipStream_t s = (hipStream_t)stream;
    int rc = 0;
    MBL_MODE_SWITCH(mode, rc = spline_solve_r<M_>(s, R, d_ops, nops, lds, d_blocks, parity_mask));
    if (rc) return rc;
    MBL_LAUNCH_CHECK();
    dim3 g((max_nf + 255) / 256, nops), b(256);
    MBL_MODE_SWITCH(mode, hipLaunchKernelGGL((k_spline_eval<M_>), g, b, 0, s, d_ops, d_blocks, parity_mask, ms));
    MBL_LAUNCH_CHECK();
    return 0;
}

int launch_iface_mixed(void *stream, int mode, const LineOp *d_ops, int nspl, int nops, int max_nc, int max_nf,
                       int max_n, const BlockDesc *d_blocks, uint32_t parity_mask, const MarkerSet &ms, double xm)
{
    if (nspl <= 0 || max_nc > SPL_FUSED_MAX_NC) return MBL_E_INVALID;
    const int threads = std::min(1024, (std::max(max_nc, max_nf) + 63) / 64 * 64);
    const int ng = spline_groups(nspl);
    const size_t lds = 2 * (3 + 9 / ng) * sizeof(double) * (size_t)max_nc;
    dim3 g(std::max(ng, (max_n + threads - 1) / threads), nspl + std::max(0, nops)), b(threads);
    hipStream_t s = (hipStream_t)stream;
    if (ng == 9) {
        MBL_MODE_SWITCH(mode, hipLaunchKernelGGL((k_iface_mixed<M_, 9>), g, b, lds, s, d_ops, nspl, d_blocks, parity_mask,
                                                 ms, xm));
    } else {
        MBL_MODE_SWITCH(mode, hipLaunchKernelGGL((k_iface_mixed<M_, 3>), g, b, lds, s, d_ops, nspl, d_blocks, parity_mask,
                                                 ms, xm));
    }
    MBL_LAUNCH_CHECK();
    return 0;
}

int launch_tri_A(void *stream, Lbd *l, Lbd *m, Lbd *u, int N)
{
    if (N <= 0) return 0;
    hipLaunchKernelGGL(k_tri_A, dim3((N + 255) / 256), dim3(256), 0, (hipStream_t)stream, l, m, u, N);
    MBL_LAUNCH_CHECK();
    return 0;
}

int launch_residual(void *stream, const BlockDesc *d_descs, const int *d_tblk, int ndesc, int total_rtiles,
                    double *partials, int mode, double umax, double LY, double *d_out)
{
    if (total_rtiles <= 0 || ndesc <= 0) return MBL_E_INVALID;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_residual_tiles, dim3(total_rtiles), dim3(256), 0, s, d_descs, d_tblk, ndesc, partials, mode,
                       umax, LY);
    MBL_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_residual_final, dim3(1), dim3(1024), 0, s, total_rtiles, (const double *)partials, d_out);
    MBL_LAUNCH_CHECK();
    return 0;
}

int launch_residual_cases(void *stream, const BlockDesc *d_descs, const int *d_tblk, int ndesc, int total_rtiles,
                          double *partials, int ncases, const int *d_case_tiles, double *d_out)
{
    if (total_rtiles <= 0 || ncases <= 0) return MBL_E_INVALID;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_residual_tiles, dim3(total_rtiles), dim3(256), 0, s, d_descs, d_tblk, ndesc, partials, 0, 0.0,
                       0.0);
    MBL_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_residual_cases, dim3(ncases), dim3(256), 0, s, d_case_tiles, (const double *)partials, d_out);
    MBL_LAUNCH_CHECK();
    return 0;
}

int launch_drag(void *stream, int mode, const DragArgs &a)
{
    hipStream_t s = (hipStream_t)stream;
    MBL_MODE_SWITCH(mode, hipLaunchKernelGGL((k_drag<M_>), dim3(1), dim3(1024), 0, s, a));
    MBL_LAUNCH_CHECK();
    return 0;
}

int launch_drag_batch(void *stream, int mode, const DragArgs *d_args, int ncases, double *out)
{
    if (ncases <= 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    MBL_MODE_SWITCH(mode, hipLaunchKernelGGL((k_drag_batch<M_>), dim3(ncases), dim3(256), 0, s, d_args, out));
    MBL_LAUNCH_CHECK();
    return 0;
}

int launch_step_diag(void *stream, int mode, const StepDiagArgs &a)
{
    if (a.list_wgs < 1 || a.ntiles < 0 || !a.rparts || !a.slot_rd || !a.slot_wr || a.slot_rd == a.slot_wr ||
        a.cap < 1 || !a.lbase || a.lstride < (size_t)a.list_wgs || (a.drag.n > 0 && !a.drag_base))
        return MBL_E_INVALID;
    const int grid = a.list_wgs + (a.drag.n > 0 ? 1 : 0);
    hipStream_t s = (hipStream_t)stream;
    MBL_MODE_SWITCH(mode, hipLaunchKernelGGL((k_step_diag<M_>), dim3(grid), dim3(256), 0, s, a));
    MBL_LAUNCH_CHECK();
    return 0;
}

int launch_ops_diag(void *stream, int mode, const LineOp *d_ops, int nops, int max_n, const BlockDesc *d_blocks,
                    uint32_t parity_mask, const MarkerSet &ms, double xm, const StepDiagArgs &a)
{
    if (nops <= 0 || max_n <= 0 || a.list_wgs < 1 || a.ntiles < 0 || !a.rparts || !a.slot_rd || !a.slot_wr ||
        a.slot_rd == a.slot_wr || a.cap < 1 || !a.lbase || a.lstride < (size_t)a.list_wgs ||
        (a.drag.n > 0 && !a.drag_base))
        return MBL_E_INVALID;
    const int opx = (max_n + 255) / 256;
    dim3 g(nops * opx + a.list_wgs + (a.drag.n > 0 ? 1 : 0)), b(256);
    hipStream_t s = (hipStream_t)stream;
    MBL_MODE_SWITCH(mode, hipLaunchKernelGGL((k_ops_diag<M_>), g, b, 0, s, d_ops, nops, opx, d_blocks, parity_mask, ms,
                                             xm, a));
    MBL_LAUNCH_CHECK();
    return 0;
}

int launch_residual_steps(void *stream, int nsteps, int nparts, size_t stride, const double *partials, double *out)
{
    if (nsteps <= 0) return 0;
    if (nparts < 1 || stride < (size_t)nparts || !partials || !out) return MBL_E_INVALID;
    hipLaunchKernelGGL(k_residual_steps, dim3(nsteps), dim3(1024), 0, (hipStream_t)stream, nparts, stride, partials,
                       out);
    MBL_LAUNCH_CHECK();
    return 0;
}

int launch_moments(void *stream, int mode, const void *const *d_planes, size_t n, int nparts, double *partials,
                   double *d_out)
{
    if (n == 0 || nparts < 1) return MBL_E_INVALID;
    hipStream_t s = (hipStream_t)stream;
    MBL_MODE_SWITCH(mode, hipLaunchKernelGGL((k_moments<M_>), dim3(nparts), dim3(256), 0, s, d_planes, n, partials));
    MBL_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_moments_final, dim3(1), dim3(256), 0, s, nparts, (const double *)partials, d_out);
    MBL_LAUNCH_CHECK();
    return 0;
}

int launch_decode(void *stream, int mode, const void *plane, int k, size_t n, double *out)
{
    if (n == 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    dim3 g((unsigned)((n + 255) / 256)), b(256);
    MBL_MODE_SWITCH(mode, hipLaunchKernelGGL((k_decode<M_>), g, b, 0, s, plane, k, n, out));
    MBL_LAUNCH_CHECK();
    return 0;
}

int launch_encode(void *stream, int mode, void *plane, int k, size_t n, const double *in)
{
    if (n == 0) return 0;
    hipStream_t s = (hipStream_t)stream;
    dim3 g((unsigned)((n + 255) / 256)), b(256);
    MBL_MODE_SWITCH(mode, hipLaunchKernelGGL((k_encode<M_>), g, b, 0, s, plane, k, n, in));
    MBL_LAUNCH_CHECK();
    return 0;
}

// Rows per paired tile (the step-n+1 values of one tile in LDS stay within 48 KiB) and threads
// per workgroup: 16 rows of 16-bit and 4-byte planes, 8 rows of 8-byte planes; 512 threads for
// 4-byte planes (16384^2 FP32: 130k-133k MLUPS against 121k-130k with 256; 8 rows 123k-127k),
// 256 for 16-bit planes (1-3-12: 165k against 163k with 512).
static int periodic2_rows(int mode)
{
    return store_elem_bytes(mode) > 4 ? 8 : 16;
}

static int periodic2_threads(int mode)
{
    return store_elem_bytes(mode) == 4 ? 512 : 256;
}

static bool periodic2_strip_fits(int nx, int ny) { return nx >= 64 * P2S_W && ny >= P2S_H && ny % P2S_H == 0; }

static bool periodic2_tile_fits(int mode, int nx, int ny)
{
    const int ty = periodic2_rows(mode);
    return nx >= P2_TX && ny >= ty && nx % P2_TX == 0 && ny % ty == 0;
}

int periodic2_fits(int mode, int nx, int ny, int form)
{
    return (form == 2 && periodic2_strip_fits(nx, ny)) || periodic2_tile_fits(mode, nx, ny);
}

// three steps per launch for 4- and 8-byte planes: 16384^2 FP32 164.6k-166.0k MLUPS against
// 123.0k-137.4k for pairs (same box), 8192^2 FP64 99.6k-100.3k against 71.8k-73.0k; 16-bit planes
// stay with pairs (1-3-12: 150k against 164k-166k, their update is VALU-bound)
int periodic3_fits(int mode, int nx, int ny)
{
    return store_elem_bytes(mode) >= 4 && periodic2_tile_fits(mode, nx, ny);
}

int launch_periodic3(void *stream, int mode, const Periodic2Args &args)
{
    if (!periodic3_fits(mode, args.nx, args.ny)) return MBL_E_INVALID;
    Periodic2Args a = args;
    const int ty = periodic2_rows(mode);
    a.tiles_x = a.nx / P2_TX;
    dim3 g(a.tiles_x * (a.ny / ty)), b(P3_THREADS);
    hipStream_t s = (hipStream_t)stream;
    if (ty == 8) {
        MBL_MODE_SWITCH(mode, hipLaunchKernelGGL((k_periodic3<M_, 8>), g, b, 0, s, a));
    } else {
        MBL_MODE_SWITCH(mode, hipLaunchKernelGGL((k_periodic3<M_, 16>), g, b, 0, s, a));
    }
    MBL_LAUNCH_CHECK();
    return 0;
}

int launch_periodic2(void *stream, int mode, const Periodic2Args &args)
{
    if (!periodic2_fits(mode, args.nx, args.ny, args.form)) return MBL_E_INVALID;
    if (args.form == 2 && periodic2_strip_fits(args.nx, args.ny)) {
        Periodic2Args a = args;
        a.tiles_x = (a.nx + 64 * P2S_W - 3) / (64 * P2S_W - 2);
        dim3 g(a.tiles_x * (a.ny / P2S_H)), b(64 * P2S_W);
        hipStream_t s = (hipStream_t)stream;
        MBL_MODE_SWITCH(mode, hipLaunchKernelGGL((k_periodic2s<M_>), g, b, 0, s, a));
        MBL_LAUNCH_CHECK();
        return 0;
    }
    Periodic2Args a = args;
    const int ty = periodic2_rows(mode), nt = periodic2_threads(mode);
    a.tiles_x = a.nx / P2_TX;
    dim3 g(a.tiles_x * (a.ny / ty)), b(nt);
    hipStream_t s = (hipStream_t)stream;
    if (ty == 8) {
        MBL_MODE_SWITCH(mode, hipLaunchKernelGGL((k_periodic2<M_, 8, 256>), g, b, 0, s, a));
    } else if (nt == 512) {
        MBL_MODE_SWITCH(mode, hipLaunchKernelGGL((k_periodic2<M_, 16, 512>), g, b, 0, s, a));
    } else {
        MBL_MODE_SWITCH(mode, hipLaunchKernelGGL((k_periodic2<M_, 16, 256>), g, b, 0, s, a));
    }
    MBL_LAUNCH_CHECK();
    return 0;
}

int launch_stream_copy(void *stream, const void *src, void *dst, size_t bytes)
{
    const size_t n16 = bytes / 16;
    if (n16 == 0 || (n16 + 255) / 256 > 0x7fffffffu) return MBL_E_INVALID;
    hipLaunchKernelGGL(k_stream_copy, dim3((unsigned)((n16 + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       (const uint4 *)src, (uint4 *)dst, n16);
    MBL_LAUNCH_CHECK();
    return 0;
}

int launch_plane_copy(void *stream, const void *src, void *dst, size_t plane_bytes)
{
    const size_t m16 = plane_bytes / 16;
    if (m16 == 0 || (m16 + 255) / 256 > 0x7fffffffu) return MBL_E_INVALID;
    hipLaunchKernelGGL(k_plane_copy<MBL_Q>, dim3((unsigned)((m16 + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       (const uint4 *)src, (uint4 *)dst, m16);
    MBL_LAUNCH_CHECK();
    return 0;
}

int launch_raster(void *stream, const RasterBlock *d_blocks, const RasterCase *d_cases, int ncases, int H, int W,
                  float *d_out)
{
    if (ncases <= 0 || H < 2 || W < 2) return MBL_E_INVALID;
    hipStream_t s = (hipStream_t)stream;
    dim3 g((unsigned)((H * W + 255) / 256), ncases), b(256);
    hipLaunchKernelGGL(k_raster, g, b, 0, s, d_blocks, d_cases, H, W, d_out);
    MBL_LAUNCH_CHECK();
    return 0;
}

int launch_init_raster(void *stream, int mode, const RasterBlock *d_blocks, int nblocks, size_t max_nodes,
                       const RasterCase *d_cases, int H, int W, const float *d_fields)
{
    if (nblocks <= 0 || H < 2 || W < 2) return MBL_E_INVALID;
    hipStream_t s = (hipStream_t)stream;
    dim3 g((unsigned)((max_nodes + 255) / 256), nblocks), b(256);
    MBL_MODE_SWITCH(mode, hipLaunchKernelGGL((k_init_raster<M_>), g, b, 0, s, d_blocks, d_cases, H, W, d_fields));
    MBL_LAUNCH_CHECK();
    return 0;
}

} // namespace mbl
