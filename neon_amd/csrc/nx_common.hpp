// nx_common.hpp -- device-side building blocks shared by the dense (dGrid)
// and block-sparse (bGrid) D3Q19 kernels for gfx950.
//
// The collision restates benchmarks/lbm-lid-driven-cavity-flow/src/
// LbmTools.h:171-282 operation by operation; the library is compiled with
// -ffp-contract=off so that binary64 results match a CPU evaluation of the
// same expression tree bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "neon_amd.h"

#pragma clang fp contract(off)

namespace nx {

// ---- lattice (D3Q19.h:23-132) --------------------------------------------
constexpr int Q = 19;
constexpr int CX[Q] = {-1, 0, 0, -1, -1, -1, -1, 0, 0, 0, 1, 0, 0, 1, 1, 1, 1, 0, 0};
constexpr int CY[Q] = {0, -1, 0, -1, 1, 0, 0, -1, -1, 0, 0, 1, 0, 1, -1, 0, 0, 1, 1};
constexpr int CZ[Q] = {0, 0, -1, 0, 0, -1, 1, -1, 1, 0, 0, 0, 1, 0, 0, 1, -1, 1, -1};
constexpr int OPP[Q] = {10, 11, 12, 13, 14, 15, 16, 17, 18, 9, 0, 1, 2, 3, 4, 5, 6, 7, 8};
__host__ __device__ constexpr double wq(int q)
{
    return q == 9 ? 1. / 3. : ((q % 10) < 3 ? 1. / 18. : 1. / 36.);
}

constexpr uint32_t kInvalidBlock = 0xFFFFFFFFu;

__device__ __forceinline__ int cls_of(uint32_t fl) { return (int)((fl >> NX_CLASS_SHIFT) & 3u); }

// ---- storage codecs -------------------------------------------------------
// P is the precision in which the pulled value is formed (the reference adds
// own + neighbour on a wall link in LbmStoreType, LbmTools.h:83-84).
// Bulk populations are handled in a working representation: f itself for
// F64 / F32 storage, the deviation d = f - w_q for the deviation storages
// (whose collision is written in d, so that FP32 arithmetic rounds relative
// to the deviations rather than to w_q; project-defined).  Wall cells always
// hold f.  A wall link adds own + wall in either representation (w_q = w_opp).
//   bulk(v, q)    stored bulk value -> working value
//   raw(v)        stored wall value -> P
//   full(v, q)    stored bulk value -> f in binary64 (mass sums)
//   enc(v, q)     working value -> stored
//   enc_full(f,q) f -> stored bulk value (fill, equilibrium start); the
//                 deviation is formed in binary64 and rounded through binary32
template <int S>
struct Stor;

template <>
struct Stor<NX_F64> {
    using T = double;
    using P = double;
    static constexpr bool kDev = false;
    static __device__ __forceinline__ P bulk(T v, int) { return v; }
    static __device__ __forceinline__ P raw(T v) { return v; }
    static __device__ __forceinline__ double full(T v, int) { return v; }
    template <typename R>
    static __device__ __forceinline__ T enc(R v, int)
    {
        return (T)v;
    }
    static __device__ __forceinline__ T enc_full(double f, int) { return f; }
};

template <>
struct Stor<NX_F32> {
    using T = float;
    using P = float;
    static constexpr bool kDev = false;
    static __device__ __forceinline__ P bulk(T v, int) { return v; }
    static __device__ __forceinline__ P raw(T v) { return v; }
    static __device__ __forceinline__ double full(T v, int) { return (double)v; }
    template <typename R>
    static __device__ __forceinline__ T enc(R v, int)
    {
        return (T)v;
    }
    static __device__ __forceinline__ T enc_full(double f, int) { return (float)f; }
};

// d = f - w_q in binary32 for bulk cells, f for wall cells
template <>
struct Stor<NX_F32DEV> {
    using T = float;
    using P = float;
    static constexpr bool kDev = true;
    static __device__ __forceinline__ P bulk(T v, int) { return v; }
    static __device__ __forceinline__ P raw(T v) { return v; }
    static __device__ __forceinline__ double full(T v, int q) { return (double)(float)wq(q) + (double)v; }
    template <typename R>
    static __device__ __forceinline__ T enc(R v, int)
    {
        return (T)v;
    }
    static __device__ __forceinline__ T enc_full(double f, int q) { return (float)(f - wq(q)); }
};

// d = f - w_q in binary16 (v_cvt_f16_f32, round to nearest even) for bulk cells
template <>
struct Stor<NX_F16DEV> {
    using T = _Float16;
    using P = float;
    static constexpr bool kDev = true;
    static __device__ __forceinline__ P bulk(T v, int) { return (float)v; }
    static __device__ __forceinline__ P raw(T v) { return (float)v; }
    static __device__ __forceinline__ double full(T v, int q)
    {
        return (double)(float)wq(q) + (double)(float)v;
    }
    template <typename R>
    static __device__ __forceinline__ T enc(R v, int)
    {
        // round to binary32 first, then to binary16: the opaque value keeps the
        // compiler from fusing the producing FMA and this conversion into one
        // v_fma_mixlo_f16 (a single rounding, different results)
        float f = (float)v;
        asm volatile("" : "+v"(f));
        return (_Float16)f;
    }
    static __device__ __forceinline__ T enc_full(double f, int q)
    {
        return (_Float16)(float)(f - wq(q));
    }
};

template <int S, int CM>
struct Real {
    using R = float;
};
template <int CM>
struct Real<NX_F64, CM> {
    using R = double;
};
template <>
struct Real<NX_F32, NX_COMPUTE_F64> {
    using R = double;
};

// ---- macroscopic + BGK (LbmTools.h:171-282) -------------------------------
template <typename R>
__device__ __forceinline__ void macroscopic(const R f[Q], R& rho, R& u0, R& u1, R& u2)
{
    const R XM1 = f[0] + f[3] + f[4] + f[5] + f[6];
    const R XP1 = f[10] + f[13] + f[14] + f[15] + f[16];
    const R X0 = f[9] + f[1] + f[2] + f[7] + f[8] + f[11] + f[12] + f[17] + f[18];
    const R YM1 = f[1] + f[3] + f[7] + f[8] + f[14];
    const R YP1 = f[4] + f[11] + f[13] + f[17] + f[18];
    const R ZM1 = f[2] + f[5] + f[7] + f[16] + f[18];
    const R ZP1 = f[6] + f[8] + f[12] + f[15] + f[17];
    rho = XM1 + XP1 + X0;
    u0 = (XP1 - XM1) / rho;
    u1 = (YP1 - YM1) / rho;
    u2 = (ZP1 - ZM1) / rho;
}

// Same expression tree in one precision R (binary64 = the reference's
// double/double path; binary32 = pure FP32 collision).
template <typename R>
__device__ __forceinline__ void bgk(const R f[Q], R omega, R out[Q])
{
    R rho, u0, u1, u2;
    macroscopic<R>(f, rho, u0, u1, u2);
    const R usqr = R(1.5) * (u0 * u0 + u1 * u1 + u2 * u2);
    const R ck[9] = {u0, u1, u2, u0 + u1, u0 - u1, u0 + u2, u0 - u2, u1 + u2, u1 - u2};
#pragma unroll
    for (int i = 0; i < 9; i++) {
        const R wi = i < 3 ? R(1. / 18.) : R(1. / 36.);
        const R eq = rho * wi * (R(1.) - R(3.) * ck[i] + R(4.5) * ck[i] * ck[i] - usqr);
        const R eqopp = eq + rho * wi * R(6.) * ck[i];
        out[i] = (R(1.) - omega) * f[i] + omega * eq;
        out[i + 10] = (R(1.) - omega) * f[i + 10] + omega * eqopp;
    }
    const R eq9 = rho * R(1. / 3.) * (R(1.) - usqr);
    out[9] = (R(1.) - omega) * f[9] + omega * eq9;
}

// The reference's float/float instantiation: double literals promote the
// eq / eqopp / out expressions to binary64; `omega * eq` is a float product.
__device__ __forceinline__ void bgk_ref(const float f[Q], float omega, float out[Q])
{
    float rho, u0, u1, u2;
    macroscopic<float>(f, rho, u0, u1, u2);
    const float usqr = (float)(1.5 * (double)(u0 * u0 + u1 * u1 + u2 * u2));
    const float ck[9] = {u0, u1, u2, u0 + u1, u0 - u1, u0 + u2, u0 - u2, u1 + u2, u1 - u2};
#pragma unroll
    for (int i = 0; i < 9; i++) {
        const double wi = i < 3 ? (1. / 18.) : (1. / 36.);
        const double c = (double)ck[i];
        const float eq = (float)((double)rho * wi * (1. - 3. * c + 4.5 * c * c - (double)usqr));
        const float eqopp = (float)((double)eq + (double)rho * wi * 6. * c);
        out[i] = (float)((1. - (double)omega) * (double)f[i] + (double)(omega * eq));
        out[i + 10] = (float)((1. - (double)omega) * (double)f[i + 10] + (double)(omega * eqopp));
    }
    const float eq9 = (float)((double)rho * (1. / 3.) * (1. - (double)usqr));
    out[9] = (float)((1. - (double)omega) * (double)f[9] + (double)(omega * eq9));
}

// Zou-He boundary cell, streamAndZouhe + collideBgkUnrolled
// (benchmarks/lbm-flow-over-sphere/src/LbmContainers.h:145-287, :351-443).
// face 1..6 = unknowns from -x, +x, -y, +y, -z, +z; the unknown and middle
// direction lists (and so the summation orders) are the reference's tables
// (:633-646).  f holds the pulled populations; the 5 unknown entries are
// replaced.  Density face: rho = value, u.n = 1 - (M + 2K) / rho, written as
// the reference does, u_axis = ((M + 2K) / rho - 1) * (+1 if the unknowns
// point to -axis, else -1).  Velocity face (textbook form): u.n = value,
// rho = (M + 2K) / (1 - u.n).
constexpr int kZhUnk[6][5] = {{10, 13, 14, 15, 16}, {0, 3, 4, 5, 6},   {11, 4, 13, 17, 18},
                              {1, 3, 7, 8, 14},     {12, 6, 8, 15, 17}, {2, 5, 7, 16, 18}};
constexpr int kZhMid[3][4] = {{1, 2, 7, 8}, {0, 2, 5, 6}, {0, 1, 3, 4}};

// One face, every direction index a compile-time constant (f stays in
// registers).
template <typename R, int FACE>
__device__ __forceinline__ void zouhe_face(R f[Q], bool velocity, R value, R omega, R out[Q])
{
    constexpr int AXIS = (FACE - 1) >> 1;
    R known = R(0), middle = R(0);
#pragma unroll
    for (int k = 0; k < 5; k++) {
        const int iu = kZhUnk[FACE - 1][k];
        known += f[iu < 9 ? iu + 10 : iu - 10];
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
        middle += f[kZhMid[AXIS][k]];
        middle += f[kZhMid[AXIS][k] + 10];
    }
    middle += f[9];
    const R sgn = kZhUnk[FACE - 1][0] < 9 ? R(1) : R(-1);
    R rho, ua;
    if (velocity) {
        ua = -sgn * value;
        rho = (middle + known * R(2)) / (R(1) - value);
    } else {
        rho = value;
        ua = (((middle + known * R(2)) / rho) - R(1)) * sgn;
    }
    const R u0 = AXIS == 0 ? ua : R(0), u1 = AXIS == 1 ? ua : R(0), u2 = AXIS == 2 ? ua : R(0);
    const R usqr = R(1.5) * (u0 * u0 + u1 * u1 + u2 * u2);
    const R ck[9] = {u0, u1, u2, u0 + u1, u0 - u1, u0 + u2, u0 - u2, u1 + u2, u1 - u2};
    R eq[Q];
#pragma unroll
    for (int i = 0; i < 9; i++) {
        const R wi = i < 3 ? R(1. / 18.) : R(1. / 36.);
        eq[i] = rho * wi * (R(1.) - R(3.) * ck[i] + R(4.5) * ck[i] * ck[i] - usqr);
        eq[i + 10] = eq[i] + rho * wi * R(6.) * ck[i];
    }
    eq[9] = rho * R(1. / 3.) * (R(1.) - usqr);
#pragma unroll
    for (int k = 0; k < 5; k++) {
        const int iu = kZhUnk[FACE - 1][k], ik = iu < 9 ? iu + 10 : iu - 10;
        f[iu] = f[ik] + eq[iu] - eq[ik];
    }
#pragma unroll
    for (int q = 0; q < Q; q++) out[q] = (R(1.) - omega) * f[q] + omega * eq[q];
}

template <typename R>
__device__ __forceinline__ void zouhe_bgk(R f[Q], int face, bool velocity, R value, R omega, R out[Q])
{
    switch (face) {
        case 1: zouhe_face<R, 1>(f, velocity, value, omega, out); break;
        case 2: zouhe_face<R, 2>(f, velocity, value, omega, out); break;
        case 3: zouhe_face<R, 3>(f, velocity, value, omega, out); break;
        case 4: zouhe_face<R, 4>(f, velocity, value, omega, out); break;
        case 5: zouhe_face<R, 5>(f, velocity, value, omega, out); break;
        default: zouhe_face<R, 6>(f, velocity, value, omega, out); break;
    }
}

// BGK with Guo et al. (2002) forcing (no reference counterpart), body
// acceleration g, force density rho*g:
//   u = sum c f / rho + g/2 ;  out = (1-omega) f + omega feq(rho,u)
//       + (1-omega/2) w rho [3 (c-u).g + 9 (c.u)(c.g)]
template <typename R>
__device__ __forceinline__ void bgk_forced(const R f[Q], R omega, R g0, R g1, R g2, R out[Q])
{
    R rho, u0, u1, u2;
    macroscopic<R>(f, rho, u0, u1, u2);
    u0 += R(0.5) * g0;
    u1 += R(0.5) * g1;
    u2 += R(0.5) * g2;
    const R usqr = R(1.5) * (u0 * u0 + u1 * u1 + u2 * u2);
    const R ug = u0 * g0 + u1 * g1 + u2 * g2;
    const R fpre = R(1) - R(0.5) * omega;
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const R w = R(wq(q));
        const R cu = R(CX[q]) * u0 + R(CY[q]) * u1 + R(CZ[q]) * u2;
        const R cg = R(CX[q]) * g0 + R(CY[q]) * g1 + R(CZ[q]) * g2;
        const R eq = rho * w * (R(1) + R(3) * cu + R(4.5) * cu * cu - usqr);
        const R src = fpre * w * rho * (R(3) * (cg - ug) + R(9) * cu * cg);
        out[q] = (R(1) - omega) * f[q] + omega * eq + src;
    }
}

// Deviation form (project-defined): d = f - w_q.  rho = 1 + sum d,
// u = sum c d * (1 / rho), and feq - w_q = w_q (rho - 1) + w_q rho (3 c.u
// + 4.5 (c.u)^2 - 1.5 u^2): the same BGK as bgk<R>, every term small.  The
// expression is written with one reciprocal and explicit fused multiply-adds
// (the oracle restates the same operations with fmaf):
//   poly = fma(c, fma(4.5, c, -3), -usqr)  (for c = c_q.u with the reference's
//          sign convention of the first nine directions), deq = fma(rho w,
//          poly, w drho), deq_opp = fma(6 rho w, c, deq),
//   d' = fma(omega, deq - d, d).
__device__ __forceinline__ float fmaR(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmaR(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <typename R>
__device__ __forceinline__ void macroscopic_dev(const R d[Q], R& drho, R& rho, R& u0, R& u1, R& u2)
{
    const R XM1 = d[0] + d[3] + d[4] + d[5] + d[6];
    const R XP1 = d[10] + d[13] + d[14] + d[15] + d[16];
    const R X0 = d[9] + d[1] + d[2] + d[7] + d[8] + d[11] + d[12] + d[17] + d[18];
    const R YM1 = d[1] + d[3] + d[7] + d[8] + d[14];
    const R YP1 = d[4] + d[11] + d[13] + d[17] + d[18];
    const R ZM1 = d[2] + d[5] + d[7] + d[16] + d[18];
    const R ZP1 = d[6] + d[8] + d[12] + d[15] + d[17];
    drho = XM1 + XP1 + X0;
    rho = R(1.) + drho;
    const R inv = R(1.) / rho;
    u0 = (XP1 - XM1) * inv;
    u1 = (YP1 - YM1) * inv;
    u2 = (ZP1 - ZM1) * inv;
}

template <typename R>
__device__ __forceinline__ void bgk_dev(const R d[Q], R omega, R out[Q])
{
    R drho, rho, u0, u1, u2;
    macroscopic_dev<R>(d, drho, rho, u0, u1, u2);
    const R usqr = R(1.5) * (u0 * u0 + u1 * u1 + u2 * u2);
    const R ck[9] = {u0, u1, u2, u0 + u1, u0 - u1, u0 + u2, u0 - u2, u1 + u2, u1 - u2};
    const R w18 = R(1. / 18.), w36 = R(1. / 36.), w3 = R(1. / 3.);
    const R dr18 = w18 * drho, dr36 = w36 * drho;
    const R rw18 = rho * w18, rw36 = rho * w36;
    const R rw18x6 = rw18 * R(6.), rw36x6 = rw36 * R(6.);
#pragma unroll
    for (int i = 0; i < 9; i++) {
        const R c = ck[i];
        const R poly = fmaR(c, fmaR(R(4.5), c, R(-3.)), -usqr);
        const R deq = fmaR(i < 3 ? rw18 : rw36, poly, i < 3 ? dr18 : dr36);
        const R deqopp = fmaR(i < 3 ? rw18x6 : rw36x6, c, deq);
        out[i] = fmaR(omega, deq - d[i], d[i]);
        out[i + 10] = fmaR(omega, deqopp - d[i + 10], d[i + 10]);
    }
    const R deq9 = fmaR(-(rho * w3), usqr, w3 * drho);
    out[9] = fmaR(omega, deq9 - d[9], d[9]);
}

template <typename R>
__device__ __forceinline__ void bgk_forced_dev(const R d[Q], R omega, R g0, R g1, R g2, R out[Q])
{
    R drho, rho, u0, u1, u2;
    macroscopic_dev<R>(d, drho, rho, u0, u1, u2);
    u0 += R(0.5) * g0;
    u1 += R(0.5) * g1;
    u2 += R(0.5) * g2;
    const R usqr = R(1.5) * (u0 * u0 + u1 * u1 + u2 * u2);
    const R ug = u0 * g0 + u1 * g1 + u2 * g2;
    const R fpre = R(1) - R(0.5) * omega;
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const R w = R(wq(q));
        const R cu = R(CX[q]) * u0 + R(CY[q]) * u1 + R(CZ[q]) * u2;
        const R cg = R(CX[q]) * g0 + R(CY[q]) * g1 + R(CZ[q]) * g2;
        const R poly = fmaR(cu, fmaR(R(4.5), cu, R(3.)), -usqr);
        const R deq = fmaR(rho * w, poly, w * drho);
        const R src = fpre * w * rho * (R(3) * (cg - ug) + R(9) * cu * cg);
        out[q] = fmaR(omega, deq - d[q], d[q]) + src;
    }
}

template <int S, int CM, bool FORCE>
__device__ __forceinline__ void collide(const typename Stor<S>::P fp[Q], double omega,
                                        const double g[3], typename Real<S, CM>::R out[Q])
{
    using R = typename Real<S, CM>::R;
    R f[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) f[q] = (R)fp[q];
    if constexpr (Stor<S>::kDev) {
        if constexpr (FORCE)
            bgk_forced_dev<R>(f, (R)omega, (R)g[0], (R)g[1], (R)g[2], out);
        else
            bgk_dev<R>(f, (R)omega, out);
    } else if constexpr (FORCE) {
        bgk_forced<R>(f, (R)omega, (R)g[0], (R)g[1], (R)g[2], out);
    } else if constexpr (CM == NX_COMPUTE_REF && S != NX_F64) {
        bgk_ref(f, (float)omega, out);
    } else {
        bgk<R>(f, (R)omega, out);
    }
}

// post-stream moments of one cell (LbmTools.h:401-433), as computed by the
// reference in LbmComputeType
template <int S, int CM>
__device__ __forceinline__ void cell_moments(int cls, const typename Stor<S>::P fp[Q],
                                             typename Real<S, CM>::R& rho,
                                             typename Real<S, CM>::R u[3])
{
    using R = typename Real<S, CM>::R;
    if (cls == NX_CELL_BULK) {
        R f[Q];
#pragma unroll
        for (int q = 0; q < Q; q++) f[q] = (R)fp[q];
        if constexpr (Stor<S>::kDev) {
            R drho;
            macroscopic_dev<R>(f, drho, rho, u[0], u[1], u[2]);
        } else {
            macroscopic<R>(f, rho, u[0], u[1], u[2]);
        }
    } else if (cls == NX_CELL_MOVINGWALL) {
        rho = R(1.0);
        // fp[0..2] hold the cell's own populations 0..2
#pragma unroll
        for (int a = 0; a < 3; a++) u[a] = (R)fp[a] / R(6. * 1. / 18.);
    } else {
        rho = R(0);
        u[0] = u[1] = u[2] = R(0);
    }
}

// ---- XCD-aware workgroup order --------------------------------------------
// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b runs on
// XCD b % 8).  Remap so that each XCD walks one contiguous range of the
// domain, keeping neighbouring rows/blocks in the same L2.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t n)
{
    const uint32_t q = n >> 3, r = n & 7u, k = b & 7u, j = b >> 3;
    return k * q + (k < r ? k : r) + j;
}

// ---- wave64 reductions with DPP ---------------------------------------------
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ double dpp_f64(double v)
{
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, CTRL, ROW_MASK, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), CTRL, ROW_MASK, 0xF, false);
    return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// Sum over the 64 lanes of a wavefront; the result is returned to all lanes.
// quad_perm butterflies -> half-row mirror -> row mirror give every lane its
// row sum; row_bcast:15 / row_bcast:31 fold the four rows into lane 63.
__device__ __forceinline__ double wave_sum_f64(double v)
{
    v += dpp_f64<0xB1>(v);        // quad_perm [1,0,3,2]
    v += dpp_f64<0x4E>(v);        // quad_perm [2,3,0,1]
    v += dpp_f64<0x141>(v);       // row_half_mirror
    v += dpp_f64<0x140>(v);       // row_mirror
    v += dpp_f64<0x142, 0xA>(v);  // row_bcast:15 into rows 1,3
    v += dpp_f64<0x143, 0xC>(v);  // row_bcast:31 into rows 2,3
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)b, 63);
    const uint32_t hi = __builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), 63);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Workgroup sum of NV doubles per thread (blockDim.x multiple of 64, <= 1024);
// thread 0 receives the totals.  Fixed order: deterministic.
template <int NV>
__device__ __forceinline__ void block_sum_f64(double v[NV], double* lds /* NV*16 */)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; k++) {
        const double s = wave_sum_f64(v[k]);
        if (lane == 0) lds[k * 16 + wave] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) {
            double s = 0;
            for (int w = 0; w < nw; w++) s += lds[k * 16 + w];
            v[k] = s;
        }
    }
}

// ---- host-side helpers (defined in nx_runtime.hip) --------------------------
int set_error(int code, const std::string& msg);
int hip_error(hipError_t e, const char* what);
// per-device reduction scratch of at least `bytes`, owned by the library
int scratch(size_t bytes, void** ptr);
// device address of the allocation-free residual's static scratch, looked up
// for the current device (d3q19_dense.hip)
int dense_noalloc_setup();
int finish_reduce(const double* partials, int n, int nv, double* out, hipStream_t s);

}  // namespace nx

#define NX_CHECK_HIP(call)                                       \
    do {                                                         \
        hipError_t nx_e_ = (call);                               \
        if (nx_e_ != hipSuccess) return nx::hip_error(nx_e_, #call); \
    } while (0)

#define NX_CHECK_LAUNCH(name)                                    \
    do {                                                         \
        hipError_t nx_e_ = hipGetLastError();                    \
        if (nx_e_ != hipSuccess) return nx::hip_error(nx_e_, name); \
    } while (0)
