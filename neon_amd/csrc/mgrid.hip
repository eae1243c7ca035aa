// mgrid.hip -- grid-refinement LBM (Neon's mGrid, depth 2 to 5, and the
// lbmMultiRes application's scheme with its BGK D3Q19 or KBC D3Q27
// collision) on a stack of block-sparse levels.
//
// Level 0 is the finest level (the domain); level l has voxels of edge 2^l,
// ceil(dim / 2^l) per axis, and a level-(l+1) voxel v covers the level-l
// voxels 2v + {0,1}^3.  Each level is a bGrid of 8^3 blocks (the library's
// builder: blocks ordered bz, by, bx, 27-entry connectivity) with AoSoA
// populations (b*Q + q)*512 + x + 8y + 64z (Q = 19 or 27), and a dense table from block
// coordinates to block index for the cross-level lookups.
//
// Everything that depends only on the geometry is resolved on the host when
// the cavity is set up: every bulk leaf voxel carries a 3-bit state per
// population (plain pull, bounce-back link, explosion link, coalescence
// link, no update), each fine-to-coarse accumulation is a gather list whose
// per-target order is the fine voxels' z, y, x order, and the coalescence
// weights are a field of every level above 0.  The kernels then do no
// searching.  One coarse step is the reference's recursion
//   step(l): collide l, store l -> l+1 | step(l-1) | stream l
//            | unless l is the top: the same once more
// (lbmMultiRes.h:141-263, baseline operation order with the fine store).
// Reference semantics restated (file:line in each kernel):
//   grid      libNeonDomain/src/domain/details/mGrid/mGrid.cpp:89-345, 466-555
//   set-up    apps/lbmMultiRes/lidDrivenCavity.h:8-78, init.h:8-86
//   collide   apps/lbmMultiRes/collide.h:8-149 (BGK), :148-284 (KBC)
//   lattice   apps/lbmMultiRes/lattice.h:14-106 (D3Q27), :108-173 (D3Q19)
//   store     apps/lbmMultiRes/store.h:98-163
//   stream    apps/lbmMultiRes/stream.h:170-225
//   helpers   apps/lbmMultiRes/util.h:19-45, mPartition_imp.h:155-235
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "block_common.hpp"

namespace nx {
namespace mres {
namespace {

// state byte of a voxel (host and device)
constexpr uint32_t kActive = 1u, kRefined = 2u;
constexpr int kClsShift = 2;
// per-population link states of a bulk leaf voxel, 3 bits each
constexpr uint64_t kPlain = 0, kWall = 1, kExpl = 2, kCoal = 3, kSkip = 4;
constexpr int kMaxDepth = 5;

__host__ __device__ inline int cls_bits(uint32_t st) { return (int)((st >> kClsShift) & 3u); }
__host__ __device__ inline bool bulk_leaf(uint32_t st)
{
    return (st & kActive) && !(st & kRefined) && cls_bits(st) == NX_CELL_BULK;
}

template <int S>
struct MS;  // storage of the refinement step: the reference's float / double fields
template <>
struct MS<NX_F64> {
    using T = double;
};
template <>
struct MS<NX_F32> {
    using T = float;
};

// ---- lattices ------------------------------------------------------------------

// D3Q27 of apps/lbmMultiRes/lattice.h:14-75 (z fastest, then y, then x, each
// over 0, -1, +1), the lattice of its default KBC collision
constexpr int K27X[27] = {0, 0, 0, 0, 0, 0, 0, 0, 0, -1, -1, -1, -1, -1, -1, -1, -1, -1, 1, 1, 1, 1, 1, 1, 1, 1, 1};
constexpr int K27Y[27] = {0, 0, 0, -1, -1, -1, 1, 1, 1, 0, 0, 0, -1, -1, -1, 1, 1, 1, 0, 0, 0, -1, -1, -1, 1, 1, 1};
constexpr int K27Z[27] = {0, -1, 1, 0, -1, 1, 0, -1, 1, 0, -1, 1, 0, -1, 1, 0, -1, 1, 0, -1, 1, 0, -1, 1, 0, -1, 1};
constexpr int K27OPP[27] = {0,  2,  1,  6,  8,  7,  3,  5,  4,  18, 20, 19, 24, 26,
                            25, 21, 23, 22, 9,  11, 10, 15, 17, 16, 12, 14, 13};
__host__ __device__ constexpr double w27(int q)
{
    return (K27X[q] != 0) + (K27Y[q] != 0) + (K27Z[q] != 0) == 0   ? 8.0 / 27.0
           : (K27X[q] != 0) + (K27Y[q] != 0) + (K27Z[q] != 0) == 1 ? 2.0 / 27.0
           : (K27X[q] != 0) + (K27Y[q] != 0) + (K27Z[q] != 0) == 2 ? 1.0 / 54.0
                                                                    : 1.0 / 216.0;
}

// KBC D3Q27 collision (collide.h:148-284) with the reference's type
// promotions: the density and velocity sums over q in order (util.h:47-62),
// the double literals of the equilibrium, the shear split and gamma
// evaluated in binary64 and stored in T, tiny = 1e-7 in gamma's denominator.
// Two forms with the same results: kbc27_kept keeps the equilibrium,
// fneq and the shear part of every population in arrays; kbc27_recomputed
// recomputes them in each of the collision's three passes -- moment flux,
// e0 / e1, output -- from copies of rho, u and usqr that the compiler may
// not look through, so only the 27 populations stay live (occupancy).  The
// same expressions give the same values.
template <typename T>
__device__ __forceinline__ T kbc27_feq(int q, T rho, T v0, T v1, T v2, T usqr)
{
    T cu = 0;
    cu += (T)K27X[q] * v0;
    cu += (T)K27Y[q] * v1;
    cu += (T)K27Z[q] * v2;
    cu = (T)((double)cu * 3.0);
    return (T)((double)rho * w27(q) * (1. + (double)cu + 0.5 * (double)cu * (double)cu - (double)usqr));
}

template <typename T>
__device__ __forceinline__ void kbc27_kept(const T f[27], T om, T out[27])
{
    const T beta = (T)((double)om * 0.5);
    const T invBeta = (T)(1.0 / (double)beta);
    const T tiny = (T)1e-7;
    T rho = 0;
#pragma unroll
    for (int q = 0; q < 27; q++) rho += f[q];
    T v0 = 0, v1 = 0, v2 = 0;
#pragma unroll
    for (int q = 0; q < 27; q++) {
        v0 += f[q] * (T)K27X[q];
        v1 += f[q] * (T)K27Y[q];
        v2 += f[q] * (T)K27Z[q];
    }
    v0 /= rho;
    v1 /= rho;
    v2 /= rho;
    const T usqr = (T)((3.0 / 2.0) * (double)(v0 * v0 + v1 * v1 + v2 * v2));
    T feq[27], fneq[27];
    T p0 = 0, p1 = 0, p2 = 0, p3 = 0, p4 = 0, p5 = 0;  // Pi: xx, xy, xz, yy, yz, zz
#pragma unroll
    for (int q = 0; q < 27; q++) {
        T cu = 0;
        cu += (T)K27X[q] * v0;
        cu += (T)K27Y[q] * v1;
        cu += (T)K27Z[q] * v2;
        cu = (T)((double)cu * 3.0);
        feq[q] = (T)((double)rho * w27(q) * (1. + (double)cu + 0.5 * (double)cu * (double)cu - (double)usqr));
        fneq[q] = f[q] - feq[q];
    }
#pragma unroll
    for (int q = 0; q < 27; q++) {
        p0 += fneq[q] * (T)(K27X[q] * K27X[q]);
        p1 += fneq[q] * (T)(K27X[q] * K27Y[q]);
        p2 += fneq[q] * (T)(K27X[q] * K27Z[q]);
        p3 += fneq[q] * (T)(K27Y[q] * K27Y[q]);
        p4 += fneq[q] * (T)(K27Y[q] * K27Z[q]);
        p5 += fneq[q] * (T)(K27Z[q] * K27Z[q]);
    }
    const T Nxz = p0 - p5, Nyz = p3 - p5;
    const T shx = (T)((2.0 * (double)Nxz - (double)Nyz) / 6.0);
    const T shy = (T)((-(double)Nxz + 2.0 * (double)Nyz) / 6.0);
    const T shz = (T)((double)(-Nxz - Nyz) / 6.0);
    const T sxy = (T)((double)p1 / 4.0), sxz = (T)((double)p2 / 4.0), syz = (T)((double)p4 / 4.0);
    T dS[27];
    T e0 = 0, e1 = 0;
#pragma unroll
    for (int q = 0; q < 27; q++) {
        const int cx = K27X[q], cy = K27Y[q], cz = K27Z[q];
        const int nz = (cx != 0) + (cy != 0) + (cz != 0);
        T sh = (T)0;
        if (nz == 1) sh = cx ? shx : (cy ? shy : shz);
        if (nz == 2) sh = !cz ? (cx * cy > 0 ? sxy : -sxy) : (!cy ? (cx * cz > 0 ? sxz : -sxz) : (cy * cz > 0 ? syz : -syz));
        dS[q] = rho * sh;
        const T dH = fneq[q] - dS[q];
        e0 += (dS[q] * dH / feq[q]);
        e1 += (dH * dH / feq[q]);
    }
    const T gamma = (T)((double)invBeta - (2.0 - (double)invBeta) * (double)e0 / (double)(tiny + e1));
#pragma unroll
    for (int q = 0; q < 27; q++) {
        const T dH = fneq[q] - dS[q];
        out[q] = (T)((double)f[q] - (double)beta * (2.0 * (double)dS[q] + (double)(gamma * dH)));
    }
}

template <typename T>
__device__ __forceinline__ void kbc27_recomputed(const T f[27], T om, T out[27])
{
    const T beta = (T)((double)om * 0.5);
    const T invBeta = (T)(1.0 / (double)beta);
    const T tiny = (T)1e-7;
    T rho = 0;
#pragma unroll
    for (int q = 0; q < 27; q++) rho += f[q];
    T v0 = 0, v1 = 0, v2 = 0;
#pragma unroll
    for (int q = 0; q < 27; q++) {
        v0 += f[q] * (T)K27X[q];
        v1 += f[q] * (T)K27Y[q];
        v2 += f[q] * (T)K27Z[q];
    }
    v0 /= rho;
    v1 /= rho;
    v2 /= rho;
    const T usqr = (T)((3.0 / 2.0) * (double)(v0 * v0 + v1 * v1 + v2 * v2));
    T p0 = 0, p1 = 0, p2 = 0, p3 = 0, p4 = 0, p5 = 0;  // Pi: xx, xy, xz, yy, yz, zz
#pragma unroll
    for (int q = 0; q < 27; q++) {
        const T fneq = f[q] - kbc27_feq<T>(q, rho, v0, v1, v2, usqr);
        p0 += fneq * (T)(K27X[q] * K27X[q]);
        p1 += fneq * (T)(K27X[q] * K27Y[q]);
        p2 += fneq * (T)(K27X[q] * K27Z[q]);
        p3 += fneq * (T)(K27Y[q] * K27Y[q]);
        p4 += fneq * (T)(K27Y[q] * K27Z[q]);
        p5 += fneq * (T)(K27Z[q] * K27Z[q]);
    }
    const T Nxz = p0 - p5, Nyz = p3 - p5;
    const T shx = (T)((2.0 * (double)Nxz - (double)Nyz) / 6.0);
    const T shy = (T)((-(double)Nxz + 2.0 * (double)Nyz) / 6.0);
    const T shz = (T)((double)(-Nxz - Nyz) / 6.0);
    const T sxy = (T)((double)p1 / 4.0), sxz = (T)((double)p2 / 4.0), syz = (T)((double)p4 / 4.0);
    auto shear = [&](int q) -> T {
        const int cx = K27X[q], cy = K27Y[q], cz = K27Z[q];
        const int nz = (cx != 0) + (cy != 0) + (cz != 0);
        if (nz == 1) return cx ? shx : (cy ? shy : shz);
        if (nz == 2)
            return !cz ? (cx * cy > 0 ? sxy : -sxy) : (!cy ? (cx * cz > 0 ? sxz : -sxz) : (cy * cz > 0 ? syz : -syz));
        return (T)0;
    };
    T e0 = 0, e1 = 0;
    {
        T r = rho, a = v0, b = v1, c = v2, us = usqr;
        nx_pin(r);
        nx_pin(a);
        nx_pin(b);
        nx_pin(c);
        nx_pin(us);
#pragma unroll
        for (int q = 0; q < 27; q++) {
            const T feq = kbc27_feq<T>(q, r, a, b, c, us);
            const T dS = rho * shear(q);
            const T dH = (f[q] - feq) - dS;
            e0 += (dS * dH / feq);
            e1 += (dH * dH / feq);
        }
    }
    const T gamma = (T)((double)invBeta - (2.0 - (double)invBeta) * (double)e0 / (double)(tiny + e1));
    T r = rho, a = v0, b = v1, c = v2, us = usqr;
    nx_pin(r);
    nx_pin(a);
    nx_pin(b);
    nx_pin(c);
    nx_pin(us);
#pragma unroll
    for (int q = 0; q < 27; q++) {
        const T feq = kbc27_feq<T>(q, r, a, b, c, us);
        const T dS = rho * shear(q);
        const T dH = (f[q] - feq) - dS;
        out[q] = (T)((double)f[q] - (double)beta * (2.0 * (double)dS + (double)(gamma * dH)));
    }
}

// binary64: the recomputing form (127 instead of 221 VGPRs, 16% faster on
// the 240^3 grid); binary32: the form that keeps the 3 x 27 values (75
// against 121 VGPRs did not pay for the recomputation there)
template <typename T>
__device__ __forceinline__ void kbc27(const T f[27], T om, T out[27])
{
    if constexpr (sizeof(T) == 8)
        kbc27_recomputed<T>(f, om, out);
    else
        kbc27_kept<T>(f, om, out);
}

// lattice traits: populations, velocity table, rest population, collision
struct L19 {
    static constexpr int Q = 19, REST = 9, W = 1;  // W: 64-bit link words per voxel
    static constexpr __host__ __device__ int cx(int q) { return CX[q]; }
    static constexpr __host__ __device__ int cy(int q) { return CY[q]; }
    static constexpr __host__ __device__ int cz(int q) { return CZ[q]; }
    static constexpr __host__ __device__ int opp(int q) { return OPP[q]; }
    static constexpr __host__ __device__ double w(int q) { return wq(q); }
    template <typename T>
    static __device__ __forceinline__ void collide(const T f[19], T om, T out[19])
    {
        bgk<T>(f, om, out);
    }
};
struct L27 {
    static constexpr int Q = 27, REST = 0, W = 2;
    static constexpr __host__ __device__ int cx(int q) { return K27X[q]; }
    static constexpr __host__ __device__ int cy(int q) { return K27Y[q]; }
    static constexpr __host__ __device__ int cz(int q) { return K27Z[q]; }
    static constexpr __host__ __device__ int opp(int q) { return K27OPP[q]; }
    static constexpr __host__ __device__ double w(int q) { return w27(q); }
    template <typename T>
    static __device__ __forceinline__ void collide(const T f[27], T om, T out[27])
    {
        kbc27<T>(f, om, out);
    }
};

// link state of population q: 3 bits, 21 populations per 64-bit word
constexpr int kLinksPerWord = 21;

// ---- kernels ----------------------------------------------------------------

// collide.h:8-284: the lattice's collision on every bulk leaf voxel (fin ->
// fout, same voxel); a refined voxel (levels above 0) resets its fout, the
// accumulator of the fine store, to 0.
template <int S, class LT>
__global__ void __launch_bounds__(kBV)
    k_mr_collide(const typename MS<S>::T* __restrict__ fin, typename MS<S>::T* __restrict__ fout,
                 const uint32_t* __restrict__ state, double omega)
{
    using T = typename MS<S>::T;
    constexpr int NQ = LT::Q;
    const uint32_t b = xcd_remap(blockIdx.x, gridDim.x), v = threadIdx.x;  // neighbour blocks share an XCD's L2
    const uint32_t st = state[(uint64_t)b * kBV + v];
    if (!(st & kActive) || cls_bits(st) != NX_CELL_BULK) return;
    const uint64_t own = (uint64_t)b * (NQ * kBV) + v;
    if (st & kRefined) {
#pragma unroll
        for (int q = 0; q < NQ; q++) __builtin_nontemporal_store(T(0), &fout[own + (uint64_t)q * kBV]);
        return;
    }
    T f[NQ], o[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) f[q] = fin[own + (uint64_t)q * kBV];
    LT::template collide<T>(f, (T)omega, o);
#pragma unroll
    for (int q = 0; q < NQ; q++) __builtin_nontemporal_store(o[q], &fout[own + (uint64_t)q * kBV]);
}

// store.h:98-163 as a gather: entry e adds its fine sources, in order, to
// the coarse accumulator target[e].
template <int S>
__global__ void __launch_bounds__(256)
    k_mr_store(const typename MS<S>::T* __restrict__ fine, typename MS<S>::T* __restrict__ coarse,
               const uint64_t* __restrict__ target, const uint32_t* __restrict__ first,
               const uint64_t* __restrict__ src, uint32_t n)
{
    using T = typename MS<S>::T;
    const uint32_t e = blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    T acc = coarse[target[e]];
    for (uint32_t k = first[e]; k < first[e + 1]; k++) acc += fine[src[k]];
    coarse[target[e]] = acc;
}

struct UpLevel {  // the next coarser level, seen from this one (explosion)
    const void* fout;
    const uint32_t* table;
    int32_t nb[3];
};

// stream.h:170-225 (stream fused with coalescence and explosion), pull, for
// the bulk leaf voxels of one level.  The plain pull of all populations is
// issued first (upstream voxel through the connectivity row in scalar
// registers); voxels with another link state then fix those links up:
// bounce-back (own opposite + the wall's opposite), explosion (the coarse
// post-collision population of the uncle, the next level's voxel holding
// the upstream position), coalescence (the refined upstream voxel's
// accumulated fine populations times its weight); links in state "none"
// keep the value in `out`.
// COLLIDE: the pulled populations are collided at once with omega and the
// result written to `out`, and refined voxels reset their accumulator in
// `out` -- stream followed by the next collide of the same level
// (collide.h:8-284) in one pass; the results are those of the two kernels in
// sequence, bit for bit.  KEEP = false: the level has no link in state
// "none" (checked on the host), every population is stored.  `links` holds
// LT::W words per voxel slot, word k at k * n_slots.
template <int S, class LT, bool COLLIDE, bool KEEP>
__global__ void __launch_bounds__(kBV)
    k_mr_stream(const typename MS<S>::T* __restrict__ fout, typename MS<S>::T* __restrict__ out,
                const uint32_t* __restrict__ state, const uint64_t* __restrict__ links, uint64_t n_slots,
                const uint32_t* __restrict__ conn, const int32_t* __restrict__ origins,
                const float* __restrict__ ss, UpLevel up, double omega)
{
    using T = typename MS<S>::T;
    constexpr int NQ = LT::Q;
    const uint32_t b = xcd_remap(blockIdx.x, gridDim.x), v = threadIdx.x;  // neighbour blocks share an XCD's L2
    const Conn cb = load_conn(conn, b);
    const uint32_t st = state[(uint64_t)b * kBV + v];
    if ((st & kActive) && (st & kRefined) && cls_bits(st) == NX_CELL_BULK) {
        // a refined voxel: with COLLIDE the collide's reset of its
        // accumulator (collide.h:131-141); otherwise its fin, which the
        // reference never writes after the set-up's 0, back to 0 (a fused
        // pass earlier in the call left an accumulator in this buffer)
        const uint64_t own = (uint64_t)b * (NQ * kBV) + v;
#pragma unroll
        for (int q = 0; q < NQ; q++) __builtin_nontemporal_store(T(0), &out[own + (uint64_t)q * kBV]);
        return;
    }
    if (!bulk_leaf(st)) return;
    uint64_t lk[LT::W];
#pragma unroll
    for (int k = 0; k < LT::W; k++) lk[k] = links[(uint64_t)k * n_slots + (uint64_t)b * kBV + v];
    const int x = (int)(v & 7u), y = (int)((v >> 3) & 7u), z = (int)(v >> 6);
    const uint64_t own = (uint64_t)b * (NQ * kBV) + v;
    T f[NQ];
    for_q(std::make_integer_sequence<int, NQ>{}, [&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if constexpr (q == LT::REST) {
            f[q] = fout[own + (uint64_t)q * kBV];
        } else {
            uint32_t t, nv;
            upstream_off<-LT::cx(q), -LT::cy(q), -LT::cz(q)>(b, cb, x, y, z, t, nv);
            f[q] = fout[t != kInvalidBlock ? ((uint64_t)t * NQ + q) * kBV + nv : own + (uint64_t)q * kBV];
        }
    });
#pragma unroll
    for (int q = 0; q < NQ; q++) nx_pin(f[q]);
    uint32_t keep = 0;  // links whose value stays as it is in `out`
    bool any = false;
#pragma unroll
    for (int k = 0; k < LT::W; k++) any |= lk[k] != 0;
    if (any) {
        for_q(std::make_integer_sequence<int, NQ>{}, [&](auto qc) {
            constexpr int q = decltype(qc)::value;
            if constexpr (q != LT::REST) {
                const uint64_t kind = (lk[q / kLinksPerWord] >> (3 * (q % kLinksPerWord))) & 7u;
                if (kind == kPlain) return;
                uint32_t t, nv;
                upstream_off<-LT::cx(q), -LT::cy(q), -LT::cz(q)>(b, cb, x, y, z, t, nv);
                bool ok = true;
                if (kind == kWall) {
                    constexpr int o = LT::opp(q);
                    ok = t != kInvalidBlock;
                    if (ok) f[q] = fout[own + (uint64_t)o * kBV] + fout[((uint64_t)t * NQ + o) * kBV + nv];
                } else if (kind == kExpl) {
                    const int ux = (origins[3 * (uint64_t)b] + x - LT::cx(q)) >> 1;
                    const int uy = (origins[3 * (uint64_t)b + 1] + y - LT::cy(q)) >> 1;
                    const int uz = (origins[3 * (uint64_t)b + 2] + z - LT::cz(q)) >> 1;
                    const int bx = ux >> 3, by = uy >> 3, bz = uz >> 3;
                    ok = up.table && !(bx < 0 || by < 0 || bz < 0 || bx >= up.nb[0] || by >= up.nb[1] ||
                                       bz >= up.nb[2]);
                    const uint32_t tb = ok ? up.table[((uint64_t)bz * up.nb[1] + by) * up.nb[0] + bx]
                                           : kInvalidBlock;
                    ok = tb != kInvalidBlock;
                    const uint32_t slot = (uint32_t)((ux & 7) + 8 * (uy & 7) + 64 * (uz & 7));
                    if (ok) f[q] = ((const T*)up.fout)[((uint64_t)tb * NQ + q) * kBV + slot];
                } else if (kind == kCoal) {
                    ok = t != kInvalidBlock && ss;
                    if (ok) f[q] = f[q] * (T)ss[((uint64_t)t * NQ + q) * kBV + nv];
                } else {
                    ok = false;
                }
                if (!ok) {
                    keep |= 1u << q;
                    f[q] = out[own + (uint64_t)q * kBV];
                }
            }
        });
    }
    if constexpr (COLLIDE) {
        T o[NQ];
        LT::template collide<T>(f, (T)omega, o);
#pragma unroll
        for (int q = 0; q < NQ; q++) __builtin_nontemporal_store(o[q], &out[own + (uint64_t)q * kBV]);
    } else if constexpr (KEEP) {
#pragma unroll
        for (int q = 0; q < NQ; q++)
            if (!((keep >> q) & 1u)) __builtin_nontemporal_store(f[q], &out[own + (uint64_t)q * kBV]);
    } else {
#pragma unroll
        for (int q = 0; q < NQ; q++) __builtin_nontemporal_store(f[q], &out[own + (uint64_t)q * kBV]);
    }
}

// ---- host structures ------------------------------------------------------------

struct Level {
    int32_t dim[3] = {0, 0, 0};
    int32_t nb[3] = {0, 0, 0};
    uint32_t n_blocks = 0;
    std::vector<int32_t> origins;
    std::vector<uint32_t> masks, conn, table;
    std::vector<uint8_t> st;  // dense state byte per voxel
    uint32_t* d_state = nullptr;
    uint64_t* d_links = nullptr;
    uint32_t* d_conn = nullptr;
    int32_t* d_origins = nullptr;
    uint32_t* d_table = nullptr;
    float* d_ss = nullptr;  // coalescence weights (levels above 0)
    void* fin = nullptr;
    void* fout = nullptr;
    bool keep = true;  // the level has links in state "none"
    int nq = 19;       // populations per voxel

    size_t nvox() const { return (size_t)dim[0] * dim[1] * dim[2]; }
    size_t lin(int x, int y, int z) const { return (size_t)x + (size_t)dim[0] * ((size_t)y + (size_t)dim[1] * z); }
    bool inside(int x, int y, int z) const
    {
        return x >= 0 && y >= 0 && z >= 0 && x < dim[0] && y < dim[1] && z < dim[2];
    }
    bool active(int x, int y, int z) const { return inside(x, y, z) && (st[lin(x, y, z)] & kActive); }
    // AoSoA slot (block * 512 + in-block index) of an active voxel, or -1
    int64_t slot(int x, int y, int z) const
    {
        if (!active(x, y, z)) return -1;
        const uint32_t b = table[((size_t)(z >> 3) * nb[1] + (y >> 3)) * nb[0] + (x >> 3)];
        return (int64_t)b * kBV + ((x & 7) + 8 * (y & 7) + 64 * (z & 7));
    }
    // population element of an active voxel's slot
    uint64_t elem(int64_t slot, int q) const { return ((uint64_t)(slot / kBV) * nq + q) * kBV + slot % kBV; }
    size_t slots() const { return (size_t)kBV * n_blocks; }
};

// the lattice on the host: velocities, opposites, weights, link words
struct HostLat {
    int Q = 19, rest = 9, W = 1;
    int c[27][3] = {};
    int opp[27] = {};
    double w[27] = {};
};

template <class LT>
HostLat host_lat_of()
{
    HostLat h;
    h.Q = LT::Q;
    h.rest = LT::REST;
    h.W = LT::W;
    for (int q = 0; q < LT::Q; q++) {
        h.c[q][0] = LT::cx(q);
        h.c[q][1] = LT::cy(q);
        h.c[q][2] = LT::cz(q);
        h.opp[q] = LT::opp(q);
        h.w[q] = LT::w(q);
    }
    return h;
}

struct Store {  // fine-to-coarse gather list of one level pair (l -> l + 1)
    uint32_t n = 0;
    uint64_t n_src = 0;
    uint64_t* d_target = nullptr;
    uint32_t* d_first = nullptr;
    uint64_t* d_src = nullptr;
};

}  // namespace
}  // namespace mres
}  // namespace nx

struct nx_mgrid_s {
    int storage = NX_F64;
    int lattice = NX_MGRID_BGK19;
    nx::mres::HostLat lat;
    int depth = 2;
    nx::mres::Level L[nx::mres::kMaxDepth];
    nx::mres::Store store[nx::mres::kMaxDepth];
    bool ready = false;
};

using namespace nx;
using namespace nx::mres;

namespace {

int upload(void** dst, const void* src, size_t bytes)
{
    if (*dst) {
        NX_CHECK_HIP(hipFree(*dst));
        *dst = nullptr;
    }
    NX_CHECK_HIP(hipMalloc(dst, std::max<size_t>(bytes, 8)));
    if (bytes) NX_CHECK_HIP(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
    return NX_OK;
}

// isRefined (mGrid.cpp:167-199): a child of level-l voxel (x, y, z) is
// active on level l - 1 (the voxel itself need not be)
bool any_child(const Level& fine, int x, int y, int z)
{
    for (int k = 0; k < 8; k++)
        if (fine.active(2 * x + (k & 1), 2 * y + ((k >> 1) & 1), 2 * z + (k >> 2))) return true;
    return false;
}

// The mGrid constructor (mGrid.cpp:89-345): activation level by level in
// 2^3 groups, each active group activating its parent; overlap culling from
// the coarsest level down (a refined voxel whose in-domain neighbours are
// all refined goes; level l - 1 is still unculled when level l is visited);
// the strong-balance pass; then the refined flags from the final masks
// (:466-555).  Returns the number of voxels the strong-balance pass added.
uint64_t activate(const int32_t dim[3], int depth, const uint8_t* const* lams, Level* L)
{
    for (int a = 0; a < 3; a++) L[0].dim[a] = dim[a];
    for (int l = 1; l < depth; l++)
        for (int a = 0; a < 3; a++) L[l].dim[a] = (L[l - 1].dim[a] + 1) / 2;
    for (int l = 0; l < depth; l++) L[l].st.assign(L[l].nvox(), 0);
    for (int l = 0; l < depth; l++) {
        Level& lv = L[l];
        const uint8_t* lam = lams[l];
        for (int gz = 0; gz < (lv.dim[2] + 1) / 2; gz++)
            for (int gy = 0; gy < (lv.dim[1] + 1) / 2; gy++)
                for (int gx = 0; gx < (lv.dim[0] + 1) / 2; gx++) {
                    bool any = false;
                    for (int k = 0; k < 8; k++) {
                        const int x = 2 * gx + (k & 1), y = 2 * gy + ((k >> 1) & 1), z = 2 * gz + (k >> 2);
                        if (lv.inside(x, y, z) && (lv.st[lv.lin(x, y, z)] || lam[lv.lin(x, y, z)])) any = true;
                    }
                    if (!any) continue;
                    for (int k = 0; k < 8; k++) {
                        const int x = 2 * gx + (k & 1), y = 2 * gy + ((k >> 1) & 1), z = 2 * gz + (k >> 2);
                        if (lv.inside(x, y, z)) lv.st[lv.lin(x, y, z)] = (uint8_t)kActive;
                    }
                    if (l + 1 < depth) L[l + 1].st[L[l + 1].lin(gx, gy, gz)] = (uint8_t)kActive;
                }
    }
    for (int l = depth - 1; l > 0; l--) {
        Level& c = L[l];
        std::vector<uint8_t> refined(c.nvox(), 0);
        for (int z = 0; z < c.dim[2]; z++)
            for (int y = 0; y < c.dim[1]; y++)
                for (int x = 0; x < c.dim[0]; x++) refined[c.lin(x, y, z)] = any_child(L[l - 1], x, y, z);
        for (int z = 0; z < c.dim[2]; z++)
            for (int y = 0; y < c.dim[1]; y++)
                for (int x = 0; x < c.dim[0]; x++) {
                    const size_t i = c.lin(x, y, z);
                    if (!c.st[i] || !refined[i]) continue;
                    bool all = true;
                    for (int d = 0; d < 27 && all; d++) {
                        if (d == 13) continue;
                        const int nx_ = x + d % 3 - 1, ny = y + (d / 3) % 3 - 1, nz = z + d / 9 - 1;
                        if (c.inside(nx_, ny, nz) && !refined[c.lin(nx_, ny, nz)]) all = false;
                    }
                    if (all) c.st[i] = 0;
                }
    }
    // strong balance: the finest active voxel covering each in-domain
    // neighbour position of an active voxel of level l lies on level l or
    // l + 1; found on l_n >= l + 2, the voxel of level l_n - 1 covering it
    // is activated alone and the search continues upwards; passes repeat
    // until one activates nothing and see their own activations
    uint64_t added = 0;
    for (bool again = true; again;) {
        again = false;
        for (int l = 0; l < depth; l++) {
            const Level& lv = L[l];
            for (int bz = 0; bz < (lv.dim[2] + 1) / 2; bz++)
                for (int by = 0; by < (lv.dim[1] + 1) / 2; by++)
                    for (int bx = 0; bx < (lv.dim[0] + 1) / 2; bx++)
                        for (int k8 = 0; k8 < 8; k8++) {
                            const int x = 2 * bx + (k8 & 1), y = 2 * by + ((k8 >> 1) & 1), z = 2 * bz + (k8 >> 2);
                            if (!lv.active(x, y, z)) continue;
                            for (int d = 0; d < 27; d++) {
                                if (d == 13) continue;
                                int p[3] = {x + d % 3 - 1, y + (d / 3) % 3 - 1, z + d / 9 - 1};
                                if (!lv.inside(p[0], p[1], p[2])) continue;
                                int prev[3] = {0, 0, 0};
                                for (int ln = l; ln < depth; ln++) {
                                    if (L[ln].active(p[0], p[1], p[2])) {
                                        if (ln <= l + 1) break;
                                        uint8_t& s = L[ln - 1].st[L[ln - 1].lin(prev[0], prev[1], prev[2])];
                                        added += (s & kActive) ? 0 : 1;
                                        s = (uint8_t)kActive;
                                        again = true;
                                    }
                                    for (int a = 0; a < 3; a++) {
                                        prev[a] = p[a];
                                        p[a] >>= 1;
                                    }
                                }
                            }
                        }
        }
    }
    for (int l = 1; l < depth; l++) {
        Level& c = L[l];
        for (int z = 0; z < c.dim[2]; z++)
            for (int y = 0; y < c.dim[1]; y++)
                for (int x = 0; x < c.dim[0]; x++) {
                    uint8_t& s = c.st[c.lin(x, y, z)];
                    if ((s & kActive) && any_child(L[l - 1], x, y, z)) s |= (uint8_t)kRefined;
                }
    }
    return added;
}

int build_level(Level& lv)
{
    std::vector<uint8_t> act(lv.nvox());
    for (size_t i = 0; i < act.size(); i++) act[i] = lv.st[i] & kActive;
    nx_bgrid g = nullptr;
    if (int e = nx_bgrid_create(lv.dim, act.data(), 0, &g)) return e;
    uint64_t n_act = 0;
    nx_bgrid_info(g, &lv.n_blocks, &n_act);
    lv.origins.resize(3 * (size_t)lv.n_blocks);
    lv.masks.resize(16 * (size_t)lv.n_blocks);
    lv.conn.resize(27 * (size_t)lv.n_blocks);
    nx_bgrid_tables(g, lv.origins.data(), lv.masks.data(), lv.conn.data());
    nx_bgrid_destroy(g);
    for (int a = 0; a < 3; a++) lv.nb[a] = (lv.dim[a] + 7) / 8;
    lv.table.assign((size_t)lv.nb[0] * lv.nb[1] * lv.nb[2], kInvalidBlock);
    for (uint32_t b = 0; b < lv.n_blocks; b++)
        lv.table[((size_t)(lv.origins[3 * b + 2] >> 3) * lv.nb[1] + (lv.origins[3 * b + 1] >> 3)) * lv.nb[0] +
                 (lv.origins[3 * b] >> 3)] = b;
    return NX_OK;
}

// uncleOffset (util.h:19-37): offset, from the parent of x, of the parent of x + d
inline int uoff(int x, int d)
{
    const int s = (x & 1) + d;
    return s <= 0 ? s : s - 1;
}

// The class of every leaf (cls[l] per level-l voxel; nullptr: the lid-driven
// cavity of lidDrivenCavity.h:8-78, leaf voxels of level 0 on the domain
// faces bounce-back, those with y = Y-1 the lid) and its populations: w_q
// (bulk), 0 (bounce-back), -6 w_q (c_q . uw) for moving-wall cells (the lid,
// or the inlet of flowOverShape.h:62-104), evaluated as the reference does in
// T with binary64 constants.
template <typename T>
void fill_cells(nx_mgrid_s& g, const uint8_t* const* cls_in, const double uw[3], std::vector<T>* pop)
{
    const HostLat& L = g.lat;
    for (int l = 0; l < g.depth; l++) {
        Level& lv = g.L[l];
        pop[l].assign((size_t)L.Q * kBV * lv.n_blocks, T(0));
        for (int z = 0; z < lv.dim[2]; z++)
            for (int y = 0; y < lv.dim[1]; y++)
                for (int x = 0; x < lv.dim[0]; x++) {
                    const size_t i = lv.lin(x, y, z);
                    if (!(lv.st[i] & kActive)) continue;
                    int cls = NX_CELL_BULK;
                    const int64_t s = lv.slot(x, y, z);
                    if (!(lv.st[i] & kRefined)) {
                        if (cls_in)
                            cls = cls_in[l][i];
                        else if (l == 0 && (x == 0 || y == 0 || z == 0 || x == lv.dim[0] - 1 ||
                                            y == lv.dim[1] - 1 || z == lv.dim[2] - 1))
                            cls = y == lv.dim[1] - 1 ? NX_CELL_MOVINGWALL : NX_CELL_BOUNCEBACK;
                        for (int q = 0; q < L.Q; q++) {
                            T p = (T)L.w[q];
                            if (cls != NX_CELL_BULK) p = 0;
                            if (cls == NX_CELL_MOVINGWALL) {
                                p = (T)((double)p + (double)L.c[q][0] * uw[0]);
                                p = (T)((double)p + (double)L.c[q][1] * uw[1]);
                                p = (T)((double)p + (double)L.c[q][2] * uw[2]);
                                p = (T)((double)p * (-6. * L.w[q]));
                            }
                            pop[l][lv.elem(s, q)] = p;
                        }
                    }
                    lv.st[i] = (uint8_t)((lv.st[i] & (kActive | kRefined)) | (cls << kClsShift));
                }
    }
}

// hasParent (mPartition_imp.h:172-208): the memory block (8^3) of level
// l + 1 holding the parent of level-l voxel (x, y, z) exists
bool has_parent(const nx_mgrid_s& g, int l, int x, int y, int z)
{
    if (l + 1 >= g.depth) return false;
    const Level& c = g.L[l + 1];
    const int bx = (x >> 1) >> 3, by = (y >> 1) >> 3, bz = (z >> 1) >> 3;
    if (bx >= c.nb[0] || by >= c.nb[1] || bz >= c.nb[2]) return false;
    return c.table[((size_t)bz * c.nb[1] + by) * c.nb[0] + bx] != kInvalidBlock;
}

// Link states of every bulk leaf voxel of level l (stream.h:170-225 decided
// per link), 3 bits per population, W words per voxel slot (word k at
// k * slots); returns the number of links in state "none".
uint64_t links_of(const nx_mgrid_s& g, int l, std::vector<uint64_t>& out)
{
    uint64_t n_skip = 0;
    const Level& lv = g.L[l];
    const HostLat& L = g.lat;
    out.assign((size_t)L.W * lv.slots(), 0);
    for (int z = 0; z < lv.dim[2]; z++)
        for (int y = 0; y < lv.dim[1]; y++)
            for (int x = 0; x < lv.dim[0]; x++) {
                const uint8_t st = lv.st[lv.lin(x, y, z)];
                if (!bulk_leaf(st)) continue;
                const bool parent = has_parent(g, l, x, y, z);
                uint64_t lk[2] = {0, 0};
                for (int q = 0; q < L.Q; q++) {
                    if (q == L.rest) continue;  // own population: plain
                    const int nx_ = x - L.c[q][0], ny = y - L.c[q][1], nz = z - L.c[q][2];
                    uint64_t kind = kSkip;
                    if (lv.active(nx_, ny, nz)) {
                        const uint8_t ns = lv.st[lv.lin(nx_, ny, nz)];
                        if (ns & kRefined)
                            kind = kCoal;
                        else
                            kind = cls_bits(ns) == NX_CELL_BULK ? kPlain : kWall;
                    } else if (parent) {
                        const Level& c = g.L[l + 1];
                        const int ux = (x >> 1) + uoff(x, -L.c[q][0]), uy = (y >> 1) + uoff(y, -L.c[q][1]),
                                  uz = (z >> 1) + uoff(z, -L.c[q][2]);
                        if (c.active(ux, uy, uz)) kind = kExpl;
                    }
                    lk[q / kLinksPerWord] |= kind << (3 * (q % kLinksPerWord));
                    n_skip += kind == kSkip;
                }
                for (int k = 0; k < L.W; k++) out[(size_t)k * lv.slots() + lv.slot(x, y, z)] = lk[k];
            }
    return n_skip;
}

int check_level(const nx_mgrid_s* g, int32_t level)
{
    if (!g) return set_error(NX_ERR_INVALID, "null mgrid");
    if (level < 0 || level >= g->depth)
        return set_error(NX_ERR_INVALID, "mgrid: level " + std::to_string(level) + " outside [0, " +
                                             std::to_string(g->depth) + ")");
    return NX_OK;
}

}  // namespace

extern "C" int nx_mgrid_create_levels(const int32_t dim[3], int32_t depth, const uint8_t* const* lams,
                                      int32_t storage, int32_t lattice, nx_mgrid* out)
{
    if (!dim || !lams || !out) return set_error(NX_ERR_INVALID, "nx_mgrid_create_levels: null argument");
    if (depth < 2 || depth > kMaxDepth)
        return set_error(NX_ERR_INVALID, "nx_mgrid_create_levels: depth in [2, " + std::to_string(kMaxDepth) + "]");
    for (int l = 0; l < depth; l++)
        if (!lams[l]) return set_error(NX_ERR_INVALID, "nx_mgrid_create_levels: null lambda array");
    for (int a = 0; a < 3; a++)
        if (dim[a] < 2 || dim[a] > (1 << 14))
            return set_error(NX_ERR_INVALID, "nx_mgrid_create_levels: dimensions in [2, 16384]");
    if (storage != NX_F64 && storage != NX_F32)
        return set_error(NX_ERR_UNSUPPORTED, "nx_mgrid_create: storage NX_F64 or NX_F32 (the reference's types)");
    if (lattice != NX_MGRID_BGK19 && lattice != NX_MGRID_KBC27)
        return set_error(NX_ERR_INVALID, "nx_mgrid_create_levels: lattice NX_MGRID_BGK19 or NX_MGRID_KBC27");
    auto* g = new nx_mgrid_s;
    g->storage = storage;
    g->lattice = lattice;
    g->lat = lattice == NX_MGRID_KBC27 ? host_lat_of<L27>() : host_lat_of<L19>();
    g->depth = depth;
    for (int l = 0; l < depth; l++) g->L[l].nq = g->lat.Q;
    activate(dim, depth, lams, g->L);
    for (int l = 0; l < depth; l++)
        if (int e = build_level(g->L[l])) {
            nx_mgrid_destroy(g);
            return e;
        }
    *out = g;
    return NX_OK;
}

extern "C" int nx_mgrid_create(const int32_t dim[3], const uint8_t* lam0, const uint8_t* lam1, int32_t storage,
                               nx_mgrid* out)
{
    if (!lam0 || !lam1) return set_error(NX_ERR_INVALID, "nx_mgrid_create: null argument");
    const uint8_t* lams[2] = {lam0, lam1};
    return nx_mgrid_create_levels(dim, 2, lams, storage, NX_MGRID_BGK19, out);
}

extern "C" int nx_mgrid_depth(nx_mgrid g, int32_t* depth, int32_t* dims)
{
    if (!g) return set_error(NX_ERR_INVALID, "null mgrid");
    if (depth) *depth = g->depth;
    if (dims)
        for (int l = 0; l < g->depth; l++)
            for (int a = 0; a < 3; a++) dims[3 * l + a] = g->L[l].dim[a];
    return NX_OK;
}

extern "C" int nx_mgrid_populations(nx_mgrid g, int32_t* q)
{
    if (!g || !q) return set_error(NX_ERR_INVALID, "nx_mgrid_populations: null argument");
    *q = g->lat.Q;
    return NX_OK;
}

extern "C" int nx_mgrid_info(nx_mgrid g, uint32_t* n_blocks, uint64_t* n_active, uint64_t* n_leaf,
                             uint64_t* n_store, int32_t* fine_fused)
{
    if (!g) return set_error(NX_ERR_INVALID, "null mgrid");
    if (fine_fused) {  // bit l: level l's stream runs fused with its next collide
        *fine_fused = 0;
        for (int l = 0; l < g->depth; l++)
            if (g->ready && !g->L[l].keep) *fine_fused |= 1 << l;
    }
    uint64_t ns = 0;
    for (int l = 0; l < g->depth; l++) {
        uint64_t a = 0, lf = 0;
        for (uint8_t s : g->L[l].st) {
            a += (s & kActive) ? 1 : 0;
            lf += ((s & kActive) && !(s & kRefined)) ? 1 : 0;
        }
        if (n_blocks) n_blocks[l] = g->L[l].n_blocks;
        if (n_active) n_active[l] = a;
        if (n_leaf) n_leaf[l] = lf;
        ns += g->store[l].n_src;
    }
    if (n_store) *n_store = ns;
    return NX_OK;
}

extern "C" int nx_mgrid_state_level(nx_mgrid g, int32_t level, uint8_t* st)
{
    if (int e = check_level(g, level)) return e;
    if (!st) return set_error(NX_ERR_INVALID, "nx_mgrid_state_level: null output");
    std::copy(g->L[level].st.begin(), g->L[level].st.end(), st);
    return NX_OK;
}

extern "C" int nx_mgrid_state(nx_mgrid g, uint8_t* st0, uint8_t* st1)
{
    if (!g) return set_error(NX_ERR_INVALID, "null mgrid");
    if (st0) std::copy(g->L[0].st.begin(), g->L[0].st.end(), st0);
    if (st1) std::copy(g->L[1].st.begin(), g->L[1].st.end(), st1);
    return NX_OK;
}

namespace {

int setup_cells(nx_mgrid g, const uint8_t* const* cls, const double uw[3])
{
    const int D = g->depth;
    const size_t es = g->storage == NX_F64 ? 8 : 4;
    // populations
    std::vector<double> p64[kMaxDepth];
    std::vector<float> p32[kMaxDepth];
    if (g->storage == NX_F64)
        fill_cells<double>(*g, cls, uw, p64);
    else
        fill_cells<float>(*g, cls, uw, p32);
    for (int l = 0; l < D; l++) {
        Level& lv = g->L[l];
        const size_t elems = (size_t)g->lat.Q * kBV * lv.n_blocks;
        if (int e = upload((void**)&lv.d_conn, lv.conn.data(), lv.conn.size() * 4)) return e;
        if (int e = upload((void**)&lv.d_origins, lv.origins.data(), lv.origins.size() * 4)) return e;
        if (int e = upload((void**)&lv.d_table, lv.table.data(), lv.table.size() * 4)) return e;
        const void* src = g->storage == NX_F64 ? (const void*)p64[l].data() : (const void*)p32[l].data();
        if (int e = upload(&lv.fin, src, elems * es)) return e;
        if (int e = upload(&lv.fout, src, elems * es)) return e;
        std::vector<uint32_t> st((size_t)kBV * lv.n_blocks, 0);
        for (int z = 0; z < lv.dim[2]; z++)
            for (int y = 0; y < lv.dim[1]; y++)
                for (int x = 0; x < lv.dim[0]; x++) {
                    const int64_t s = lv.slot(x, y, z);
                    if (s >= 0) st[s] = lv.st[lv.lin(x, y, z)];
                }
        if (int e = upload((void**)&lv.d_state, st.data(), st.size() * 4)) return e;
        std::vector<uint64_t> lk;
        const uint64_t n_skip = links_of(*g, l, lk);
        lv.keep = n_skip != 0;
        if (int e = upload((void**)&lv.d_links, lk.data(), lk.size() * 8)) return e;
    }
    // fine store gather lists (store.h:98-163) and sumStore counts
    // (init.h:8-55) of every level pair: fine voxels in z, y, x order, q
    // ascending.  The counts take every active fine voxel with a parent,
    // the store only those without children whose csChild is active.
    for (int l = 0; l + 1 < D; l++) {
        const Level& f = g->L[l];
        const Level& c = g->L[l + 1];
        std::vector<float> cnt((size_t)g->lat.Q * kBV * c.n_blocks, 0.f);
        std::vector<std::pair<uint64_t, uint64_t>> tl;  // (target element, source element)
        for (int z = 0; z < f.dim[2]; z++)
            for (int y = 0; y < f.dim[1]; y++)
                for (int x = 0; x < f.dim[0]; x++) {
                    if (!f.active(x, y, z) || !has_parent(*g, l, x, y, z)) continue;
                    const bool leaf = !(f.st[f.lin(x, y, z)] & kRefined);
                    const int px = x >> 1, py = y >> 1, pz = z >> 1;
                    for (int q = 0; q < g->lat.Q; q++) {
                        if (q == g->lat.rest) continue;
                        const int* cq = g->lat.c[q];
                        const int ux = uoff(x, cq[0]), uy = uoff(y, cq[1]), uz = uoff(z, cq[2]);
                        if (f.active(x + ux, y + uy, z + uz)) continue;
                        if (!c.active(px + ux, py + uy, pz + uz)) continue;
                        const int sx = ux - cq[0], sy = uy - cq[1], sz = uz - cq[2];
                        const int64_t cs = c.slot(px + sx, py + sy, pz + sz);
                        if (cs < 0) continue;
                        cnt[c.elem(cs, q)] += 1.f;
                        if (!leaf || !f.active(x + sx, y + sy, z + sz)) continue;
                        tl.emplace_back(c.elem(cs, q), f.elem(f.slot(x, y, z), q));
                    }
                }
        for (float& v : cnt)
            if ((double)v > 0.001) v = g->storage == NX_F64 ? (float)(1.0 / (2.0 * (double)v)) : 1.f / (2.f * v);
        if (int e = upload((void**)&g->L[l + 1].d_ss, cnt.data(), cnt.size() * 4)) return e;
        std::stable_sort(tl.begin(), tl.end(),
                         [](const std::pair<uint64_t, uint64_t>& a, const std::pair<uint64_t, uint64_t>& b) {
                             return a.first < b.first;
                         });
        std::vector<uint64_t> target, src;
        std::vector<uint32_t> first;
        for (size_t k = 0; k < tl.size(); k++) {
            if (k == 0 || tl[k].first != tl[k - 1].first) {
                target.push_back(tl[k].first);
                first.push_back((uint32_t)k);
            }
            src.push_back(tl[k].second);
        }
        first.push_back((uint32_t)tl.size());
        Store& sl = g->store[l];
        sl.n = (uint32_t)target.size();
        sl.n_src = src.size();
        if (int e = upload((void**)&sl.d_target, target.data(), target.size() * 8)) return e;
        if (int e = upload((void**)&sl.d_first, first.data(), first.size() * 4)) return e;
        if (int e = upload((void**)&sl.d_src, src.data(), src.size() * 8)) return e;
    }
    g->ready = true;
    return NX_OK;
}

}  // namespace

extern "C" int nx_mgrid_cavity(nx_mgrid g, double ulb)
{
    if (!g) return set_error(NX_ERR_INVALID, "null mgrid");
    if (!std::isfinite(ulb)) return set_error(NX_ERR_INVALID, "nx_mgrid_cavity: non-finite lid speed");
    const double uw[3] = {ulb, 0., 0.};
    return setup_cells(g, nullptr, uw);
}

extern "C" int nx_mgrid_setup_cells(nx_mgrid g, const uint8_t* const* cls, const double u_wall[3])
{
    if (!g || !cls || !u_wall) return set_error(NX_ERR_INVALID, "nx_mgrid_setup_cells: null argument");
    for (int l = 0; l < g->depth; l++) {
        if (!cls[l]) return set_error(NX_ERR_INVALID, "nx_mgrid_setup_cells: null class array");
        const Level& lv = g->L[l];
        for (size_t i = 0; i < lv.nvox(); i++)
            if ((lv.st[i] & kActive) && !(lv.st[i] & kRefined) && cls[l][i] != NX_CELL_BULK &&
                cls[l][i] != NX_CELL_BOUNCEBACK && cls[l][i] != NX_CELL_MOVINGWALL)
                return set_error(NX_ERR_INVALID, "nx_mgrid_setup_cells: classes NX_CELL_BULK, NX_CELL_BOUNCEBACK or "
                                                 "NX_CELL_MOVINGWALL");
    }
    for (int a = 0; a < 3; a++)
        if (!std::isfinite(u_wall[a])) return set_error(NX_ERR_INVALID, "nx_mgrid_setup_cells: non-finite velocity");
    return setup_cells(g, cls, u_wall);
}

namespace {

// One coarse step as the sequence of the reference's recursion
// (lbmMultiRes.h:141-263): per level l from the top, collide l and store
// l -> l+1, the step of level l-1, stream l, and below the top level the
// same once more.
struct Op {
    uint8_t kind;   // 0 collide, 1 store, 2 stream
    uint8_t level;
    bool last;      // the level's last stream of the coarse step
};

void expand(int l, int depth, std::vector<Op>& ops)
{
    for (int half = 0; half < 2; half++) {
        ops.push_back({0, (uint8_t)l, false});
        if (l + 1 < depth) ops.push_back({1, (uint8_t)l, false});
        if (l != 0) expand(l - 1, depth, ops);
        ops.push_back({2, (uint8_t)l, false});
        if (l == depth - 1) break;
    }
}

template <int S, class LT>
struct Stepper {
    using T = typename MS<S>::T;
    nx_mgrid_s& g;
    double omega0;
    hipStream_t st;

    // computeOmega (util.h:40-45): omega0 is the coarsest level's
    double omega(int l) const
    {
        const int il = g.depth - l - 1;
        return 2 * omega0 / (std::scalbln(1.0, il + 1) + (1. - std::scalbln(1.0, il)) * omega0);
    }
    UpLevel up(int l) const
    {
        if (l + 1 >= g.depth) return UpLevel{nullptr, nullptr, {0, 0, 0}};
        const Level& c = g.L[l + 1];
        return UpLevel{c.fout, c.d_table, {c.nb[0], c.nb[1], c.nb[2]}};
    }
    void collide(int l) const
    {
        const Level& lv = g.L[l];
        if (lv.n_blocks)
            hipLaunchKernelGGL((k_mr_collide<S, LT>), dim3(lv.n_blocks), dim3(kBV), 0, st, (const T*)lv.fin,
                               (T*)lv.fout, lv.d_state, omega(l));
    }
    void store(int l) const
    {
        const Store& sl = g.store[l];
        if (l + 1 < g.depth && sl.n)
            hipLaunchKernelGGL((k_mr_store<S>), dim3((sl.n + 255) / 256), dim3(256), 0, st, (const T*)g.L[l].fout,
                               (T*)g.L[l + 1].fout, sl.d_target, sl.d_first, sl.d_src, sl.n);
    }
    template <bool COLLIDE, bool KEEP>
    void stream(int l) const
    {
        const Level& lv = g.L[l];
        if (lv.n_blocks)
            hipLaunchKernelGGL((k_mr_stream<S, LT, COLLIDE, KEEP>), dim3(lv.n_blocks), dim3(kBV), 0, st,
                               (const T*)lv.fout, (T*)lv.fin, lv.d_state, lv.d_links, (uint64_t)lv.slots(),
                               lv.d_conn, lv.d_origins, (const float*)lv.d_ss, up(l), COLLIDE ? omega(l) : 0.0);
    }

    // n coarse steps.  A level's stream and its next collide touch no other
    // level and nothing of the level runs between them (the operations in
    // between are the coarser levels' stream, collide and store), so on a
    // level without links in state "none" they run as one pass (fout -> fin
    // buffer, then the two buffers trade roles) and the collide is skipped
    // when its turn comes; the store after it then reads the fused pass's
    // output.  The last stream of each level in the call is not fused, so
    // the fields between calls are those of the unfused sequence.
    void run(const std::vector<Op>& ops, int n) const
    {
        bool done[kMaxDepth] = {};  // the level's next collide already ran in a fused pass
        for (int k = 0; k < n; k++)
            for (const Op& op : ops) {
                const int l = op.level;
                Level& lv = g.L[l];
                if (op.kind == 0) {
                    if (done[l])
                        done[l] = false;
                    else
                        collide(l);
                } else if (op.kind == 1) {
                    store(l);
                } else if (!lv.keep && !(op.last && k == n - 1)) {
                    stream<true, false>(l);
                    std::swap(lv.fin, lv.fout);
                    done[l] = true;
                } else if (lv.keep) {
                    stream<false, true>(l);
                } else {
                    stream<false, false>(l);
                }
            }
    }
};

}  // namespace

extern "C" int nx_mgrid_step(nx_mgrid g, double omega0, int32_t n_steps, nx_stream s)
{
    if (!g || n_steps < 0) return set_error(NX_ERR_INVALID, "nx_mgrid_step: bad arguments");
    if (!g->ready) return set_error(NX_ERR_NOT_INIT, "nx_mgrid_step: call nx_mgrid_cavity first");
    if (!(omega0 > 0.0 && omega0 < 2.0)) return set_error(NX_ERR_INVALID, "nx_mgrid_step: omega0 in (0, 2)");
    std::vector<Op> ops;
    expand(g->depth - 1, g->depth, ops);
    for (int l = 0; l < g->depth; l++)
        for (size_t i = ops.size(); i-- > 0;)
            if (ops[i].kind == 2 && ops[i].level == l) {
                ops[i].last = true;
                break;
            }
    const hipStream_t st = (hipStream_t)s;
    if (g->lattice == NX_MGRID_KBC27) {
        if (g->storage == NX_F64)
            Stepper<NX_F64, L27>{*g, omega0, st}.run(ops, n_steps);
        else
            Stepper<NX_F32, L27>{*g, omega0, st}.run(ops, n_steps);
    } else {
        if (g->storage == NX_F64)
            Stepper<NX_F64, L19>{*g, omega0, st}.run(ops, n_steps);
        else
            Stepper<NX_F32, L19>{*g, omega0, st}.run(ops, n_steps);
    }
    NX_CHECK_LAUNCH("mgrid step");
    return NX_OK;
}

extern "C" int nx_mgrid_fields(nx_mgrid g, int32_t level, int32_t which, void* dense)
{
    if (int e = check_level(g, level)) return e;
    if ((which != 0 && which != 1) || !dense) return set_error(NX_ERR_INVALID, "nx_mgrid_fields: bad arguments");
    if (!g->ready) return set_error(NX_ERR_NOT_INIT, "nx_mgrid_fields: call nx_mgrid_cavity first");
    const Level& lv = g->L[level];
    const size_t es = g->storage == NX_F64 ? 8 : 4, n = lv.nvox();
    std::vector<uint8_t> buf((size_t)lv.nq * kBV * lv.n_blocks * es);
    NX_CHECK_HIP(hipDeviceSynchronize());
    if (!buf.empty())
        NX_CHECK_HIP(hipMemcpy(buf.data(), which ? lv.fout : lv.fin, buf.size(), hipMemcpyDeviceToHost));
    for (int z = 0; z < lv.dim[2]; z++)
        for (int y = 0; y < lv.dim[1]; y++)
            for (int x = 0; x < lv.dim[0]; x++) {
                const int64_t s = lv.slot(x, y, z);
                if (s < 0) continue;
                for (int q = 0; q < lv.nq; q++)
                    std::memcpy((uint8_t*)dense + ((size_t)q * n + lv.lin(x, y, z)) * es,
                                buf.data() + lv.elem(s, q) * es, es);
            }
    return NX_OK;
}

extern "C" int nx_mgrid_destroy(nx_mgrid g)
{
    if (!g) return NX_OK;
    for (auto& lv : g->L)
        for (void* p : {(void*)lv.d_state, (void*)lv.d_links, (void*)lv.d_conn, (void*)lv.d_origins,
                        (void*)lv.d_table, (void*)lv.d_ss, lv.fin, lv.fout})
            if (p) hipFree(p);
    for (auto& sl : g->store)
        for (void* p : {(void*)sl.d_target, (void*)sl.d_first, (void*)sl.d_src})
            if (p) hipFree(p);
    delete g;
    return NX_OK;
}
