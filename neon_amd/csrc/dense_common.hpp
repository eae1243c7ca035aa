// dense_common.hpp -- dense-grid (dGrid) geometry and the pull stream shared
// by the dense kernels (d3q19_dense.hip, d3q19_container.hip).
#pragma once

#include <algorithm>
#include <type_traits>

#include "nx_common.hpp"

namespace nx {
namespace dense {


constexpr int kWG = 256;

struct DGeom {
    int32_t X, Y, Z;
    uint32_t XY;
    uint32_t n;       // cells
    uint64_t plane;   // cells per population plane
    int32_t periodic;
};

inline DGeom make_geom(const nx_dense_desc* d)
{
    DGeom g;
    g.X = d->dim[0];
    g.Y = d->dim[1];
    g.Z = d->dim[2];
    g.XY = (uint32_t)g.X * (uint32_t)g.Y;
    g.n = g.XY * (uint32_t)g.Z;
    g.plane = g.n;
    g.periodic = d->periodic;
    return g;
}

struct Par {
    double omega;
    double g[3];
};

// The common wall value of each population at the x ends of a box (raw
// storage bits), for the edge-speculating steps.
struct XWall {
    uint64_t v[Q];
};

template <typename T>
__device__ __forceinline__ uint64_t raw_bits(T v)
{
    if constexpr (sizeof(T) == 8) return __builtin_bit_cast(uint64_t, v);
    else if constexpr (sizeof(T) == 4) return __builtin_bit_cast(uint32_t, v);
    else return __builtin_bit_cast(uint16_t, v);
}

template <typename T>
__device__ __forceinline__ T from_raw(uint64_t u)
{
    if constexpr (sizeof(T) == 8) return __builtin_bit_cast(T, u);
    else if constexpr (sizeof(T) == 4) return __builtin_bit_cast(T, (uint32_t)u);
    else return __builtin_bit_cast(T, (uint16_t)u);
}

// Upstream index of population q of cell i (x - c_q), with periodic wrap.
// Returns false when the upstream cell lies outside the grid.
__device__ __forceinline__ bool upstream(const DGeom& g, int x, int y, int z, int q, uint32_t& j)
{
    int a = x - CX[q], b = y - CY[q], c = z - CZ[q];
    if (g.periodic & NX_PERIODIC_X) a = a < 0 ? a + g.X : (a >= g.X ? a - g.X : a);
    if (g.periodic & NX_PERIODIC_Y) b = b < 0 ? b + g.Y : (b >= g.Y ? b - g.Y : b);
    if (g.periodic & NX_PERIODIC_Z) c = c < 0 ? c + g.Z : (c >= g.Z ? c - g.Z : c);
    j = (uint32_t)a + (uint32_t)g.X * (uint32_t)b + g.XY * (uint32_t)c;
    return a >= 0 && b >= 0 && c >= 0 && a < g.X && b < g.Y && c < g.Z;
}

__device__ __forceinline__ void coords(const DGeom& g, uint32_t i, int& x, int& y, int& z)
{
    z = (int)(i / g.XY);
    const uint32_t r = i - (uint32_t)z * g.XY;
    y = (int)(r / (uint32_t)g.X);
    x = (int)(r - (uint32_t)y * (uint32_t)g.X);
}

// Pull stream of one bulk cell (LbmTools.h:78-168): for q != 9,
//   wall bit q set:  f_q = fin(x, opp q) + fin(x - c_q, opp q)
//   otherwise:       f_q = fin(x - c_q, q)
// A masked link whose upstream cell does not exist contributes 0
// (NX_FLAG_OPEN; the reference leaves that value undefined, NghData.h:13-16).
template <int S, bool PERIODIC>
__device__ __forceinline__ void pull(const typename Stor<S>::T* __restrict__ fin, const DGeom& g,
                                     uint32_t i, uint32_t fl, typename Stor<S>::P f[Q])
{
    using T = typename Stor<S>::T;
    using P = typename Stor<S>::P;
    const uint32_t m = fl & NX_MASK_BITS;
    uint32_t jj[Q];
    uint32_t missing = 0;
    if (PERIODIC || (fl & NX_FLAG_OPEN)) {
        int x, y, z;
        coords(g, i, x, y, z);
#pragma unroll
        for (int q = 0; q < Q; q++) {
            if (q == 9) continue;
            uint32_t j;
            const bool ok = upstream(g, x, y, z, q, j);
            jj[q] = ok ? j : i;
            missing |= ok ? 0u : (1u << q);
        }
    } else {
#pragma unroll
        for (int q = 0; q < Q; q++) {
            if (q == 9) continue;
            const int off = CX[q] + CY[q] * g.X + CZ[q] * (int)g.XY;
            // flags from the wall-mask kernel keep this in range; the clamp
            // keeps any caller-supplied flag word from addressing outside
            jj[q] = min(i - (uint32_t)off, g.n - 1u);
        }
    }
    T nb[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) {
        if (q == 9) {
            nb[9] = fin[9 * g.plane + i];
            continue;
        }
        const bool w = (m >> q) & 1u;
        const uint64_t src = w ? (uint64_t)OPP[q] : (uint64_t)q;
        nb[q] = fin[src * g.plane + jj[q]];
    }
    if (m) {
        // own populations of the masked links: one batch of loads, combined
        // with selects (a branch per link would wait once per link)
        T own[Q];
#pragma unroll
        for (int q = 0; q < Q; q++)
            if (q != 9) own[q] = fin[(uint64_t)OPP[q] * g.plane + i];
#pragma unroll
        for (int q = 0; q < Q; q++) {
            if (q == 9) continue;
            const bool w = (m >> q) & 1u;
            const P up = ((missing >> q) & 1u) ? P(0) : Stor<S>::raw(nb[q]);
            f[q] = w ? Stor<S>::bulk(own[q], OPP[q]) + up : Stor<S>::bulk(nb[q], q);
        }
    } else {
#pragma unroll
        for (int q = 0; q < Q; q++)
            if (q != 9) f[q] = Stor<S>::bulk(nb[q], q);
    }
    f[9] = Stor<S>::bulk(nb[9], 9);
}

template <int LP, typename T>
__device__ __forceinline__ T ldp(const T* p)
{
    if constexpr (LP == 1) return __builtin_nontemporal_load(p);
    else return *p;
}

// Empty asm that consumes a value: keeps a load ahead of the branch that
// follows (LLVM otherwise sinks loads into the block that uses them, which
// serialises the flag load and the population loads).
template <typename T>
__device__ __forceinline__ void pin(T& v)
{
    if constexpr (sizeof(T) <= 4) {
        asm volatile("" : "+v"(v));
    } else {
        constexpr int W = (int)(sizeof(T) / 4);
        uint32_t w[W];
        __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
        for (int k = 0; k < W; k++) asm volatile("" : "+v"(w[k]));
        __builtin_memcpy(&v, w, sizeof(T));
    }
}

inline int check(const nx_dense_desc* d)
{
    if (!d) return set_error(NX_ERR_INVALID, "null descriptor");
    for (int a = 0; a < 3; a++)
        if (d->dim[a] < 1) return set_error(NX_ERR_INVALID, "grid dimensions must be >= 1");
    const uint64_t n = (uint64_t)d->dim[0] * d->dim[1] * d->dim[2];
    if (n >= (1ull << 32) - kWG) return set_error(NX_ERR_INVALID, "grid too large (>= 2^32 cells)");
    if (d->storage < NX_F64 || d->storage > NX_F16DEV)
        return set_error(NX_ERR_INVALID, "unknown storage type");
    if (d->compute < NX_COMPUTE_F64 || d->compute > NX_COMPUTE_REF)
        return set_error(NX_ERR_INVALID, "unknown compute mode");
    if ((d->storage == NX_F32DEV || d->storage == NX_F16DEV) && d->compute != NX_COMPUTE_F32)
        return set_error(NX_ERR_UNSUPPORTED, "deviation storage requires NX_COMPUTE_F32");
    if (d->storage == NX_F64 && d->compute == NX_COMPUTE_F32)
        return set_error(NX_ERR_UNSUPPORTED, "binary64 storage requires binary64 compute");
    return NX_OK;
}

inline uint32_t nwg(const DGeom& g) { return (g.n + kWG - 1) / kWG; }

template <int V>
using IC = std::integral_constant<int, V>;


// Calls f(integral_constant<S>, integral_constant<CM>) for the (storage,
// compute) pair of d.
template <typename F>
inline int by_type(const nx_dense_desc* d, F&& f)
{
    switch (d->storage) {
        case NX_F64: f(IC<NX_F64>{}, IC<NX_COMPUTE_F64>{}); return NX_OK;
        case NX_F32:
            if (d->compute == NX_COMPUTE_F64) f(IC<NX_F32>{}, IC<NX_COMPUTE_F64>{});
            else if (d->compute == NX_COMPUTE_REF) f(IC<NX_F32>{}, IC<NX_COMPUTE_REF>{});
            else f(IC<NX_F32>{}, IC<NX_COMPUTE_F32>{});
            return NX_OK;
        case NX_F32DEV: f(IC<NX_F32DEV>{}, IC<NX_COMPUTE_F32>{}); return NX_OK;
        case NX_F16DEV: f(IC<NX_F16DEV>{}, IC<NX_COMPUTE_F32>{}); return NX_OK;
    }
    return set_error(NX_ERR_INVALID, "unknown storage type");
}

}  // namespace dense
}  // namespace nx
