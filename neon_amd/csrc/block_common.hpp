// block_common.hpp -- device helpers shared by the kernels on 8^3-block
// grids (the bGrid step and the two-level refinement): the connectivity row
// in scalar registers, the upstream voxel of a population through it
// (bPartition_imp.h:109-124, 218-309), compile-time population loops, and a
// pin that keeps a load ahead of the branch using it.
#pragma once

#include <type_traits>
#include <utility>

#include "nx_common.hpp"

// pins a loaded value ahead of the branch that follows (see dense_common.hpp)
template <typename T>
__device__ __forceinline__ void nx_pin(T& v)
{
    if constexpr (sizeof(T) <= 4) {
        asm volatile("" : "+v"(v));
    } else {
        uint32_t w[sizeof(T) / 4];
        __builtin_memcpy(w, &v, sizeof(T));
#pragma unroll
        for (unsigned k = 0; k < sizeof(T) / 4; k++) asm volatile("" : "+v"(w[k]));
        __builtin_memcpy(&v, w, sizeof(T));
    }
}

namespace nx {

constexpr int kBV = 512;  // voxels per block

// The block's connectivity row, read unconditionally once per workgroup (a
// uniform address: scalar loads into SGPRs).  Reading it inside a per-lane
// select would make every entry a per-lane load that each population load
// then waits for.
struct Conn {
    uint32_t c[27];
};

__device__ __forceinline__ Conn load_conn(const uint32_t* __restrict__ conn, uint32_t b)
{
    Conn r;
#pragma unroll
    for (int d = 0; d < 27; d++) r.c[d] = conn[(uint64_t)b * 27 + d];
    return r;
}

// Neighbour voxel (block t, in-block index nv) of voxel (x,y,z) of block b at
// offset (s0, s1, s2), each in {-1, 0, 1}: the block is b unless the offset
// leaves it across a face, an edge or a corner, then the connectivity row's
// entry 13 + s0 + 3 s1 + 9 s2 restricted to the axes crossed.
template <int s0, int s1, int s2>
__device__ __forceinline__ void upstream_off(uint32_t b, const Conn& cn, int x, int y, int z, uint32_t& t,
                                             uint32_t& nv)
{
    const bool c0 = s0 != 0 && x == (s0 < 0 ? 0 : 7);
    const bool c1 = s1 != 0 && y == (s1 < 0 ? 0 : 7);
    const bool c2 = s2 != 0 && z == (s2 < 0 ? 0 : 7);
    t = b;
    if constexpr (s0 != 0) t = c0 ? cn.c[13 + s0] : t;
    if constexpr (s1 != 0) t = c1 ? cn.c[13 + 3 * s1] : t;
    if constexpr (s2 != 0) t = c2 ? cn.c[13 + 9 * s2] : t;
    if constexpr (s0 != 0 && s1 != 0) t = (c0 && c1) ? cn.c[13 + s0 + 3 * s1] : t;
    if constexpr (s0 != 0 && s2 != 0) t = (c0 && c2) ? cn.c[13 + s0 + 9 * s2] : t;
    if constexpr (s1 != 0 && s2 != 0) t = (c1 && c2) ? cn.c[13 + 3 * s1 + 9 * s2] : t;
    if constexpr (s0 != 0 && s1 != 0 && s2 != 0) t = (c0 && c1 && c2) ? cn.c[13 + s0 + 3 * s1 + 9 * s2] : t;
    nv = (uint32_t)(((x + s0) & 7) + 8 * ((y + s1) & 7) + 64 * ((z + s2) & 7));
}

// the upstream voxel of D3Q19 population q (offset -c_q)
template <int q>
__device__ __forceinline__ void upstream(uint32_t b, const Conn& cn, int x, int y, int z, uint32_t& t, uint32_t& nv)
{
    upstream_off<-CX[q], -CY[q], -CZ[q]>(b, cn, x, y, z, t, nv);
}

template <int q>
struct QI {
    static constexpr int value = q;
};

template <int... qs, typename F>
__device__ __forceinline__ void for_q(std::integer_sequence<int, qs...>, F&& f)
{
    (f(QI<qs>{}), ...);
}
using AllQ = std::make_integer_sequence<int, Q>;

}  // namespace nx
