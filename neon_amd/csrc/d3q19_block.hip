// d3q19_block.hip -- D3Q19 kernels on the block-sparse grid (the bGrid path,
// 8^3 blocks).
//
// Layout: AoSoA populations, index (b*19 + q)*512 + x + 8y + 64z
// (bPartition_imp.h:95-107); flags (b*512 + v); active bitmask 16 words per
// block; 27-entry block connectivity (bGrid_imp.h:139-172).
//
// One workgroup of 512 threads per block, thread v = x + 8y + 64z, so each
// wavefront owns one z-plane of the block and every population store is 64
// contiguous values.  The block's neighbourhood is staged once per
// workgroup: the own bitmask in LDS, and the connectivity row read with
// scalar loads (uniform per workgroup).  Per population the neighbour block
// of a lane is chosen with compile-time candidate directions -- a D3Q19
// velocity has at most two non-zero components, so at most three neighbour
// blocks (x-face, y/z-face, edge) can be involved -- and selected with
// v_cndmask, no per-lane table lookup.  This restates the neighbour rule of
// bPartition::helpGetNghIdx / helpNghPitch (bPartition_imp.h:109-124,
// 218-309).
#include <algorithm>
#include <utility>
#include <vector>
#include <type_traits>

#include "block_common.hpp"

namespace nx {
namespace {

constexpr uint32_t kBStride = Q * kBV;

struct BGeom {
    int32_t X, Y, Z;
    int32_t periodic;
    uint32_t n_blocks;
    const int32_t* origins;
    const uint32_t* masks;
    const uint32_t* conn;
};

BGeom make_bgeom(const nx_block_desc* d)
{
    return BGeom{d->dim[0],   d->dim[1], d->dim[2], d->periodic,
                 d->n_blocks, d->origins, d->masks, d->conn};
}

struct Par {
    double omega;
    double g[3];
};

__device__ __forceinline__ bool bit_of(const uint32_t* w, uint32_t v)
{
    return (w[v >> 5] >> (v & 31u)) & 1u;
}

// validity of (t, nv) as seen from block b (own mask in LDS)
__device__ __forceinline__ bool valid_voxel(uint32_t b, uint32_t t, uint32_t nv,
                                            const uint32_t* smask, const uint32_t* __restrict__ masks)
{
    if (t == kInvalidBlock) return false;
    if (t == b) return bit_of(smask, nv);
    return (masks[(uint64_t)t * 16 + (nv >> 5)] >> (nv & 31u)) & 1u;
}


// Pull stream of one bulk voxel (LbmTools.h:78-168); masked links whose
// upstream voxel is absent or inactive contribute 0 (NX_FLAG_OPEN).
template <int S>
__device__ __forceinline__ void pull(const typename Stor<S>::T* __restrict__ fin, const BGeom& g,
                                     uint32_t b, const Conn& cb,
                                     const uint32_t* smask, uint32_t v, uint32_t fl,
                                     typename Stor<S>::P f[Q])
{
    using T = typename Stor<S>::T;
    using P = typename Stor<S>::P;
    const int x = (int)(v & 7u), y = (int)((v >> 3) & 7u), z = (int)(v >> 6);
    const uint32_t m = fl & NX_MASK_BITS;
    const bool open = (fl & NX_FLAG_OPEN) != 0;
    const uint64_t own = (uint64_t)b * kBStride + v;
    for_q(AllQ{}, [&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if constexpr (q == 9) {
            f[9] = Stor<S>::bulk(fin[own + 9 * kBV], 9);
        } else {
            uint32_t t, nv;
            upstream<q>(b, cb, x, y, z, t, nv);
            const bool w = (m >> q) & 1u;
            // an absent block can only be reached through inconsistent flags;
            // treat it as a missing link rather than addressing through it
            bool ok = t != kInvalidBlock;
            if (w && open) ok = valid_voxel(b, t, nv, smask, g.masks);
            const uint32_t src = w ? (uint32_t)OPP[q] : (uint32_t)q;
            const uint64_t a = ok ? (uint64_t)t * kBStride + src * kBV + nv : own;
            T val = fin[a];
            if (!ok) val = T(0);
            f[q] = w ? Stor<S>::raw(val) : Stor<S>::bulk(val, q);
        }
    });
    if (m) {
        for_q(AllQ{}, [&](auto qc) {
            constexpr int q = decltype(qc)::value;
            if constexpr (q != 9) {
                if ((m >> q) & 1u) {
                    const P o = Stor<S>::bulk(fin[own + (uint64_t)OPP[q] * kBV], OPP[q]);
                    f[q] = o + f[q];
                }
            }
        });
    }
}

// Stage the block's bitmask in LDS; returns whether voxel v is active.
__device__ __forceinline__ bool stage(const BGeom& g, uint32_t b, uint32_t* smask)
{
    if (threadIdx.x < 16) smask[threadIdx.x] = g.masks[(uint64_t)b * 16 + threadIdx.x];
    __syncthreads();
    return bit_of(smask, threadIdx.x);
}

// The rest of the step of one active bulk voxel once its 19 streamed values
// val[q] (upstream voxel's population q, raw storage) are loaded: masked
// links fixed up with one batch of loads (the wall's value only where the
// link can have an upstream voxel: no NX_FLAG_OPEN, or NX_FLAG_MIXED), then
// the collision and the 19 stores.
template <int S, int CM, bool FORCE>
__device__ __forceinline__ void bulk_voxel_out(const typename Stor<S>::T* __restrict__ fin, const BGeom& g,
                                               const Par& p, uint32_t b, const Conn& cb, int x, int y,
                                               int z, uint64_t own, uint32_t fl,
                                               const typename Stor<S>::T val[Q], typename Stor<S>::T res[Q])
{
    using T = typename Stor<S>::T;
    using P = typename Stor<S>::P;
    using R = typename Real<S, CM>::R;
    P f[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) f[q] = Stor<S>::bulk(val[q], q);
    const uint32_t m = fl & NX_MASK_BITS;
    if (m) {
        // loads only for the masked links, issued as one batch (predicated
        // loads, no wait between them); the wall's value only where the
        // link can have an upstream voxel (not NX_FLAG_OPEN, or MIXED)
        const bool open = (fl & NX_FLAG_OPEN) != 0, mixed = (fl & NX_FLAG_MIXED) != 0;
        const bool need_nb = !open || mixed;
        T ownv[Q], nbv[Q];
        uint32_t okm = 0;
        for_q(AllQ{}, [&](auto qc) {
            constexpr int q = decltype(qc)::value;
            if constexpr (q != 9) {
                const bool w = (m >> q) & 1u;
                ownv[q] = w ? fin[own + (uint64_t)OPP[q] * kBV] : T(0);
                bool ok = false;
                uint32_t t = kInvalidBlock, nv = 0;
                if (w && need_nb) {
                    upstream<q>(b, cb, x, y, z, t, nv);
                    ok = !open && t != kInvalidBlock;
                    if (mixed)
                        ok = t != kInvalidBlock &&
                             ((g.masks[(uint64_t)t * 16 + (nv >> 5)] >> (nv & 31u)) & 1u);
                }
                okm |= (ok ? 1u : 0u) << q;
                nbv[q] = ok ? fin[(uint64_t)t * kBStride + (uint32_t)OPP[q] * kBV + nv] : T(0);
            }
        });
#pragma unroll
        for (int q = 0; q < Q; q++) {
            if (q == 9) continue;
            const P up = ((okm >> q) & 1u) ? Stor<S>::raw(nbv[q]) : P(0);
            const P fixed = Stor<S>::bulk(ownv[q], OPP[q]) + up;
            f[q] = ((m >> q) & 1u) ? fixed : f[q];
        }
    }
    R out[Q];
    collide<S, CM, FORCE>(f, p.omega, p.g, out);
#pragma unroll
    for (int q = 0; q < Q; q++) res[q] = Stor<S>::template enc<R>(out[q], q);
}

template <int S, int CM, bool FORCE>
__device__ __forceinline__ void bulk_voxel(const typename Stor<S>::T* __restrict__ fin,
                                           typename Stor<S>::T* __restrict__ fout, const BGeom& g,
                                           const Par& p, uint32_t b, const Conn& cb, int x, int y,
                                           int z, uint64_t own, uint32_t fl,
                                           const typename Stor<S>::T val[Q])
{
    typename Stor<S>::T res[Q];
    bulk_voxel_out<S, CM, FORCE>(fin, g, p, b, cb, x, y, z, own, fl, val, res);
#pragma unroll
    for (int q = 0; q < Q; q++) __builtin_nontemporal_store(res[q], &fout[own + q * kBV]);
}

// Step of one block per workgroup.  The 19 population loads of an active
// voxel are issued together with its flag word (their addresses depend only
// on the block's connectivity row, read with scalar loads); inactive voxels
// load nothing.  Masked links are fixed up afterwards in one batch of loads.
// No cache line of fout is written partially: an inactive voxel whose
// 128-byte line also holds active voxels receives 0 (the inactive value of
// the reference's population fields, RunCavityTwoPop.cu:56-57); lines with
// no active voxel are not written; active non-bulk voxels are untouched.
// The flag bits select the wall-link case without per-link lookups:
// NX_FLAG_OPEN without NX_FLAG_MIXED -- every masked link is missing (porous
// media: the link contributes the voxel's own opposite population only);
// no NX_FLAG_OPEN -- every masked link has an upstream voxel; NX_FLAG_MIXED
// -- per-link validity from the neighbour block's active mask.
template <int S, int CM, bool FORCE>
__global__ void __launch_bounds__(kBV)
    k_step(const typename Stor<S>::T* __restrict__ fin, typename Stor<S>::T* __restrict__ fout,
           const uint32_t* __restrict__ flags, BGeom g, Par p, int remap)
{
    using T = typename Stor<S>::T;
    const uint32_t b = remap ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x;
    const uint32_t v = threadIdx.x;
    const int x = (int)(v & 7u), y = (int)((v >> 3) & 7u), z = (int)(v >> 6);
    const uint64_t own = (uint64_t)b * kBStride + v;
    // active bits of this voxel's z-plane (64 voxels) and the connectivity
    // row: both uniform per wavefront, loaded before the population loads
    const uint64_t plane_bits =
        (uint64_t)g.masks[(uint64_t)b * 16 + 2 * (uint32_t)z] |
        ((uint64_t)g.masks[(uint64_t)b * 16 + 2 * (uint32_t)z + 1] << 32);
    const Conn cb = load_conn(g.conn, b);
    const bool active = (plane_bits >> (v & 63u)) & 1u;
    // voxels sharing this voxel's 128-byte line of each population plane
    constexpr uint32_t kLine = 128u / sizeof(T);
    const uint64_t line_bits = kLine >= 64 ? ~0ull : ((1ull << kLine) - 1ull) << ((v & 63u) & ~(kLine - 1u));
    const bool line_any = (plane_bits & line_bits) != 0;
    uint32_t fl = 0;
    T val[Q];
    if (active) {
        fl = flags[(uint64_t)b * kBV + v];
        for_q(AllQ{}, [&](auto qc) {
            constexpr int q = decltype(qc)::value;
            if constexpr (q == 9) {
                val[9] = fin[own + 9 * kBV];
            } else {
                uint32_t t, nv;
                upstream<q>(b, cb, x, y, z, t, nv);
                const uint64_t a =
                    t != kInvalidBlock ? (uint64_t)t * kBStride + (uint32_t)q * kBV + nv : own;
                val[q] = fin[a];
            }
        });
#pragma unroll
        for (int q = 0; q < Q; q++) nx_pin(val[q]);
        nx_pin(fl);
    }
    if (active && cls_of(fl) == NX_CELL_BULK) {
        bulk_voxel<S, CM, FORCE>(fin, fout, g, p, b, cb, x, y, z, own, fl, val);
    } else if (!active && line_any) {
        // inactive voxel in a line that also holds active voxels: write the
        // inactive value so the line is written whole
#pragma unroll
        for (int q = 0; q < Q; q++) __builtin_nontemporal_store(T(0), &fout[own + q * kBV]);
    }
}

// ---- 2-byte storage: two x-adjacent voxels per lane --------------------------
//
// One workgroup of 256 threads per block, lane t owning voxels 2t and 2t + 1
// (x even and x + 1 of one row), so a wavefront covers two z-planes and each
// population is one 4-byte load per lane: at the pair's upstream voxels when
// both lie in one block (2-byte aligned when the velocity has an x
// component; unaligned dword loads are part of the amdhsa target), else --
// the pair at the block's x face -- the aligned pair holding the one
// upstream voxel inside, plus one 2-byte load from the neighbour block.  The
// arithmetic is bulk_voxel_out per voxel, as in k_step, so the two steps
// agree bit for bit.  Stores: one 4-byte store per population where both
// voxels are written, else 2-byte stores for the written one (an active
// non-bulk voxel is not written).
template <int S, int CM, bool FORCE>
__global__ void __launch_bounds__(kBV / 2)
    k_step_pair(const typename Stor<S>::T* __restrict__ fin, typename Stor<S>::T* __restrict__ fout,
                const uint32_t* __restrict__ flags, BGeom g, Par p, int remap)
{
    using T = typename Stor<S>::T;
    static_assert(sizeof(T) == 2, "two voxels per 4-byte word");
    typedef uint32_t u32a2 __attribute__((aligned(2)));
    const uint32_t b = remap ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x;
    const uint32_t v0 = 2u * threadIdx.x;
    const int x0 = (int)(v0 & 7u), y = (int)((v0 >> 3) & 7u), z = (int)(v0 >> 6);
    const uint64_t own0 = (uint64_t)b * kBStride + v0;
    const uint64_t plane_bits =
        (uint64_t)g.masks[(uint64_t)b * 16 + 2 * (uint32_t)z] |
        ((uint64_t)g.masks[(uint64_t)b * 16 + 2 * (uint32_t)z + 1] << 32);
    const Conn cb = load_conn(g.conn, b);
    const uint32_t lb = v0 & 63u;
    const bool act0 = (plane_bits >> lb) & 1u, act1 = (plane_bits >> (lb + 1u)) & 1u;
    const bool line_any = plane_bits != 0;  // 2-byte storage: the z-plane is one 128-byte line
    uint32_t fl0 = 0, fl1 = 0;
    T a0[Q], a1[Q];
    if (act0 || act1) {
        const uint64_t fw = *(const uint64_t*)(flags + (uint64_t)b * kBV + v0);
        fl0 = (uint32_t)fw;
        fl1 = (uint32_t)(fw >> 32);
        for_q(AllQ{}, [&](auto qc) {
            constexpr int q = decltype(qc)::value;
            uint32_t t0, n0, t1, n1;
            upstream<q>(b, cb, x0, y, z, t0, n0);
            upstream<q>(b, cb, x0 + 1, y, z, t1, n1);
            // the dword's first voxel and which half each voxel takes
            uint64_t wa;
            bool split = false;
            if constexpr (CX[q] == 0) {
                wa = t0 != kInvalidBlock ? (uint64_t)t0 * kBStride + (uint32_t)q * kBV + n0 : own0 + 9u * kBV;
            } else if constexpr (CX[q] > 0) {
                // sources x0 - 1, x0: at x0 = 0 the first lies in the -x block
                split = x0 == 0;
                const uint32_t tw = split ? t1 : t0, nw = split ? n1 : n0;
                wa = tw != kInvalidBlock ? (uint64_t)tw * kBStride + (uint32_t)q * kBV + nw : own0 + 9u * kBV;
            } else {
                // sources x0 + 1, x0 + 2: at x0 = 6 the second lies in the +x block
                split = x0 == 6;
                const uint32_t nw = split ? n0 - 1u : n0;
                wa = t0 != kInvalidBlock ? (uint64_t)t0 * kBStride + (uint32_t)q * kBV + nw : own0 + 9u * kBV;
            }
            const uint32_t wv = *(const u32a2*)(fin + wa);
            const T lo = __builtin_bit_cast(T, (uint16_t)(wv & 0xFFFFu));
            const T hi = __builtin_bit_cast(T, (uint16_t)(wv >> 16));
            if constexpr (CX[q] == 0) {
                a0[q] = lo;
                a1[q] = hi;
            } else if constexpr (CX[q] > 0) {
                T e = lo;
                if (split) e = t0 != kInvalidBlock ? fin[(uint64_t)t0 * kBStride + (uint32_t)q * kBV + n0] : T(0);
                a0[q] = split ? e : lo;
                a1[q] = split ? lo : hi;
            } else {
                T e = hi;
                if (split) e = t1 != kInvalidBlock ? fin[(uint64_t)t1 * kBStride + (uint32_t)q * kBV + n1] : T(0);
                a0[q] = split ? hi : lo;
                a1[q] = split ? e : hi;
            }
        });
#pragma unroll
        for (int q = 0; q < Q; q++) {
            nx_pin(a0[q]);
            nx_pin(a1[q]);
        }
        nx_pin(fl0);
        nx_pin(fl1);
    }
    const bool bulk0 = act0 && cls_of(fl0) == NX_CELL_BULK, bulk1 = act1 && cls_of(fl1) == NX_CELL_BULK;
    // which voxels are written: bulk ones (their new values) and inactive
    // ones in a line that holds active voxels (the inactive value 0)
    const bool w0 = bulk0 || (!act0 && line_any), w1 = bulk1 || (!act1 && line_any);
    if (!w0 && !w1) return;
    T r0[Q], r1[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) r0[q] = r1[q] = T(0);
    if (bulk0) bulk_voxel_out<S, CM, FORCE>(fin, g, p, b, cb, x0, y, z, own0, fl0, a0, r0);
    if (bulk1) bulk_voxel_out<S, CM, FORCE>(fin, g, p, b, cb, x0 + 1, y, z, own0 + 1u, fl1, a1, r1);
    if (w0 && w1) {
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const uint32_t wv = (uint32_t)__builtin_bit_cast(uint16_t, r0[q]) |
                                ((uint32_t)__builtin_bit_cast(uint16_t, r1[q]) << 16);
            __builtin_nontemporal_store(wv, (uint32_t*)(fout + own0 + (uint64_t)q * kBV));
        }
    } else if (w0) {
#pragma unroll
        for (int q = 0; q < Q; q++) __builtin_nontemporal_store(r0[q], &fout[own0 + (uint64_t)q * kBV]);
    } else {
#pragma unroll
        for (int q = 0; q < Q; q++) __builtin_nontemporal_store(r1[q], &fout[own0 + 1u + (uint64_t)q * kBV]);
    }
}

// ---- the step with the block neighbourhood staged in LDS --------------------
//
// One workgroup per tile: the up to four blocks with origins (16i + 8a,
// 16j + 8c, 8k), a, c in {0, 1}, member a + 2c (absent members are
// kInvalidBlock in the tile table).  For each population with an x or y
// component (16 of 19) the workgroup reads the members' planes whole -- two
// x-adjacent voxels per lane, every 128-byte line once -- and the tile's halo
// -- the face and edge voxels of the neighbour blocks that stream into the
// tile, one lane per face voxel of the tile -- and writes each value into LDS
// at the voxel it streams to (the upstream rule of bPartition_imp.h:218-309
// applied from the source side).  After one barrier every voxel finds its
// streamed values at its own LDS index.  A face shared by two members is
// fetched once instead of once per block: per 64 voxels about 32 line
// requests instead of 43 at 2-byte storage.  The rest population and the two
// z populations are read in place (their sources are whole lines already),
// with the flag words, before the staging loads; at 4-byte storage so are
// the eight populations with a z component (LDS: 64 KiB per workgroup).  In a tile with an absent
// member, a voxel whose unmasked link starts in that member's extent (only a
// periodic wrap makes this happen: the upstream block is then another one)
// reads that value through its block's connectivity row.  The arithmetic
// after the loads is bulk_voxel, as in k_step, so both steps agree bit for
// bit.
constexpr int kTileB = 4;                 // blocks per tile (2 x 2 x 1)
constexpr int kTileV = kTileB * kBV;      // voxels per tile
constexpr int kStaged = 16;               // most populations staged in LDS

// LDS slot of population q (-1: read in place).  2-byte storage stages every
// population with an x or y component (64 KiB); 4-byte storage the eight in
// the x-y plane only, so that two workgroups still fit one CU's LDS.
template <int S>
constexpr int tiled_slot(int q)
{
    if constexpr (sizeof(typename Stor<S>::T) == 2) {
        constexpr int s2[Q] = {0, 1, -1, 2, 3, 4, 5, 6, 7, -1, 8, 9, -1, 10, 11, 12, 13, 14, 15};
        return s2[q];
    } else {
        constexpr int s4[Q] = {0, 1, -1, 2, 3, -1, -1, -1, -1, -1, 4, 5, -1, 6, 7, -1, -1, -1, -1};
        return s4[q];
    }
}

template <int S>
constexpr int tiled_slots()
{
    return sizeof(typename Stor<S>::T) == 2 ? 16 : 8;
}

template <int S>
using TiledStaged = typename std::conditional<
    sizeof(typename Stor<S>::T) == 2,
    std::integer_sequence<int, 0, 1, 3, 4, 5, 6, 7, 8, 10, 11, 13, 14, 15, 16, 17, 18>,
    std::integer_sequence<int, 0, 1, 3, 4, 10, 11, 13, 14>>::type;

// the halo voxel of population q loaded by lane l of wavefront w: wavefronts
// 0-1 the tile's x face (member row c = w), 2-3 its y face (member column
// a = w - 2), 4-7 its z face (member w - 4); in-block destination (hx, hy, hz)
template <int q>
__device__ __forceinline__ bool halo_of(uint32_t w, uint32_t l, uint32_t& m, int& hx, int& hy, int& hz)
{
    constexpr int cx = CX[q], cy = CY[q], cz = CZ[q];
    if (w < 2u) {
        if constexpr (cx == 0) return false;
        m = (cx > 0 ? 0u : 1u) + 2u * w;
        hx = cx > 0 ? 0 : 7;
        hy = (int)(l & 7u);
        hz = (int)(l >> 3);
        return true;
    }
    if (w < 4u) {
        if constexpr (cy == 0) return false;
        m = (w - 2u) + 2u * (cy > 0 ? 0u : 1u);
        hx = (int)(l & 7u);
        hy = cy > 0 ? 0 : 7;
        hz = (int)(l >> 3);
        return true;
    }
    if constexpr (cz == 0) return false;
    m = w - 4u;
    hx = (int)(l & 7u);
    hy = (int)(l >> 3);
    hz = cz > 0 ? 0 : 7;
    return true;
}

// the member block ids of a tile as four values (selected by value, never
// through an indexed private array, which would live in scratch memory)
struct Members {
    uint32_t m0, m1, m2, m3;
    template <int k>
    __device__ __forceinline__ uint32_t at() const
    {
        if constexpr (k == 0) return m0;
        if constexpr (k == 1) return m1;
        if constexpr (k == 2) return m2;
        return m3;
    }
};

__device__ __forceinline__ uint32_t sel4(uint32_t m, uint32_t a, uint32_t b, uint32_t c, uint32_t d)
{
    return m == 0 ? a : m == 1 ? b : m == 2 ? c : d;
}

// Loads of the staged populations qs -- member planes (lane pair index
// t & 255 of members mh and 2 + mh) and one halo voxel per lane -- then each
// value written to the LDS index of the voxel it streams to.
template <int S, int... qs>
__device__ __forceinline__ void stage_batch(std::integer_sequence<int, qs...>,
                                            const typename Stor<S>::T* __restrict__ fin,
                                            typename Stor<S>::T* stage, Members mid, uint32_t w,
                                            uint32_t mh, uint32_t l, uint32_t hb_neg, uint32_t hb_pos,
                                            const Conn& hc_neg, const Conn& hc_pos)
{
    using T = typename Stor<S>::T;
    typedef T T2 __attribute__((ext_vector_type(2)));
    const uint32_t pr = threadIdx.x & 255u;
    T2 pv[kStaged][2];  // indexed by slot: only the batch's slots are live
    T hv[kStaged];
    for_q(std::integer_sequence<int, qs...>{}, [&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int sl = tiled_slot<S>(q);
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const uint32_t b = j ? (mh ? mid.m3 : mid.m2) : (mh ? mid.m1 : mid.m0);
            pv[sl][j] = b != kInvalidBlock
                            ? *(const T2*)(fin + (uint64_t)b * kBStride + (uint32_t)q * kBV + 2u * pr)
                            : T2(T(0));
        }
        hv[sl] = T(0);
        uint32_t m = 0;
        int hx = 0, hy = 0, hz = 0;
        if (halo_of<q>(w, l, m, hx, hy, hz)) {
            // the member m on the tile face: on the negative side of the
            // wavefront's axis for a positive velocity component
            const bool neg = w < 2u ? CX[q] > 0 : (w < 4u ? CY[q] > 0 : CZ[q] > 0);
            const uint32_t b = neg ? hb_neg : hb_pos;
            if (b != kInvalidBlock) {
                uint32_t tb, nv;
                upstream<q>(b, neg ? hc_neg : hc_pos, hx, hy, hz, tb, nv);
                if (tb != kInvalidBlock) hv[sl] = fin[(uint64_t)tb * kBStride + (uint32_t)q * kBV + nv];
            }
        }
    });
    for_q(std::integer_sequence<int, qs...>{}, [&](auto qc) {
        constexpr int sl = tiled_slot<S>(decltype(qc)::value);
        nx_pin(hv[sl]);
#pragma unroll
        for (int j = 0; j < 2; j++) {
            T a0 = pv[sl][j].x, a1 = pv[sl][j].y;
            nx_pin(a0);
            nx_pin(a1);
            pv[sl][j].x = a0;
            pv[sl][j].y = a1;
        }
    });
    for_q(std::integer_sequence<int, qs...>{}, [&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int sl = tiled_slot<S>(q);
        constexpr int cx = CX[q], cy = CY[q], cz = CZ[q];
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const uint32_t m = 2u * (uint32_t)j + mh;
#pragma unroll
            for (int e = 0; e < 2; e++) {
                const uint32_t v = 2u * pr + (uint32_t)e;
                const int dx = 8 * (int)(m & 1u) + (int)(v & 7u) + cx;
                const int dy = 8 * (int)(m >> 1) + (int)((v >> 3) & 7u) + cy;
                const int dz = (int)(v >> 6) + cz;
                if (dx >= 0 && dx < 16 && dy >= 0 && dy < 16 && dz >= 0 && dz < 8)
                    stage[sl * kTileV + ((dx >> 3) + 2 * (dy >> 3)) * kBV + (dx & 7) + 8 * (dy & 7) + 64 * dz] =
                        e ? pv[sl][j].y : pv[sl][j].x;
            }
        }
        uint32_t m = 0;
        int hx = 0, hy = 0, hz = 0;
        if (halo_of<q>(w, l, m, hx, hy, hz))
            stage[sl * kTileV + m * kBV + (uint32_t)(hx + 8 * hy + 64 * hz)] = hv[sl];
    });
}

template <int S, int CM, bool FORCE>
__global__ void __launch_bounds__(kBV)
    k_step_tiled(const typename Stor<S>::T* __restrict__ fin, typename Stor<S>::T* __restrict__ fout,
                 const uint32_t* __restrict__ flags, const uint32_t* __restrict__ tiles, BGeom g,
                 Par p, int remap)
{
    using T = typename Stor<S>::T;
    __shared__ T stage[tiled_slots<S>() * kTileV];
    const uint32_t tile = remap ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x;
    const Members mid{tiles[(uint64_t)tile * kTileB], tiles[(uint64_t)tile * kTileB + 1],
                      tiles[(uint64_t)tile * kTileB + 2], tiles[(uint64_t)tile * kTileB + 3]};
    const uint32_t t = threadIdx.x;
    const int x = (int)(t & 7u), y = (int)((t >> 3) & 7u), z = (int)(t >> 6);
    // wavefront index and member half, uniform per wavefront
    const uint32_t w = __builtin_amdgcn_readfirstlane(t >> 6);
    const uint32_t mh = __builtin_amdgcn_readfirstlane(t >> 8);
    const uint32_t l = t & 63u;
    // the members whose tile-face voxels this wavefront's halo lanes serve
    // (x face: row w; y face: column w - 2; z face: member w - 4), for a
    // source beyond the negative and the positive side, with their
    // connectivity rows read up front (block 0 stands in for an absent one)
    const uint32_t hm_neg = w < 2u ? 2u * w : (w < 4u ? w - 2u : w - 4u);
    const uint32_t hm_pos = w < 2u ? 2u * w + 1u : (w < 4u ? w : w - 4u);
    const uint32_t hb_neg = sel4(hm_neg, mid.m0, mid.m1, mid.m2, mid.m3);
    const uint32_t hb_pos = sel4(hm_pos, mid.m0, mid.m1, mid.m2, mid.m3);
    const Conn hc_neg = load_conn(g.conn, hb_neg == kInvalidBlock ? 0u : hb_neg);
    const Conn hc_pos = load_conn(g.conn, hb_pos == kInvalidBlock ? 0u : hb_pos);

    // in place, every member: the z-plane's active bits, the flag word and
    // the rest and z populations of voxel t
    uint64_t pbits[kTileB];
    uint32_t fl[kTileB];
    T vd[kTileB][Q];  // only the unstaged populations are live
    for_q(std::make_integer_sequence<int, kTileB>{}, [&](auto kc) {
        constexpr int k = decltype(kc)::value;
        pbits[k] = 0;
        fl[k] = 0;
#pragma unroll
        for (int q = 0; q < Q; q++) vd[k][q] = T(0);
        const uint32_t b = mid.template at<k>();
        const Conn cb = load_conn(g.conn, b == kInvalidBlock ? 0u : b);  // ahead of the branches
        if (b == kInvalidBlock) return;
        pbits[k] = (uint64_t)g.masks[(uint64_t)b * 16 + 2 * (uint32_t)z] |
                   ((uint64_t)g.masks[(uint64_t)b * 16 + 2 * (uint32_t)z + 1] << 32);
        if ((pbits[k] >> l) & 1u) {
            const uint64_t own = (uint64_t)b * kBStride + t;
            fl[k] = flags[(uint64_t)b * kBV + t];
            for_q(AllQ{}, [&](auto qc) {
                constexpr int q = decltype(qc)::value;
                if constexpr (q == 9) {
                    vd[k][9] = fin[own + 9 * kBV];
                } else if constexpr (tiled_slot<S>(q) < 0) {
                    uint32_t tb, nv;
                    upstream<q>(b, cb, x, y, z, tb, nv);
                    vd[k][q] = fin[tb != kInvalidBlock ? (uint64_t)tb * kBStride + (uint32_t)q * kBV + nv : own];
                }
            });
        }
    });
#pragma unroll
    for (int k = 0; k < kTileB; k++) {
        nx_pin(fl[k]);
#pragma unroll
        for (int q = 0; q < Q; q++)
            if (tiled_slot<S>(q) < 0) nx_pin(vd[k][q]);
    }
    stage_batch<S>(TiledStaged<S>{}, fin, stage, mid, w, mh, l, hb_neg, hb_pos, hc_neg, hc_pos);
    __syncthreads();

    constexpr uint32_t kLine = 128u / sizeof(T);
    const uint64_t line_bits = kLine >= 64 ? ~0ull : ((1ull << kLine) - 1ull) << (l & ~(kLine - 1u));
    const bool incomplete = mid.m0 == kInvalidBlock || mid.m1 == kInvalidBlock || mid.m2 == kInvalidBlock ||
                            mid.m3 == kInvalidBlock;
    for_q(std::make_integer_sequence<int, kTileB>{}, [&](auto kc) {
        constexpr int k = decltype(kc)::value;
        const uint32_t b = mid.template at<k>();
        if (b == kInvalidBlock) return;
        const uint64_t own = (uint64_t)b * kBStride + t;
        const bool active = (pbits[k] >> l) & 1u;
        if (active && cls_of(fl[k]) == NX_CELL_BULK) {
            T val[Q];
#pragma unroll
            for (int q = 0; q < Q; q++)
                val[q] = tiled_slot<S>(q) >= 0 ? stage[tiled_slot<S>(q) * kTileV + k * kBV + t] : vd[k][q];
            const Conn cb = load_conn(g.conn, b);
            if (incomplete) {
                // a source inside the tile's extent but in an absent member
                // was not staged: where the link is not masked (a periodic
                // wrap makes the upstream block another one) read it through
                // this block's connectivity row
                const uint32_t mk = fl[k] & NX_MASK_BITS;
                for_q(AllQ{}, [&](auto qc) {
                    constexpr int q = decltype(qc)::value;
                    if constexpr (tiled_slot<S>(q) >= 0) {
                        const int sx = 8 * (k & 1) + x - CX[q], sy = 8 * (k >> 1) + y - CY[q], sz = z - CZ[q];
                        if (sx >= 0 && sx < 16 && sy >= 0 && sy < 16 && sz >= 0 && sz < 8 &&
                            !((mk >> q) & 1u) &&
                            sel4((uint32_t)((sx >> 3) + 2 * (sy >> 3)), mid.m0, mid.m1, mid.m2, mid.m3) ==
                                kInvalidBlock) {
                            uint32_t tb, nv;
                            upstream<q>(b, cb, x, y, z, tb, nv);
                            val[q] = fin[tb != kInvalidBlock ? (uint64_t)tb * kBStride + (uint32_t)q * kBV + nv
                                                             : own];
                        }
                    }
                });
            }
            bulk_voxel<S, CM, FORCE>(fin, fout, g, p, b, cb, x, y, z, own, fl[k], val);
        } else if (!active && (pbits[k] & line_bits) != 0) {
#pragma unroll
            for (int q = 0; q < Q; q++) __builtin_nontemporal_store(T(0), &fout[own + q * kBV]);
        }
    });
}

__global__ void __launch_bounds__(kBV) k_cavity_flags(uint32_t* flags, BGeom g)
{
    __shared__ uint32_t smask[16];
    const uint32_t b = blockIdx.x, v = threadIdx.x;
    const bool act = stage(g, b, smask);
    int cls = NX_CELL_UNDEFINED;
    if (act) {
        const int x = g.origins[3 * (uint64_t)b] + (int)(v & 7u);
        const int y = g.origins[3 * (uint64_t)b + 1] + (int)((v >> 3) & 7u);
        const int z = g.origins[3 * (uint64_t)b + 2] + (int)(v >> 6);
        cls = NX_CELL_BULK;
        if (x == 0 || x == g.X - 1 || y == 0 || y == g.Y - 1 || z == 0 || z == g.Z - 1) {
            cls = NX_CELL_BOUNCEBACK;
            if (y == g.Y - 1) cls = NX_CELL_MOVINGWALL;
        }
    }
    flags[(uint64_t)b * kBV + v] = (uint32_t)cls << NX_CLASS_SHIFT;
}

template <int S>
__global__ void __launch_bounds__(kBV)
    k_fill(typename Stor<S>::T* pop, const uint32_t* flags, BGeom g, double ulb, double l0,
           double l1, double l2)
{
    __shared__ uint32_t smask[16];
    const uint32_t b = blockIdx.x, v = threadIdx.x;
    const bool act = stage(g, b, smask);
    const int cls = act ? cls_of(flags[(uint64_t)b * kBV + v]) : NX_CELL_UNDEFINED;
    const uint64_t own = (uint64_t)b * kBStride + v;
#pragma unroll
    for (int q = 0; q < Q; q++) {
        typename Stor<S>::T val;
        if (cls == NX_CELL_BULK) {
            val = Stor<S>::enc_full(wq(q), q);
        } else if (cls == NX_CELL_MOVINGWALL) {
            const double w = -6. * wq(q) * ulb * (CX[q] * l0 + CY[q] * l1 + CZ[q] * l2);
            val = (typename Stor<S>::T)(typename Stor<S>::P)w;
        } else {
            val = (typename Stor<S>::T)0.;
        }
        pop[own + q * kBV] = val;
    }
}

// computeWallNghMask on the block grid; absent or inactive upstream voxels
// count as non-bulk (their NghData default is CellType::undefined).
__global__ void __launch_bounds__(kBV)
    k_wall_mask(const uint32_t* fin, uint32_t* fout, BGeom g)
{
    __shared__ uint32_t smask[16];
    const uint32_t b = blockIdx.x, v = threadIdx.x;
    const bool act = stage(g, b, smask);
    const uint64_t i = (uint64_t)b * kBV + v;
    const uint32_t fl = fin[i];
    if (!act || cls_of(fl) != NX_CELL_BULK) {
        fout[i] = fl;
        return;
    }
    const Conn cb = load_conn(g.conn, b);
    const int x = (int)(v & 7u), y = (int)((v >> 3) & 7u), z = (int)(v >> 6);
    uint32_t m = 0, open = 0, closed = 0;
    for_q(AllQ{}, [&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if constexpr (q != 9) {
            uint32_t t, nv;
            upstream<q>(b, cb, x, y, z, t, nv);
            if (!valid_voxel(b, t, nv, smask, g.masks)) {
                m |= 1u << q;
                open = NX_FLAG_OPEN;
            } else if (cls_of(fin[(uint64_t)t * kBV + nv]) != NX_CELL_BULK) {
                m |= 1u << q;
                closed = 1;
            }
        }
    });
    fout[i] = ((uint32_t)NX_CELL_BULK << NX_CLASS_SHIFT) | m | open |
              (open && closed ? NX_FLAG_MIXED : 0u);
}

template <int S, int CM>
__global__ void __launch_bounds__(kBV)
    k_moments(const typename Stor<S>::T* __restrict__ fin, const uint32_t* __restrict__ flags,
              void* rho_, void* u_, BGeom g)
{
    using R = typename Real<S, CM>::R;
    using M = typename std::conditional<S == NX_F64, double, float>::type;
    __shared__ uint32_t smask[16];
    const uint32_t b = blockIdx.x, v = threadIdx.x;
    const bool act = stage(g, b, smask);
    const uint32_t fl = act ? flags[(uint64_t)b * kBV + v] : (uint32_t)NX_CELL_UNDEFINED << 24;
    const int cls = cls_of(fl);
    typename Stor<S>::P f[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) f[q] = 0;
    const uint64_t own = (uint64_t)b * kBStride + v;
    if (cls == NX_CELL_BULK) pull<S>(fin, g, b, load_conn(g.conn, b), smask, v, fl, f);
    if (cls == NX_CELL_MOVINGWALL)
        for (int a = 0; a < 3; a++) f[a] = Stor<S>::raw(fin[own + a * kBV]);
    R r, uu[3];
    cell_moments<S, CM>(cls, f, r, uu);
    M* rho = (M*)rho_;
    M* u = (M*)u_;
    rho[(uint64_t)b * kBV + v] = (M)r;
    for (int a = 0; a < 3; a++) u[((uint64_t)b * 3 + a) * kBV + v] = (M)uu[a];
}

constexpr int kRedGrid = 2048;

template <int S>
__global__ void __launch_bounds__(kBV)
    k_mass(const typename Stor<S>::T* __restrict__ fin, const uint32_t* __restrict__ flags,
           BGeom g, double* partials)
{
    __shared__ double lds[16];
    double acc[1] = {0.0};
    const uint32_t v = threadIdx.x;
    for (uint32_t b = blockIdx.x; b < g.n_blocks; b += gridDim.x) {
        const bool act = (g.masks[(uint64_t)b * 16 + (v >> 5)] >> (v & 31u)) & 1u;
        if (!act || cls_of(flags[(uint64_t)b * kBV + v]) != NX_CELL_BULK) continue;
        const uint64_t own = (uint64_t)b * kBStride + v;
        double r = 0;
#pragma unroll
        for (int q = 0; q < Q; q++) r += Stor<S>::full(fin[own + q * kBV], q);
        acc[0] += r;
    }
    block_sum_f64<1>(acc, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = acc[0];
}

template <int S, int CM>
__global__ void __launch_bounds__(kBV)
    k_residual(const typename Stor<S>::T* __restrict__ fin, const uint32_t* __restrict__ flags,
               float* __restrict__ u_prev, BGeom g, double* partials)
{
    using R = typename Real<S, CM>::R;
    __shared__ double lds[32];
    __shared__ uint32_t smask[16];
    double acc[2] = {0.0, 0.0};
    const uint32_t v = threadIdx.x;
    for (uint32_t b = blockIdx.x; b < g.n_blocks; b += gridDim.x) {
        __syncthreads();  // smask reuse across iterations
        const bool act = stage(g, b, smask);
        if (!act) continue;
        const uint32_t fl = flags[(uint64_t)b * kBV + v];
        if (cls_of(fl) != NX_CELL_BULK) continue;
        typename Stor<S>::P f[Q];
        pull<S>(fin, g, b, load_conn(g.conn, b), smask, v, fl, f);
        R r, uu[3];
        cell_moments<S, CM>(NX_CELL_BULK, f, r, uu);
        for (int a = 0; a < 3; a++) {
            const uint64_t k = ((uint64_t)b * 3 + a) * kBV + v;
            const float un = (float)uu[a];
            const double du = (double)un - (double)u_prev[k];
            acc[0] += du * du;
            acc[1] += (double)un * (double)un;
            u_prev[k] = un;
        }
    }
    block_sum_f64<2>(acc, lds);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = acc[0];
        partials[2 * blockIdx.x + 1] = acc[1];
    }
}

int check(const nx_block_desc* d)
{
    if (!d) return set_error(NX_ERR_INVALID, "null descriptor");
    if (!d->masks || !d->conn || !d->origins)
        return set_error(NX_ERR_INVALID, "block descriptor without device tables");
    if (d->n_blocks == 0) return set_error(NX_ERR_INVALID, "empty block grid");
    if (d->n_blocks > 0x7FFFFFFFu) return set_error(NX_ERR_INVALID, "too many blocks");
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

template <int V>
using IC = std::integral_constant<int, V>;

// 1: contiguous range of blocks per XCD (nx_d3q19_set_block_remap)
int g_block_remap = 1;
// 1: 2-byte storage steps two voxels per lane (nx_d3q19_set_block_pairs)
int g_block_pairs = 0;

template <typename F>
int by_type(const nx_block_desc* d, F&& f)
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

}  // namespace
}  // namespace nx

using namespace nx;

extern "C" size_t nx_pop_bytes_block(const nx_block_desc* d)
{
    if (!d) return 0;
    const size_t n = (size_t)d->n_blocks * kBStride;
    switch (d->storage) {
        case NX_F64: return n * 8;
        case NX_F32:
        case NX_F32DEV: return n * 4;
        case NX_F16DEV: return n * 2;
    }
    return 0;
}

extern "C" int nx_cavity_flags_block(const nx_block_desc* d, uint32_t* flags, nx_stream s)
{
    if (int e = check(d)) return e;
    if (!flags) return set_error(NX_ERR_INVALID, "null flags");
    hipLaunchKernelGGL(k_cavity_flags, dim3(d->n_blocks), dim3(kBV), 0, (hipStream_t)s, flags,
                       make_bgeom(d));
    NX_CHECK_LAUNCH("k_cavity_flags(block)");
    return NX_OK;
}

extern "C" int nx_d3q19_fill_block(const nx_block_desc* d, void* pop, const uint32_t* flags,
                                   double ulb, const double ulid[3], nx_stream s)
{
    if (int e = check(d)) return e;
    if (!pop || !flags || !ulid) return set_error(NX_ERR_INVALID, "null pointer");
    const BGeom g = make_bgeom(d);
    const hipStream_t st = (hipStream_t)s;
    if (int e = by_type(d, [&](auto sc, auto) {
            constexpr int S = decltype(sc)::value;
            hipLaunchKernelGGL((k_fill<S>), dim3(g.n_blocks), dim3(kBV), 0, st,
                               (typename Stor<S>::T*)pop, flags, g, ulb, ulid[0], ulid[1], ulid[2]);
        }))
        return e;
    NX_CHECK_LAUNCH("k_fill(block)");
    return NX_OK;
}

extern "C" int nx_d3q19_wall_mask_block(const nx_block_desc* d, const uint32_t* flags_in,
                                        uint32_t* flags_out, nx_stream s)
{
    if (int e = check(d)) return e;
    if (!flags_in || !flags_out) return set_error(NX_ERR_INVALID, "null flags");
    hipLaunchKernelGGL(k_wall_mask, dim3(d->n_blocks), dim3(kBV), 0, (hipStream_t)s, flags_in,
                       flags_out, make_bgeom(d));
    NX_CHECK_LAUNCH("k_wall_mask(block)");
    return NX_OK;
}

extern "C" int nx_d3q19_step_block(const nx_block_desc* d, const void* fin, void* fout,
                                   const uint32_t* flags, const nx_lbm_params* p, nx_stream s)
{
    if (int e = check(d)) return e;
    if (!fin || !fout || !flags || !p) return set_error(NX_ERR_INVALID, "null pointer");
    if (fin == fout)
        return set_error(NX_ERR_INVALID, "fin and fout must differ (two-population scheme)");
    const BGeom g = make_bgeom(d);
    const Par par{p->omega, {p->force[0], p->force[1], p->force[2]}};
    const bool force = p->force[0] != 0. || p->force[1] != 0. || p->force[2] != 0.;
    if (force && d->compute == NX_COMPUTE_REF)
        return set_error(NX_ERR_UNSUPPORTED,
                         "body force requires NX_COMPUTE_F32 or NX_COMPUTE_F64");
    const hipStream_t st = (hipStream_t)s;
    if (int e = by_type(d, [&](auto sc, auto cc) {
            constexpr int S = decltype(sc)::value;
            constexpr int CM = decltype(cc)::value;
            using T = typename Stor<S>::T;
            if constexpr (sizeof(T) == 2) {
                if (g_block_pairs) {
                    if (force)
                        hipLaunchKernelGGL((k_step_pair<S, CM, true>), dim3(g.n_blocks), dim3(kBV / 2), 0, st,
                                           (const T*)fin, (T*)fout, flags, g, par, g_block_remap);
                    else
                        hipLaunchKernelGGL((k_step_pair<S, CM, false>), dim3(g.n_blocks), dim3(kBV / 2), 0, st,
                                           (const T*)fin, (T*)fout, flags, g, par, g_block_remap);
                    return;
                }
            }
            if (force)
                hipLaunchKernelGGL((k_step<S, CM, true>), dim3(g.n_blocks), dim3(kBV), 0, st,
                                   (const T*)fin, (T*)fout, flags, g, par, g_block_remap);
            else
                hipLaunchKernelGGL((k_step<S, CM, false>), dim3(g.n_blocks), dim3(kBV), 0, st,
                                   (const T*)fin, (T*)fout, flags, g, par, g_block_remap);
        }))
        return e;
    NX_CHECK_LAUNCH("k_step(block)");
    return NX_OK;
}

struct nx_block_tiles_s {
    uint32_t n_blocks = 0;
    uint32_t n_tiles = 0;
    uint32_t* d_tiles = nullptr;      // device, 4 x n_tiles block ids
    const int32_t* origins = nullptr;  // the block tables the table was built from
    const uint32_t* conn = nullptr;
};

extern "C" int nx_block_tiles_create(const nx_block_desc* d, nx_block_tiles* out)
{
    if (int e = check(d)) return e;
    if (!out) return set_error(NX_ERR_INVALID, "null output handle");
    *out = nullptr;
    const uint32_t n = d->n_blocks;
    std::vector<int32_t> org((size_t)3 * n);
    if (hipMemcpy(org.data(), d->origins, org.size() * sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess)
        return set_error(NX_ERR_HIP, "reading the block origins");
    // (tile key, member, block id), tiles ordered z, y, x like the blocks
    const uint64_t ntx = ((uint64_t)d->dim[0] + 15) / 16, nty = ((uint64_t)d->dim[1] + 15) / 16;
    std::vector<std::pair<uint64_t, uint32_t>> key(n);
    for (uint32_t b = 0; b < n; b++) {
        const int32_t* o = &org[(size_t)3 * b];
        if (o[0] < 0 || o[1] < 0 || o[2] < 0 || (o[0] & 7) || (o[1] & 7) || (o[2] & 7))
            return set_error(NX_ERR_INVALID, "block origin not on the 8-voxel lattice");
        const uint64_t tk = ((uint64_t)(o[2] >> 3) * nty + (uint64_t)(o[1] >> 4)) * ntx + (uint64_t)(o[0] >> 4);
        const uint32_t member = (uint32_t)((o[0] >> 3) & 1) + 2u * (uint32_t)((o[1] >> 3) & 1);
        key[b] = {tk * 4 + member, b};
    }
    std::sort(key.begin(), key.end());
    std::vector<uint32_t> tab;
    for (size_t i = 0; i < key.size(); i++) {
        if (i && key[i].first == key[i - 1].first)
            return set_error(NX_ERR_INVALID, "two blocks with one origin");
        if (i == 0 || (key[i].first >> 2) != (key[i - 1].first >> 2))
            tab.insert(tab.end(), kTileB, kInvalidBlock);
        tab[tab.size() - kTileB + (key[i].first & 3)] = key[i].second;
    }
    auto* t = new nx_block_tiles_s;
    t->n_blocks = n;
    t->origins = d->origins;
    t->conn = d->conn;
    t->n_tiles = (uint32_t)(tab.size() / kTileB);
    if (hipMalloc((void**)&t->d_tiles, tab.size() * sizeof(uint32_t)) != hipSuccess) {
        delete t;
        return set_error(NX_ERR_OOM, "tile table");
    }
    if (hipMemcpy(t->d_tiles, tab.data(), tab.size() * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(t->d_tiles);
        delete t;
        return set_error(NX_ERR_HIP, "uploading the tile table");
    }
    *out = t;
    return NX_OK;
}

extern "C" int nx_block_tiles_info(nx_block_tiles t, uint32_t* n_tiles)
{
    if (!t || !n_tiles) return set_error(NX_ERR_INVALID, "null pointer");
    *n_tiles = t->n_tiles;
    return NX_OK;
}

extern "C" int nx_block_tiles_destroy(nx_block_tiles t)
{
    if (!t) return NX_OK;
    if (t->d_tiles) (void)hipFree(t->d_tiles);
    delete t;
    return NX_OK;
}

extern "C" int nx_d3q19_step_block_tiled(const nx_block_desc* d, nx_block_tiles tiles, const void* fin,
                                         void* fout, const uint32_t* flags, const nx_lbm_params* p,
                                         nx_stream s)
{
    if (int e = check(d)) return e;
    if (!tiles || !fin || !fout || !flags || !p) return set_error(NX_ERR_INVALID, "null pointer");
    if (tiles->n_blocks != d->n_blocks || tiles->origins != d->origins || tiles->conn != d->conn)
        return set_error(NX_ERR_INVALID, "tile table built for another block grid");
    if (fin == fout)
        return set_error(NX_ERR_INVALID, "fin and fout must differ (two-population scheme)");
    if (d->storage == NX_F64)
        return set_error(NX_ERR_UNSUPPORTED, "the LDS-staged step holds 2- and 4-byte storage");
    const BGeom g = make_bgeom(d);
    const Par par{p->omega, {p->force[0], p->force[1], p->force[2]}};
    const bool force = p->force[0] != 0. || p->force[1] != 0. || p->force[2] != 0.;
    if (force && d->compute == NX_COMPUTE_REF)
        return set_error(NX_ERR_UNSUPPORTED,
                         "body force requires NX_COMPUTE_F32 or NX_COMPUTE_F64");
    const hipStream_t st = (hipStream_t)s;
    if (int e = by_type(d, [&](auto sc, auto cc) {
            constexpr int S = decltype(sc)::value;
            constexpr int CM = decltype(cc)::value;
            if constexpr (S != NX_F64) {
                using T = typename Stor<S>::T;
                if (force)
                    hipLaunchKernelGGL((k_step_tiled<S, CM, true>), dim3(tiles->n_tiles), dim3(kBV), 0, st,
                                       (const T*)fin, (T*)fout, flags, tiles->d_tiles, g, par, g_block_remap);
                else
                    hipLaunchKernelGGL((k_step_tiled<S, CM, false>), dim3(tiles->n_tiles), dim3(kBV), 0, st,
                                       (const T*)fin, (T*)fout, flags, tiles->d_tiles, g, par, g_block_remap);
            }
        }))
        return e;
    NX_CHECK_LAUNCH("k_step_tiled(block)");
    return NX_OK;
}

extern "C" int nx_d3q19_moments_block(const nx_block_desc* d, const void* fin,
                                      const uint32_t* flags, void* rho, void* u, nx_stream s)
{
    if (int e = check(d)) return e;
    if (!fin || !flags || !rho || !u) return set_error(NX_ERR_INVALID, "null pointer");
    const BGeom g = make_bgeom(d);
    const hipStream_t st = (hipStream_t)s;
    if (int e = by_type(d, [&](auto sc, auto cc) {
            constexpr int S = decltype(sc)::value;
            constexpr int CM = decltype(cc)::value;
            hipLaunchKernelGGL((k_moments<S, CM>), dim3(g.n_blocks), dim3(kBV), 0, st,
                               (const typename Stor<S>::T*)fin, flags, rho, u, g);
        }))
        return e;
    NX_CHECK_LAUNCH("k_moments(block)");
    return NX_OK;
}

extern "C" int nx_d3q19_mass_block(const nx_block_desc* d, const void* fin, const uint32_t* flags,
                                   double* out, nx_stream s)
{
    if (int e = check(d)) return e;
    if (!fin || !flags || !out) return set_error(NX_ERR_INVALID, "null pointer");
    const BGeom g = make_bgeom(d);
    const int nb = (int)std::min<uint32_t>(g.n_blocks, kRedGrid);
    void* part = nullptr;
    if (int e = scratch(sizeof(double) * nb, &part)) return e;
    const hipStream_t st = (hipStream_t)s;
    if (int e = by_type(d, [&](auto sc, auto) {
            constexpr int S = decltype(sc)::value;
            hipLaunchKernelGGL((k_mass<S>), dim3(nb), dim3(kBV), 0, st,
                               (const typename Stor<S>::T*)fin, flags, g, (double*)part);
        }))
        return e;
    NX_CHECK_LAUNCH("k_mass(block)");
    return finish_reduce((const double*)part, nb, 1, out, st);
}

extern "C" int nx_d3q19_residual_block(const nx_block_desc* d, const void* fin,
                                       const uint32_t* flags, float* u_prev, double* out,
                                       nx_stream s)
{
    if (int e = check(d)) return e;
    if (!fin || !flags || !u_prev || !out) return set_error(NX_ERR_INVALID, "null pointer");
    const BGeom g = make_bgeom(d);
    const int nb = (int)std::min<uint32_t>(g.n_blocks, kRedGrid);
    void* part = nullptr;
    if (int e = scratch(2 * sizeof(double) * nb, &part)) return e;
    const hipStream_t st = (hipStream_t)s;
    if (int e = by_type(d, [&](auto sc, auto cc) {
            constexpr int S = decltype(sc)::value;
            constexpr int CM = decltype(cc)::value;
            hipLaunchKernelGGL((k_residual<S, CM>), dim3(nb), dim3(kBV), 0, st,
                               (const typename Stor<S>::T*)fin, flags, u_prev, g, (double*)part);
        }))
        return e;
    NX_CHECK_LAUNCH("k_residual(block)");
    return finish_reduce((const double*)part, nb, 2, out, st);
}

extern "C" int nx_d3q19_set_block_remap(int on)
{
    g_block_remap = on ? 1 : 0;
    return NX_OK;
}

extern "C" int nx_d3q19_set_block_pairs(int on)
{
    g_block_pairs = on ? 1 : 0;
    return NX_OK;
}
