// mgrid.hip -- two-level grid refinement LBM (Neon's mGrid with depth 2 and
// the lbmMultiRes application's BGK D3Q19 scheme) on a stack of two
// block-sparse levels.
//
// Level 0 is the finest level (the domain); a level-1 voxel v covers the
// level-0 voxels 2v + {0,1}^3.  Each level is a bGrid of 8^3 blocks (the
// library's builder: blocks ordered bz, by, bx, 27-entry connectivity) with
// AoSoA populations (b*19 + q)*512 + x + 8y + 64z, and a dense table from
// block coordinates to block index for the cross-level lookups.
//
// Everything that depends only on the geometry is resolved on the host when
// the cavity is set up: every bulk leaf voxel carries a 2-bit state per
// population (plain pull, bounce-back link, interface link, no update), the
// fine-to-coarse accumulation is a gather list whose per-target order is the
// fine voxels' z, y, x order, and the coalescence weights are a field of
// level 1.  The kernels then do no searching: per coarse step
//   collide L1 | 2 x (collide L0, store L0 -> L1, stream L0) | stream L1
// (lbmMultiRes.h:141-263 for two levels, baseline operation order with the
// fine store).  Reference semantics restated (file:line in each kernel):
//   grid      libNeonDomain/src/domain/details/mGrid/mGrid.cpp:89-264
//   set-up    apps/lbmMultiRes/lidDrivenCavity.h:8-78, init.h:8-86
//   collide   apps/lbmMultiRes/collide.h:8-149
//   store     apps/lbmMultiRes/store.h:98-163
//   stream    apps/lbmMultiRes/stream.h:170-225
//   helpers   apps/lbmMultiRes/util.h:19-45, mPartition_imp.h:186-235
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
// per-population link states of a bulk leaf voxel
constexpr uint64_t kPlain = 0, kWall = 1, kIface = 2, kSkip = 3;

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

// ---- kernels ----------------------------------------------------------------

// collide.h:8-149: BGK on every bulk leaf voxel (fin -> fout, same voxel);
// a refined voxel of level 1 resets its fout, the accumulator of the fine
// store, to 0.
template <int S>
__global__ void __launch_bounds__(kBV)
    k_mr_collide(const typename MS<S>::T* __restrict__ fin, typename MS<S>::T* __restrict__ fout,
                 const uint32_t* __restrict__ state, double omega)
{
    using T = typename MS<S>::T;
    const uint32_t b = blockIdx.x, v = threadIdx.x;
    const uint32_t st = state[(uint64_t)b * kBV + v];
    if (!(st & kActive) || cls_bits(st) != NX_CELL_BULK) return;
    const uint64_t own = (uint64_t)b * (Q * kBV) + v;
    if (st & kRefined) {
#pragma unroll
        for (int q = 0; q < Q; q++) __builtin_nontemporal_store(T(0), &fout[own + (uint64_t)q * kBV]);
        return;
    }
    T f[Q], o[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) f[q] = fin[own + (uint64_t)q * kBV];
    bgk<T>(f, (T)omega, o);
#pragma unroll
    for (int q = 0; q < Q; q++) __builtin_nontemporal_store(o[q], &fout[own + (uint64_t)q * kBV]);
}

// store.h:98-163 as a gather: entry e adds its fine sources, in order, to
// the level-1 accumulator target[e].
template <int S>
__global__ void __launch_bounds__(256)
    k_mr_store(const typename MS<S>::T* __restrict__ fout0, typename MS<S>::T* __restrict__ fout1,
               const uint64_t* __restrict__ target, const uint32_t* __restrict__ first,
               const uint64_t* __restrict__ src, uint32_t n)
{
    using T = typename MS<S>::T;
    const uint32_t e = blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    T acc = fout1[target[e]];
    for (uint32_t k = first[e]; k < first[e + 1]; k++) acc += fout0[src[k]];
    fout1[target[e]] = acc;
}

struct UpLevel {  // the coarser level, seen from level 0 (explosion)
    const void* fout;
    const uint32_t* table;
    int32_t nb[3];
};

// stream.h:170-225 (stream fused with coalescence and explosion), pull, for
// the bulk leaf voxels of one level.  The plain pull of all 19 populations
// is issued first (upstream voxel through the connectivity row in scalar
// registers); voxels with another link state then fix those links up:
// bounce-back (own opposite + the wall's opposite), LEVEL 0 interface links
// read the coarse post-collision population of the uncle, the level-1 voxel
// holding the upstream position (explosion), LEVEL 1 interface links scale
// the refined upstream voxel's accumulated fine populations by its weight
// (coalescence); links in state "none" keep the value in `out`.
// COLLIDE: the pulled populations are collided at once with omega and the
// result written to `out` -- stream followed by the next collide of the
// same level (collide.h:8-149) in one pass; the results are those of the two
// kernels in sequence, bit for bit.  KEEP = false: the level has no link in
// state "none" (checked on the host), every population is stored.
template <int S, int LEVEL, bool COLLIDE, bool KEEP>
__global__ void __launch_bounds__(kBV)
    k_mr_stream(const typename MS<S>::T* __restrict__ fout, typename MS<S>::T* __restrict__ out,
                const uint32_t* __restrict__ state, const uint64_t* __restrict__ links,
                const uint32_t* __restrict__ conn, const int32_t* __restrict__ origins,
                const float* __restrict__ ss, UpLevel up, double omega)
{
    using T = typename MS<S>::T;
    const uint32_t b = blockIdx.x, v = threadIdx.x;
    const Conn cb = load_conn(conn, b);
    const uint32_t st = state[(uint64_t)b * kBV + v];
    if (!bulk_leaf(st)) return;
    const uint64_t lk = links[(uint64_t)b * kBV + v];
    const int x = (int)(v & 7u), y = (int)((v >> 3) & 7u), z = (int)(v >> 6);
    const uint64_t own = (uint64_t)b * (Q * kBV) + v;
    T f[Q];
    for_q(AllQ{}, [&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if constexpr (q == 9) {
            f[9] = fout[own + 9 * kBV];
        } else {
            uint32_t t, nv;
            upstream<q>(b, cb, x, y, z, t, nv);
            f[q] = fout[t != kInvalidBlock ? ((uint64_t)t * Q + q) * kBV + nv : own + (uint64_t)q * kBV];
        }
    });
#pragma unroll
    for (int q = 0; q < Q; q++) nx_pin(f[q]);
    uint32_t keep = 0;  // links whose value stays as it is in `out`
    if (lk != 0) {
        for_q(AllQ{}, [&](auto qc) {
            constexpr int q = decltype(qc)::value;
            if constexpr (q != 9) {
                const uint64_t kind = (lk >> (2 * q)) & 3u;
                if (kind == kPlain) return;
                uint32_t t, nv;
                upstream<q>(b, cb, x, y, z, t, nv);
                bool ok = true;
                if (kind == kWall) {
                    constexpr int o = OPP[q];
                    ok = t != kInvalidBlock;
                    if (ok) f[q] = fout[own + (uint64_t)o * kBV] + fout[((uint64_t)t * Q + o) * kBV + nv];
                } else if (kind == kIface) {
                    if constexpr (LEVEL == 0) {
                        const int ux = (origins[3 * (uint64_t)b] + x - CX[q]) >> 1;
                        const int uy = (origins[3 * (uint64_t)b + 1] + y - CY[q]) >> 1;
                        const int uz = (origins[3 * (uint64_t)b + 2] + z - CZ[q]) >> 1;
                        const int bx = ux >> 3, by = uy >> 3, bz = uz >> 3;
                        ok = !(bx < 0 || by < 0 || bz < 0 || bx >= up.nb[0] || by >= up.nb[1] || bz >= up.nb[2]);
                        const uint32_t tb = ok ? up.table[((uint64_t)bz * up.nb[1] + by) * up.nb[0] + bx]
                                               : kInvalidBlock;
                        ok = tb != kInvalidBlock;
                        const uint32_t slot = (uint32_t)((ux & 7) + 8 * (uy & 7) + 64 * (uz & 7));
                        if (ok) f[q] = ((const T*)up.fout)[((uint64_t)tb * Q + q) * kBV + slot];
                    } else {
                        ok = t != kInvalidBlock;
                        if (ok) f[q] = f[q] * (T)ss[((uint64_t)t * Q + q) * kBV + nv];
                    }
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
        T o[Q];
        bgk<T>(f, (T)omega, o);
#pragma unroll
        for (int q = 0; q < Q; q++) __builtin_nontemporal_store(o[q], &out[own + (uint64_t)q * kBV]);
    } else if constexpr (KEEP) {
#pragma unroll
        for (int q = 0; q < Q; q++)
            if (!((keep >> q) & 1u)) __builtin_nontemporal_store(f[q], &out[own + (uint64_t)q * kBV]);
    } else {
#pragma unroll
        for (int q = 0; q < Q; q++) __builtin_nontemporal_store(f[q], &out[own + (uint64_t)q * kBV]);
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
    void* fin = nullptr;
    void* fout = nullptr;

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
    static uint64_t elem(int64_t slot, int q) { return ((uint64_t)(slot / kBV) * Q + q) * kBV + slot % kBV; }
};

}  // namespace
}  // namespace mres
}  // namespace nx

struct nx_mgrid_s {
    int storage = NX_F64;
    nx::mres::Level L[2];
    std::vector<uint8_t> parent_block;  // level-1 8^3 block (dense block coords) holds an active voxel
    float* d_ss = nullptr;
    uint32_t n_store = 0;
    uint64_t* d_target = nullptr;
    uint32_t* d_first = nullptr;
    uint64_t* d_src = nullptr;
    uint64_t n_store_src = 0;
    bool fuse_fine = false;   // level 0 without "none" links: stream fused with the next collide
    bool keep_coarse = true;  // level 1 has "none" links
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

// Activation of both levels and overlap culling, mGrid.cpp:89-264 for depth
// 2 (the strong-balance pass, :266-345, cannot change a two-level grid).
void activate(const int32_t dim[3], const uint8_t* lam0, const uint8_t* lam1, Level& f, Level& c)
{
    auto group = [](Level& lv, const uint8_t* lam, int gx, int gy, int gz) {
        bool any = false;
        for (int k = 0; k < 8; k++) {
            const int x = 2 * gx + (k & 1), y = 2 * gy + ((k >> 1) & 1), z = 2 * gz + (k >> 2);
            if (lv.inside(x, y, z) && (lv.st[lv.lin(x, y, z)] || lam[lv.lin(x, y, z)])) any = true;
        }
        if (any)
            for (int k = 0; k < 8; k++) {
                const int x = 2 * gx + (k & 1), y = 2 * gy + ((k >> 1) & 1), z = 2 * gz + (k >> 2);
                if (lv.inside(x, y, z)) lv.st[lv.lin(x, y, z)] = (uint8_t)kActive;
            }
        return any;
    };
    for (int a = 0; a < 3; a++) {
        f.dim[a] = dim[a];
        c.dim[a] = (dim[a] + 1) / 2;
    }
    f.st.assign(f.nvox(), 0);
    c.st.assign(c.nvox(), 0);
    // level 0: a group (one level-1 voxel) with an active voxel activates its
    // voxels and its parent
    for (int gz = 0; gz < c.dim[2]; gz++)
        for (int gy = 0; gy < c.dim[1]; gy++)
            for (int gx = 0; gx < c.dim[0]; gx++)
                if (group(f, lam0, gx, gy, gz)) c.st[c.lin(gx, gy, gz)] = (uint8_t)kActive;
    // level 1, the top level: groups of 2^3 level-1 voxels
    for (int gz = 0; gz < (c.dim[2] + 1) / 2; gz++)
        for (int gy = 0; gy < (c.dim[1] + 1) / 2; gy++)
            for (int gx = 0; gx < (c.dim[0] + 1) / 2; gx++) group(c, lam1, gx, gy, gz);
    std::vector<uint8_t> refined(c.nvox(), 0);
    for (int z = 0; z < c.dim[2]; z++)
        for (int y = 0; y < c.dim[1]; y++)
            for (int x = 0; x < c.dim[0]; x++) {
                bool r = false;
                for (int k = 0; k < 8 && !r; k++)
                    r = f.active(2 * x + (k & 1), 2 * y + ((k >> 1) & 1), 2 * z + (k >> 2));
                refined[c.lin(x, y, z)] = r;
            }
    // culling: a refined voxel whose 26 in-domain neighbours are refined goes
    std::vector<uint8_t> keep = c.st;
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
                if (all) keep[i] = 0;
            }
    for (size_t i = 0; i < c.nvox(); i++) c.st[i] = keep[i] ? (uint8_t)(kActive | (refined[i] ? kRefined : 0u)) : 0;
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

template <typename T>
void fill_cavity(nx_mgrid_s& g, double ulb, std::vector<T> pop[2])
{
    // lidDrivenCavity.h:8-78: leaf voxels of level 0 on the domain faces are
    // bounce-back, those with y = Y-1 the lid; populations w_q / 0 /
    // -6 w_q (c_q . ulid), ulid = (ulb, 0, 0), evaluated as the reference
    // does in T with binary64 constants
    for (int l = 0; l < 2; l++) {
        Level& lv = g.L[l];
        pop[l].assign(19 * (size_t)kBV * lv.n_blocks, T(0));
        for (int z = 0; z < lv.dim[2]; z++)
            for (int y = 0; y < lv.dim[1]; y++)
                for (int x = 0; x < lv.dim[0]; x++) {
                    const size_t i = lv.lin(x, y, z);
                    if (!(lv.st[i] & kActive)) continue;
                    int cls = NX_CELL_BULK;
                    const int64_t s = lv.slot(x, y, z);
                    if (!(lv.st[i] & kRefined)) {
                        if (l == 0 && (x == 0 || y == 0 || z == 0 || x == lv.dim[0] - 1 || y == lv.dim[1] - 1 ||
                                       z == lv.dim[2] - 1))
                            cls = y == lv.dim[1] - 1 ? NX_CELL_MOVINGWALL : NX_CELL_BOUNCEBACK;
                        for (int q = 0; q < Q; q++) {
                            T p = (T)wq(q);
                            if (cls != NX_CELL_BULK) p = 0;
                            if (cls == NX_CELL_MOVINGWALL) {
                                p = (T)((double)p + (double)CX[q] * ulb);
                                p = (T)((double)p + (double)CY[q] * 0.);
                                p = (T)((double)p + (double)CZ[q] * 0.);
                                p = (T)((double)p * (-6. * wq(q)));
                            }
                            pop[l][Level::elem(s, q)] = p;
                        }
                    }
                    lv.st[i] = (uint8_t)((lv.st[i] & (kActive | kRefined)) | (cls << kClsShift));
                }
    }
}

bool has_parent(const nx_mgrid_s& g, int x, int y, int z)
{
    // mPartition_imp.h:199-208: the level-1 memory block (8^3) holding the
    // parent voxel exists
    const Level& c = g.L[1];
    const int bx = (x >> 1) >> 3, by = (y >> 1) >> 3, bz = (z >> 1) >> 3;
    if (bx >= c.nb[0] || by >= c.nb[1] || bz >= c.nb[2]) return false;
    return c.table[((size_t)bz * c.nb[1] + by) * c.nb[0] + bx] != kInvalidBlock;
}

// Link states of every bulk leaf voxel (stream.h:170-225 decided per link);
// returns the number of links in state "none".
uint64_t links_of(const nx_mgrid_s& g, int l, std::vector<uint64_t>& out)
{
    uint64_t n_skip = 0;
    const Level& lv = g.L[l];
    const Level& c = g.L[1];
    out.assign((size_t)kBV * lv.n_blocks, 0);
    for (int z = 0; z < lv.dim[2]; z++)
        for (int y = 0; y < lv.dim[1]; y++)
            for (int x = 0; x < lv.dim[0]; x++) {
                const uint8_t st = lv.st[lv.lin(x, y, z)];
                if (!bulk_leaf(st)) continue;
                const bool parent = l == 0 && has_parent(g, x, y, z);
                uint64_t lk = 0;
                for (int q = 0; q < Q; q++) {
                    if (q == 9) continue;  // own population: plain
                    const int nx_ = x - CX[q], ny = y - CY[q], nz = z - CZ[q];
                    uint64_t kind = kSkip;
                    if (lv.active(nx_, ny, nz)) {
                        const uint8_t ns = lv.st[lv.lin(nx_, ny, nz)];
                        if (ns & kRefined)
                            kind = kIface;  // coalescence (level 1 only)
                        else
                            kind = cls_bits(ns) == NX_CELL_BULK ? kPlain : kWall;
                    } else if (parent) {
                        const int ux = (x >> 1) + uoff(x, -CX[q]), uy = (y >> 1) + uoff(y, -CY[q]),
                                  uz = (z >> 1) + uoff(z, -CZ[q]);
                        if (c.active(ux, uy, uz)) kind = kIface;  // explosion
                    }
                    lk |= kind << (2 * q);
                    n_skip += kind == kSkip;
                }
                out[lv.slot(x, y, z)] = lk;
            }
    return n_skip;
}

}  // namespace

extern "C" int nx_mgrid_create(const int32_t dim[3], const uint8_t* lam0, const uint8_t* lam1, int32_t storage,
                               nx_mgrid* out)
{
    if (!dim || !lam0 || !lam1 || !out) return set_error(NX_ERR_INVALID, "nx_mgrid_create: null argument");
    for (int a = 0; a < 3; a++)
        if (dim[a] < 2 || dim[a] > (1 << 14))
            return set_error(NX_ERR_INVALID, "nx_mgrid_create: dimensions in [2, 16384]");
    if (storage != NX_F64 && storage != NX_F32)
        return set_error(NX_ERR_UNSUPPORTED, "nx_mgrid_create: storage NX_F64 or NX_F32 (the reference's types)");
    auto* g = new nx_mgrid_s;
    g->storage = storage;
    activate(dim, lam0, lam1, g->L[0], g->L[1]);
    for (int l = 0; l < 2; l++)
        if (int e = build_level(g->L[l])) {
            nx_mgrid_destroy(g);
            return e;
        }
    *out = g;
    return NX_OK;
}

extern "C" int nx_mgrid_info(nx_mgrid g, uint32_t n_blocks[2], uint64_t n_active[2], uint64_t n_leaf[2],
                             uint64_t* n_store, int32_t* fine_fused)
{
    if (g && fine_fused) *fine_fused = g->fuse_fine ? 1 : 0;
    if (!g) return set_error(NX_ERR_INVALID, "null mgrid");
    for (int l = 0; l < 2; l++) {
        uint64_t a = 0, lf = 0;
        for (uint8_t s : g->L[l].st) {
            a += (s & kActive) ? 1 : 0;
            lf += ((s & kActive) && !(s & kRefined)) ? 1 : 0;
        }
        if (n_blocks) n_blocks[l] = g->L[l].n_blocks;
        if (n_active) n_active[l] = a;
        if (n_leaf) n_leaf[l] = lf;
    }
    if (n_store) *n_store = g->n_store_src;
    return NX_OK;
}

extern "C" int nx_mgrid_state(nx_mgrid g, uint8_t* st0, uint8_t* st1)
{
    if (!g) return set_error(NX_ERR_INVALID, "null mgrid");
    if (st0) std::copy(g->L[0].st.begin(), g->L[0].st.end(), st0);
    if (st1) std::copy(g->L[1].st.begin(), g->L[1].st.end(), st1);
    return NX_OK;
}

extern "C" int nx_mgrid_cavity(nx_mgrid g, double ulb)
{
    if (!g) return set_error(NX_ERR_INVALID, "null mgrid");
    if (!std::isfinite(ulb)) return set_error(NX_ERR_INVALID, "nx_mgrid_cavity: non-finite lid speed");
    Level& f = g->L[0];
    Level& c = g->L[1];
    const size_t elems[2] = {19 * (size_t)kBV * f.n_blocks, 19 * (size_t)kBV * c.n_blocks};
    const size_t es = g->storage == NX_F64 ? 8 : 4;
    // populations
    std::vector<double> p64[2];
    std::vector<float> p32[2];
    if (g->storage == NX_F64)
        fill_cavity<double>(*g, ulb, p64);
    else
        fill_cavity<float>(*g, ulb, p32);
    for (int l = 0; l < 2; l++) {
        Level& lv = g->L[l];
        if (int e = upload((void**)&lv.d_conn, lv.conn.data(), lv.conn.size() * 4)) return e;
        if (int e = upload((void**)&lv.d_origins, lv.origins.data(), lv.origins.size() * 4)) return e;
        if (int e = upload((void**)&lv.d_table, lv.table.data(), lv.table.size() * 4)) return e;
        const void* src = g->storage == NX_F64 ? (const void*)p64[l].data() : (const void*)p32[l].data();
        if (int e = upload(&g->L[l].fin, src, elems[l] * es)) return e;
        if (int e = upload(&g->L[l].fout, src, elems[l] * es)) return e;
        std::vector<uint32_t> st((size_t)kBV * g->L[l].n_blocks, 0);
        for (int z = 0; z < lv.dim[2]; z++)
            for (int y = 0; y < lv.dim[1]; y++)
                for (int x = 0; x < lv.dim[0]; x++) {
                    const int64_t s = lv.slot(x, y, z);
                    if (s >= 0) st[s] = lv.st[lv.lin(x, y, z)];
                }
        if (int e = upload((void**)&lv.d_state, st.data(), st.size() * 4)) return e;
        std::vector<uint64_t> lk;
        const uint64_t n_skip = links_of(*g, l, lk);
        if (l == 0) g->fuse_fine = n_skip == 0;
        if (l == 1) g->keep_coarse = n_skip != 0;
        if (int e = upload((void**)&lv.d_links, lk.data(), lk.size() * 8)) return e;
    }
    // fine store gather lists (store.h:98-163) and sumStore counts
    // (init.h:8-55): fine voxels in z, y, x order, q ascending
    std::vector<float> cnt(elems[1], 0.f);
    std::vector<std::pair<uint64_t, uint64_t>> tl;  // (target element, source element)
    for (int z = 0; z < f.dim[2]; z++)
        for (int y = 0; y < f.dim[1]; y++)
            for (int x = 0; x < f.dim[0]; x++) {
                if (!f.active(x, y, z) || !has_parent(*g, x, y, z)) continue;
                const int px = x >> 1, py = y >> 1, pz = z >> 1;
                for (int q = 0; q < Q; q++) {
                    if (q == 9) continue;
                    const int ux = uoff(x, CX[q]), uy = uoff(y, CY[q]), uz = uoff(z, CZ[q]);
                    if (f.active(x + ux, y + uy, z + uz)) continue;
                    if (!c.active(px + ux, py + uy, pz + uz)) continue;
                    const int sx = ux - CX[q], sy = uy - CY[q], sz = uz - CZ[q];
                    const int64_t cs = c.slot(px + sx, py + sy, pz + sz);
                    if (cs < 0) continue;
                    cnt[Level::elem(cs, q)] += 1.f;  // initSumStore has no csChild test
                    if (!f.active(x + sx, y + sy, z + sz)) continue;
                    tl.emplace_back(Level::elem(cs, q), Level::elem(f.slot(x, y, z), q));
                }
            }
    for (float& v : cnt)
        if ((double)v > 0.001) v = g->storage == NX_F64 ? (float)(1.0 / (2.0 * (double)v)) : 1.f / (2.f * v);
    if (int e = upload((void**)&g->d_ss, cnt.data(), cnt.size() * 4)) return e;
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
    g->n_store = (uint32_t)target.size();
    g->n_store_src = src.size();
    if (int e = upload((void**)&g->d_target, target.data(), target.size() * 8)) return e;
    if (int e = upload((void**)&g->d_first, first.data(), first.size() * 4)) return e;
    if (int e = upload((void**)&g->d_src, src.data(), src.size() * 8)) return e;
    g->ready = true;
    return NX_OK;
}

namespace {

template <int S>
int launch_step(nx_mgrid_s& g, double omega0, hipStream_t st)
{
    using T = typename MS<S>::T;
    Level& f = g.L[0];
    Level& c = g.L[1];
    // computeOmega (util.h:40-45): omega0 is the coarse level's
    const double om1 = 2 * omega0 / (std::scalbln(1.0, 1) + (1. - std::scalbln(1.0, 0)) * omega0);
    const double om0 = 2 * omega0 / (std::scalbln(1.0, 2) + (1. - std::scalbln(1.0, 1)) * omega0);
    const UpLevel none{nullptr, nullptr, {0, 0, 0}};
    const UpLevel up{c.fout, c.d_table, {c.nb[0], c.nb[1], c.nb[2]}};
    auto store = [&](const void* fout0) {
        if (g.n_store)
            hipLaunchKernelGGL((k_mr_store<S>), dim3((g.n_store + 255) / 256), dim3(256), 0, st, (const T*)fout0,
                               (T*)c.fout, g.d_target, g.d_first, g.d_src, g.n_store);
    };
    if (c.n_blocks)
        hipLaunchKernelGGL((k_mr_collide<S>), dim3(c.n_blocks), dim3(kBV), 0, st, (const T*)c.fin, (T*)c.fout,
                           c.d_state, om1);
    if (f.n_blocks && !g.fuse_fine) {
        for (int sub = 0; sub < 2; sub++) {
            hipLaunchKernelGGL((k_mr_collide<S>), dim3(f.n_blocks), dim3(kBV), 0, st, (const T*)f.fin,
                               (T*)f.fout, f.d_state, om0);
            store(f.fout);
            hipLaunchKernelGGL((k_mr_stream<S, 0, false, true>), dim3(f.n_blocks), dim3(kBV), 0, st,
                               (const T*)f.fout, (T*)f.fin, f.d_state, f.d_links, f.d_conn, f.d_origins,
                               (const float*)nullptr, up, 0.0);
        }
    } else if (f.n_blocks) {
        // level 0, two sub-steps: collide (fin -> fout), store, then stream
        // fused with the second collide (fout -> fin buffer, which now holds
        // post-collision values), store, stream (-> the old fout buffer);
        // afterwards the two buffers have traded roles.  Only when no bulk
        // leaf of level 0 has a link that keeps its previous value (state
        // "none"): that value is gone once the fused pass has written.
        hipLaunchKernelGGL((k_mr_collide<S>), dim3(f.n_blocks), dim3(kBV), 0, st, (const T*)f.fin, (T*)f.fout,
                           f.d_state, om0);
        store(f.fout);
        hipLaunchKernelGGL((k_mr_stream<S, 0, true, false>), dim3(f.n_blocks), dim3(kBV), 0, st, (const T*)f.fout,
                           (T*)f.fin, f.d_state, f.d_links, f.d_conn, f.d_origins, (const float*)nullptr, up, om0);
        store(f.fin);
        hipLaunchKernelGGL((k_mr_stream<S, 0, false, false>), dim3(f.n_blocks), dim3(kBV), 0, st, (const T*)f.fin,
                           (T*)f.fout, f.d_state, f.d_links, f.d_conn, f.d_origins, (const float*)nullptr, up, 0.0);
        std::swap(f.fin, f.fout);
    }
    if (c.n_blocks && g.keep_coarse)
        hipLaunchKernelGGL((k_mr_stream<S, 1, false, true>), dim3(c.n_blocks), dim3(kBV), 0, st, (const T*)c.fout,
                           (T*)c.fin, c.d_state, c.d_links, c.d_conn, c.d_origins, (const float*)g.d_ss, none, 0.0);
    else if (c.n_blocks)
        hipLaunchKernelGGL((k_mr_stream<S, 1, false, false>), dim3(c.n_blocks), dim3(kBV), 0, st, (const T*)c.fout,
                           (T*)c.fin, c.d_state, c.d_links, c.d_conn, c.d_origins, (const float*)g.d_ss, none, 0.0);
    NX_CHECK_LAUNCH("mgrid step");
    return NX_OK;
}

}  // namespace

extern "C" int nx_mgrid_step(nx_mgrid g, double omega0, int32_t n_steps, nx_stream s)
{
    if (!g || n_steps < 0) return set_error(NX_ERR_INVALID, "nx_mgrid_step: bad arguments");
    if (!g->ready) return set_error(NX_ERR_NOT_INIT, "nx_mgrid_step: call nx_mgrid_cavity first");
    if (!(omega0 > 0.0 && omega0 < 2.0)) return set_error(NX_ERR_INVALID, "nx_mgrid_step: omega0 in (0, 2)");
    for (int k = 0; k < n_steps; k++) {
        const int e = g->storage == NX_F64 ? launch_step<NX_F64>(*g, omega0, (hipStream_t)s)
                                           : launch_step<NX_F32>(*g, omega0, (hipStream_t)s);
        if (e) return e;
    }
    return NX_OK;
}

extern "C" int nx_mgrid_fields(nx_mgrid g, int32_t level, int32_t which, void* dense)
{
    if (!g || (level != 0 && level != 1) || (which != 0 && which != 1) || !dense)
        return set_error(NX_ERR_INVALID, "nx_mgrid_fields: bad arguments");
    if (!g->ready) return set_error(NX_ERR_NOT_INIT, "nx_mgrid_fields: call nx_mgrid_cavity first");
    const Level& lv = g->L[level];
    const size_t es = g->storage == NX_F64 ? 8 : 4, n = lv.nvox();
    std::vector<uint8_t> buf(19 * (size_t)kBV * lv.n_blocks * es);
    NX_CHECK_HIP(hipDeviceSynchronize());
    if (!buf.empty())
        NX_CHECK_HIP(hipMemcpy(buf.data(), which ? lv.fout : lv.fin, buf.size(), hipMemcpyDeviceToHost));
    for (int z = 0; z < lv.dim[2]; z++)
        for (int y = 0; y < lv.dim[1]; y++)
            for (int x = 0; x < lv.dim[0]; x++) {
                const int64_t s = lv.slot(x, y, z);
                if (s < 0) continue;
                for (int q = 0; q < Q; q++)
                    std::memcpy((uint8_t*)dense + ((size_t)q * n + lv.lin(x, y, z)) * es,
                                buf.data() + Level::elem(s, q) * es, es);
            }
    return NX_OK;
}

extern "C" int nx_mgrid_destroy(nx_mgrid g)
{
    if (!g) return NX_OK;
    for (auto& lv : g->L)
        for (void* p : {(void*)lv.d_state, (void*)lv.d_links, (void*)lv.d_conn, (void*)lv.d_origins,
                        (void*)lv.d_table, lv.fin, lv.fout})
            if (p) hipFree(p);
    for (void* p : {(void*)g->d_ss, (void*)g->d_target, (void*)g->d_first, (void*)g->d_src})
        if (p) hipFree(p);
    delete g;
    return NX_OK;
}
