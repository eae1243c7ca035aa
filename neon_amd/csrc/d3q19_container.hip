// d3q19_container.hip -- the dense LBM step as a reusable launch plan.
//
// The counterpart of LbmContainers::iteration (benchmarks/
// lbm-lid-driven-cavity-flow/src/LbmTools.h:284-325), which returns a
// Neon::set::Container that each time step runs (Containter.h:25-35,
// LbmIteration.h:56-61).  The container binds (fin, fout, flags, omega, g)
// and, once at creation, builds a 1-bit-per-cell bitmap of the interior
// cells -- bulk cells with no wall link whose 18 upstream cells need no
// periodic wrap -- with a wave64 ballot per 64 cells.
//
// A step is one launch of k_fused on the caller's stream:
//   phase 1  interior lanes: the 19 population loads are issued before the
//            bitmap test (pinned ahead of the branch; their addresses do not
//            depend on per-cell data), then collide and store.  Addresses are
//            a uniform per-plane base plus a 32-bit lane offset.
//   other    the remaining lanes of the same wavefront: flag word, then for
//            bulk cells the full pull (wall links, moving-wall term, missing
//            upstream cells, periodic wrap) and collide; wall cells take
//            their own value from fin.  Its loads hit the cache lines the
//            interior lanes of the wavefront fetched.
// Every lane of a wavefront stores all 19 populations at once.  Leaving the
// wall lanes unwritten makes the first and last cache line of every grid row
// a partial write, which HBM serves as read-modify-write (measured on
// MI355X: 27% slower at 256^3).  Wall cells are never changed by a step, so
// copying fin's wall values is exact when both fields hold the same wall
// values -- the reference initialises both fields identically
// (RunCavityTwoPop.cu:168-206).  The container checks this once at
// creation; if any wall value differs it leaves wall cells unwritten.
// Both paths evaluate the same expression tree as nx_d3q19_step_dense, so
// the populations agree bit for bit.
#include <cstring>
#include <vector>

#include "dense_common.hpp"

struct nx_container_s {
    nx_dense_desc d;
    nx::dense::DGeom g;
    const void* fin;
    void* fout;
    const uint32_t* flags;
    nx::dense::Par p;
    bool force;
    uint32_t* interior = nullptr;  // device bitmap, ceil(n/32) words
    uint64_t n_boundary = 0;       // bulk cells on the full-pull path
    uint64_t n_interior = 0;
    int copy_walls = 0;            // 1: fin and fout wall values are identical
    // 1: in the rows away from the y/z faces (2 <= y <= Y-3, 2 <= z <= Z-3)
    // every x = 0 and x = X-1 cell is a non-bulk cell holding the same value
    // per population (xwall.v), in both fields
    int xwall_uniform = 0;
    nx::dense::XWall xwall{};
    // graph of 2 * kPairSteps steps alternating this container and `pair_with`
    hipGraphExec_t pair_graph = nullptr;
    const nx_container_s* pair_with = nullptr;
    nx_tuning pair_tuning[2] = {};
};

namespace nx {
namespace dense {
namespace {

constexpr int kCWG = 256;  // classification workgroup (4 wavefronts)

__device__ __forceinline__ bool is_interior(const DGeom& g, uint32_t i, uint32_t fl)
{
    if (cls_of(fl) != NX_CELL_BULK || (fl & (NX_MASK_BITS | NX_FLAG_OPEN | NX_ZH_BITS))) return false;
    if (!g.periodic) return true;
    int x, y, z;
    coords(g, i, x, y, z);
    if ((g.periodic & NX_PERIODIC_X) && (x == 0 || x == g.X - 1)) return false;
    if ((g.periodic & NX_PERIODIC_Y) && (y == 0 || y == g.Y - 1)) return false;
    if ((g.periodic & NX_PERIODIC_Z) && (z == 0 || z == g.Z - 1)) return false;
    return true;
}

// Interior bitmap (one ballot word pair per wavefront), and the numbers of
// interior and other bulk cells per workgroup.
__global__ void __launch_bounds__(kCWG)
    k_classify(const uint32_t* __restrict__ flags, DGeom g, uint32_t* __restrict__ bits,
               uint32_t* __restrict__ wg_count, uint64_t* __restrict__ wg_interior)
{
    __shared__ uint32_t cnt[kCWG / 64];
    __shared__ uint32_t icnt[kCWG / 64];
    const uint32_t i = blockIdx.x * kCWG + threadIdx.x;
    const uint32_t fl = i < g.n ? flags[i] : 0u;
    const bool in = i < g.n && is_interior(g, i, fl);
    const bool bd = i < g.n && cls_of(fl) == NX_CELL_BULK && !in;
    const uint64_t bi = __ballot(in);
    const uint64_t bb = __ballot(bd);
    const uint64_t bz = __ballot(bd && (fl & NX_ZH_BITS));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // wavefront w of workgroup b covers cells 64*(4b+w) .. +63: words 2*(4b+w), +1
    const uint32_t word = 2u * (blockIdx.x * (kCWG / 64) + wave);
    if (lane == 0 && word < (g.n + 31) / 32) bits[word] = (uint32_t)bi;
    if (lane == 1 && word + 1 < (g.n + 31) / 32) bits[word + 1] = (uint32_t)(bi >> 32);
    if (lane == 0) {
        // bit 31: the wavefront holds a Zou-He cell (counts are <= 256)
        cnt[wave] = (uint32_t)__popcll(bb) | (bz ? 0x80000000u : 0u);
        icnt[wave] = (uint32_t)__popcll(bi);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t c = 0, ic = 0;
        uint32_t zh = 0;
        for (int w = 0; w < kCWG / 64; w++) {
            c += cnt[w] & 0x7FFFFFFFu;
            zh |= cnt[w] & 0x80000000u;
            ic += icnt[w];
        }
        c |= zh;
        wg_count[blockIdx.x] = c;
        wg_interior[blockIdx.x] = ic;
    }
}

// Load of element j of a population plane: uniform 64-bit plane base plus a
// 32-bit byte offset per lane (global_load saddr form; planes < 4 GiB).
template <int LP, typename T>
__device__ __forceinline__ T ld_plane(const T* plane_base, uint32_t j)
{
    return ldp<LP>((const T*)((const char*)plane_base + (uint32_t)(j * (uint32_t)sizeof(T))));
}

template <typename T>
__device__ __forceinline__ void st_plane(T* plane_base, uint32_t j, T v)
{
    __builtin_nontemporal_store(v, (T*)((char*)plane_base + (uint32_t)(j * (uint32_t)sizeof(T))));
}

// 1 in *mismatch when a non-bulk cell holds different values in a and b.
template <typename T>
__global__ void __launch_bounds__(kCWG)
    k_walls_equal(const T* __restrict__ a, const T* __restrict__ b,
                  const uint32_t* __restrict__ flags, DGeom g, uint32_t* mismatch)
{
    const uint32_t i = blockIdx.x * kCWG + threadIdx.x;
    bool diff = false;
    if (i < g.n && cls_of(flags[i]) != NX_CELL_BULK)
        for (int q = 0; q < Q; q++) {
            using U = typename std::conditional<
                sizeof(T) == 8, uint64_t,
                typename std::conditional<sizeof(T) == 4, uint32_t, uint16_t>::type>::type;
            diff = diff || __builtin_bit_cast(U, a[q * g.plane + i]) !=
                               __builtin_bit_cast(U, b[q * g.plane + i]);
        }
    if (__ballot(diff) && (threadIdx.x & 63) == 0) mismatch[blockIdx.x] = 1u;
}

// The x-end cells of the rows away from the y/z faces (x in {0, X-1},
// 2 <= y <= Y-3, 2 <= z <= Z-3): 1 in *mismatch unless every one is a
// non-bulk cell holding, per population, the value of the first of them,
// whose values go to ref[0..18].
template <typename T>
__global__ void __launch_bounds__(kCWG)
    k_xwall_check(const T* __restrict__ f, const uint32_t* __restrict__ flags, DGeom g,
                  uint32_t* mismatch, uint64_t* ref)
{
    const uint32_t t = blockIdx.x * kCWG + threadIdx.x;
    const uint32_t ny = (uint32_t)g.Y - 4u, nz = (uint32_t)g.Z - 4u;
    const uint32_t i0 = 2u * (uint32_t)g.X + 2u * g.XY;
    if (t < Q) ref[t] = raw_bits(f[(uint64_t)t * g.plane + i0]);
    if (t >= 2u * ny * nz) return;
    const uint32_t r = t >> 1;
    const uint32_t x = (t & 1u) ? (uint32_t)g.X - 1u : 0u;
    const uint32_t i = x + (uint32_t)g.X * (2u + r % ny) + g.XY * (2u + r / ny);
    bool bad = cls_of(flags[i]) == NX_CELL_BULK;
    for (int q = 0; q < Q; q++)
        bad = bad || raw_bits(f[(uint64_t)q * g.plane + i]) != raw_bits(f[(uint64_t)q * g.plane + i0]);
    if (bad) *mismatch = 1u;
}

// WG threads, CPT cells per thread (cells base + c*WG), LP = load policy
// (1: non-temporal), copy_walls as above.
template <int S, int CM, bool FORCE, bool PER, int WG, int CPT, int LP>
__global__ void __launch_bounds__(WG)
    k_fused(const typename Stor<S>::T* __restrict__ fin, typename Stor<S>::T* __restrict__ fout,
            const uint32_t* __restrict__ flags, const uint32_t* __restrict__ bits, DGeom g, Par p,
            int remap, int copy_walls)
{
    using T = typename Stor<S>::T;
    using P = typename Stor<S>::P;
    using R = typename Real<S, CM>::R;
    const uint32_t wg = remap ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x;
    const uint32_t base = wg * (uint32_t)(WG * CPT) + threadIdx.x;
    uint32_t word[CPT];
    T v[CPT][Q];
#pragma unroll
    for (int c = 0; c < CPT; c++) {
        const uint32_t ic = min(base + (uint32_t)(c * WG), g.n - 1u);
        word[c] = bits[ic >> 5];
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const int off = CX[q] + CY[q] * g.X + CZ[q] * (int)g.XY;
            const uint32_t j = q == 9 ? ic : min(ic - (uint32_t)off, g.n - 1u);
            v[c][q] = ld_plane<LP>(fin + q * g.plane, j);
        }
    }
#pragma unroll
    for (int c = 0; c < CPT; c++)
#pragma unroll
        for (int q = 0; q < Q; q++) pin(v[c][q]);
    // the interior word travels with the population loads (left to the
    // compiler it is fetched after they complete: one more memory latency)
#pragma unroll
    for (int c = 0; c < CPT; c++) pin(word[c]);
#pragma unroll
    for (int c = 0; c < CPT; c++) {
        const uint32_t i = base + (uint32_t)(c * WG);
        if (i >= g.n) continue;
        bool store = true;
        if ((word[c] >> (i & 31u)) & 1u) {
            P f[Q];
#pragma unroll
            for (int q = 0; q < Q; q++) f[q] = Stor<S>::bulk(v[c][q], q);
            R out[Q];
            collide<S, CM, FORCE>(f, p.omega, p.g, out);
#pragma unroll
            for (int q = 0; q < Q; q++) v[c][q] = Stor<S>::template enc<R>(out[q], q);
        } else {
            const uint32_t fl = flags[i];
            if (cls_of(fl) == NX_CELL_BULK) {
                P f[Q];
                pull<S, PER>(fin, g, i, fl, f);
                R out[Q];
                collide<S, CM, FORCE>(f, p.omega, p.g, out);
#pragma unroll
                for (int q = 0; q < Q; q++) v[c][q] = Stor<S>::template enc<R>(out[q], q);
            } else if (copy_walls) {
#pragma unroll
                for (int q = 0; q < Q; q++) v[c][q] = fin[q * g.plane + i];
            } else {
                store = false;
            }
        }
        if (store) {
#pragma unroll
            for (int q = 0; q < Q; q++) st_plane(fout + q * g.plane, i, v[c][q]);
        }
    }
}

// Vector variant: each lane owns VX consecutive cells of a row (X % VX == 0,
// VX * sizeof(T) = 16 bytes), so every population is one 16-byte load at
// i - off_q (dword aligned, which multi-dword global loads accept) and one
// aligned 16-byte store at i.  A lane whose VX cells are
// all interior takes the vector path; any other lane runs the per-cell path
// of k_fused for each of its cells and stores per cell.
template <int S, int CM, bool FORCE, bool PER, int WG, int VX, int LP>
__global__ void __launch_bounds__(WG)
    k_fused_vx(const typename Stor<S>::T* __restrict__ fin, typename Stor<S>::T* __restrict__ fout,
               const uint32_t* __restrict__ flags, const uint32_t* __restrict__ bits, DGeom g, Par p,
               int remap, int copy_walls)
{
    using T = typename Stor<S>::T;
    using P = typename Stor<S>::P;
    using R = typename Real<S, CM>::R;
    typedef T VL __attribute__((ext_vector_type(VX), aligned(sizeof(T))));  // dword-aligned load
    typedef T VS __attribute__((ext_vector_type(VX)));                      // aligned store
    static_assert(VX * sizeof(T) == 16 && sizeof(T) >= 4, "16-byte lanes of dword elements");
    const uint32_t wg = remap ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x;
    const uint32_t i0 = (wg * (uint32_t)WG + threadIdx.x) * (uint32_t)VX;
    if (i0 >= g.n) return;
    const uint32_t word = bits[i0 >> 5];
    VL v[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int off = CX[q] + CY[q] * g.X + CZ[q] * (int)g.XY;
        // clamped into the plane: only lanes whose cells are all interior
        // (so every upstream cell exists) use these values
        const int64_t jr = (int64_t)i0 - off;
        const uint32_t jj = jr < 0 ? 0u : (jr > (int64_t)(g.n - VX) ? g.n - (uint32_t)VX : (uint32_t)jr);
        const VL* src = (const VL*)((const char*)(fin + q * g.plane) + (uint32_t)(jj * (uint32_t)sizeof(T)));
        v[q] = LP ? __builtin_nontemporal_load(src) : *src;
    }
#pragma unroll
    for (int q = 0; q < Q; q++)
#pragma unroll
        for (int c = 0; c < VX; c++) {
            T t = v[q][c];
            pin(t);
            v[q][c] = t;
        }
    constexpr uint32_t kAll = (1u << VX) - 1u;
    const uint32_t nib = (word >> (i0 & 31u)) & kAll;
    if (nib == kAll) {
#pragma unroll
        for (int c = 0; c < VX; c++) {
            P f[Q];
#pragma unroll
            for (int q = 0; q < Q; q++) f[q] = Stor<S>::bulk(v[q][c], q);
            R out[Q];
            collide<S, CM, FORCE>(f, p.omega, p.g, out);
#pragma unroll
            for (int q = 0; q < Q; q++) v[q][c] = Stor<S>::template enc<R>(out[q], q);
        }
#pragma unroll
        for (int q = 0; q < Q; q++) {
            VS o;
#pragma unroll
            for (int c = 0; c < VX; c++) o[c] = v[q][c];
            __builtin_nontemporal_store(o, (VS*)(fout + q * g.plane + i0));
        }
        return;
    }
#pragma unroll 1
    for (int c = 0; c < VX; c++) {
        const uint32_t i = i0 + (uint32_t)c;
        T o[Q];
        bool store = true;
        P f[Q];
        bool bulk = false;
        if ((nib >> c) & 1u) {
#pragma unroll
            for (int q = 0; q < Q; q++) f[q] = Stor<S>::bulk(v[q][c], q);
            bulk = true;
        } else {
            const uint32_t fl = flags[i];
            if (cls_of(fl) == NX_CELL_BULK) {
                pull<S, PER>(fin, g, i, fl, f);
                bulk = true;
            } else if (copy_walls) {
#pragma unroll
                for (int q = 0; q < Q; q++) o[q] = fin[q * g.plane + i];
            } else {
                store = false;
            }
        }
        if (bulk) {
            R out[Q];
            collide<S, CM, FORCE>(f, p.omega, p.g, out);
#pragma unroll
            for (int q = 0; q < Q; q++) o[q] = Stor<S>::template enc<R>(out[q], q);
        }
        if (store) {
#pragma unroll
            for (int q = 0; q < Q; q++) st_plane(fout + q * g.plane, i, o[q]);
        }
    }
}

// Single-collision variant: every lane loads its flag word with the 19
// speculative loads; a bulk lane with masked links then fetches, in one
// batch of predicated loads, its own opposite population and the wall cell's
// opposite population of each masked link, a wall lane its own values; one
// collision serves every lane (a separate boundary path would run, and wait,
// in every wavefront that holds a row end).  Lanes whose masked links lack
// an upstream cell (NX_FLAG_OPEN) or wrap a periodic face take the general
// pull.
// With BITMAP the lane reads the container's interior bit instead of its
// flag word; only non-interior lanes then load their flags (a dependent
// load, 4 bytes per cell less traffic).
template <int S, int CM, bool FORCE, bool PER, int WG, int LP, bool BITMAP>
__global__ void __launch_bounds__(WG)
    k_fused_fix(const typename Stor<S>::T* __restrict__ fin, typename Stor<S>::T* __restrict__ fout,
                const uint32_t* __restrict__ flags, const uint32_t* __restrict__ bits, DGeom g, Par p,
                int remap, int copy_walls)
{
    using T = typename Stor<S>::T;
    using P = typename Stor<S>::P;
    using R = typename Real<S, CM>::R;
    const uint32_t wg = remap ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x;
    const uint32_t i = wg * (uint32_t)WG + threadIdx.x;
    const uint32_t ic = min(i, g.n - 1u);
    uint32_t fl = BITMAP ? bits[ic >> 5] : flags[ic];
    T v[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int off = CX[q] + CY[q] * g.X + CZ[q] * (int)g.XY;
        const uint32_t j = q == 9 ? ic : min(ic - (uint32_t)off, g.n - 1u);
        v[q] = ld_plane<LP>(fin + q * g.plane, j);
    }
#pragma unroll
    for (int q = 0; q < Q; q++) pin(v[q]);
    pin(fl);
    if (i >= g.n) return;
    if constexpr (BITMAP)
        fl = ((fl >> (i & 31u)) & 1u) ? ((uint32_t)NX_CELL_BULK << NX_CLASS_SHIFT) : flags[i];
    const bool bulk = cls_of(fl) == NX_CELL_BULK;
    const uint32_t m = bulk ? (fl & NX_MASK_BITS) : 0u;
    bool general = bulk && (fl & NX_FLAG_OPEN);
    if (PER && bulk && !general) {
        int x, y, z;
        coords(g, i, x, y, z);
        general = ((g.periodic & NX_PERIODIC_X) && (x == 0 || x == g.X - 1)) ||
                  ((g.periodic & NX_PERIODIC_Y) && (y == 0 || y == g.Y - 1)) ||
                  ((g.periodic & NX_PERIODIC_Z) && (z == 0 || z == g.Z - 1));
    }
    T a[Q], w[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) {
        if (q == 9) continue;
        const bool mq = (m >> q) & 1u;
        const int off = CX[q] + CY[q] * g.X + CZ[q] * (int)g.XY;
        const uint64_t po = (uint64_t)OPP[q] * g.plane;
        a[q] = ((mq && !general) || (!bulk && copy_walls)) ? fin[po + i] : T(0);
        w[q] = (mq && !general) ? fin[po + min(i - (uint32_t)off, g.n - 1u)] : T(0);
    }
    P f[Q];
    if (general) {
        pull<S, PER>(fin, g, i, fl, f);
    } else {
#pragma unroll
        for (int q = 0; q < Q; q++)
            f[q] = (q != 9 && ((m >> q) & 1u)) ? Stor<S>::bulk(a[q], OPP[q]) + Stor<S>::raw(w[q])
                                               : Stor<S>::bulk(v[q], q);
    }
    R out[Q];
    collide<S, CM, FORCE>(f, p.omega, p.g, out);
    if (!bulk && !copy_walls) return;
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const T o = bulk ? Stor<S>::template enc<R>(out[q], q) : (q == 9 ? v[9] : a[OPP[q]]);
        st_plane(fout + q * g.plane, i, o);
    }
}

// Edge-speculating variant.  A lane whose cell lies in the two outermost
// layers of the box (any coordinate in {0, 1, N-2, N-1}: the wall layer and
// the cells beside it) fetches, in the same batch as the 19 streaming loads,
// its flag word, its own 18 moving populations and the 18 upstream values of
// the opposite populations, so a box wall -- bounce-back links of the cell
// beside it, or the copy of a wall cell -- costs no second memory round trip.
// Every lane stores, so each wavefront writes whole cache lines (a row whose
// wall lanes skip their stores costs partial-line writes).  Other lanes read
// the interior bitmap; a non-interior lane off the box edge (an obstacle
// inside the box) fetches its flags and takes the general pull.  The
// arithmetic is that of k_fused_fix, so results agree bit for bit.
//
// xw = 1 (checked when the container is made: every x-end cell of the rows
// away from the y/z faces is a non-bulk cell holding the values cw): in
// those rows the x = 0 and x = X-1 lanes load nothing and store cw, and the
// x = 1 / x = X-2 lanes fetch only the five populations pulled from the wall
// column (c_x = +1 / -1); further links there take the general pull.
template <int S, int CM, bool FORCE, bool PER, int WG, int LP>
__global__ void __launch_bounds__(WG)
    k_fused_edge(const typename Stor<S>::T* __restrict__ fin, typename Stor<S>::T* __restrict__ fout,
                 const uint32_t* __restrict__ flags, const uint32_t* __restrict__ bits, DGeom g, Par p,
                 int remap, int copy_walls, int xw, XWall cw)
{
    using T = typename Stor<S>::T;
    using P = typename Stor<S>::P;
    using R = typename Real<S, CM>::R;
    // masks of the populations with c_x = +1 and c_x = -1
    constexpr uint32_t MP = (1u << 10) | (1u << 13) | (1u << 14) | (1u << 15) | (1u << 16);
    constexpr uint32_t MM = (1u << 0) | (1u << 3) | (1u << 4) | (1u << 5) | (1u << 6);
    const uint32_t wg = remap ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x;
    const uint32_t i = wg * (uint32_t)WG + threadIdx.x;
    const uint32_t ic = min(i, g.n - 1u);
    int x, y, z;
    coords(g, ic, x, y, z);
    const bool yz = y <= 1 || y >= g.Y - 2 || z <= 1 || z >= g.Z - 2;
    const bool edge = yz || x <= 1 || x >= g.X - 2;
    const bool xlo = x == 1;
    const bool known_wall = xw && !yz && (x == 0 || x == g.X - 1);
    const bool part = xw && !yz && (xlo || x == g.X - 2);
    const bool full = edge && !known_wall && !part;
    uint32_t word = bits[ic >> 5];
    uint32_t fl = 0u;
    T v[Q], a[Q], w[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int off = CX[q] + CY[q] * g.X + CZ[q] * (int)g.XY;
        const uint32_t j = q == 9 ? ic : min(ic - (uint32_t)off, g.n - 1u);
        v[q] = ld_plane<LP>(fin + q * g.plane, j);
        a[q] = T(0);
        w[q] = T(0);
    }
    if (edge && !known_wall) {
        fl = flags[ic];
        // a part lane needs only the populations pulled from the wall column
        // (c_x = +1 at x = 1, c_x = -1 at x = X-2): its loads for the other
        // x direction re-read its own population 9 (a line the streaming load
        // fetches anyway, values unused), and it skips the c_x = 0 ones
#pragma unroll
        for (int q = 0; q < Q; q++) {
            if (q == 9 || CX[q] == 0) continue;
            const bool need = !part || (CX[q] == 1 && xlo) || (CX[q] == -1 && !xlo);
            const int off = CX[q] + CY[q] * g.X + CZ[q] * (int)g.XY;
            const T* po = fin + (uint64_t)(need ? OPP[q] : 9) * g.plane;
            a[q] = ld_plane<LP>(po, ic);
            w[q] = ld_plane<LP>(po, need ? min(ic - (uint32_t)off, g.n - 1u) : ic);
        }
        if (!part) {
#pragma unroll
            for (int q = 0; q < Q; q++) {
                if (q == 9 || CX[q] != 0) continue;
                const int off = CX[q] + CY[q] * g.X + CZ[q] * (int)g.XY;
                const T* po = fin + (uint64_t)OPP[q] * g.plane;
                a[q] = ld_plane<LP>(po, ic);
                w[q] = ld_plane<LP>(po, min(ic - (uint32_t)off, g.n - 1u));
            }
        }
    }
#pragma unroll
    for (int q = 0; q < Q; q++) {
        pin(v[q]);
        pin(a[q]);
        pin(w[q]);
    }
    pin(word);
    pin(fl);
    if (i >= g.n) return;
    const bool interior = (word >> (i & 31u)) & 1u;
    if (interior) fl = (uint32_t)NX_CELL_BULK << NX_CLASS_SHIFT;
    else if (!edge) fl = flags[i];
    const bool bulk = cls_of(fl) == NX_CELL_BULK;
    const uint32_t m = (bulk && !interior) ? (fl & NX_MASK_BITS) : 0u;
    bool general = bulk && !interior && (!edge || (fl & NX_FLAG_OPEN) || (part && (m & ~(xlo ? MP : MM))));
    if (PER && bulk && !interior && !general)
        general = ((g.periodic & NX_PERIODIC_X) && (x == 0 || x == g.X - 1)) ||
                  ((g.periodic & NX_PERIODIC_Y) && (y == 0 || y == g.Y - 1)) ||
                  ((g.periodic & NX_PERIODIC_Z) && (z == 0 || z == g.Z - 1));
    if (!bulk && !copy_walls) return;
    if (!bulk && !full && !known_wall) {
#pragma unroll
        for (int q = 0; q < Q; q++)
            if (q != 9) a[q] = fin[(uint64_t)OPP[q] * g.plane + i];
    }
    P f[Q];
    if (general) {
        pull<S, PER>(fin, g, i, fl, f);
    } else {
#pragma unroll
        for (int q = 0; q < Q; q++)
            f[q] = (q != 9 && ((m >> q) & 1u)) ? Stor<S>::bulk(a[q], OPP[q]) + Stor<S>::raw(w[q])
                                               : Stor<S>::bulk(v[q], q);
    }
    R out[Q];
    collide<S, CM, FORCE>(f, p.omega, p.g, out);
    // one store per population for every lane (wall lanes store their copy)
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const T o = bulk ? Stor<S>::template enc<R>(out[q], q)
                         : known_wall ? from_raw<T>(cw.v[q]) : (q == 9 ? v[9] : a[OPP[q]]);
        st_plane(fout + q * g.plane, i, o);
    }
}

// Launch configuration of the container step per (storage, compute),
// process-wide.  Defaults: the fastest configuration measured on the 256^3
// cavity on MI355X -- the single-collision kernel for 8- and 4-byte storage
// (FP32 35.4k MLUPS with the interior bitmap, 35.3k without, 32.9k for the
// bitmap + full-pull kernel; FP64 17.2k), the edge-speculating kernel for
// binary16 (65.6k with uniform x-end walls, vs 51.9k for the bitmap +
// full-pull kernel); 4- and 8-byte storage keep the bitmap kernels, whose
// lower register count matters more there (FP32: 35.2k either way).
nx_tuning g_tuning[4][3] = {
    {{256, 1, 0, 0, 1, 1}, {256, 1, 0, 0, 1, 1}, {256, 1, 0, 0, 1, 1}},
    {{256, 1, 0, 0, 1, 2}, {256, 1, 0, 0, 1, 2}, {256, 1, 0, 0, 1, 2}},
    {{256, 1, 0, 0, 1, 1}, {256, 1, 0, 0, 1, 1}, {256, 1, 0, 0, 1, 1}},
    {{256, 1, 0, 0, 1, 3}, {256, 1, 0, 0, 1, 3}, {256, 1, 0, 0, 1, 3}},
};

template <int S, int CM>
constexpr bool tunable()
{
    return (S == NX_F64 && CM == NX_COMPUTE_F64) || (S != NX_F64 && CM == NX_COMPUTE_F32);
}

// cells per lane of the vector variant for a storage type: 16-byte lanes of
// 4- or 8-byte elements (always dword aligned); binary16 keeps one cell per
// lane (a shifted run of halves is only 2-byte aligned)
constexpr int vx_of(int storage) { return storage == NX_F64 ? 2 : (storage == NX_F16DEV ? 1 : 4); }

bool valid_tuning(const nx_tuning& t, int storage)
{
    return (t.wg == 256 || t.wg == 512) && (t.cpt == 1 || t.cpt == 2) &&
           (t.load_policy == 0 || t.load_policy == 1) && (t.xcd_remap == 0 || t.xcd_remap == 1) &&
           (t.vx == 1 || t.vx == vx_of(storage)) && t.boundary >= 0 && t.boundary <= 3 &&
           !(t.boundary != 0 && t.vx != 1);
}

template <int S, int CM, bool FORCE, bool PER>
void launch_fused(const nx_tuning& t, const nx_container_s& c, hipStream_t s)
{
    using T = typename Stor<S>::T;
    auto go = [&](auto wgc, auto cptc, auto lpc) {
        constexpr int WG = decltype(wgc)::value, CPT = decltype(cptc)::value,
                      LP = decltype(lpc)::value;
        const uint32_t nb = (c.g.n + WG * CPT - 1) / (WG * CPT);
        hipLaunchKernelGGL((k_fused<S, CM, FORCE, PER, WG, CPT, LP>), dim3(nb), dim3(WG), 0, s,
                           (const T*)c.fin, (T*)c.fout, c.flags, c.interior, c.g, c.p,
                           t.xcd_remap, c.copy_walls);
    };
    if (t.boundary >= 1 && t.boundary <= 3) {
        auto go_fix = [&](auto wgc, auto lpc) {
            constexpr int WG = decltype(wgc)::value, LP = decltype(lpc)::value;
            const uint32_t nb = (c.g.n + WG - 1) / WG;
            if (t.boundary == 3)
                hipLaunchKernelGGL((k_fused_edge<S, CM, FORCE, PER, WG, LP>), dim3(nb), dim3(WG), 0, s,
                                   (const T*)c.fin, (T*)c.fout, c.flags, c.interior, c.g, c.p,
                                   t.xcd_remap, c.copy_walls, c.xwall_uniform, c.xwall);
            else if (t.boundary == 2)
                hipLaunchKernelGGL((k_fused_fix<S, CM, FORCE, PER, WG, LP, true>), dim3(nb), dim3(WG), 0, s,
                                   (const T*)c.fin, (T*)c.fout, c.flags, c.interior, c.g, c.p,
                                   t.xcd_remap, c.copy_walls);
            else
                hipLaunchKernelGGL((k_fused_fix<S, CM, FORCE, PER, WG, LP, false>), dim3(nb), dim3(WG), 0, s,
                                   (const T*)c.fin, (T*)c.fout, c.flags, c.interior, c.g, c.p,
                                   t.xcd_remap, c.copy_walls);
        };
        if constexpr (tunable<S, CM>() && !FORCE && !PER) {
            auto by_lp = [&](auto wgc) {
                if (t.load_policy) go_fix(wgc, IC<1>{});
                else go_fix(wgc, IC<0>{});
            };
            if (t.wg == 512) by_lp(IC<512>{});
            else by_lp(IC<256>{});
        } else {
            go_fix(IC<256>{}, IC<0>{});
        }
        return;
    }
    constexpr int VX = vx_of(S);
    if constexpr (VX > 1) {
    if (t.vx == VX && c.g.X % VX == 0) {
        auto go_vx = [&](auto wgc, auto lpc) {
            constexpr int WG = decltype(wgc)::value, LP = decltype(lpc)::value;
            const uint32_t lanes = c.g.n / VX;
            const uint32_t nb = (lanes + WG - 1) / WG;
            hipLaunchKernelGGL((k_fused_vx<S, CM, FORCE, PER, WG, VX, LP>), dim3(nb), dim3(WG), 0, s,
                               (const T*)c.fin, (T*)c.fout, c.flags, c.interior, c.g, c.p,
                               t.xcd_remap, c.copy_walls);
        };
        if constexpr (tunable<S, CM>() && !FORCE && !PER) {
            auto by_lp = [&](auto wgc) {
                if (t.load_policy) go_vx(wgc, IC<1>{});
                else go_vx(wgc, IC<0>{});
            };
            if (t.wg == 512) by_lp(IC<512>{});
            else by_lp(IC<256>{});
        } else {
            go_vx(IC<256>{}, IC<0>{});
        }
        return;
    }
    }
    if constexpr (tunable<S, CM>() && !FORCE && !PER) {
        auto by_lp = [&](auto wgc, auto cptc) {
            if (t.load_policy) go(wgc, cptc, IC<1>{});
            else go(wgc, cptc, IC<0>{});
        };
        auto by_cpt = [&](auto wgc) {
            if (t.cpt == 2) by_lp(wgc, IC<2>{});
            else by_lp(wgc, IC<1>{});
        };
        if (t.wg == 512) by_cpt(IC<512>{});
        else by_cpt(IC<256>{});
    } else {
        go(IC<256>{}, IC<1>{}, IC<0>{});
    }
}

int run_container(const nx_container_s& c, const nx_tuning& t, hipStream_t s)
{
    if (int e = by_type(&c.d, [&](auto sc, auto cc) {
            constexpr int S = decltype(sc)::value;
            constexpr int CM = decltype(cc)::value;
            const bool per = c.g.periodic != 0;
            if (c.force) {
                if (per) launch_fused<S, CM, true, true>(t, c, s);
                else launch_fused<S, CM, true, false>(t, c, s);
            } else {
                if (per) launch_fused<S, CM, false, true>(t, c, s);
                else launch_fused<S, CM, false, false>(t, c, s);
            }
        }))
        return e;
    NX_CHECK_LAUNCH("k_fused");
    return NX_OK;
}

}  // namespace
}  // namespace dense
}  // namespace nx

using namespace nx;
using namespace nx::dense;

extern "C" int nx_d3q19_iteration_dense(const nx_dense_desc* d, const void* fin, void* fout,
                                        const uint32_t* flags, const nx_lbm_params* p,
                                        nx_container* out)
{
    if (int e = check(d)) return e;
    if (!fin || !fout || !flags || !p || !out) return set_error(NX_ERR_INVALID, "null pointer");
    if ((uint64_t)d->dim[0] * d->dim[1] * d->dim[2] * 8 >= (1ull << 32))
        return set_error(NX_ERR_UNSUPPORTED, "container: population planes must be < 4 GiB");
    if (fin == fout)
        return set_error(NX_ERR_INVALID, "fin and fout must differ (two-population scheme)");
    const bool force = p->force[0] != 0. || p->force[1] != 0. || p->force[2] != 0.;
    if (force && d->compute == NX_COMPUTE_REF)
        return set_error(NX_ERR_UNSUPPORTED,
                         "body force requires NX_COMPUTE_F32 or NX_COMPUTE_F64");
    auto* c = new nx_container_s;
    c->d = *d;
    c->g = make_geom(d);
    c->fin = fin;
    c->fout = fout;
    c->flags = flags;
    c->p = Par{p->omega, {p->force[0], p->force[1], p->force[2]}};
    c->force = force;
    const uint32_t nwg_c = (c->g.n + kCWG - 1) / kCWG;
    const size_t words = ((size_t)c->g.n + 31) / 32 + 2;
    uint32_t* d_cnt = nullptr;
    uint64_t* d_icnt = nullptr;
    auto fail = [&](int code) {
        if (d_cnt) hipFree(d_cnt);
        if (d_icnt) hipFree(d_icnt);
        if (c->interior) hipFree(c->interior);
        delete c;
        return code;
    };
    hipError_t e;
    if ((e = hipMalloc(&c->interior, words * 4)) != hipSuccess) return fail(hip_error(e, "hipMalloc"));
    if ((e = hipMemset(c->interior, 0, words * 4)) != hipSuccess) return fail(hip_error(e, "hipMemset"));
    if ((e = hipMalloc(&d_cnt, (size_t)nwg_c * 4)) != hipSuccess) return fail(hip_error(e, "hipMalloc"));
    if ((e = hipMalloc(&d_icnt, (size_t)nwg_c * 8)) != hipSuccess) return fail(hip_error(e, "hipMalloc"));
    hipLaunchKernelGGL(k_classify, dim3(nwg_c), dim3(kCWG), 0, 0, flags, c->g, c->interior, d_cnt,
                       d_icnt);
    if ((e = hipGetLastError()) != hipSuccess) return fail(hip_error(e, "k_classify"));
    std::vector<uint32_t> cnt(nwg_c);
    std::vector<uint64_t> icnt(nwg_c);
    if ((e = hipMemcpy(cnt.data(), d_cnt, (size_t)nwg_c * 4, hipMemcpyDeviceToHost)) != hipSuccess)
        return fail(hip_error(e, "hipMemcpy"));
    if ((e = hipMemcpy(icnt.data(), d_icnt, (size_t)nwg_c * 8, hipMemcpyDeviceToHost)) != hipSuccess)
        return fail(hip_error(e, "hipMemcpy"));
    uint64_t total = 0, itotal = 0;
    bool zouhe = false;
    for (uint32_t k = 0; k < nwg_c; k++) {
        total += cnt[k] & 0x7FFFFFFFu;
        zouhe = zouhe || (cnt[k] & 0x80000000u);
        itotal += icnt[k];
    }
    if (zouhe)
        return fail(set_error(NX_ERR_UNSUPPORTED,
                              "container: the flags hold Zou-He cells; step them with "
                              "nx_d3q19_step_dense_zouhe"));
    c->n_boundary = total;
    c->n_interior = itotal;
    // do both fields hold the same wall values?
    if ((e = hipMemset(d_cnt, 0, (size_t)nwg_c * 4)) != hipSuccess) return fail(hip_error(e, "hipMemset"));
    by_type(d, [&](auto sc, auto) {
        constexpr int S = decltype(sc)::value;
        using T = typename Stor<S>::T;
        hipLaunchKernelGGL((k_walls_equal<T>), dim3(nwg_c), dim3(kCWG), 0, 0, (const T*)fin,
                           (const T*)fout, flags, c->g, d_cnt);
    });
    if ((e = hipGetLastError()) != hipSuccess) return fail(hip_error(e, "k_walls_equal"));
    if ((e = hipMemcpy(cnt.data(), d_cnt, (size_t)nwg_c * 4, hipMemcpyDeviceToHost)) != hipSuccess)
        return fail(hip_error(e, "hipMemcpy"));
    c->copy_walls = 1;
    for (uint32_t k = 0; k < nwg_c; k++)
        if (cnt[k]) c->copy_walls = 0;
    // uniform x-end walls (the edge-speculating kernel's short path)
    if (c->copy_walls && d->dim[0] >= 4 && d->dim[1] >= 5 && d->dim[2] >= 5) {
        const uint32_t items = 2u * (d->dim[1] - 4) * (d->dim[2] - 4);
        // d_icnt (8 bytes per classification workgroup) is free again: the
        // mismatch word and, when it holds >= 19 words, the reference values
        uint64_t* d_ref = nullptr;
        if (nwg_c >= Q) d_ref = d_icnt;
        else if ((e = hipMalloc(&d_ref, Q * sizeof(uint64_t))) != hipSuccess)
            return fail(hip_error(e, "hipMalloc"));
        auto done = [&](int code) {
            if (d_ref != d_icnt) hipFree(d_ref);
            return code == NX_OK ? code : fail(code);
        };
        if ((e = hipMemset(d_cnt, 0, 4)) != hipSuccess) return done(hip_error(e, "hipMemset"));
        by_type(d, [&](auto sc, auto) {
            constexpr int S = decltype(sc)::value;
            using T = typename Stor<S>::T;
            hipLaunchKernelGGL((k_xwall_check<T>), dim3((items + kCWG - 1) / kCWG), dim3(kCWG), 0, 0,
                               (const T*)fin, flags, c->g, d_cnt, d_ref);
        });
        if ((e = hipGetLastError()) != hipSuccess) return done(hip_error(e, "k_xwall_check"));
        uint32_t bad = 1;
        if ((e = hipMemcpy(&bad, d_cnt, 4, hipMemcpyDeviceToHost)) != hipSuccess)
            return done(hip_error(e, "hipMemcpy"));
        if ((e = hipMemcpy(c->xwall.v, d_ref, sizeof(c->xwall.v), hipMemcpyDeviceToHost)) != hipSuccess)
            return done(hip_error(e, "hipMemcpy"));
        c->xwall_uniform = bad ? 0 : 1;
        done(NX_OK);
    }
    hipFree(d_cnt);
    hipFree(d_icnt);
    *out = c;
    return NX_OK;
}

extern "C" int nx_container_run(nx_container c, nx_stream s)
{
    if (!c) return set_error(NX_ERR_INVALID, "null container");
    return run_container(*c, g_tuning[c->d.storage][c->d.compute], (hipStream_t)s);
}

extern "C" int nx_container_info(nx_container c, uint64_t* n_interior, uint64_t* n_boundary,
                                 int* copies_walls)
{
    if (!c) return set_error(NX_ERR_INVALID, "null container");
    if (n_interior) *n_interior = c->n_interior;
    if (n_boundary) *n_boundary = c->n_boundary;
    if (copies_walls) *copies_walls = c->copy_walls;
    return NX_OK;
}

extern "C" int nx_container_destroy(nx_container c)
{
    if (!c) return NX_OK;
    if (c->pair_graph) hipGraphExecDestroy(c->pair_graph);
    if (c->interior) hipFree(c->interior);
    delete c;
    return NX_OK;
}

// Ping-pong steps a, b, a, b, ... (LbmIterationD3Q19::run called n times,
// LbmIteration.h:56-69).  On a created stream, runs of 2 * kPairSteps steps
// replay one hipGraph (built at the first call for this pair and launch
// configuration, kept in `a`): one launch instead of 64 kernel launches.
// The default stream cannot be captured: there every step is launched.
namespace {
constexpr int kPairSteps = 32;
}

extern "C" int nx_container_run_pair(nx_container a, nx_container b, int32_t n_steps, nx_stream s)
{
    if (!a || !b) return set_error(NX_ERR_INVALID, "null container");
    if (n_steps < 0) return set_error(NX_ERR_INVALID, "n_steps must be >= 0");
    if (a->d.storage != b->d.storage || a->d.compute != b->d.compute || a->fin != b->fout ||
        a->fout != b->fin)
        return set_error(NX_ERR_INVALID, "run_pair: b must step b's input back into a's input");
    const hipStream_t st = (hipStream_t)s;
    const nx_tuning& t = g_tuning[a->d.storage][a->d.compute];
    int32_t done = 0;
    if (st != nullptr && n_steps >= 2 * kPairSteps) {
        const bool same = a->pair_graph && a->pair_with == b &&
                          !memcmp(&a->pair_tuning[0], &t, sizeof(t));
        if (!same) {
            if (a->pair_graph) hipGraphExecDestroy(a->pair_graph);
            a->pair_graph = nullptr;
            hipGraph_t g;
            NX_CHECK_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            int e = 0;
            for (int k = 0; k < kPairSteps && !e; k++) {
                e = run_container(*a, t, st);
                if (!e) e = run_container(*b, t, st);
            }
            const hipError_t ec = hipStreamEndCapture(st, &g);
            if (e) {
                if (ec == hipSuccess) hipGraphDestroy(g);
                return e;
            }
            NX_CHECK_HIP(ec);
            const hipError_t ei = hipGraphInstantiate(&a->pair_graph, g, nullptr, nullptr, 0);
            hipGraphDestroy(g);
            NX_CHECK_HIP(ei);
            a->pair_with = b;
            a->pair_tuning[0] = t;
        }
        for (; done + 2 * kPairSteps <= n_steps; done += 2 * kPairSteps)
            NX_CHECK_HIP(hipGraphLaunch(a->pair_graph, st));
    }
    for (; done < n_steps; done++)
        if (int e = run_container(done % 2 == 0 ? *a : *b, t, st)) return e;
    return NX_OK;
}

extern "C" int nx_d3q19_get_tuning_dense(const nx_dense_desc* d, nx_tuning* t)
{
    if (int e = check(d)) return e;
    if (!t) return set_error(NX_ERR_INVALID, "null pointer");
    *t = g_tuning[d->storage][d->compute];
    return NX_OK;
}

extern "C" int nx_d3q19_set_tuning_dense(const nx_dense_desc* d, const nx_tuning* t)
{
    if (int e = check(d)) return e;
    if (!t || !valid_tuning(*t, d->storage))
        return set_error(NX_ERR_INVALID,
                         "tuning: wg in {256,512}, cpt in {1,2}, load_policy and xcd_remap in "
                         "{0,1}, vx = 1 or 16 bytes of cells (2 f64, 4 f32), boundary in {0,1,2,3} "
                         "(1 and 2 need vx = 1)");
    g_tuning[d->storage][d->compute] = *t;
    return NX_OK;
}

// Times every launch configuration of the container's step (fout is
// overwritten, fin only read) and keeps the fastest for its (storage, compute).
extern "C" int nx_container_autotune(nx_container c, nx_stream s, nx_tuning* best, float* best_ms)
{
    if (!c) return set_error(NX_ERR_INVALID, "null container");
    nx_tuning& slot = g_tuning[c->d.storage][c->d.compute];
    hipEvent_t e0, e1;
    NX_CHECK_HIP(hipEventCreate(&e0));
    NX_CHECK_HIP(hipEventCreate(&e1));
    nx_tuning bt = slot;
    float bms = 1e30f;
    const int reps = 10;
    for (int kv = 0; kv < 5; kv++)
    if (kv != 1 || vx_of(c->d.storage) > 1)
    for (int wg : {256, 512})
        for (int cpt : {1, 2})
            for (int lp : {0, 1})
                for (int rm : {0, 1}) {
                    const int vx = kv == 1 ? vx_of(c->d.storage) : 1;
                    const int bd = kv >= 2 ? kv - 1 : 0;
                    if ((vx > 1 || bd) && cpt == 2) continue;  // one cell (or lane of VX cells) per thread
                    const nx_tuning t{wg, cpt, lp, rm, vx, bd};
                    for (int r = 0; r < 3; r++)
                        if (int e = run_container(*c, t, (hipStream_t)s)) return e;
                    NX_CHECK_HIP(hipEventRecord(e0, (hipStream_t)s));
                    for (int r = 0; r < reps; r++)
                        if (int e = run_container(*c, t, (hipStream_t)s)) return e;
                    NX_CHECK_HIP(hipEventRecord(e1, (hipStream_t)s));
                    NX_CHECK_HIP(hipEventSynchronize(e1));
                    float ms = 0;
                    NX_CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
                    ms /= reps;
                    if (ms < bms) {
                        bms = ms;
                        bt = t;
                    }
                }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    slot = bt;
    if (best) *best = bt;
    if (best_ms) *best_ms = bms;
    return NX_OK;
}
