// d3q19_batch.hip -- many independent dense cavities stepped together.
//
// No counterpart in the reference (its driver runs one grid,
// RunCavityTwoPop.cu:17-276); each member is exactly a dGrid partition with
// the reference's step, so a member's populations equal a standalone
// nx_d3q19_step_dense run bit for bit.
//
// Memory: one arena.  Member k owns cells [off_k, off_k + n_k) of the flag
// array and the 19 population planes of n_k cells starting at element
// 19*off_k of each of the two population arenas (the dGrid SoA layout of that
// member).  Work: the members are cut into tiles of WG consecutive cells; a
// persistent kernel (a fixed grid of workgroups) walks the tile table and
// skips tiles of converged members.
//
// Convergence: every K steps, r_k = sqrt(sum |u - u_prev|^2 / sum |u|^2) over
// the member's bulk cells (post-stream moments, FP64 sums: per-tile partials,
// then one workgroup per member in tile order -- deterministic), u_prev <- u;
// a member with r_k < tol stops stepping.  K steps + the residual + the
// convergence update are captured once in a hipGraph and replayed.
#include <algorithm>
#include <cmath>
#include <vector>

#include "dense_common.hpp"

namespace nx {
namespace batch {

using dense::DGeom;
using dense::Par;

struct Member {
    uint64_t pop_base;   // element offset of population plane 0 in an arena
    uint32_t cell_base;  // offset of the member's first cell (flags, u_prev)
    uint32_t tile_first;
    uint32_t n_tiles;
    int32_t X, Y, Z;
    uint32_t XY, n;
    double omega;
};

}  // namespace batch
}  // namespace nx

struct nx_batch_s {
    int32_t storage = NX_F32;
    uint32_t n_members = 0;
    uint32_t n_tiles = 0;
    uint64_t n_cells = 0;
    std::vector<nx::batch::Member> members;
    std::vector<uint32_t> tile_member;
    // device
    void* pop[2] = {nullptr, nullptr};
    uint32_t* flags = nullptr;
    nx::batch::Member* d_members = nullptr;
    uint32_t* d_tile_member = nullptr;
    uint32_t* active = nullptr;      // 1 while the member is stepping
    uint32_t* steps_done = nullptr;  // per member
    uint32_t* n_active = nullptr;    // device counter written by the update
    double* partials = nullptr;      // 2 per tile
    double* res = nullptr;           // 2 per member
    float* u_prev = nullptr;         // 3 per cell
    // launch configuration
    int wg = 256;
    int wgs_per_cu = 0;   // one workgroup per tile (the autotuner's choice for FP32, below)
    int xcd_chunks = 0;
    int edge = 0;         // edge-speculating step (set per storage type, autotuned)
    int n_cu = 256;
    // captured graph
    hipGraphExec_t graph = nullptr;
    int graph_k = 0;
    double graph_tol = -1;
    hipStream_t graph_stream = nullptr;
};

namespace nx {
namespace batch {
namespace {

constexpr int kTile = 256;

// EDGE: a lane in the two outer layers of its cavity (the wall layer and the
// cells beside it) fetches its own 18 moving populations and the 18 upstream
// values of the opposite populations with the streaming loads, so the cavity
// walls cost no second memory round trip; cells beside the boxes inside the
// cavity still fetch what they lack after their flag word arrives.
template <int S, bool EDGE>
__device__ __forceinline__ void step_cell(const typename Stor<S>::T* __restrict__ fin,
                                          typename Stor<S>::T* __restrict__ fout,
                                          const uint32_t* __restrict__ flags, const DGeom& g,
                                          const Member& M, uint32_t i)
{
    using T = typename Stor<S>::T;
    using P = typename Stor<S>::P;
    using R = typename Real<S, NX_COMPUTE_F32>::R;
    const uint32_t ic = min(i, g.n - 1u);
    uint32_t fl = flags[ic];
    T v[Q], a[Q], w[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int off = CX[q] + CY[q] * g.X + CZ[q] * (int)g.XY;
        const uint32_t j = q == 9 ? ic : min(ic - (uint32_t)off, g.n - 1u);
        v[q] = fin[q * g.plane + j];
        a[q] = T(0);
        w[q] = T(0);
    }
    bool edge = false;
    if constexpr (EDGE) {
        int x, y, z;
        dense::coords(g, ic, x, y, z);
        edge = x <= 1 || x >= g.X - 2 || y <= 1 || y >= g.Y - 2 || z <= 1 || z >= g.Z - 2;
        if (edge) {
#pragma unroll
            for (int q = 0; q < Q; q++) {
                if (q == 9) continue;
                const int off = CX[q] + CY[q] * g.X + CZ[q] * (int)g.XY;
                const uint64_t po = (uint64_t)OPP[q] * g.plane;
                a[q] = fin[po + ic];
                w[q] = fin[po + min(ic - (uint32_t)off, g.n - 1u)];
            }
        }
#pragma unroll
        for (int q = 0; q < Q; q++) {
            dense::pin(a[q]);
            dense::pin(w[q]);
        }
    }
#pragma unroll
    for (int q = 0; q < Q; q++) dense::pin(v[q]);
    dense::pin(fl);
    if (i >= g.n) return;
    const double g0[3] = {0, 0, 0};
    // Every lane runs one collision: in a cavity of 32..96 cells per row each
    // wavefront holds wall and wall-adjacent lanes, so a separate boundary
    // path would run (and wait) in every wavefront.  One batch of predicated
    // loads fetches what the speculative loads lack:
    //   masked link q of a bulk cell:  a[q] = own f_opp(q), w[q] = f_opp(q) of
    //                                  the wall cell at x - c_q
    //   wall cell:                     a[q] = own f_opp(q) for every q (copied)
    // Only a masked link without upstream cell (NX_FLAG_OPEN) takes the
    // general pull.
    const bool bulk = cls_of(fl) == NX_CELL_BULK;
    const uint32_t m = bulk ? (fl & NX_MASK_BITS) : 0u;
    const bool open = bulk && (fl & NX_FLAG_OPEN);
    if (!edge) {
#pragma unroll
        for (int q = 0; q < Q; q++) {
            if (q == 9) continue;
            const bool mq = (m >> q) & 1u;
            const int off = CX[q] + CY[q] * g.X + CZ[q] * (int)g.XY;
            const uint64_t po = (uint64_t)OPP[q] * g.plane;
            a[q] = (mq || !bulk) ? fin[po + i] : T(0);
            w[q] = (mq && !open) ? fin[po + min(i - (uint32_t)off, g.n - 1u)] : T(0);
        }
    }
    P f[Q];
    if (open) {
        dense::pull<S, false>(fin, g, i, fl, f);
    } else {
#pragma unroll
        for (int q = 0; q < Q; q++)
            f[q] = (q != 9 && ((m >> q) & 1u)) ? Stor<S>::bulk(a[q], OPP[q]) + Stor<S>::raw(w[q])
                                               : Stor<S>::bulk(v[q], q);
    }
    R out[Q];
    collide<S, (S == NX_F64 ? NX_COMPUTE_F64 : NX_COMPUTE_F32), false>(f, M.omega, g0, out);
    // wall cells keep their values (the batch fills both arenas alike)
#pragma unroll
    for (int q = 0; q < Q; q++)
        v[q] = bulk ? Stor<S>::template enc<R>(out[q], q) : (q == 9 ? v[9] : a[OPP[q]]);
#pragma unroll
    for (int q = 0; q < Q; q++) __builtin_nontemporal_store(v[q], &fout[q * g.plane + i]);
}

__device__ __forceinline__ DGeom geom_of(const Member& M)
{
    DGeom g;
    g.X = M.X;
    g.Y = M.Y;
    g.Z = M.Z;
    g.XY = M.XY;
    g.n = M.n;
    g.plane = M.n;
    g.periodic = 0;
    return g;
}

constexpr uint32_t kChunkCells = 65536;  // cells per XCD chunk of tiles

// Persistent step: a fixed grid walks the tile table.  With xcd_chunks the
// tiles are dealt to the 8 XCDs in chunks of kChunkCells consecutive cells
// (chunk c -> XCD c % 8; workgroup w runs on XCD w % 8, the hardware's
// round-robin dispatch) and the workgroups of one XCD walk their chunks in
// order: the tiles one L2 serves at a time are contiguous in the arena, so
// the z-neighbour planes of a member (at most 36 tiles away for 96^3) are
// read through the same L2.  Dealt tile by tile instead, a tile and its z
// neighbours sit on different XCDs and every plane is fetched ~3 times.
template <int S, int WG, bool EDGE>
__global__ void __launch_bounds__(WG)
    k_batch_step(const typename Stor<S>::T* __restrict__ fin, typename Stor<S>::T* __restrict__ fout,
                 const uint32_t* __restrict__ flags, const Member* __restrict__ members,
                 const uint32_t* __restrict__ tile_member, const uint32_t* __restrict__ active,
                 uint32_t n_tiles, int xcd_chunks)
{
    auto tile = [&](uint32_t t) {
        const uint32_t m = tile_member[t];
        if (!active[m]) return;
        const Member M = members[m];
        const DGeom g = geom_of(M);
        const uint32_t i = (t - M.tile_first) * (uint32_t)WG + threadIdx.x;
        step_cell<S, EDGE>(fin + M.pop_base, fout + M.pop_base, flags + M.cell_base, g, M, i);
    };
    if (xcd_chunks == 2) {
        // one workgroup per tile: each XCD takes one contiguous range of
        // tiles in order (the dense step's remap), so a tile's z neighbours
        // are read through the L2 that fetched them
        const uint32_t t = xcd_remap(blockIdx.x, gridDim.x);
        if (t < n_tiles) tile(t);
    } else if (xcd_chunks) {
        constexpr uint32_t C = kChunkCells / (uint32_t)WG;  // tiles per chunk
        const uint32_t xcd = blockIdx.x & 7u, G8 = gridDim.x >> 3;
        for (uint32_t k = blockIdx.x >> 3;; k += G8) {
            const uint32_t c = xcd + 8u * (k / C);
            if (c * C >= n_tiles) break;  // this XCD's chunks are done
            const uint32_t t = c * C + k % C;
            if (t < n_tiles) tile(t);
        }
    } else {
        for (uint32_t t = blockIdx.x; t < n_tiles; t += gridDim.x) tile(t);
    }
}

// Residual partials of one tile (0 for converged members).
template <int S>
__global__ void __launch_bounds__(kTile)
    k_batch_residual(const typename Stor<S>::T* __restrict__ fin, const uint32_t* __restrict__ flags,
                     const Member* __restrict__ members, const uint32_t* __restrict__ tile_member,
                     const uint32_t* __restrict__ active, float* __restrict__ u_prev,
                     double* __restrict__ partials, uint32_t n_tiles, int tile)
{
    __shared__ double lds[32];
    for (uint32_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        const uint32_t m = tile_member[t];
        double acc[2] = {0.0, 0.0};
        if (active[m]) {
            const Member M = members[m];
            const DGeom g = geom_of(M);
            const uint32_t i0 = (t - M.tile_first) * (uint32_t)tile;
            for (int c = threadIdx.x; c < tile; c += kTile) {
                const uint32_t i = i0 + (uint32_t)c;
                if (i >= g.n) break;
                const uint32_t fl = flags[M.cell_base + i];
                if (cls_of(fl) != NX_CELL_BULK) continue;
                typename Stor<S>::P f[Q];
                dense::pull<S, false>(fin + M.pop_base, g, i, fl, f);
                typename Real<S, NX_COMPUTE_F32>::R r, uu[3];
                cell_moments<S, (S == NX_F64 ? NX_COMPUTE_F64 : NX_COMPUTE_F32)>(NX_CELL_BULK, f, r, uu);
                float* up = u_prev + 3ull * M.cell_base;
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    const float un = (float)uu[a];
                    const double du = (double)un - (double)up[(uint64_t)a * g.n + i];
                    acc[0] += du * du;
                    acc[1] += (double)un * (double)un;
                    up[(uint64_t)a * g.n + i] = un;
                }
            }
        }
        block_sum_f64<2>(acc, lds);
        if (threadIdx.x == 0) {
            partials[2 * t] = acc[0];
            partials[2 * t + 1] = acc[1];
        }
        __syncthreads();
    }
}

// One workgroup per member: sum its tiles' partials in order, then retire it
// when converged.  steps_done advances by the K steps just taken.
__global__ void __launch_bounds__(256)
    k_batch_update(const Member* __restrict__ members, const double* __restrict__ partials,
                   double* __restrict__ res, uint32_t* __restrict__ active,
                   uint32_t* __restrict__ steps_done, int k_steps, double tol2)
{
    __shared__ double lds[32];
    const uint32_t m = blockIdx.x;
    const Member M = members[m];
    double acc[2] = {0.0, 0.0};
    for (uint32_t t = threadIdx.x; t < M.n_tiles; t += 256) {
        acc[0] += partials[2 * (M.tile_first + t)];
        acc[1] += partials[2 * (M.tile_first + t) + 1];
    }
    block_sum_f64<2>(acc, lds);
    if (threadIdx.x == 0 && active[m]) {
        res[2 * m] = acc[0];
        res[2 * m + 1] = acc[1];
        steps_done[m] += (uint32_t)k_steps;
        // converged, or diverged (a non-finite sum: the member's flow blew up)
        const bool finite = acc[0] - acc[0] == 0.0 && acc[1] - acc[1] == 0.0;
        if (!finite || acc[0] < tol2 * acc[1]) active[m] = 0;
    }
}

// Number of members still stepping (one workgroup).
__global__ void __launch_bounds__(256) k_count_active(const uint32_t* active, uint32_t n, uint32_t* out)
{
    __shared__ uint32_t s[256];
    uint32_t c = 0;
    for (uint32_t m = threadIdx.x; m < n; m += 256) c += active[m];
    s[threadIdx.x] = c;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = s[0];
}

// Populations of bulk cells from (rho, u): the reference's equilibrium
// expression (LbmTools.h:216-234) in binary64; u_prev <- u.
template <int S>
__global__ void __launch_bounds__(kTile)
    k_batch_init(typename Stor<S>::T* __restrict__ p0, typename Stor<S>::T* __restrict__ p1,
                 const uint32_t* __restrict__ flags, const Member* __restrict__ members,
                 const uint32_t* __restrict__ tile_member, const float* __restrict__ rho_in,
                 const float* __restrict__ u_in, float* __restrict__ u_prev, uint32_t n_tiles,
                 int tile)
{
    for (uint32_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        const Member M = members[tile_member[t]];
        for (int cc = threadIdx.x; cc < tile; cc += kTile) {
            const uint32_t i = (t - M.tile_first) * (uint32_t)tile + (uint32_t)cc;
            if (i >= M.n) break;
            const uint64_t c = M.cell_base + (uint64_t)i;
            float* up = u_prev + 3ull * M.cell_base;
            if (cls_of(flags[c]) != NX_CELL_BULK) {
                for (int a = 0; a < 3; a++) up[(uint64_t)a * M.n + i] = 0.f;
                continue;
            }
            const double rho = rho_in ? (double)rho_in[c] : 1.0;
            double u[3];
            for (int a = 0; a < 3; a++) {
                u[a] = u_in ? (double)u_in[3ull * M.cell_base + (uint64_t)a * M.n + i] : 0.0;
                up[(uint64_t)a * M.n + i] = (float)u[a];
            }
            const double usqr = 1.5 * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
            for (int q = 0; q < Q; q++) {
                const double cu = CX[q] * u[0] + CY[q] * u[1] + CZ[q] * u[2];
                const double feq = rho * wq(q) * (1. + 3. * cu + 4.5 * cu * cu - usqr);
                const auto v = Stor<S>::enc_full(feq, q);
                p0[M.pop_base + (uint64_t)q * M.n + i] = v;
                p1[M.pop_base + (uint64_t)q * M.n + i] = v;
            }
        }
    }
}

// Consistent initialisation after an equilibrium start (after Mei, Luo,
// Lallemand & d'Humieres 2006): the equilibrium part of every bulk cell is
// moved to the prescribed velocity u*, keeping the cell's density and its
// non-equilibrium part,  f <- f - feq(rho, u(f)) + feq(rho, u*), in binary64
// (the difference is the same for f and for d = f - w_q).
template <int S>
__global__ void __launch_bounds__(kTile)
    k_batch_constrain(typename Stor<S>::T* __restrict__ p, const uint32_t* __restrict__ flags,
                      const Member* __restrict__ members, const uint32_t* __restrict__ tile_member,
                      const uint32_t* __restrict__ active, const float* __restrict__ u_in,
                      uint32_t n_tiles, int tile)
{
    for (uint32_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        const uint32_t m = tile_member[t];
        if (!active[m]) continue;
        const Member M = members[m];
        for (int cc = threadIdx.x; cc < tile; cc += kTile) {
            const uint32_t i = (t - M.tile_first) * (uint32_t)tile + (uint32_t)cc;
            if (i >= M.n) break;
            const uint64_t c = M.cell_base + (uint64_t)i;
            if (cls_of(flags[c]) != NX_CELL_BULK) continue;
            double f[Q], rho = 0.0, j[3] = {0.0, 0.0, 0.0};
            for (int q = 0; q < Q; q++) {
                f[q] = (double)Stor<S>::bulk(p[M.pop_base + (uint64_t)q * M.n + i], q);
                rho += f[q];
                j[0] += CX[q] * f[q];
                j[1] += CY[q] * f[q];
                j[2] += CZ[q] * f[q];
            }
            if constexpr (Stor<S>::kDev) rho += 1.0;
            double u[3], us[3];
            for (int a = 0; a < 3; a++) {
                u[a] = j[a] / rho;
                us[a] = (double)u_in[3ull * M.cell_base + (uint64_t)a * M.n + i];
            }
            const double uu = 1.5 * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
            const double uus = 1.5 * (us[0] * us[0] + us[1] * us[1] + us[2] * us[2]);
            for (int q = 0; q < Q; q++) {
                const double cu = CX[q] * u[0] + CY[q] * u[1] + CZ[q] * u[2];
                const double cus = CX[q] * us[0] + CY[q] * us[1] + CZ[q] * us[2];
                const double d = rho * wq(q) * (3. * (cus - cu) + 4.5 * (cus * cus - cu * cu) - (uus - uu));
                p[M.pop_base + (uint64_t)q * M.n + i] = Stor<S>::template enc<double>(f[q] + d, q);
            }
        }
    }
}

template <int V>
using IC = std::integral_constant<int, V>;

template <typename F>
int by_storage(int storage, F&& f)
{
    switch (storage) {
        case NX_F64: f(IC<NX_F64>{}); return NX_OK;
        case NX_F32: f(IC<NX_F32>{}); return NX_OK;
        case NX_F32DEV: f(IC<NX_F32DEV>{}); return NX_OK;
        case NX_F16DEV: f(IC<NX_F16DEV>{}); return NX_OK;
    }
    return set_error(NX_ERR_INVALID, "unknown storage type");
}

size_t elem_size(int storage)
{
    return storage == NX_F64 ? 8 : (storage == NX_F16DEV ? 2 : 4);
}

// Persistent grid: wgs_per_cu workgroups per CU, at most one per tile;
// wgs_per_cu == 0 asks for one workgroup per tile.  The product is taken in
// 64 bits so that no count of CUs and workgroups can wrap.
unsigned grid_size(uint32_t n_tiles, int n_cu, int wgs_per_cu)
{
    if (wgs_per_cu <= 0) return std::max(1u, n_tiles);
    const uint64_t cap = (uint64_t)std::max(n_cu, 1) * (uint64_t)wgs_per_cu;
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(n_tiles, cap));
}

unsigned grid_of(const nx_batch_s& b)
{
    return grid_size(b.n_tiles, b.n_cu, b.wgs_per_cu);
}

// How tiles reach the XCDs: 0 dealt round robin (grid-stride), 1 persistent
// grid in XCD chunks (needs a multiple of the 8 XCDs), 2 one workgroup per
// tile remapped so that each XCD takes a contiguous range of tiles.
int chunked(const nx_batch_s& b)
{
    const unsigned G = grid_of(b);
    if (!b.xcd_chunks) return 0;
    if (G >= b.n_tiles) return 2;
    return G % 8 == 0 ? 1 : 0;
}

int launch_steps(nx_batch_s& b, int k_steps, int parity0, hipStream_t s)
{
    for (int k = 0; k < k_steps; k++) {
        const int par = (parity0 + k) & 1;
        const void* fin = b.pop[par];
        void* fout = b.pop[par ^ 1];
        if (int e = by_storage(b.storage, [&](auto sc) {
                constexpr int S = decltype(sc)::value;
                using T = typename Stor<S>::T;
                auto go = [&](auto wgc, auto edgec) {
                    constexpr int WG = decltype(wgc)::value;
                    constexpr bool E = decltype(edgec)::value;
                    hipLaunchKernelGGL((k_batch_step<S, WG, E>), dim3(grid_of(b)), dim3(WG), 0, s,
                                       (const T*)fin, (T*)fout, b.flags, b.d_members,
                                       b.d_tile_member, b.active, b.n_tiles, chunked(b));
                };
                if (b.wg == 512) {
                    if (b.edge) go(std::integral_constant<int, 512>{}, std::true_type{});
                    else go(std::integral_constant<int, 512>{}, std::false_type{});
                } else {
                    if (b.edge) go(std::integral_constant<int, 256>{}, std::true_type{});
                    else go(std::integral_constant<int, 256>{}, std::false_type{});
                }
            }))
            return e;
    }
    NX_CHECK_LAUNCH("k_batch_step");
    return NX_OK;
}

int launch_check(nx_batch_s& b, int k_steps, double tol, hipStream_t s)
{
    if (int e = by_storage(b.storage, [&](auto sc) {
            constexpr int S = decltype(sc)::value;
            using T = typename Stor<S>::T;
            hipLaunchKernelGGL((k_batch_residual<S>), dim3(std::min<uint32_t>(b.n_tiles, 4096)),
                               dim3(kTile), 0, s, (const T*)b.pop[0], b.flags, b.d_members,
                               b.d_tile_member, b.active, b.u_prev, b.partials, b.n_tiles,
                               b.wg);
        }))
        return e;
    hipLaunchKernelGGL(k_batch_update, dim3(b.n_members), dim3(256), 0, s, b.d_members,
                       b.partials, b.res, b.active, b.steps_done, k_steps, tol * tol);
    hipLaunchKernelGGL(k_count_active, dim3(1), dim3(256), 0, s, b.active, b.n_members, b.n_active);
    NX_CHECK_LAUNCH("k_batch_residual/update");
    return NX_OK;
}

// Re-tile the members for a workgroup size (tiles hold wg cells).
void build_tiles(nx_batch_s& b, int wg)
{
    b.tile_member.clear();
    for (uint32_t k = 0; k < b.n_members; k++) {
        auto& M = b.members[k];
        M.tile_first = (uint32_t)b.tile_member.size();
        M.n_tiles = (M.n + (uint32_t)wg - 1) / (uint32_t)wg;
        for (uint32_t t = 0; t < M.n_tiles; t++) b.tile_member.push_back(k);
    }
    b.n_tiles = (uint32_t)b.tile_member.size();
}

int upload_tiles(nx_batch_s& b)
{
    if (b.d_tile_member) hipFree(b.d_tile_member);
    if (b.partials) hipFree(b.partials);
    b.d_tile_member = nullptr;
    b.partials = nullptr;
    NX_CHECK_HIP(hipMalloc(&b.d_tile_member, (size_t)b.n_tiles * 4));
    NX_CHECK_HIP(hipMalloc(&b.partials, (size_t)b.n_tiles * 16));
    NX_CHECK_HIP(hipMemcpy(b.d_tile_member, b.tile_member.data(), (size_t)b.n_tiles * 4,
                           hipMemcpyHostToDevice));
    NX_CHECK_HIP(hipMemcpy(b.d_members, b.members.data(), b.members.size() * sizeof(Member),
                           hipMemcpyHostToDevice));
    return NX_OK;
}

void drop_graph(nx_batch_s& b)
{
    if (b.graph) hipGraphExecDestroy(b.graph);
    b.graph = nullptr;
    b.graph_k = 0;
}

}  // namespace
}  // namespace batch
}  // namespace nx

using namespace nx;
using namespace nx::batch;

extern "C" int nx_batch_destroy(nx_batch b)
{
    if (!b) return NX_OK;
    drop_graph(*b);
    for (void* p : {b->pop[0], b->pop[1], (void*)b->flags, (void*)b->d_members,
                    (void*)b->d_tile_member, (void*)b->active, (void*)b->steps_done,
                    (void*)b->n_active, (void*)b->partials, (void*)b->res, (void*)b->u_prev})
        if (p) hipFree(p);
    if (b->graph_stream) hipStreamDestroy(b->graph_stream);
    delete b;
    return NX_OK;
}

extern "C" int nx_batch_create(uint32_t n, const int32_t* dims, const uint8_t* classes,
                               const double* omega, const double* ulb, const double* ulid,
                               int32_t storage, nx_batch* out)
{
    if (!n || !dims || !classes || !omega || !ulb || !ulid || !out)
        return set_error(NX_ERR_INVALID, "nx_batch_create: null argument");
    if (storage < NX_F64 || storage > NX_F16DEV)
        return set_error(NX_ERR_INVALID, "unknown storage type");
    auto* b = new nx_batch_s;
    b->storage = storage;
    // measured on 512 cavities of 32..96 cells (the autotuner's choices):
    // binary16 takes the edge-speculating step with 32 workgroups per CU
    // dealt tile by tile (46.1k MLUPS; 45.2k with 16 per CU, 44.3k with one
    // per tile, 32.6k with 2 per CU in XCD chunks, 19.0k without the
    // speculation), FP32 the flag-dependent fetch
    // with one workgroup per tile (32.7k; 28.5k with a persistent grid of 2
    // per CU in XCD chunks; the speculation's 36 extra registers cost more
    // than the round trip they save)
    if (storage == NX_F16DEV) {
        b->edge = 1;
        b->wgs_per_cu = 32;
        b->xcd_chunks = 0;
    }
    b->n_members = n;
    uint64_t off = 0;
    for (uint32_t k = 0; k < n; k++) {
        Member M{};
        M.X = dims[3 * k];
        M.Y = dims[3 * k + 1];
        M.Z = dims[3 * k + 2];
        if (M.X < 3 || M.Y < 3 || M.Z < 3) {
            delete b;
            return set_error(NX_ERR_INVALID, "batch member dimensions must be >= 3");
        }
        M.XY = (uint32_t)(M.X * M.Y);
        M.n = M.XY * (uint32_t)M.Z;
        M.cell_base = (uint32_t)off;
        M.pop_base = 19ull * off;
        M.omega = omega[k];
        off += M.n;
        if (off >= 0xFFFFFFFFull) {
            delete b;
            return set_error(NX_ERR_INVALID, "batch too large (>= 2^32 cells)");
        }
        b->members.push_back(M);
    }
    b->n_cells = off;
    build_tiles(*b, b->wg);
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
        b->n_cu = prop.multiProcessorCount;
    auto fail = [&](hipError_t e, const char* what) {
        nx_batch_destroy(b);
        return hip_error(e, what);
    };
    hipError_t e;
    const size_t pb = 19 * off * elem_size(storage);
    for (int k = 0; k < 2; k++)
        if ((e = hipMalloc(&b->pop[k], pb)) != hipSuccess) return fail(e, "hipMalloc(populations)");
    if ((e = hipMalloc(&b->flags, off * 4)) != hipSuccess) return fail(e, "hipMalloc(flags)");
    if ((e = hipMalloc(&b->u_prev, off * 12)) != hipSuccess) return fail(e, "hipMalloc(u_prev)");
    if ((e = hipMalloc(&b->d_members, n * sizeof(Member))) != hipSuccess) return fail(e, "hipMalloc");
    if ((e = hipMalloc(&b->active, n * 4)) != hipSuccess) return fail(e, "hipMalloc");
    if ((e = hipMalloc(&b->steps_done, n * 4)) != hipSuccess) return fail(e, "hipMalloc");
    if ((e = hipMalloc(&b->n_active, 4)) != hipSuccess) return fail(e, "hipMalloc");
    if ((e = hipMalloc(&b->res, (size_t)n * 16)) != hipSuccess) return fail(e, "hipMalloc");
    if (int r = upload_tiles(*b)) {
        nx_batch_destroy(b);
        return r;
    }
    // flags from the classes, then fill and wall mask per member (set-up)
    std::vector<uint32_t> fl(off);
    for (uint64_t i = 0; i < off; i++) fl[i] = (uint32_t)(classes[i] & 3u) << NX_CLASS_SHIFT;
    if ((e = hipMemcpy(b->flags, fl.data(), off * 4, hipMemcpyHostToDevice)) != hipSuccess)
        return fail(e, "hipMemcpy(flags)");
    for (uint32_t k = 0; k < n; k++) {
        const Member& M = b->members[k];
        nx_dense_desc d{{M.X, M.Y, M.Z}, storage,
                        storage == NX_F64 ? NX_COMPUTE_F64 : NX_COMPUTE_F32, 0};
        uint32_t* fk = b->flags + M.cell_base;
        const size_t es = elem_size(storage);
        for (int p = 0; p < 2; p++) {
            void* pk = (char*)b->pop[p] + M.pop_base * es;
            if (int r = nx_d3q19_fill_dense(&d, pk, fk, ulb[k], ulid + 3 * k, nullptr)) {
                nx_batch_destroy(b);
                return r;
            }
        }
        if (int r = nx_d3q19_wall_mask_dense(&d, fk, fk, nullptr)) {
            nx_batch_destroy(b);
            return r;
        }
    }
    std::vector<uint32_t> ones(n, 1u), zeros(n, 0u);
    if ((e = hipMemcpy(b->active, ones.data(), n * 4, hipMemcpyHostToDevice)) != hipSuccess)
        return fail(e, "hipMemcpy");
    if ((e = hipMemcpy(b->steps_done, zeros.data(), n * 4, hipMemcpyHostToDevice)) != hipSuccess)
        return fail(e, "hipMemcpy");
    if ((e = hipMemset(b->u_prev, 0, off * 12)) != hipSuccess) return fail(e, "hipMemset");
    if ((e = hipDeviceSynchronize()) != hipSuccess) return fail(e, "batch set-up");
    *out = b;
    return NX_OK;
}

extern "C" int nx_batch_info(nx_batch b, uint32_t* n_members, uint64_t* n_cells, uint32_t* n_tiles)
{
    if (!b) return set_error(NX_ERR_INVALID, "null batch");
    if (n_members) *n_members = b->n_members;
    if (n_cells) *n_cells = b->n_cells;
    if (n_tiles) *n_tiles = b->n_tiles;
    return NX_OK;
}

// Dense view of one member (its populations in arena `which`, its flags), so
// every dense entry point (moments, mass, ...) applies to it.
extern "C" int nx_batch_member(nx_batch b, uint32_t k, int which, nx_dense_desc* d, void** pop,
                               uint32_t** flags)
{
    if (!b || k >= b->n_members || (which != 0 && which != 1))
        return set_error(NX_ERR_INVALID, "nx_batch_member: bad member or arena");
    const Member& M = b->members[k];
    if (d) *d = nx_dense_desc{{M.X, M.Y, M.Z}, b->storage,
                              b->storage == NX_F64 ? NX_COMPUTE_F64 : NX_COMPUTE_F32, 0};
    if (pop) *pop = (char*)b->pop[which] + M.pop_base * elem_size(b->storage);
    if (flags) *flags = b->flags + M.cell_base;
    return NX_OK;
}

// Equilibrium start from (rho, u) on the device, concatenated over members in
// cell order (u: 3 planes per member); NULL rho -> 1, NULL u -> 0.  Resets
// every member to stepping with a zero step count.
extern "C" int nx_batch_init(nx_batch b, const float* rho, const float* u, nx_stream s)
{
    if (!b) return set_error(NX_ERR_INVALID, "null batch");
    const hipStream_t st = (hipStream_t)s;
    if (int e = by_storage(b->storage, [&](auto sc) {
            constexpr int S = decltype(sc)::value;
            using T = typename Stor<S>::T;
            hipLaunchKernelGGL((k_batch_init<S>), dim3(std::min<uint32_t>(b->n_tiles, 8192)),
                               dim3(kTile), 0, st, (T*)b->pop[0], (T*)b->pop[1], b->flags,
                               b->d_members, b->d_tile_member, rho, u, b->u_prev, b->n_tiles,
                               b->wg);
        }))
        return e;
    NX_CHECK_LAUNCH("k_batch_init");
    std::vector<uint32_t> ones(b->n_members, 1u), zeros(b->n_members, 0u);
    NX_CHECK_HIP(hipMemcpyAsync(b->active, ones.data(), b->n_members * 4, hipMemcpyHostToDevice, st));
    NX_CHECK_HIP(hipMemcpyAsync(b->steps_done, zeros.data(), b->n_members * 4, hipMemcpyHostToDevice, st));
    NX_CHECK_HIP(hipStreamSynchronize(st));
    return NX_OK;
}

// Consistent start: `iterations` times, two steps then k_batch_constrain on
// the latest populations with the prescribed velocity u (layout of
// nx_batch_init), for the members selected by `members` (host bytes, NULL:
// all); the others do not step.  Every member is stepping again afterwards;
// step counters and residual state are left untouched.
extern "C" int nx_batch_constrain(nx_batch b, const float* u, int32_t iterations,
                                  const uint8_t* members, nx_stream s)
{
    if (!b || !u || iterations < 0) return set_error(NX_ERR_INVALID, "nx_batch_constrain: bad arguments");
    const hipStream_t st = (hipStream_t)s;
    std::vector<uint32_t> sel(b->n_members, 1u), ones(b->n_members, 1u);
    if (members)
        for (uint32_t k = 0; k < b->n_members; k++) sel[k] = members[k] ? 1u : 0u;
    NX_CHECK_HIP(hipMemcpyAsync(b->active, sel.data(), b->n_members * 4, hipMemcpyHostToDevice, st));
    for (int it = 0; it < iterations; it++) {
        if (int e = launch_steps(*b, 2, 0, st)) return e;
        if (int e = by_storage(b->storage, [&](auto sc) {
                constexpr int S = decltype(sc)::value;
                using T = typename Stor<S>::T;
                hipLaunchKernelGGL((k_batch_constrain<S>), dim3(std::min<uint32_t>(b->n_tiles, 8192)),
                                   dim3(kTile), 0, st, (T*)b->pop[0], b->flags, b->d_members,
                                   b->d_tile_member, b->active, u, b->n_tiles, b->wg);
            }))
            return e;
        NX_CHECK_LAUNCH("k_batch_constrain");
    }
    NX_CHECK_HIP(hipMemcpyAsync(b->active, ones.data(), b->n_members * 4, hipMemcpyHostToDevice, st));
    NX_CHECK_HIP(hipStreamSynchronize(st));
    return NX_OK;
}

// k_steps steps of every stepping member, no convergence test (k even keeps
// the latest populations in arena 0).
extern "C" int nx_batch_step(nx_batch b, int32_t k_steps, nx_stream s)
{
    if (!b || k_steps < 0) return set_error(NX_ERR_INVALID, "nx_batch_step: bad arguments");
    if (k_steps & 1) return set_error(NX_ERR_INVALID, "nx_batch_step: k_steps must be even");
    return launch_steps(*b, k_steps, 0, (hipStream_t)s);
}

// Step until every member's residual over check_every steps is below tol or
// max_steps steps were taken.  One hipGraph holds check_every steps and the
// convergence update; it is replayed on the batch's own stream.  steps and
// residual (optional, host arrays of n) receive each member's step count and
// last residual.
extern "C" int nx_batch_run(nx_batch b, int32_t max_steps, int32_t check_every, double tol,
                            int32_t* steps, double* residual, uint32_t* n_unconverged)
{
    if (!b || check_every < 2 || (check_every & 1) || max_steps < 0 || !(tol > 0))
        return set_error(NX_ERR_INVALID,
                         "nx_batch_run: check_every must be even and >= 2, tol > 0");
    if (!b->graph_stream) NX_CHECK_HIP(hipStreamCreateWithFlags(&b->graph_stream, hipStreamNonBlocking));
    const hipStream_t st = b->graph_stream;
    // the replays run on the batch's own non-blocking stream: whatever the
    // caller queued on its streams (nx_batch_step, nx_batch_init, ...) must
    // have finished before they touch the arenas
    NX_CHECK_HIP(hipDeviceSynchronize());
    if (!b->graph || b->graph_k != check_every || b->graph_tol != tol) {
        drop_graph(*b);
        hipGraph_t g;
        NX_CHECK_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        int e1 = launch_steps(*b, check_every, 0, st);
        int e2 = e1 ? e1 : launch_check(*b, check_every, tol, st);
        hipError_t ec = hipStreamEndCapture(st, &g);
        if (e2) return e2;
        NX_CHECK_HIP(ec);
        hipError_t ei = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        NX_CHECK_HIP(ei);
        b->graph_k = check_every;
        b->graph_tol = tol;
    }
    uint32_t left = b->n_members;
    for (int32_t done = 0; done < max_steps && left > 0; done += check_every) {
        NX_CHECK_HIP(hipGraphLaunch(b->graph, st));
        NX_CHECK_HIP(hipMemcpyAsync(&left, b->n_active, 4, hipMemcpyDeviceToHost, st));
        NX_CHECK_HIP(hipStreamSynchronize(st));
    }
    if (n_unconverged) *n_unconverged = left;
    if (steps) NX_CHECK_HIP(hipMemcpy(steps, b->steps_done, b->n_members * 4, hipMemcpyDeviceToHost));
    if (residual) {
        std::vector<double> r(2 * (size_t)b->n_members);
        NX_CHECK_HIP(hipMemcpy(r.data(), b->res, r.size() * 8, hipMemcpyDeviceToHost));
        for (uint32_t k = 0; k < b->n_members; k++) {
            const bool finite = std::isfinite(r[2 * k]) && std::isfinite(r[2 * k + 1]);
            residual[k] = !finite ? std::nan("") : (r[2 * k + 1] > 0 ? std::sqrt(r[2 * k] / r[2 * k + 1]) : 0.0);
        }
    }
    return NX_OK;
}

// Launch configuration: wg in {256, 512} cells per tile, persistent grid of
// wgs_per_cu workgroups per CU (0: one workgroup per tile).
extern "C" int nx_batch_set_launch(nx_batch b, int32_t wg, int32_t wgs_per_cu)
{
    if (!b || (wg != 256 && wg != 512) || wgs_per_cu < 0)
        return set_error(NX_ERR_INVALID, "nx_batch_set_launch: wg in {256,512}, wgs_per_cu >= 0");
    if (wg != b->wg) {
        b->wg = wg;
        build_tiles(*b, wg);
        if (int e = upload_tiles(*b)) return e;
    }
    b->wgs_per_cu = wgs_per_cu;
    drop_graph(*b);
    return NX_OK;
}

// Times the step for every launch configuration (the populations advance by
// an even number of steps per configuration) and keeps the fastest.
extern "C" int nx_batch_set_variant(nx_batch b, int32_t xcd_chunks, int32_t edge)
{
    if (!b) return set_error(NX_ERR_INVALID, "null batch");
    if ((xcd_chunks != 0 && xcd_chunks != 1) || (edge != 0 && edge != 1))
        return set_error(NX_ERR_INVALID, "batch variant: xcd_chunks and edge are 0 or 1");
    if (b->xcd_chunks != xcd_chunks || b->edge != edge) drop_graph(*b);  // nx_batch_run recaptures
    b->xcd_chunks = xcd_chunks;
    b->edge = edge;
    return NX_OK;
}

extern "C" int nx_batch_grid_size(uint32_t n_tiles, int32_t n_cu, int32_t wgs_per_cu, uint32_t* grid)
{
    if (!grid || n_cu < 0 || wgs_per_cu < 0)
        return set_error(NX_ERR_INVALID, "nx_batch_grid_size: n_cu, wgs_per_cu >= 0 and a result pointer");
    *grid = grid_size(n_tiles, n_cu, wgs_per_cu);
    return NX_OK;
}

extern "C" int nx_batch_launch_grid(nx_batch b, uint32_t* grid, int32_t* xcd_chunked)
{
    if (!b) return set_error(NX_ERR_INVALID, "null batch");
    if (grid) *grid = grid_of(*b);
    if (xcd_chunked) *xcd_chunked = chunked(*b);
    return NX_OK;
}

extern "C" int nx_batch_get_launch(nx_batch b, int32_t* wg, int32_t* wgs_per_cu, int32_t* xcd_chunks,
                                   int32_t* edge)
{
    if (!b) return set_error(NX_ERR_INVALID, "null batch");
    if (wg) *wg = b->wg;
    if (wgs_per_cu) *wgs_per_cu = b->wgs_per_cu;
    if (xcd_chunks) *xcd_chunks = b->xcd_chunks;
    if (edge) *edge = b->edge;
    return NX_OK;
}

extern "C" int nx_batch_autotune(nx_batch b, nx_stream s, int32_t* best_wg, int32_t* best_wgs_per_cu,
                                 float* best_ms)
{
    if (!b) return set_error(NX_ERR_INVALID, "null batch");
    const hipStream_t st = (hipStream_t)s;
    hipEvent_t e0, e1;
    NX_CHECK_HIP(hipEventCreate(&e0));
    NX_CHECK_HIP(hipEventCreate(&e1));
    int bw = b->wg, bc = b->wgs_per_cu, bx = b->xcd_chunks, be = b->edge;
    float bms = 1e30f;
    for (int ed : {1, 0})
    for (int xc : {1, 0})
    for (int wg : {256, 512})
        for (int c : {2, 4, 8, 16, 32, 0}) {
            if (int e = nx_batch_set_launch(b, wg, c)) return e;
            b->xcd_chunks = xc;
            b->edge = ed;
            if (int e = launch_steps(*b, 2, 0, st)) return e;
            NX_CHECK_HIP(hipEventRecord(e0, st));
            if (int e = launch_steps(*b, 6, 0, st)) return e;
            NX_CHECK_HIP(hipEventRecord(e1, st));
            NX_CHECK_HIP(hipEventSynchronize(e1));
            float ms = 0;
            NX_CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
            ms /= 6;
            if (ms < bms) {
                bms = ms;
                bw = wg;
                bc = c;
                bx = xc;
                be = ed;
            }
        }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (int e = nx_batch_set_launch(b, bw, bc)) return e;
    b->xcd_chunks = bx;
    b->edge = be;
    if (best_wg) *best_wg = bw;
    if (best_wgs_per_cu) *best_wgs_per_cu = bc;
    if (best_ms) *best_ms = bms;
    return NX_OK;
}
