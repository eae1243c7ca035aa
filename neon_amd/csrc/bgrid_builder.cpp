// bgrid_builder.cpp -- host construction of the block-sparse grid structure
// for one partition: the semantics of bGrid's constructor
// (libNeonDomain/include/Neon/domain/details/bGrid/bGrid_imp.h:19-246) with
// Partitioner1D / SpanClassifier on a single device
// (tools/Partitioner1D.h:106-255, tools/partitioning/SpanClassifier.h:170-236):
//   * candidate blocks: the ceil(dim/8) lattice of 8^3 blocks;
//   * kept blocks: those with at least one active voxel inside the domain;
//   * block index: rank in bz-major, then by, then bx order;
//   * origin = 8*(bx,by,bz);
//   * bitmask: 16 uint32 words, bit (x+8y+64z) (StaticBlock.h:79-96);
//   * connectivity: 27 entries, direction i+3j+9k <-> offset (i-1,j-1,k-1),
//     0xFFFFFFFF where the neighbour block is absent (bGrid_imp.h:139-172).
// The work is split over z-slabs of blocks on host threads: one pass computes
// the per-block masks and counts, an exclusive scan over the counts in
// z-major order gives the block ranks, a second pass fills the tables.
#include <stdint.h>

#include <algorithm>
#include <string>
#include <thread>
#include <vector>

#include "neon_amd.h"

struct nx_bgrid_s {
    int32_t dim[3];
    int32_t periodic;
    uint32_t n_blocks = 0;
    uint64_t n_active = 0;
    std::vector<int32_t> origins;
    std::vector<uint32_t> masks;
    std::vector<uint32_t> conn;
};

namespace nx {
int set_error(int code, const std::string& msg);
}

namespace {

template <typename F>
void parallel_slabs(int n, F&& f)
{
    unsigned nt = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
    nt = std::min<unsigned>(nt, (unsigned)std::max(1, n));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; t++)
        th.emplace_back([&, t] {
            for (int k = (int)t; k < n; k += (int)nt) f(k);
        });
    for (auto& x : th) x.join();
}

template <typename Active>
int build(const int32_t dim[3], Active&& active, int32_t periodic, nx_bgrid* out)
{
    if (!out) return nx::set_error(NX_ERR_INVALID, "null output handle");
    for (int a = 0; a < 3; a++)
        if (dim[a] < 1) return nx::set_error(NX_ERR_INVALID, "bGrid dimensions must be >= 1");
    for (int a = 0; a < 3; a++)
        if ((periodic >> a) & 1)
            if (dim[a] % 8)
                return nx::set_error(NX_ERR_INVALID,
                                     "periodic bGrid axes must be a multiple of the block size 8");
    const int nbx = (dim[0] + 7) / 8, nby = (dim[1] + 7) / 8, nbz = (dim[2] + 7) / 8;
    const size_t nb3 = (size_t)nbx * nby * nbz;
    if (nb3 >= 0xFFFFFFFFull) return nx::set_error(NX_ERR_INVALID, "bGrid too large");

    // pass 1: bitmask and active count of every candidate block
    std::vector<uint32_t> bits(nb3 * 16, 0u);
    std::vector<uint32_t> cnt(nb3, 0u);
    parallel_slabs(nbz, [&](int bz) {
        for (int by = 0; by < nby; by++)
            for (int bx = 0; bx < nbx; bx++) {
                const size_t l = (size_t)bx + (size_t)nbx * ((size_t)by + (size_t)nby * bz);
                uint32_t* w = &bits[16 * l];
                uint32_t c = 0;
                for (int z = 0; z < 8; z++) {
                    const int gz = bz * 8 + z;
                    if (gz >= dim[2]) break;
                    for (int y = 0; y < 8; y++) {
                        const int gy = by * 8 + y;
                        if (gy >= dim[1]) break;
                        for (int x = 0; x < 8; x++) {
                            const int gx = bx * 8 + x;
                            if (gx >= dim[0]) break;
                            if (active(gx, gy, gz)) {
                                const int v = x + 8 * y + 64 * z;
                                w[v >> 5] |= 1u << (v & 31);
                                c++;
                            }
                        }
                    }
                }
                cnt[l] = c;
            }
    });

    // exclusive scan in z-major order -> block rank
    std::vector<uint32_t> rank(nb3);
    uint32_t n = 0;
    uint64_t nact = 0;
    for (size_t l = 0; l < nb3; l++) {
        if (cnt[l]) {
            rank[l] = n++;
            nact += cnt[l];
        } else {
            rank[l] = 0xFFFFFFFFu;
        }
    }

    nx_bgrid_s* g = new nx_bgrid_s;
    for (int a = 0; a < 3; a++) g->dim[a] = dim[a];
    g->periodic = periodic;
    g->n_blocks = n;
    g->n_active = nact;
    g->origins.resize(3 * (size_t)n);
    g->masks.resize(16 * (size_t)n);
    g->conn.resize(27 * (size_t)n);

    // pass 2: tables
    parallel_slabs(nbz, [&](int bz) {
        for (int by = 0; by < nby; by++)
            for (int bx = 0; bx < nbx; bx++) {
                const size_t l = (size_t)bx + (size_t)nbx * ((size_t)by + (size_t)nby * bz);
                const uint32_t r = rank[l];
                if (r == 0xFFFFFFFFu) continue;
                g->origins[3 * (size_t)r + 0] = 8 * bx;
                g->origins[3 * (size_t)r + 1] = 8 * by;
                g->origins[3 * (size_t)r + 2] = 8 * bz;
                std::copy(&bits[16 * l], &bits[16 * l] + 16, &g->masks[16 * (size_t)r]);
                for (int k = 0; k < 3; k++)
                    for (int j = 0; j < 3; j++)
                        for (int i = 0; i < 3; i++) {
                            int x = bx + i - 1, y = by + j - 1, z = bz + k - 1;
                            if (periodic & NX_PERIODIC_X) x = (x + nbx) % nbx;
                            if (periodic & NX_PERIODIC_Y) y = (y + nby) % nby;
                            if (periodic & NX_PERIODIC_Z) z = (z + nbz) % nbz;
                            uint32_t t = 0xFFFFFFFFu;
                            if (x >= 0 && y >= 0 && z >= 0 && x < nbx && y < nby && z < nbz)
                                t = rank[(size_t)x + (size_t)nbx * ((size_t)y + (size_t)nby * z)];
                            g->conn[27 * (size_t)r + i + 3 * j + 9 * k] = t;
                        }
            }
    });
    *out = g;
    return NX_OK;
}

}  // namespace

extern "C" {

int nx_bgrid_create(const int32_t dim[3], const uint8_t* active, int32_t periodic, nx_bgrid* g)
{
    if (!dim || !active) return nx::set_error(NX_ERR_INVALID, "null pointer");
    const size_t X = (size_t)dim[0], XY = (size_t)dim[0] * (size_t)dim[1];
    return build(
        dim, [&](int x, int y, int z) { return active[(size_t)x + X * y + XY * z] != 0; }, periodic,
        g);
}

int nx_bgrid_create_cb(const int32_t dim[3], int (*active)(int32_t, int32_t, int32_t, void*),
                       void* user, int32_t periodic, nx_bgrid* g)
{
    if (!dim || !active) return nx::set_error(NX_ERR_INVALID, "null pointer");
    return build(
        dim, [&](int x, int y, int z) { return active(x, y, z, user) != 0; }, periodic, g);
}

int nx_bgrid_info(nx_bgrid g, uint32_t* n_blocks, uint64_t* n_active)
{
    if (!g) return nx::set_error(NX_ERR_INVALID, "null handle");
    if (n_blocks) *n_blocks = g->n_blocks;
    if (n_active) *n_active = g->n_active;
    return NX_OK;
}

int nx_bgrid_tables(nx_bgrid g, int32_t* origins, uint32_t* masks, uint32_t* conn)
{
    if (!g) return nx::set_error(NX_ERR_INVALID, "null handle");
    if (origins) std::copy(g->origins.begin(), g->origins.end(), origins);
    if (masks) std::copy(g->masks.begin(), g->masks.end(), masks);
    if (conn) std::copy(g->conn.begin(), g->conn.end(), conn);
    return NX_OK;
}

int nx_bgrid_destroy(nx_bgrid g)
{
    delete g;
    return NX_OK;
}

}  // extern "C"
