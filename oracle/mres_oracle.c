/* mres_oracle.c -- CPU restatement of Neon's grid-refinement LBM
 * (apps/lbmMultiRes, BGK D3Q19 variant) on a stack of `depth` levels
 * (2 <= depth <= OR_MRES_MAX_DEPTH).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the library's nx_mgrid_* path.
 * Nothing in the product links or calls this file.
 *
 * Two lattices: BGK D3Q19 (the app's BGK variant, the cavity benchmark's
 * lattice and arithmetic) and KBC D3Q27 (its default; collide.h:148-284,
 * lattice.h:14-106), selected by `lat` (OR_MRES_BGK19 / OR_MRES_KBC27).
 *
 * Level 0 is the finest level (the domain dim[3]); level l has voxels of
 * edge 2^l and ceil(dim / 2^l) of them per axis; a level-(l+1) voxel v
 * covers the level-l voxels 2v + {0,1}^3.  Storage is dense per level, SoA
 * [q][z][y][x], one byte of state per voxel: active, refined (has
 * children), class.  Every function follows the reference operation by
 * operation, in the reference's loop order where the result depends on it
 * (the strong-balance pass, the fine-to-coarse accumulation).
 *
 * Reference files restated:
 *   grid:     libNeonDomain/src/domain/details/mGrid/mGrid.cpp:89-156
 *             (activation per 2^3 group, parent marking), :159-263 (overlap
 *             culling, coarsest level first), :266-345 (strong balance) and
 *             :466-555 (a voxel has children iff one of its eight children
 *             is active in the final masks)
 *   neighbours: mPartition_imp.h:155-235 (children, parent, uncle),
 *             apps/lbmMultiRes/util.h:19-37 (uncleOffset),
 *             util.h:40-45 (computeOmega)
 *   set-up:   apps/lbmMultiRes/lidDrivenCavity.h:8-78 (cavity),
 *             init.h:8-86 (sumStore)
 *   step:     apps/lbmMultiRes/lbmMultiRes.h:141-263 (recursion),
 *             collide.h:8-149 (BGK, accumulator reset),
 *             store.h:98-163 (storeFine),
 *             stream.h:170-225 (stream fused with coalescence and explosion)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "lbm_oracle.h"

/* D3Q19 of apps/lbmMultiRes/lattice.h:108-173 (the BGK lattice): the same
 * order as the cavity benchmark's D3Q19 */
static const int MC[19][3] = {{-1, 0, 0}, {0, -1, 0}, {0, 0, -1}, {-1, -1, 0}, {-1, 1, 0},
                              {-1, 0, -1}, {-1, 0, 1}, {0, -1, -1}, {0, -1, 1}, {0, 0, 0},
                              {1, 0, 0},   {0, 1, 0},   {0, 0, 1},   {1, 1, 0},   {1, -1, 0},
                              {1, 0, 1},   {1, 0, -1},  {0, 1, 1},   {0, 1, -1}};
static const int MOPP[19] = {10, 11, 12, 13, 14, 15, 16, 17, 18, 9, 0, 1, 2, 3, 4, 5, 6, 7, 8};
static double mw(int q) { return q == 9 ? 1. / 3. : ((q % 10) < 3 ? 1. / 18. : 1. / 36.); }

/* D3Q27 of apps/lbmMultiRes/lattice.h:14-75 (the KBC lattice): z fastest,
 * then y, then x, each over 0, -1, +1 */
static const int KC[27][3] = {
    {0, 0, 0},    {0, 0, -1},  {0, 0, 1},   {0, -1, 0},  {0, -1, -1}, {0, -1, 1},  {0, 1, 0},
    {0, 1, -1},   {0, 1, 1},   {-1, 0, 0},  {-1, 0, -1}, {-1, 0, 1},  {-1, -1, 0}, {-1, -1, -1},
    {-1, -1, 1},  {-1, 1, 0},  {-1, 1, -1}, {-1, 1, 1},  {1, 0, 0},   {1, 0, -1},  {1, 0, 1},
    {1, -1, 0},   {1, -1, -1}, {1, -1, 1},  {1, 1, 0},   {1, 1, -1},  {1, 1, 1}};
static const int KOPP[27] = {0,  2,  1,  6,  8,  7,  3,  5,  4,  18, 20, 19, 24, 26,
                             25, 21, 23, 22, 9,  11, 10, 15, 17, 16, 12, 14, 13};
/* weights by the number of nonzero velocity components (lattice.h:47-75) */
static double kw(int q)
{
    const int n = (KC[q][0] != 0) + (KC[q][1] != 0) + (KC[q][2] != 0);
    return n == 0 ? 8.0 / 27.0 : n == 1 ? 2.0 / 27.0 : n == 2 ? 1.0 / 54.0 : 1.0 / 216.0;
}

typedef struct {
    int Q;
    int rest; /* index of the rest population */
    const int (*c)[3];
    const int* opp;
    double (*w)(int);
} mr_lat;

static const mr_lat LAT19 = {19, 9, MC, MOPP, mw};
static const mr_lat LAT27 = {27, 0, KC, KOPP, kw};
static const mr_lat* lat_of(int lat) { return lat == OR_MRES_KBC27 ? &LAT27 : &LAT19; }

/* per-voxel state bits */
#define MR_ACTIVE 1u
#define MR_REFINED 2u
#define MR_CLS_SHIFT 2 /* 2 bits: OR_BOUNCEBACK, OR_MOVINGWALL, OR_BULK */

static inline size_t lin(const int32_t d[3], int x, int y, int z)
{
    return (size_t)x + (size_t)d[0] * ((size_t)y + (size_t)d[1] * (size_t)z);
}
static inline int inside(const int32_t d[3], int x, int y, int z)
{
    return x >= 0 && y >= 0 && z >= 0 && x < d[0] && y < d[1] && z < d[2];
}
static inline size_t nvox(const int32_t d[3]) { return (size_t)d[0] * d[1] * d[2]; }

/* level l: ceil(dim / 2^l) voxels per axis (a voxel is in the domain when
 * its level-0 origin v 2^l is, mGrid.cpp:105-108) */
void or_mres_dims(const int32_t dim[3], int depth, int32_t* dims)
{
    for (int a = 0; a < 3; a++) dims[a] = dim[a];
    for (int l = 1; l < depth; l++)
        for (int a = 0; a < 3; a++) dims[3 * l + a] = (dims[3 * (l - 1) + a] + 1) / 2;
}

static inline int act(const int32_t d[3], const uint8_t* st, int x, int y, int z)
{
    return inside(d, x, y, z) && (st[lin(d, x, y, z)] & MR_ACTIVE);
}

/* isRefined (mGrid.cpp:167-199): one of the eight children of level-l voxel
 * (x, y, z) is active on level l - 1 (the voxel itself need not be) */
static int has_child(const int32_t dc[3], const uint8_t* stc, int x, int y, int z)
{
    for (int k = 0; k < 8; k++)
        if (act(dc, stc, 2 * x + (k & 1), 2 * y + ((k >> 1) & 1), 2 * z + (k >> 2))) return 1;
    return 0;
}

/* mGrid constructor (mGrid.cpp:89-345, :466-555).  lam[l]: the level-l
 * activation lambda at every level-l voxel (evaluated by the caller at the
 * voxel's level-0 origin v 2^l).  Outputs the state byte of every voxel of
 * every level; returns the number of voxels the strong-balance pass added. */
int or_mres_build(const int32_t dim[3], int depth, const uint8_t* const* lam, uint8_t* const* st)
{
    int32_t dims[3 * OR_MRES_MAX_DEPTH];
    or_mres_dims(dim, depth, dims);
    for (int l = 0; l < depth; l++) memset(st[l], 0, nvox(dims + 3 * l));
    /* activation (:89-156): level by level from the finest, groups of 2^3
     * voxels; a group with an active voxel (set earlier by a finer level or
     * by its lambda) activates all of its voxels inside the domain and its
     * parent */
    for (int l = 0; l < depth; l++) {
        const int32_t* d = dims + 3 * l;
        const int32_t g[3] = {(d[0] + 1) / 2, (d[1] + 1) / 2, (d[2] + 1) / 2};
        for (int gz = 0; gz < g[2]; gz++)
            for (int gy = 0; gy < g[1]; gy++)
                for (int gx = 0; gx < g[0]; gx++) {
                    int any = 0;
                    for (int k = 0; k < 8; k++) {
                        const int x = 2 * gx + (k & 1), y = 2 * gy + ((k >> 1) & 1), z = 2 * gz + (k >> 2);
                        if (inside(d, x, y, z) && (st[l][lin(d, x, y, z)] || lam[l][lin(d, x, y, z)])) any = 1;
                    }
                    if (!any) continue;
                    for (int k = 0; k < 8; k++) {
                        const int x = 2 * gx + (k & 1), y = 2 * gy + ((k >> 1) & 1), z = 2 * gz + (k >> 2);
                        if (inside(d, x, y, z)) st[l][lin(d, x, y, z)] = MR_ACTIVE;
                    }
                    if (l < depth - 1) st[l + 1][lin(dims + 3 * (l + 1), gx, gy, gz)] = MR_ACTIVE;
                }
    }
    /* overlap culling (:159-263), coarsest level first: an active voxel with
     * an active child whose in-domain neighbours all have an active child
     * is removed (level l - 1 is still unculled when level l is visited) */
    for (int l = depth - 1; l > 0; l--) {
        const int32_t* d = dims + 3 * l;
        const int32_t* dc = dims + 3 * (l - 1);
        const size_t n = nvox(d);
        uint8_t* refined = (uint8_t*)calloc(n, 1);
        for (int z = 0; z < d[2]; z++)
            for (int y = 0; y < d[1]; y++)
                for (int x = 0; x < d[0]; x++) refined[lin(d, x, y, z)] = (uint8_t)has_child(dc, st[l - 1], x, y, z);
        for (int z = 0; z < d[2]; z++)
            for (int y = 0; y < d[1]; y++)
                for (int x = 0; x < d[0]; x++) {
                    const size_t i = lin(d, x, y, z);
                    if (!(st[l][i] & MR_ACTIVE) || !refined[i]) continue;
                    int deactivate = 1;
                    for (int k = -1; k <= 1; k++)
                        for (int j = -1; j <= 1; j++)
                            for (int h = -1; h <= 1; h++) {
                                if (!h && !j && !k) continue;
                                if (inside(d, x + h, y + j, z + k) && !refined[lin(d, x + h, y + j, z + k)])
                                    deactivate = 0;
                            }
                    if (deactivate) st[l][i] = 0;
                }
        free(refined);
    }
    /* strong balance (:266-345): for every active voxel of level l and each
     * of its 26 neighbour positions inside the domain, look for the finest
     * active voxel covering the position from level l upwards; found on
     * level l or l + 1 it is fine, found on level l_n >= l + 2 the voxel of
     * level l_n - 1 covering the position is activated (on its own), and the
     * search goes on upwards.  Repeated until a pass activates nothing; the
     * passes see their own activations in loop order. */
    int added = 0, again = 1;
    while (again) {
        again = 0;
        for (int l = 0; l < depth; l++) {
            const int32_t* d = dims + 3 * l;
            const int32_t g[3] = {(d[0] + 1) / 2, (d[1] + 1) / 2, (d[2] + 1) / 2};
            for (int bz = 0; bz < g[2]; bz++)
                for (int by = 0; by < g[1]; by++)
                    for (int bx = 0; bx < g[0]; bx++)
                        for (int k8 = 0; k8 < 8; k8++) {
                            const int x = 2 * bx + (k8 & 1), y = 2 * by + ((k8 >> 1) & 1), z = 2 * bz + (k8 >> 2);
                            if (!act(d, st[l], x, y, z)) continue;
                            for (int k = -1; k <= 1; k++)
                                for (int j = -1; j <= 1; j++)
                                    for (int h = -1; h <= 1; h++) {
                                        if (!h && !j && !k) continue;
                                        int px = x + h, py = y + j, pz = z + k;
                                        if (!inside(d, px, py, pz)) continue;
                                        int qx = 0, qy = 0, qz = 0; /* the previous level's position */
                                        for (int ln = l; ln < depth; ln++) {
                                            const int32_t* dn = dims + 3 * ln;
                                            if (act(dn, st[ln], px, py, pz)) {
                                                if (ln == l || ln == l + 1) break;
                                                uint8_t* s = &st[ln - 1][lin(dims + 3 * (ln - 1), qx, qy, qz)];
                                                if (!(*s & MR_ACTIVE)) added++;
                                                *s = MR_ACTIVE;
                                                again = 1;
                                            }
                                            qx = px;
                                            qy = py;
                                            qz = pz;
                                            px >>= 1;
                                            py >>= 1;
                                            pz >>= 1;
                                        }
                                    }
                        }
        }
    }
    /* refined flags from the final masks (:497-545) */
    for (int l = 1; l < depth; l++) {
        const int32_t* d = dims + 3 * l;
        for (int z = 0; z < d[2]; z++)
            for (int y = 0; y < d[1]; y++)
                for (int x = 0; x < d[0]; x++) {
                    const size_t i = lin(d, x, y, z);
                    if ((st[l][i] & MR_ACTIVE) && has_child(dims + 3 * (l - 1), st[l - 1], x, y, z))
                        st[l][i] |= MR_REFINED;
                }
    }
    return added;
}

/* ---- set-up ------------------------------------------------------------- */

/* uncleOffset (util.h:19-37): offset, from the parent of voxel x, of the
 * parent of x + q; per axis s = (x mod 2) + q -> s <= 0 ? s : s - 1 */
static inline int uoff(int x, int q)
{
    const int s = (x & 1) + q;
    return s <= 0 ? s : s - 1;
}

/* the 8^3 memory blocks of a level that hold an active voxel */
static uint8_t* memory_blocks(const int32_t d[3], const uint8_t* st, int32_t nb[3])
{
    for (int a = 0; a < 3; a++) nb[a] = (d[a] + 7) / 8;
    uint8_t* b = (uint8_t*)calloc(nvox(nb), 1);
    for (int z = 0; z < d[2]; z++)
        for (int y = 0; y < d[1]; y++)
            for (int x = 0; x < d[0]; x++)
                if (st[lin(d, x, y, z)] & MR_ACTIVE) b[lin(nb, x >> 3, y >> 3, z >> 3)] = 1;
    return b;
}

/* hasParent (mPartition_imp.h:172-208): the memory block (8^3 voxels,
 * mPartition.h:13) of the next level that holds the parent voxel exists,
 * i.e. has an active voxel -- not that the parent voxel itself is active */
static inline int parent_block(const int32_t nbu[3], const uint8_t* blk_up, int px, int py, int pz)
{
    return inside(nbu, px >> 3, py >> 3, pz >> 3) && blk_up[lin(nbu, px >> 3, py >> 3, pz >> 3)];
}

static inline void put(void* p, int s, size_t i, double v)
{
    if (s == OR_F64)
        ((double*)p)[i] = v;
    else
        ((float*)p)[i] = (float)v;
}
static inline double get(const void* p, int s, size_t i)
{
    return s == OR_F64 ? ((const double*)p)[i] : (double)((const float*)p)[i];
}

/* Lid-driven cavity (lidDrivenCavity.h:8-78): every active voxel is bulk
 * with zero populations; leaf voxels of level 0 on the domain faces are
 * bounce-back, those with y = Y-1 the moving lid; leaf voxels hold w_q
 * (bulk), 0 (bounce-back) or -6 w_q (c_q . ulid) (lid), both fields.
 * ulid = (ulb, 0, 0).  sumStore of levels 1.. after init.h:8-86: every
 * active voxel of level l with a parent counts, for each moving q, one towards
 * the level-(l+1) voxel that will later pull its population q, then
 * ss = 1 / (2 count).  ss[0] is not used.  Arrays hold Q populations per
 * voxel, Q = 19 or 27 after the lattice. */
void or_mres_setup(const int32_t dim[3], int depth, int lat, int storage, double ulb, uint8_t* const* st,
                   void* const* fin, void* const* fout, float* const* ss)
{
    const double uw[3] = {ulb, 0., 0.};
    or_mres_setup_cells(dim, depth, lat, storage, NULL, uw, st, fin, fout, ss);
}

/* The set-up with the cell class of every level-l leaf given by cls[l]
 * (OR_BULK / OR_BOUNCEBACK / OR_MOVINGWALL per level-l voxel, x fastest;
 * cls == NULL: the lid-driven cavity) and the velocity uw of the moving-wall
 * cells (the lid, or the inlet of flowOverShape.h:62-104, whose cells the
 * stream treats alike).  Populations: w_q (bulk), 0 (bounce-back),
 * -6 w_q (c_q . uw) (moving wall), accumulated over the three components in
 * T as the reference does. */
void or_mres_setup_cells(const int32_t dim[3], int depth, int lat, int storage, const uint8_t* const* cls_in,
                         const double uw[3], uint8_t* const* st, void* const* fin, void* const* fout,
                         float* const* ss)
{
    const mr_lat* L = lat_of(lat);
    const int Q = L->Q;
    int32_t dims[3 * OR_MRES_MAX_DEPTH];
    or_mres_dims(dim, depth, dims);
    for (int lvl = 0; lvl < depth; lvl++) {
        const int32_t* d = dims + 3 * lvl;
        const size_t n = nvox(d);
        for (int z = 0; z < d[2]; z++)
            for (int y = 0; y < d[1]; y++)
                for (int x = 0; x < d[0]; x++) {
                    const size_t i = lin(d, x, y, z);
                    if (!(st[lvl][i] & MR_ACTIVE)) continue;
                    int cls = OR_BULK;
                    for (int q = 0; q < Q; q++) {
                        put(fin[lvl], storage, (size_t)q * n + i, 0.);
                        put(fout[lvl], storage, (size_t)q * n + i, 0.);
                    }
                    if (!(st[lvl][i] & MR_REFINED)) {
                        if (cls_in) {
                            cls = cls_in[lvl][i];
                        } else if (lvl == 0 && (x == 0 || x == d[0] - 1 || y == 0 || y == d[1] - 1 || z == 0 ||
                                                z == d[2] - 1)) {
                            cls = OR_BOUNCEBACK;
                            if (y == d[1] - 1) cls = OR_MOVINGWALL;
                        }
                        for (int q = 0; q < Q; q++) {
                            double v;
                            if (storage == OR_F64) {
                                double p = L->w(q);
                                if (cls == OR_BOUNCEBACK) p = 0;
                                if (cls == OR_MOVINGWALL) {
                                    p = 0;
                                    p += (double)L->c[q][0] * uw[0];
                                    p += (double)L->c[q][1] * uw[1];
                                    p += (double)L->c[q][2] * uw[2];
                                    p *= -6. * L->w(q);
                                }
                                v = p;
                            } else {
                                float p = (float)L->w(q);
                                if (cls == OR_BOUNCEBACK) p = 0;
                                if (cls == OR_MOVINGWALL) {
                                    p = 0;
                                    p = (float)((double)p + (double)L->c[q][0] * uw[0]);
                                    p = (float)((double)p + (double)L->c[q][1] * uw[1]);
                                    p = (float)((double)p + (double)L->c[q][2] * uw[2]);
                                    p = (float)((double)p * (-6. * L->w(q)));
                                }
                                v = (double)p;
                            }
                            put(fout[lvl], storage, (size_t)q * n + i, v);
                            put(fin[lvl], storage, (size_t)q * n + i, v);
                        }
                    }
                    st[lvl][i] = (uint8_t)((st[lvl][i] & (MR_ACTIVE | MR_REFINED)) | (cls << MR_CLS_SHIFT));
                }
    }
    /* sumStore counts (init.h:8-55), level l -> l + 1: no test of the fine
     * voxel's children or of the coarse target's child */
    for (int l = 1; l < depth; l++) memset(ss[l], 0, (size_t)Q * nvox(dims + 3 * l) * sizeof(float));
    for (int l = 0; l + 1 < depth; l++) {
        const int32_t* d = dims + 3 * l;
        const int32_t* du = dims + 3 * (l + 1);
        const size_t nu = nvox(du);
        int32_t nbu[3];
        uint8_t* blk = memory_blocks(du, st[l + 1], nbu);
        for (int z = 0; z < d[2]; z++)
            for (int y = 0; y < d[1]; y++)
                for (int x = 0; x < d[0]; x++) {
                    if (!(st[l][lin(d, x, y, z)] & MR_ACTIVE)) continue;
                    const int px = x >> 1, py = y >> 1, pz = z >> 1;
                    if (!parent_block(nbu, blk, px, py, pz)) continue; /* hasParent */
                    for (int q = 0; q < Q; q++) {
                        if (q == L->rest) continue;
                        const int ux = uoff(x, L->c[q][0]), uy = uoff(y, L->c[q][1]), uz = uoff(z, L->c[q][2]);
                        /* the same-level voxel along the uncle offset: a proxy
                         * for "the uncle's region is not refined" */
                        if (act(d, st[l], x + ux, y + uy, z + uz)) continue;
                        if (!act(du, st[l + 1], px + ux, py + uy, pz + uz)) continue;
                        const int cx = px + ux - L->c[q][0], cy = py + uy - L->c[q][1], cz = pz + uz - L->c[q][2];
                        if (!act(du, st[l + 1], cx, cy, cz)) continue;
                        ss[l + 1][(size_t)q * nu + lin(du, cx, cy, cz)] += 1.f;
                    }
                }
        free(blk);
    }
    /* reciprocal (init.h:58-80): 1.f / (refFactor * ss) with refFactor of
     * the field's type T */
    for (int l = 1; l < depth; l++)
        for (size_t i = 0; i < (size_t)Q * nvox(dims + 3 * l); i++)
            if ((double)ss[l][i] > 0.001) {
                if (storage == OR_F64)
                    ss[l][i] = (float)(1.0 / (2.0 * (double)ss[l][i]));
                else
                    ss[l][i] = 1.f / (2.0f * ss[l][i]);
            }
}

/* computeOmega (util.h:40-45) */
double or_mres_omega(double omega0, int level, int num_levels)
{
    const int il = num_levels - level - 1;
    return 2 * omega0 / (scalbln(1.0, il + 1) + (1. - scalbln(1.0, il)) * omega0);
}

/* ---- step --------------------------------------------------------------- */

/* BGK D3Q19 of collide.h:8-149 on voxel i (n voxels per population) in
 * T = double / float: the moment sums and the collision in the cavity
 * benchmark's association order, all constants of type T. */
static void bgk19_voxel(int s, size_t n, size_t i, const void* fin, void* fout, double omega)
{
    if (s == OR_F64) {
        double f[19], o[19];
        for (int q = 0; q < 19; q++) f[q] = ((const double*)fin)[(size_t)q * n + i];
        const double XM1 = f[0] + f[3] + f[4] + f[5] + f[6];
        const double XP1 = f[10] + f[13] + f[14] + f[15] + f[16];
        const double X0 = f[9] + f[1] + f[2] + f[7] + f[8] + f[11] + f[12] + f[17] + f[18];
        const double YM1 = f[1] + f[3] + f[7] + f[8] + f[14];
        const double YP1 = f[4] + f[11] + f[13] + f[17] + f[18];
        const double ZM1 = f[2] + f[5] + f[7] + f[16] + f[18];
        const double ZP1 = f[6] + f[8] + f[12] + f[15] + f[17];
        const double rho = XM1 + XP1 + X0;
        const double u0 = (XP1 - XM1) / rho, u1 = (YP1 - YM1) / rho, u2 = (ZP1 - ZM1) / rho;
        const double usqr = 1.5 * (u0 * u0 + u1 * u1 + u2 * u2);
        const double ck[9] = {u0, u1, u2, u0 + u1, u0 - u1, u0 + u2, u0 - u2, u1 + u2, u1 - u2};
        for (int k = 0; k < 9; k++) {
            const double w = k < 3 ? 1. / 18. : 1. / 36.;
            const double eq = rho * w * (1. - 3. * ck[k] + 4.5 * ck[k] * ck[k] - usqr);
            const double eqo = eq + rho * w * 6. * ck[k];
            o[k] = (1. - omega) * f[k] + omega * eq;
            o[k + 10] = (1. - omega) * f[k + 10] + omega * eqo;
        }
        o[9] = (1. - omega) * f[9] + omega * (rho * (1. / 3.) * (1. - usqr));
        for (int q = 0; q < 19; q++) ((double*)fout)[(size_t)q * n + i] = o[q];
    } else {
        const float om = (float)omega;
        float f[19], o[19];
        for (int q = 0; q < 19; q++) f[q] = ((const float*)fin)[(size_t)q * n + i];
        const float XM1 = f[0] + f[3] + f[4] + f[5] + f[6];
        const float XP1 = f[10] + f[13] + f[14] + f[15] + f[16];
        const float X0 = f[9] + f[1] + f[2] + f[7] + f[8] + f[11] + f[12] + f[17] + f[18];
        const float YM1 = f[1] + f[3] + f[7] + f[8] + f[14];
        const float YP1 = f[4] + f[11] + f[13] + f[17] + f[18];
        const float ZM1 = f[2] + f[5] + f[7] + f[16] + f[18];
        const float ZP1 = f[6] + f[8] + f[12] + f[15] + f[17];
        const float rho = XM1 + XP1 + X0;
        const float u0 = (XP1 - XM1) / rho, u1 = (YP1 - YM1) / rho, u2 = (ZP1 - ZM1) / rho;
        const float usqr = 1.5f * (u0 * u0 + u1 * u1 + u2 * u2);
        const float ck[9] = {u0, u1, u2, u0 + u1, u0 - u1, u0 + u2, u0 - u2, u1 + u2, u1 - u2};
        for (int k = 0; k < 9; k++) {
            const float w = k < 3 ? (float)(1. / 18.) : (float)(1. / 36.);
            const float eq = rho * w * (1.f - 3.f * ck[k] + 4.5f * ck[k] * ck[k] - usqr);
            const float eqo = eq + rho * w * 6.f * ck[k];
            o[k] = (1.f - om) * f[k] + om * eq;
            o[k + 10] = (1.f - om) * f[k + 10] + om * eqo;
        }
        o[9] = (1.f - om) * f[9] + om * (rho * (float)(1. / 3.) * (1.f - usqr));
        for (int q = 0; q < 19; q++) ((float*)fout)[(size_t)q * n + i] = o[q];
    }
}

/* KBC D3Q27 of collide.h:148-284 on voxel i, for T = double and float with
 * the reference's type promotions: velocity() sums f c over q in order
 * (util.h:47-62), the equilibrium and the shear split evaluate their double
 * literals in binary64 and store in T, e0 / e1 / gamma / out follow the
 * source's expressions term by term. */
#define KBC_BODY(T)                                                                                      \
    do {                                                                                                 \
        const T* fi = (const T*)fin;                                                                     \
        T* fo = (T*)fout;                                                                                \
        const T om = (T)omega;                                                                           \
        const T beta = (T)(om * 0.5);                                                                    \
        const T invBeta = (T)(1.0 / beta);                                                               \
        const T tiny = (T)1e-7;                                                                          \
        T ins[27], feq[27], fneq[27], dS[27];                                                            \
        for (int q = 0; q < 27; q++) ins[q] = fi[(size_t)q * n + i];                                     \
        T rho = 0;                                                                                       \
        for (int q = 0; q < 27; q++) rho += ins[q];                                                      \
        T v[3] = {0, 0, 0};                                                                              \
        for (int q = 0; q < 27; q++)                                                                     \
            for (int d = 0; d < 3; d++) v[d] += ins[q] * (T)KC[q][d];                                    \
        for (int d = 0; d < 3; d++) v[d] /= rho;                                                         \
        const T usqr = (T)((3.0 / 2.0) * (double)(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));            \
        for (int q = 0; q < 27; q++) {                                                                   \
            T cu = 0;                                                                                    \
            for (int d = 0; d < 3; d++) cu += (T)KC[q][d] * v[d];                                        \
            cu = (T)((double)cu * 3.0);                                                                  \
            feq[q] = (T)((double)rho * kw(q) *                                                           \
                         (1. + (double)cu + 0.5 * (double)cu * (double)cu - (double)usqr));              \
            fneq[q] = ins[q] - feq[q];                                                                   \
        }                                                                                                \
        T Pi[6] = {0, 0, 0, 0, 0, 0};                                                                    \
        for (int q = 0; q < 27; q++) {                                                                   \
            const int m[6] = {KC[q][0] * KC[q][0], KC[q][0] * KC[q][1], KC[q][0] * KC[q][2],             \
                              KC[q][1] * KC[q][1], KC[q][1] * KC[q][2], KC[q][2] * KC[q][2]};            \
            for (int k = 0; k < 6; k++) Pi[k] += fneq[q] * (T)m[k];                                      \
        }                                                                                                \
        const T Nxz = Pi[0] - Pi[5], Nyz = Pi[3] - Pi[5];                                                \
        T e0 = 0, e1 = 0;                                                                                \
        for (int q = 0; q < 27; q++) {                                                                   \
            T sh;                                                                                        \
            const int cx = KC[q][0], cy = KC[q][1], cz = KC[q][2];                                       \
            const int nz = (cx != 0) + (cy != 0) + (cz != 0);                                            \
            if (nz == 1 && cx)                                                                           \
                sh = (T)((2.0 * (double)Nxz - (double)Nyz) / 6.0);                                       \
            else if (nz == 1 && cy)                                                                      \
                sh = (T)((-(double)Nxz + 2.0 * (double)Nyz) / 6.0);                                      \
            else if (nz == 1 && cz)                                                                      \
                sh = (T)((double)(-Nxz - Nyz) / 6.0); /* a T subtraction */                                \
            else if (nz == 2 && !cz)                                                                     \
                sh = (T)((double)(cx * cy > 0 ? Pi[1] : -Pi[1]) / 4.0);                                  \
            else if (nz == 2 && !cy)                                                                     \
                sh = (T)((double)(cx * cz > 0 ? Pi[2] : -Pi[2]) / 4.0);                                  \
            else if (nz == 2 && !cx)                                                                     \
                sh = (T)((double)(cy * cz > 0 ? Pi[4] : -Pi[4]) / 4.0);                                  \
            else                                                                                         \
                sh = (T)0;                                                                               \
            dS[q] = rho * sh;                                                                            \
            const T dH = fneq[q] - dS[q];                                                                \
            e0 += (dS[q] * dH / feq[q]);                                                                 \
            e1 += (dH * dH / feq[q]);                                                                    \
        }                                                                                                \
        const T gamma = (T)((double)invBeta - (2.0 - (double)invBeta) * (double)e0 / (double)(tiny + e1)); \
        for (int q = 0; q < 27; q++) {                                                                   \
            const T dH = fneq[q] - dS[q];                                                                \
            fo[(size_t)q * n + i] =                                                                      \
                (T)((double)ins[q] - (double)beta * (2.0 * (double)dS[q] + (double)(gamma * dH)));      \
        }                                                                                                \
    } while (0)

static void kbc27_voxel(int s, size_t n, size_t i, const void* fin, void* fout, double omega)
{
    if (s == OR_F64)
        KBC_BODY(double);
    else
        KBC_BODY(float);
}

/* collide.h:8-284 for every bulk active voxel of a level: the lattice's
 * collision on leaves; refined voxels of levels above 0 reset their fout
 * (the accumulator of storeFine) to 0. */
static void collide_level(const mr_lat* L, const int32_t d[3], int lvl, int s, const uint8_t* st, const void* fin,
                          void* fout, double omega)
{
    const size_t n = nvox(d);
#pragma omp parallel for schedule(static) num_threads(or_get_threads())
    for (int z = 0; z < d[2]; z++)
        for (int y = 0; y < d[1]; y++)
            for (int x = 0; x < d[0]; x++) {
                const size_t i = lin(d, x, y, z);
                const uint8_t sv = st[i];
                if (!(sv & MR_ACTIVE) || ((sv >> MR_CLS_SHIFT) & 3) != OR_BULK) continue;
                if (sv & MR_REFINED) {
                    if (lvl != 0)
                        for (int q = 0; q < L->Q; q++) put(fout, s, (size_t)q * n + i, 0.);
                    continue;
                }
                if (L->Q == 27)
                    kbc27_voxel(s, n, i, fin, fout, omega);
                else
                    bgk19_voxel(s, n, i, fin, fout, omega);
            }
}

/* storeFine (store.h:98-163) from level l into level l + 1: every active
 * voxel of level l without children and with a parent adds, for each moving
 * q, its post-collision population q to the level-(l+1) voxel that the
 * coarse leaf downstream will pull it from.  Serial, voxels in z, y, x order
 * and q ascending: the order of the additions is the order of this loop
 * (the reference adds with atomics in an unspecified order). */
static void store_fine(const mr_lat* L, const int32_t* dims, int l, int s, const uint8_t* const* st,
                       const void* fout_l, void* fout_u)
{
    const int32_t* d = dims + 3 * l;
    const int32_t* du = dims + 3 * (l + 1);
    const size_t n = nvox(d), nu = nvox(du);
    int32_t nbu[3];
    uint8_t* blk = memory_blocks(du, st[l + 1], nbu);
    for (int z = 0; z < d[2]; z++)
        for (int y = 0; y < d[1]; y++)
            for (int x = 0; x < d[0]; x++) {
                const size_t i = lin(d, x, y, z);
                if (!(st[l][i] & MR_ACTIVE) || (st[l][i] & MR_REFINED)) continue;
                const int px = x >> 1, py = y >> 1, pz = z >> 1;
                if (!parent_block(nbu, blk, px, py, pz)) continue; /* hasParent */
                for (int q = 0; q < L->Q; q++) {
                    if (q == L->rest) continue;
                    const int* c = L->c[q];
                    const int ux = uoff(x, c[0]), uy = uoff(y, c[1]), uz = uoff(z, c[2]);
                    if (act(d, st[l], x + ux, y + uy, z + uz)) continue;
                    if (!act(du, st[l + 1], px + ux, py + uy, pz + uz)) continue; /* uncle */
                    const int sx = ux - c[0], sy = uy - c[1], sz = uz - c[2];
                    if (!act(du, st[l + 1], px + sx, py + sy, pz + sz)) continue; /* cs */
                    if (!act(d, st[l], x + sx, y + sy, z + sz)) continue;         /* csChild */
                    const size_t t = (size_t)q * nu + lin(du, px + sx, py + sy, pz + sz);
                    const double v = get(fout_l, s, (size_t)q * n + i);
                    if (s == OR_F64)
                        ((double*)fout_u)[t] += v;
                    else
                        ((float*)fout_u)[t] += (float)v;
                }
            }
    free(blk);
}

/* Stream fused with coalescence and explosion (stream.h:170-225), pull,
 * for the bulk leaf voxels of level l: from a same-level neighbour without
 * children (bounce-back with the own opposite population when it is not
 * bulk); from the next level's neighbour (explosion) when the neighbour is
 * absent and the voxel has a parent; from a neighbour with children
 * (coalescence) scaled by its sumStore weight. */
static void stream_level(const mr_lat* L, const int32_t* dims, int depth, int l, int s, const uint8_t* const* st,
                         const void* fout_l, const void* fout_up, const float* ss, void* fin_l)
{
    const int32_t* d = dims + 3 * l;
    const size_t n = nvox(d);
    const int top = l == depth - 1;
    const int32_t* du = top ? NULL : dims + 3 * (l + 1);
    const size_t nu = top ? 0 : nvox(du);
    int32_t nbu[3] = {0, 0, 0};
    uint8_t* blk = top ? NULL : memory_blocks(du, st[l + 1], nbu);
#pragma omp parallel for schedule(static) num_threads(or_get_threads())
    for (int z = 0; z < d[2]; z++)
        for (int y = 0; y < d[1]; y++)
            for (int x = 0; x < d[0]; x++) {
                const size_t i = lin(d, x, y, z);
                const uint8_t sv = st[l][i];
                if (!(sv & MR_ACTIVE) || ((sv >> MR_CLS_SHIFT) & 3) != OR_BULK || (sv & MR_REFINED)) continue;
                const int has_parent = !top && parent_block(nbu, blk, x >> 1, y >> 1, z >> 1);
                for (int q = 0; q < L->Q; q++) {
                    const int dx = -L->c[q][0], dy = -L->c[q][1], dz = -L->c[q][2];
                    const int nx = x + dx, ny = y + dy, nz = z + dz;
                    const int valid = act(d, st[l], nx, ny, nz);
                    const size_t j = valid ? lin(d, nx, ny, nz) : 0;
                    const int ngh_refined = valid && (st[l][j] & MR_REFINED);
                    if (!ngh_refined) {
                        if (valid) {
                            if (((st[l][j] >> MR_CLS_SHIFT) & 3) == OR_BULK) {
                                put(fin_l, s, (size_t)q * n + i, get(fout_l, s, (size_t)q * n + j));
                            } else {
                                const int o = L->opp[q];
                                if (s == OR_F64)
                                    ((double*)fin_l)[(size_t)q * n + i] =
                                        ((const double*)fout_l)[(size_t)o * n + i] +
                                        ((const double*)fout_l)[(size_t)o * n + j];
                                else
                                    ((float*)fin_l)[(size_t)q * n + i] =
                                        ((const float*)fout_l)[(size_t)o * n + i] +
                                        ((const float*)fout_l)[(size_t)o * n + j];
                            }
                        } else if (has_parent && q != L->rest) {
                            const int ux = (x >> 1) + uoff(x, dx), uy = (y >> 1) + uoff(y, dy),
                                      uz = (z >> 1) + uoff(z, dz);
                            if (act(du, st[l + 1], ux, uy, uz))
                                put(fin_l, s, (size_t)q * n + i,
                                    get(fout_up, s, (size_t)q * nu + lin(du, ux, uy, uz)));
                        }
                    } else if (q != L->rest) {
                        const float w = ss[(size_t)q * n + j];
                        if (s == OR_F64)
                            ((double*)fin_l)[(size_t)q * n + i] =
                                ((const double*)fout_l)[(size_t)q * n + j] * (double)w;
                        else
                            ((float*)fin_l)[(size_t)q * n + i] = ((const float*)fout_l)[(size_t)q * n + j] * w;
                    }
                }
            }
    free(blk);
}

typedef struct {
    const mr_lat* L;
    const int32_t* dims;
    int depth, s;
    const uint8_t* const* st;
    void* const* fin;
    void* const* fout;
    const float* const* ss;
    double omega0;
} mr_ctx;

static void collide_store(const mr_ctx* c, int l)
{
    collide_level(c->L, c->dims + 3 * l, l, c->s, c->st[l], c->fin[l], c->fout[l],
                  or_mres_omega(c->omega0, l, c->depth));
    if (l != c->depth - 1) store_fine(c->L, c->dims, l, c->s, c->st, c->fout[l], c->fout[l + 1]);
}

static void stream_of(const mr_ctx* c, int l)
{
    stream_level(c->L, c->dims, c->depth, l, c->s, c->st, c->fout[l], l + 1 < c->depth ? c->fout[l + 1] : NULL,
                 c->ss[l], c->fin[l]);
}

/* nonUniformTimestepRecursive (lbmMultiRes.h:141-263) with the fine store
 * and the stream fused with coalescence and explosion, unfused collision */
static void recurse(const mr_ctx* c, int l)
{
    collide_store(c, l);
    if (l != 0) recurse(c, l - 1);
    stream_of(c, l);
    if (l == c->depth - 1) return;
    collide_store(c, l);
    if (l != 0) recurse(c, l - 1);
    stream_of(c, l);
}

/* One step of the coarsest level; omega0 is its relaxation. */
void or_mres_step(const int32_t dim[3], int depth, int lat, int s, const uint8_t* const* st, void* const* fin,
                  void* const* fout, const float* const* ss, double omega0)
{
    int32_t dims[3 * OR_MRES_MAX_DEPTH];
    or_mres_dims(dim, depth, dims);
    const mr_ctx c = {lat_of(lat), dims, depth, s, st, fin, fout, ss, omega0};
    recurse(&c, depth - 1);
}

/* The KBC D3Q27 collision of one voxel alone (for its closed-form checks):
 * fin, fout hold 27 values of the storage type. */
void or_kbc27_collide(int storage, const void* fin, void* fout, double omega)
{
    kbc27_voxel(storage, 1, 0, fin, fout, omega);
}
