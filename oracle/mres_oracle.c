/* mres_oracle.c -- CPU restatement of Neon's two-level grid-refinement LBM
 * (apps/lbmMultiRes, BGK D3Q19 variant) on a stack of two levels.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the library's nx_mgrid_* path.
 * Nothing in the product links or calls this file.
 *
 * Level 0 is the finest level (the domain dim[3]), level 1 has half the
 * resolution (a level-1 voxel v covers the level-0 voxels 2v + {0,1}^3).
 * Storage is dense per level, SoA [q][z][y][x], one byte of state per voxel:
 * active, refined (has children), class.  Every function follows the
 * reference operation by operation, in the reference's loop order where the
 * result depends on it (the fine-to-coarse accumulation).
 *
 * Reference files restated:
 *   grid:     libNeonDomain/src/domain/details/mGrid/mGrid.cpp:89-264
 *             (activation per 2^3 group, parent marking, overlap culling;
 *             the strong-balance pass, :266-345, cannot change a two-level
 *             grid) and :466-555 (a voxel has children iff one of its eight
 *             children is active)
 *   neighbours: mPartition_imp.h:186-235 (parent, uncle),
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

void or_mres_dims(const int32_t dim[3], int32_t dim1[3])
{
    for (int a = 0; a < 3; a++) dim1[a] = (dim[a] + 1) / 2;
}

/* mGrid constructor for depth 2 (mGrid.cpp:89-264).  lam0: the level-0
 * activation lambda at every level-0 voxel; lam1: the level-1 lambda at
 * every level-1 voxel (evaluated by the caller at the voxel's level-0
 * origin 2v).  Outputs the state byte of every voxel of both levels. */
void or_mres_build(const int32_t dim[3], const uint8_t* lam0, const uint8_t* lam1, uint8_t* st0,
                   uint8_t* st1)
{
    int32_t d1[3];
    or_mres_dims(dim, d1);
    const size_t n0 = (size_t)dim[0] * dim[1] * dim[2], n1 = (size_t)d1[0] * d1[1] * d1[2];
    memset(st0, 0, n0);
    memset(st1, 0, n1);
    /* level 0: groups of 2^3 voxels (one level-1 voxel each, :89-156): a
     * group with an active voxel activates all of its voxels inside the
     * domain and its parent */
    for (int gz = 0; gz < d1[2]; gz++)
        for (int gy = 0; gy < d1[1]; gy++)
            for (int gx = 0; gx < d1[0]; gx++) {
                int any = 0;
                for (int k = 0; k < 8; k++) {
                    const int x = 2 * gx + (k & 1), y = 2 * gy + ((k >> 1) & 1), z = 2 * gz + (k >> 2);
                    if (inside(dim, x, y, z) && (st0[lin(dim, x, y, z)] || lam0[lin(dim, x, y, z)])) any = 1;
                }
                if (!any) continue;
                for (int k = 0; k < 8; k++) {
                    const int x = 2 * gx + (k & 1), y = 2 * gy + ((k >> 1) & 1), z = 2 * gz + (k >> 2);
                    if (inside(dim, x, y, z)) st0[lin(dim, x, y, z)] = MR_ACTIVE;
                }
                st1[lin(d1, gx, gy, gz)] = MR_ACTIVE;
            }
    /* level 1 (the top level): groups of 2^3 level-1 voxels; a voxel is in
     * the domain when its level-0 origin 2v is */
    const int32_t d2[3] = {(d1[0] + 1) / 2, (d1[1] + 1) / 2, (d1[2] + 1) / 2};
    for (int gz = 0; gz < d2[2]; gz++)
        for (int gy = 0; gy < d2[1]; gy++)
            for (int gx = 0; gx < d2[0]; gx++) {
                int any = 0;
                for (int k = 0; k < 8; k++) {
                    const int x = 2 * gx + (k & 1), y = 2 * gy + ((k >> 1) & 1), z = 2 * gz + (k >> 2);
                    if (inside(d1, x, y, z) && (st1[lin(d1, x, y, z)] || lam1[lin(d1, x, y, z)])) any = 1;
                }
                if (!any) continue;
                for (int k = 0; k < 8; k++) {
                    const int x = 2 * gx + (k & 1), y = 2 * gy + ((k >> 1) & 1), z = 2 * gz + (k >> 2);
                    if (inside(d1, x, y, z)) st1[lin(d1, x, y, z)] = MR_ACTIVE;
                }
            }
    /* a level-1 voxel is refined when one of its children is active
     * (mGrid.cpp:176-205; :497-545 for the child link) */
    uint8_t* refined = (uint8_t*)calloc(n1, 1);
    for (int z = 0; z < d1[2]; z++)
        for (int y = 0; y < d1[1]; y++)
            for (int x = 0; x < d1[0]; x++) {
                int r = 0;
                for (int k = 0; k < 8; k++) {
                    const int cx = 2 * x + (k & 1), cy = 2 * y + ((k >> 1) & 1), cz = 2 * z + (k >> 2);
                    if (inside(dim, cx, cy, cz) && (st0[lin(dim, cx, cy, cz)] & MR_ACTIVE)) r = 1;
                }
                refined[lin(d1, x, y, z)] = (uint8_t)r;
            }
    /* overlap culling (:209-263): an active refined voxel whose 26
     * neighbours inside the domain are all refined is removed */
    for (int z = 0; z < d1[2]; z++)
        for (int y = 0; y < d1[1]; y++)
            for (int x = 0; x < d1[0]; x++) {
                const size_t i = lin(d1, x, y, z);
                if (!(st1[i] & MR_ACTIVE) || !refined[i]) continue;
                int deactivate = 1;
                for (int k = -1; k <= 1; k++)
                    for (int j = -1; j <= 1; j++)
                        for (int l = -1; l <= 1; l++) {
                            if (!l && !j && !k) continue;
                            const int nx = x + l, ny = y + j, nz = z + k;
                            /* inside test on the level-0 origin of the neighbour */
                            if (2 * nx < 0 || 2 * ny < 0 || 2 * nz < 0 || 2 * nx >= dim[0] ||
                                2 * ny >= dim[1] || 2 * nz >= dim[2])
                                continue;
                            if (!refined[lin(d1, nx, ny, nz)]) deactivate = 0;
                        }
                if (deactivate) st1[i] = 0;
            }
    for (size_t i = 0; i < n1; i++)
        if ((st1[i] & MR_ACTIVE) && refined[i]) st1[i] |= MR_REFINED;
    free(refined);
}

/* ---- set-up ------------------------------------------------------------- */

/* uncleOffset (util.h:19-37): offset, from the parent of voxel x, of the
 * parent of x + q; per axis s = (x mod 2) + q -> s <= 0 ? s : s - 1 */
static inline int uoff(int x, int q)
{
    const int s = (x & 1) + q;
    return s <= 0 ? s : s - 1;
}

/* level-1 voxel active (getUncle's mask test, mPartition_imp.h:213-226) */
static inline int act1(const int32_t d1[3], const uint8_t* st1, int x, int y, int z)
{
    return inside(d1, x, y, z) && (st1[lin(d1, x, y, z)] & MR_ACTIVE);
}
static inline int act0(const int32_t d0[3], const uint8_t* st0, int x, int y, int z)
{
    return inside(d0, x, y, z) && (st0[lin(d0, x, y, z)] & MR_ACTIVE);
}

/* hasParent (mPartition_imp.h:199-208): the level-1 memory block (8^3
 * level-1 voxels, mPartition.h:13) that holds the parent voxel exists, i.e.
 * has an active voxel -- not that the parent voxel itself is active */
static int parent_block(const int32_t d1[3], const uint8_t* st1, int px, int py, int pz)
{
    const int bx = px >> 3, by = py >> 3, bz = pz >> 3;
    for (int z = 8 * bz; z < 8 * bz + 8; z++)
        for (int y = 8 * by; y < 8 * by + 8; y++)
            for (int x = 8 * bx; x < 8 * bx + 8; x++)
                if (act1(d1, st1, x, y, z)) return 1;
    return 0;
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
 * ulid = (ulb, 0, 0).  sumStore of level 1 after init.h:8-86: every level-0
 * voxel with a parent counts, for each q != 9, one towards the coarse voxel
 * that will later pull its population q, then ss = 1 / (2 count). */
void or_mres_setup(const int32_t dim[3], int storage, double ulb, uint8_t* st0, uint8_t* st1,
                   void* fin0, void* fout0, void* fin1, void* fout1, float* ss1)
{
    int32_t d1[3];
    or_mres_dims(dim, d1);
    const size_t n0 = (size_t)dim[0] * dim[1] * dim[2], n1 = (size_t)d1[0] * d1[1] * d1[2];
    for (int lvl = 0; lvl < 2; lvl++) {
        const int32_t* d = lvl ? d1 : dim;
        uint8_t* st = lvl ? st1 : st0;
        void* fi = lvl ? fin1 : fin0;
        void* fo = lvl ? fout1 : fout0;
        const size_t n = lvl ? n1 : n0;
        for (int z = 0; z < d[2]; z++)
            for (int y = 0; y < d[1]; y++)
                for (int x = 0; x < d[0]; x++) {
                    const size_t i = lin(d, x, y, z);
                    if (!(st[i] & MR_ACTIVE)) continue;
                    int cls = OR_BULK;
                    for (int q = 0; q < 19; q++) {
                        put(fi, storage, (size_t)q * n + i, 0.);
                        put(fo, storage, (size_t)q * n + i, 0.);
                    }
                    if (!(st[i] & MR_REFINED)) {
                        if (lvl == 0 && (x == 0 || x == d[0] - 1 || y == 0 || y == d[1] - 1 || z == 0 ||
                                         z == d[2] - 1)) {
                            cls = OR_BOUNCEBACK;
                            if (y == d[1] - 1) cls = OR_MOVINGWALL;
                        }
                        for (int q = 0; q < 19; q++) {
                            double v;
                            if (storage == OR_F64) {
                                double p = mw(q);
                                if (cls == OR_BOUNCEBACK) p = 0;
                                if (cls == OR_MOVINGWALL) {
                                    p = 0;
                                    p += (double)MC[q][0] * ulb;
                                    p += (double)MC[q][1] * 0.;
                                    p += (double)MC[q][2] * 0.;
                                    p *= -6. * mw(q);
                                }
                                v = p;
                            } else {
                                float p = (float)mw(q);
                                if (cls == OR_BOUNCEBACK) p = 0;
                                if (cls == OR_MOVINGWALL) {
                                    p = 0;
                                    p = (float)((double)p + (double)MC[q][0] * ulb);
                                    p = (float)((double)p + (double)MC[q][1] * 0.);
                                    p = (float)((double)p + (double)MC[q][2] * 0.);
                                    p = (float)((double)p * (-6. * mw(q)));
                                }
                                v = (double)p;
                            }
                            put(fo, storage, (size_t)q * n + i, v);
                            put(fi, storage, (size_t)q * n + i, v);
                        }
                    }
                    st[i] = (uint8_t)((st[i] & (MR_ACTIVE | MR_REFINED)) | (cls << MR_CLS_SHIFT));
                }
    }
    /* sumStore counts (init.h:8-55) */
    memset(ss1, 0, 19 * n1 * sizeof(float));
    for (int z = 0; z < dim[2]; z++)
        for (int y = 0; y < dim[1]; y++)
            for (int x = 0; x < dim[0]; x++) {
                if (!(st0[lin(dim, x, y, z)] & MR_ACTIVE)) continue;
                const int px = x >> 1, py = y >> 1, pz = z >> 1;
                if (!parent_block(d1, st1, px, py, pz)) continue; /* hasParent */
                for (int q = 0; q < 19; q++) {
                    if (q == 9) continue;
                    const int ux = uoff(x, MC[q][0]), uy = uoff(y, MC[q][1]), uz = uoff(z, MC[q][2]);
                    /* the same-level voxel along the uncle offset: a proxy for
                     * "the uncle's region is not refined" */
                    if (act0(dim, st0, x + ux, y + uy, z + uz)) continue;
                    if (!act1(d1, st1, px + ux, py + uy, pz + uz)) continue;
                    const int cx = px + ux - MC[q][0], cy = py + uy - MC[q][1], cz = pz + uz - MC[q][2];
                    if (!act1(d1, st1, cx, cy, cz)) continue;
                    ss1[(size_t)q * n1 + lin(d1, cx, cy, cz)] += 1.f;
                }
            }
    /* reciprocal (init.h:58-80), level 1: 1.f / (refFactor * ss) with
     * refFactor of the field's type T */
    for (size_t i = 0; i < 19 * n1; i++)
        if ((double)ss1[i] > 0.001) {
            if (storage == OR_F64)
                ss1[i] = (float)(1.0 / (2.0 * (double)ss1[i]));
            else
                ss1[i] = 1.f / (2.0f * ss1[i]);
        }
}

/* computeOmega (util.h:40-45) */
double or_mres_omega(double omega0, int level, int num_levels)
{
    const int il = num_levels - level - 1;
    return 2 * omega0 / (scalbln(1.0, il + 1) + (1. - scalbln(1.0, il)) * omega0);
}

/* ---- step --------------------------------------------------------------- */

/* BGK of collide.h:8-149 in T = double / float: the moment sums and the
 * collision in the cavity benchmark's association order, all constants of
 * type T.  Refined voxels of level 1 reset their fout (the accumulator of
 * storeFine) to 0. */
static void collide_level(const int32_t d[3], int lvl, int s, const uint8_t* st, const void* fin,
                          void* fout, double omega)
{
    const size_t n = (size_t)d[0] * d[1] * d[2];
#pragma omp parallel for schedule(static) num_threads(or_get_threads())
    for (int z = 0; z < d[2]; z++)
        for (int y = 0; y < d[1]; y++)
            for (int x = 0; x < d[0]; x++) {
                const size_t i = lin(d, x, y, z);
                const uint8_t sv = st[i];
                if (!(sv & MR_ACTIVE) || ((sv >> MR_CLS_SHIFT) & 3) != OR_BULK) continue;
                if (sv & MR_REFINED) {
                    if (lvl != 0)
                        for (int q = 0; q < 19; q++) put(fout, s, (size_t)q * n + i, 0.);
                    continue;
                }
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
}

/* storeFine (store.h:98-163): every level-0 voxel with a parent adds, for
 * each q != 9, its post-collision population q to the level-1 voxel that
 * the coarse leaf downstream will pull it from.  Serial, voxels in z, y, x
 * order and q ascending: the order of the additions is the order of this
 * loop (the reference adds with atomics in an unspecified order). */
static void store_fine(const int32_t dim[3], int s, const uint8_t* st0, const uint8_t* st1,
                       const void* fout0, void* fout1)
{
    int32_t d1[3];
    or_mres_dims(dim, d1);
    const size_t n0 = (size_t)dim[0] * dim[1] * dim[2], n1 = (size_t)d1[0] * d1[1] * d1[2];
    for (int z = 0; z < dim[2]; z++)
        for (int y = 0; y < dim[1]; y++)
            for (int x = 0; x < dim[0]; x++) {
                const size_t i = lin(dim, x, y, z);
                if (!(st0[i] & MR_ACTIVE)) continue;
                const int px = x >> 1, py = y >> 1, pz = z >> 1;
                if (!parent_block(d1, st1, px, py, pz)) continue; /* hasParent */
                for (int q = 0; q < 19; q++) {
                    if (q == 9) continue;
                    const int ux = uoff(x, MC[q][0]), uy = uoff(y, MC[q][1]), uz = uoff(z, MC[q][2]);
                    if (act0(dim, st0, x + ux, y + uy, z + uz)) continue;
                    if (!act1(d1, st1, px + ux, py + uy, pz + uz)) continue; /* uncle */
                    const int sx = ux - MC[q][0], sy = uy - MC[q][1], sz = uz - MC[q][2];
                    if (!act1(d1, st1, px + sx, py + sy, pz + sz)) continue; /* cs */
                    if (!act0(dim, st0, x + sx, y + sy, z + sz)) continue;   /* csChild */
                    const size_t t = (size_t)q * n1 + lin(d1, px + sx, py + sy, pz + sz);
                    const double v = get(fout0, s, (size_t)q * n0 + i);
                    if (s == OR_F64)
                        ((double*)fout1)[t] += v;
                    else
                        ((float*)fout1)[t] += (float)v;
                }
            }
}

/* Stream fused with coalescence and explosion (stream.h:170-225), pull,
 * for the bulk leaf voxels of one level: from a same-level neighbour that
 * is not refined (bounce-back with the own opposite population when it is
 * not bulk); from the parent level's neighbour (explosion) when the
 * neighbour is absent and the voxel has a parent; from a refined neighbour
 * (coalescence) scaled by its sumStore weight. */
static void stream_level(const int32_t dim[3], int lvl, int s, const uint8_t* st0, const uint8_t* st1,
                         const void* fout_l, const void* fout_up, const float* ss, void* fin_l)
{
    int32_t d1[3];
    or_mres_dims(dim, d1);
    const int32_t* d = lvl ? d1 : dim;
    const uint8_t* st = lvl ? st1 : st0;
    const size_t n = (size_t)d[0] * d[1] * d[2], n1 = (size_t)d1[0] * d1[1] * d1[2];
#pragma omp parallel for schedule(static) num_threads(or_get_threads())
    for (int z = 0; z < d[2]; z++)
        for (int y = 0; y < d[1]; y++)
            for (int x = 0; x < d[0]; x++) {
                const size_t i = lin(d, x, y, z);
                const uint8_t sv = st[i];
                if (!(sv & MR_ACTIVE) || ((sv >> MR_CLS_SHIFT) & 3) != OR_BULK || (sv & MR_REFINED)) continue;
                const int has_parent = lvl == 0 && parent_block(d1, st1, x >> 1, y >> 1, z >> 1);
                for (int q = 0; q < 19; q++) {
                    const int dx = -MC[q][0], dy = -MC[q][1], dz = -MC[q][2];
                    const int nx = x + dx, ny = y + dy, nz = z + dz;
                    const int valid = inside(d, nx, ny, nz) && (st[lin(d, nx, ny, nz)] & MR_ACTIVE);
                    const size_t j = valid ? lin(d, nx, ny, nz) : 0;
                    const int ngh_refined = valid && (st[j] & MR_REFINED);
                    if (!ngh_refined) {
                        if (valid) {
                            if (((st[j] >> MR_CLS_SHIFT) & 3) == OR_BULK) {
                                put(fin_l, s, (size_t)q * n + i, get(fout_l, s, (size_t)q * n + j));
                            } else {
                                const int o = MOPP[q];
                                if (s == OR_F64)
                                    ((double*)fin_l)[(size_t)q * n + i] =
                                        ((const double*)fout_l)[(size_t)o * n + i] +
                                        ((const double*)fout_l)[(size_t)o * n + j];
                                else
                                    ((float*)fin_l)[(size_t)q * n + i] =
                                        ((const float*)fout_l)[(size_t)o * n + i] +
                                        ((const float*)fout_l)[(size_t)o * n + j];
                            }
                        } else if (has_parent && q != 9) {
                            const int ux = (x >> 1) + uoff(x, dx), uy = (y >> 1) + uoff(y, dy),
                                      uz = (z >> 1) + uoff(z, dz);
                            if (act1(d1, st1, ux, uy, uz))
                                put(fin_l, s, (size_t)q * n + i,
                                    get(fout_up, s, (size_t)q * n1 + lin(d1, ux, uy, uz)));
                        }
                    } else if (q != 9) {
                        const float w = ss[(size_t)q * n + j];
                        if (s == OR_F64)
                            ((double*)fin_l)[(size_t)q * n + i] =
                                ((const double*)fout_l)[(size_t)q * n + j] * (double)w;
                        else
                            ((float*)fin_l)[(size_t)q * n + i] = ((const float*)fout_l)[(size_t)q * n + j] * w;
                    }
                }
            }
}

/* One step of the coarse level (nonUniformTimestepRecursive for two
 * levels, lbmMultiRes.h:141-263, baseline operation order with the fine
 * store): collide level 1; twice (collide level 0, store level 0, stream
 * level 0); stream level 1.  omega0 is the coarse level's relaxation. */
void or_mres_step(const int32_t dim[3], int s, const uint8_t* st0, const uint8_t* st1, void* fin0,
                  void* fout0, void* fin1, void* fout1, const float* ss1, double omega0)
{
    int32_t d1[3];
    or_mres_dims(dim, d1);
    const double om1 = or_mres_omega(omega0, 1, 2), om0 = or_mres_omega(omega0, 0, 2);
    collide_level(d1, 1, s, st1, fin1, fout1, om1);
    for (int sub = 0; sub < 2; sub++) {
        collide_level(dim, 0, s, st0, fin0, fout0, om0);
        store_fine(dim, s, st0, st1, fout0, fout1);
        stream_level(dim, 0, s, st0, st1, fout0, fout1, NULL, fin0);
    }
    stream_level(dim, 1, s, st0, st1, fout1, NULL, ss1, fin1);
}
