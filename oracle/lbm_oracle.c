/*
 * lbm_oracle.c -- CPU oracle for the D3Q19 lattice-Boltzmann step.
 *
 * TEST INFRASTRUCTURE ONLY: tests/ use it as the checker, bench.py times it
 * as the CPU baseline.  It restates the reference's arithmetic operation
 * by operation (same association order, same literal promotions), so that
 * a binary64 GPU kernel compiled without FMA contraction matches it bit
 * for bit.  Reference citations are into Autodesk/Neon:
 *   lattice         benchmarks/lbm-lid-driven-cavity-flow/src/D3Q19.h:23-132
 *   pull stream     .../src/LbmTools.h:78-168
 *   moments         .../src/LbmTools.h:171-195
 *   BGK collision   .../src/LbmTools.h:198-282, driver :284-325
 *   wall mask       .../src/LbmTools.h:327-377
 *   rho/u export    .../src/LbmTools.h:379-437
 *   cavity set-up   .../src/RunCavityTwoPop.cu:159-242
 *   dGrid layout    libNeonDomain/include/Neon/domain/details/dGrid/dField_imp.h:67-87
 *   bGrid layout    .../details/bGrid/bPartition_imp.h:95-124, 218-309
 *   bGrid builder   .../details/bGrid/bGrid_imp.h:19-246,
 *                   .../tools/partitioning/SpanClassifier.h:170-236
 * Project-defined parts with no reference counterpart (marked below):
 * the value of an invalid neighbour on a wall link (0), periodic wrap,
 * Guo body forcing, deviation storage, binary16 rounding.
 * Zou-He boundary cells follow
 *   benchmarks/lbm-flow-over-sphere/src/LbmContainers.h:145-287 (streamAndZouhe),
 *   :351-443 (collideBgkUnrolled with the given rho, u), tables :633-646;
 * the velocity variant is the textbook one (the reference reads its
 * prescribed velocity as a density there, :128-133 and :230).
 */
#include "lbm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* D3Q19.h:23-44 (velocities), :62-82 (opposites), :112-132 (weights) */
static const int32_t C[19][3] = {
    {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}, {-1, -1, 0}, {-1, 1, 0}, {-1, 0, -1}, {-1, 0, 1},
    {0, -1, -1}, {0, -1, 1}, {0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0},
    {1, -1, 0}, {1, 0, 1}, {1, 0, -1}, {0, 1, 1}, {0, 1, -1}};
static const double W[19] = {1. / 18., 1. / 18., 1. / 18., 1. / 36., 1. / 36., 1. / 36., 1. / 36.,
                             1. / 36., 1. / 36., 1. / 3.,  1. / 18., 1. / 18., 1. / 18., 1. / 36.,
                             1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36.};
static const int32_t OPP[19] = {10, 11, 12, 13, 14, 15, 16, 17, 18, 9,
                                0,  1,  2,  3,  4,  5,  6,  7,  8};

void or_lattice(int32_t c[19][3], double w[19], int32_t opp[19])
{
    for (int q = 0; q < 19; q++) {
        for (int a = 0; a < 3; a++) c[q][a] = C[q][a];
        w[q] = W[q];
        opp[q] = OPP[q];
    }
}

static int g_threads = 0;
void or_set_threads(int n) { g_threads = n; }
int or_get_threads(void)
{
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}

/* ---- binary16, from the IEEE 754 definition (round to nearest even) ---- */
uint16_t or_f32_to_f16(float f)
{
    uint32_t x;
    memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t ax = x & 0x7FFFFFFFu;
    if (ax >= 0x7F800000u) /* inf / nan */
        return (uint16_t)(sign | 0x7C00u | (ax > 0x7F800000u ? 0x200u | ((ax >> 13) & 0x3FFu) : 0u));
    if (ax >= 0x477FF000u) /* >= 65520 rounds to inf */
        return (uint16_t)(sign | 0x7C00u);
    if (ax < 0x38800000u) { /* below 2^-14: half subnormal or zero */
        if (ax <= 0x33000000u) return (uint16_t)sign; /* <= 2^-25 (tie to even 0) */
        const uint32_t e = ax >> 23;
        const uint32_t m = (ax & 0x7FFFFFu) | 0x800000u;
        const int shift = 126 - (int)e; /* value * 2^24 = m * 2^(e-126) */
        uint32_t q = m >> shift;
        const uint32_t rem = m & ((1u << shift) - 1u);
        const uint32_t half = 1u << (shift - 1);
        if (rem > half || (rem == half && (q & 1u))) q++;
        return (uint16_t)(sign | q);
    }
    uint32_t q = (ax - 0x38000000u) >> 13;
    const uint32_t rem = ax & 0x1FFFu;
    if (rem > 0x1000u || (rem == 0x1000u && (q & 1u))) q++;
    return (uint16_t)(sign | q);
}

float or_f16_to_f32(uint16_t h)
{
    const uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
    uint32_t e = ((uint32_t)h >> 10) & 0x1Fu;
    uint32_t m = (uint32_t)h & 0x3FFu;
    uint32_t x;
    if (e == 0) {
        if (m == 0) {
            x = sign;
        } else {
            e = 113;
            while (!(m & 0x400u)) {
                m <<= 1;
                e--;
            }
            m &= 0x3FFu;
            x = sign | (e << 23) | (m << 13);
        }
    } else if (e == 31) {
        x = sign | 0x7F800000u | (m << 13);
    } else {
        x = sign | ((e + 112u) << 23) | (m << 13);
    }
    float f;
    memcpy(&f, &x, 4);
    return f;
}

/* ---- storage access ----------------------------------------------------- */
static inline float wf(int q) { return (float)W[q]; }

/* stored value, exactly, as double */
static inline double ld(const void* p, int s, size_t i)
{
    switch (s) {
        case OR_F64: return ((const double*)p)[i];
        case OR_F32:
        case OR_F32DEV: return (double)((const float*)p)[i];
        default: return (double)or_f16_to_f32(((const uint16_t*)p)[i]);
    }
}

static inline int is_dev(int s) { return s == OR_F32DEV || s == OR_F16DEV; }

/* working value of a bulk cell's population: f for F64/F32 storage, the
 * stored deviation d = f - w_q for the deviation storages (project-defined:
 * their collision is written in d) */
static inline double dec_bulk(const void* p, int s, size_t i, int q)
{
    (void)q;
    return ld(p, s, i);
}

/* f of a bulk cell in binary64 (mass sums) */
static inline double dec_full(const void* p, int s, size_t i, int q)
{
    if (!is_dev(s)) return ld(p, s, i);
    return (double)wf(q) + ld(p, s, i);
}

/* own + neighbour on a wall link, added in the storage precision
 * (LbmTools.h:83-84 adds two LbmStoreType values) */
static inline double add_p(int s, double a, double b)
{
    if (s == OR_F64) return a + b;
    return (double)((float)a + (float)b);
}

/* store a working value (bulk) or f (wall) */
static inline void st(void* p, int s, size_t i, double v)
{
    switch (s) {
        case OR_F64: ((double*)p)[i] = v; break;
        case OR_F32:
        case OR_F32DEV: ((float*)p)[i] = (float)v; break;
        default: ((uint16_t*)p)[i] = or_f32_to_f16((float)v); break;
    }
}

/* store f of a bulk cell: deviation storages keep f - w_q, formed in
 * binary64 and rounded through binary32 */
static inline void st_full(void* p, int s, size_t i, int q, double f)
{
    st(p, s, i, is_dev(s) ? (double)(float)(f - W[q]) : f);
}

/* ---- macroscopic + BGK (LbmTools.h:171-282) ------------------------------ */
#define DEF_MACRO(NAME, R)                                                          \
    static inline void NAME(const R f[19], R* rho, R u[3])                          \
    {                                                                               \
        const R XM1 = f[0] + f[3] + f[4] + f[5] + f[6];                             \
        const R XP1 = f[10] + f[13] + f[14] + f[15] + f[16];                        \
        const R X0 = f[9] + f[1] + f[2] + f[7] + f[8] + f[11] + f[12] + f[17] + f[18]; \
        const R YM1 = f[1] + f[3] + f[7] + f[8] + f[14];                            \
        const R YP1 = f[4] + f[11] + f[13] + f[17] + f[18];                         \
        const R ZM1 = f[2] + f[5] + f[7] + f[16] + f[18];                           \
        const R ZP1 = f[6] + f[8] + f[12] + f[15] + f[17];                          \
        *rho = XM1 + XP1 + X0;                                                      \
        u[0] = (XP1 - XM1) / *rho;                                                  \
        u[1] = (YP1 - YM1) / *rho;                                                  \
        u[2] = (ZP1 - ZM1) / *rho;                                                  \
    }
DEF_MACRO(macro_d, double)
DEF_MACRO(macro_f, float)

/* binary64 collision: LbmTools.h:209-281 with LbmComputeType = double */
static void collide_f64(const double f[19], double omega, double out[19])
{
    double rho, u[3];
    macro_d(f, &rho, u);
    const double usqr = 1.5 * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    const double ck[9] = {u[0],        u[1],        u[2],        u[0] + u[1], u[0] - u[1],
                          u[0] + u[2], u[0] - u[2], u[1] + u[2], u[1] - u[2]};
    for (int i = 0; i < 9; i++) {
        const double wi = i < 3 ? (1. / 18.) : (1. / 36.);
        const double eq = rho * wi * (1. - 3. * ck[i] + 4.5 * ck[i] * ck[i] - usqr);
        const double eqopp = eq + rho * wi * 6. * ck[i];
        out[i] = (1. - omega) * f[i] + omega * eq;
        out[i + 10] = (1. - omega) * f[i + 10] + omega * eqopp;
    }
    const double eq9 = rho * (1. / 3.) * (1. - usqr);
    out[9] = (1. - omega) * f[9] + omega * eq9;
}

/* binary32 collision: same expression tree, every operation in binary32 */
static void collide_f32(const float f[19], float omega, float out[19])
{
    float rho, u[3];
    macro_f(f, &rho, u);
    const float usqr = 1.5f * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    const float ck[9] = {u[0],        u[1],        u[2],        u[0] + u[1], u[0] - u[1],
                         u[0] + u[2], u[0] - u[2], u[1] + u[2], u[1] - u[2]};
    const float w18 = (float)(1. / 18.), w36 = (float)(1. / 36.), w3 = (float)(1. / 3.);
    for (int i = 0; i < 9; i++) {
        const float wi = i < 3 ? w18 : w36;
        const float eq = rho * wi * (1.f - 3.f * ck[i] + 4.5f * ck[i] * ck[i] - usqr);
        const float eqopp = eq + rho * wi * 6.f * ck[i];
        out[i] = (1.f - omega) * f[i] + omega * eq;
        out[i + 10] = (1.f - omega) * f[i + 10] + omega * eqopp;
    }
    const float eq9 = rho * w3 * (1.f - usqr);
    out[9] = (1.f - omega) * f[9] + omega * eq9;
}

/* The reference's float/float instantiation: float moments; double literals
 * promote eq/eqopp/out expressions to binary64, results stored to float;
 * `omega * eq` multiplies two floats (LbmTools.h:216-254, 276-279, 312). */
static void collide_ref(const float f[19], float omega, float out[19])
{
    float rho, u[3];
    macro_f(f, &rho, u);
    const float usqr = (float)(1.5 * (double)(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]));
    const float ck[9] = {u[0],        u[1],        u[2],        u[0] + u[1], u[0] - u[1],
                         u[0] + u[2], u[0] - u[2], u[1] + u[2], u[1] - u[2]};
    for (int i = 0; i < 9; i++) {
        const double wi = i < 3 ? (1. / 18.) : (1. / 36.);
        const double c = (double)ck[i];
        const float eq = (float)((double)rho * wi * (1. - 3. * c + 4.5 * c * c - (double)usqr));
        const float eqopp = (float)((double)eq + (double)rho * wi * 6. * c);
        out[i] = (float)((1. - (double)omega) * (double)f[i] + (double)(omega * eq));
        out[i + 10] = (float)((1. - (double)omega) * (double)f[i + 10] + (double)(omega * eqopp));
    }
    const float eq9 = (float)((double)rho * (1. / 3.) * (1. - (double)usqr));
    out[9] = (float)((1. - (double)omega) * (double)f[9] + (double)(omega * eq9));
}

/* Project-defined (no reference counterpart): BGK with Guo et al. (2002)
 * forcing, body acceleration g, force density F = rho g.
 *   u   = sum_q c_q f_q / rho + g/2
 *   out = (1-omega) f + omega feq(rho,u)
 *         + (1-omega/2) w_q rho [3 (c_q-u).g + 9 (c_q.u)(c_q.g)]        */
#define DEF_FORCED(NAME, R)                                                        \
    static void NAME(const R f[19], R omega, const R g[3], R out[19])              \
    {                                                                              \
        R rho, u[3];                                                               \
        NAME##_macro(f, &rho, u);                                                  \
        u[0] += (R)0.5 * g[0];                                                     \
        u[1] += (R)0.5 * g[1];                                                     \
        u[2] += (R)0.5 * g[2];                                                     \
        const R usqr = (R)1.5 * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);         \
        const R ug = u[0] * g[0] + u[1] * g[1] + u[2] * g[2];                      \
        const R fpre = (R)1 - (R)0.5 * omega;                                      \
        for (int q = 0; q < 19; q++) {                                             \
            const R wq = (R)W[q];                                                  \
            const R cu = (R)C[q][0] * u[0] + (R)C[q][1] * u[1] + (R)C[q][2] * u[2]; \
            const R cg = (R)C[q][0] * g[0] + (R)C[q][1] * g[1] + (R)C[q][2] * g[2]; \
            const R eq = rho * wq * ((R)1 + (R)3 * cu + (R)4.5 * cu * cu - usqr);  \
            const R src = fpre * wq * rho * ((R)3 * (cg - ug) + (R)9 * cu * cg);   \
            out[q] = ((R)1 - omega) * f[q] + omega * eq + src;                     \
        }                                                                          \
    }
#define forced_d_macro macro_d
#define forced_f_macro macro_f
DEF_FORCED(forced_d, double)
DEF_FORCED(forced_f, float)

/* Deviation form (project-defined), binary32: d = f - w_q, rho = 1 + sum d,
 * u = sum c d * (1 / rho), feq - w_q = w_q (rho-1) + w_q rho (3 c.u +
 * 4.5 (c.u)^2 - 1.5 u^2), written with one reciprocal and fused
 * multiply-adds: the operations of the device's bgk_dev / bgk_forced_dev. */
static inline void macro_dev(const float d[19], float* drho, float* rho, float u[3])
{
    const float XM1 = d[0] + d[3] + d[4] + d[5] + d[6];
    const float XP1 = d[10] + d[13] + d[14] + d[15] + d[16];
    const float X0 = d[9] + d[1] + d[2] + d[7] + d[8] + d[11] + d[12] + d[17] + d[18];
    const float YM1 = d[1] + d[3] + d[7] + d[8] + d[14];
    const float YP1 = d[4] + d[11] + d[13] + d[17] + d[18];
    const float ZM1 = d[2] + d[5] + d[7] + d[16] + d[18];
    const float ZP1 = d[6] + d[8] + d[12] + d[15] + d[17];
    *drho = XM1 + XP1 + X0;
    *rho = 1.f + *drho;
    const float inv = 1.f / *rho;
    u[0] = (XP1 - XM1) * inv;
    u[1] = (YP1 - YM1) * inv;
    u[2] = (ZP1 - ZM1) * inv;
}

static void collide_dev(const float d[19], float omega, float out[19])
{
    float drho, rho, u[3];
    macro_dev(d, &drho, &rho, u);
    const float usqr = 1.5f * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    const float ck[9] = {u[0],        u[1],        u[2],        u[0] + u[1], u[0] - u[1],
                         u[0] + u[2], u[0] - u[2], u[1] + u[2], u[1] - u[2]};
    const float w18 = (float)(1. / 18.), w36 = (float)(1. / 36.), w3 = (float)(1. / 3.);
    const float dr18 = w18 * drho, dr36 = w36 * drho;
    const float rw18 = rho * w18, rw36 = rho * w36;
    const float rw18x6 = rw18 * 6.f, rw36x6 = rw36 * 6.f;
    for (int i = 0; i < 9; i++) {
        const float c = ck[i];
        const float poly = fmaf(c, fmaf(4.5f, c, -3.f), -usqr);
        const float deq = fmaf(i < 3 ? rw18 : rw36, poly, i < 3 ? dr18 : dr36);
        const float deqopp = fmaf(i < 3 ? rw18x6 : rw36x6, c, deq);
        out[i] = fmaf(omega, deq - d[i], d[i]);
        out[i + 10] = fmaf(omega, deqopp - d[i + 10], d[i + 10]);
    }
    const float deq9 = fmaf(-(rho * w3), usqr, w3 * drho);
    out[9] = fmaf(omega, deq9 - d[9], d[9]);
}

static void forced_dev(const float d[19], float omega, const float g[3], float out[19])
{
    float drho, rho, u[3];
    macro_dev(d, &drho, &rho, u);
    u[0] += 0.5f * g[0];
    u[1] += 0.5f * g[1];
    u[2] += 0.5f * g[2];
    const float usqr = 1.5f * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    const float ug = u[0] * g[0] + u[1] * g[1] + u[2] * g[2];
    const float fpre = 1.f - 0.5f * omega;
    for (int q = 0; q < 19; q++) {
        const float wq = (float)W[q];
        const float cu = (float)C[q][0] * u[0] + (float)C[q][1] * u[1] + (float)C[q][2] * u[2];
        const float cg = (float)C[q][0] * g[0] + (float)C[q][1] * g[1] + (float)C[q][2] * g[2];
        const float poly = fmaf(cu, fmaf(4.5f, cu, 3.f), -usqr);
        const float deq = fmaf(rho * wq, poly, wq * drho);
        const float src = fpre * wq * rho * (3.f * (cg - ug) + 9.f * cu * cg);
        out[q] = fmaf(omega, deq - d[q], d[q]) + src;
    }
}

/* collide one cell: f holds pulled values (exactly representable in the
 * pull precision), out receives compute-precision results as double */
static void collide_any(int s, int cmode, const double f[19], double omega, const double g[3],
                        int forced, double out[19])
{
    float ff[19], of[19];
    if (is_dev(s)) {
        for (int q = 0; q < 19; q++) ff[q] = (float)f[q];
        if (forced) {
            const float gf[3] = {(float)g[0], (float)g[1], (float)g[2]};
            forced_dev(ff, (float)omega, gf, of);
        } else {
            collide_dev(ff, (float)omega, of);
        }
        for (int q = 0; q < 19; q++) out[q] = (double)of[q];
        return;
    }
    if (cmode == OR_COMPUTE_F64 || s == OR_F64) {
        if (forced)
            forced_d(f, omega, g, out);
        else
            collide_f64(f, omega, out);
        return;
    }
    for (int q = 0; q < 19; q++) ff[q] = (float)f[q];
    if (forced) {
        const float gf[3] = {(float)g[0], (float)g[1], (float)g[2]};
        forced_f(ff, (float)omega, gf, of);
    } else if (cmode == OR_COMPUTE_REF) {
        collide_ref(ff, (float)omega, of);
    } else {
        collide_f32(ff, (float)omega, of);
    }
    for (int q = 0; q < 19; q++) out[q] = (double)of[q];
}

static inline int cls_of(uint32_t fl) { return (int)((fl >> OR_CLASS_SHIFT) & 3u); }

/* ===================== dense grid (dGrid) ================================ */
typedef struct {
    int X, Y, Z, per;
    size_t plane;
} geo_t;

static geo_t geo_of(const or_dense* d)
{
    geo_t g = {d->dim[0], d->dim[1], d->dim[2], d->periodic, 0};
    g.plane = (size_t)g.X * g.Y * g.Z;
    return g;
}

/* index of cell (x,y,z) + (dx,dy,dz); periodic axes wrap (project-defined:
 * dPartition.h:241-244 has no periodic support); returns 0 when outside */
static inline int nidx(const geo_t* g, int x, int y, int z, int dx, int dy, int dz, size_t* out)
{
    int a = x + dx, b = y + dy, c = z + dz;
    if (g->per & 1) a = (a + g->X) % g->X;
    if (g->per & 2) b = (b + g->Y) % g->Y;
    if (g->per & 4) c = (c + g->Z) % g->Z;
    if (a < 0 || b < 0 || c < 0 || a >= g->X || b >= g->Y || c >= g->Z) return 0;
    *out = (size_t)a + (size_t)g->X * ((size_t)b + (size_t)g->Y * (size_t)c);
    return 1;
}

void or_cavity_flags_dense(const or_dense* d, uint32_t* flags)
{
    const int X = d->dim[0], Y = d->dim[1], Z = d->dim[2];
    for (int z = 0; z < Z; z++)
        for (int y = 0; y < Y; y++)
            for (int x = 0; x < X; x++) {
                int cls = OR_BULK;
                if (x == 0 || x == X - 1 || y == 0 || y == Y - 1 || z == 0 || z == Z - 1) {
                    cls = OR_BOUNCEBACK;
                    if (y == Y - 1) cls = OR_MOVINGWALL;
                }
                flags[(size_t)x + (size_t)X * ((size_t)y + (size_t)Y * z)] =
                    (uint32_t)cls << OR_CLASS_SHIFT;
            }
}

static double fill_value(int cls, int q, double ulb, const double ulid[3])
{
    if (cls == OR_BULK) return W[q];
    if (cls == OR_MOVINGWALL)
        return -6. * W[q] * ulb * (C[q][0] * ulid[0] + C[q][1] * ulid[1] + C[q][2] * ulid[2]);
    return 0.;
}

void or_fill_dense(const or_dense* d, void* pop, const uint32_t* flags, double ulb,
                   const double ulid[3])
{
    const geo_t g = geo_of(d);
    for (size_t i = 0; i < g.plane; i++) {
        const int cls = cls_of(flags[i]);
        for (int q = 0; q < 19; q++) {
            const double v = fill_value(cls, q, ulb, ulid);
            if (cls == OR_BULK)
                st_full(pop, d->storage, q * g.plane + i, q, v);
            else
                st(pop, d->storage, q * g.plane + i, v);
        }
    }
}

/* bit q <=> upstream cell x - c_q is not bulk (or does not exist) */
void or_wall_mask_dense(const or_dense* d, const uint32_t* in, uint32_t* out)
{
    const geo_t g = geo_of(d);
    uint32_t* tmp = (uint32_t*)malloc(g.plane * sizeof(uint32_t));
    for (int z = 0; z < g.Z; z++)
        for (int y = 0; y < g.Y; y++)
            for (int x = 0; x < g.X; x++) {
                const size_t i = (size_t)x + (size_t)g.X * ((size_t)y + (size_t)g.Y * z);
                uint32_t fl = in[i];
                if (cls_of(fl) != OR_BULK) {
                    tmp[i] = fl;
                    continue;
                }
                uint32_t m = 0, open = 0, closed = 0;
                for (int q = 0; q < 19; q++) {
                    if (q == 9) continue;
                    size_t j;
                    if (!nidx(&g, x, y, z, -C[q][0], -C[q][1], -C[q][2], &j)) {
                        m |= 1u << q;
                        open = OR_FLAG_OPEN;
                    } else if (cls_of(in[j]) != OR_BULK) {
                        m |= 1u << q;
                        closed = 1;
                    }
                }
                tmp[i] = ((uint32_t)OR_BULK << OR_CLASS_SHIFT) | m | open |
                         (open && closed ? OR_FLAG_MIXED : 0u) | (fl & OR_ZH_BITS);
            }
    memcpy(out, tmp, g.plane * sizeof(uint32_t));
    free(tmp);
}

/* pull stream for one bulk cell (LbmTools.h:78-168 generalised to q) */
static void pull_dense(const geo_t* g, int s, const void* fin, uint32_t fl, int x, int y, int z,
                       size_t i, double f[19])
{
    for (int q = 0; q < 19; q++) {
        if (q == 9) {
            f[9] = dec_bulk(fin, s, 9 * g->plane + i, 9);
            continue;
        }
        size_t j;
        const int valid = nidx(g, x, y, z, -C[q][0], -C[q][1], -C[q][2], &j);
        if (fl & (1u << q)) {
            const int o = OPP[q];
            const double own = dec_bulk(fin, s, o * g->plane + i, o);
            /* project-defined: a missing neighbour contributes 0 */
            const double nb = valid ? ld(fin, s, o * g->plane + j) : 0.;
            f[q] = add_p(s, own, nb);
        } else {
            f[q] = dec_bulk(fin, s, q * g->plane + j, q);
        }
    }
}

void or_step_dense(const or_dense* d, const void* fin, void* fout, const uint32_t* flags,
                   double omega, const double g3[3])
{
    const geo_t g = geo_of(d);
    const double zero[3] = {0, 0, 0};
    const double* gg = g3 ? g3 : zero;
    const int forced = gg[0] != 0. || gg[1] != 0. || gg[2] != 0.;
    /* cells whose 18 upstream neighbours are bulk: plain offsets (same
     * arithmetic as pull_dense, without the per-link bounds logic) */
    long off[19];
    for (int q = 0; q < 19; q++)
        off[q] = (long)C[q][0] + (long)g.X * ((long)C[q][1] + (long)g.Y * (long)C[q][2]);
    const int fast64 = d->storage == OR_F64 && !forced;
    const int fast32 = d->storage == OR_F32 && d->compute == OR_COMPUTE_F32 && !forced;
#pragma omp parallel for schedule(static) num_threads(or_get_threads())
    for (int z = 0; z < g.Z; z++)
        for (int y = 0; y < g.Y; y++)
            for (int x = 0; x < g.X; x++) {
                const size_t i = (size_t)x + (size_t)g.X * ((size_t)y + (size_t)g.Y * z);
                const uint32_t fl = flags[i];
                if (cls_of(fl) != OR_BULK) continue;
                double f[19], o[19];
                if (fast64 && !(fl & OR_MASK_BITS) && !g.per) {
                    const double* F = (const double*)fin;
                    double* G = (double*)fout;
                    for (int q = 0; q < 19; q++) f[q] = F[q * g.plane + (size_t)((long)i - off[q])];
                    collide_f64(f, omega, o);
                    for (int q = 0; q < 19; q++) G[q * g.plane + i] = o[q];
                    continue;
                }
                if (fast32 && !(fl & OR_MASK_BITS) && !g.per) {
                    const float* F = (const float*)fin;
                    float* G = (float*)fout;
                    float ff[19], of[19];
                    for (int q = 0; q < 19; q++) ff[q] = F[q * g.plane + (size_t)((long)i - off[q])];
                    collide_f32(ff, (float)omega, of);
                    for (int q = 0; q < 19; q++) G[q * g.plane + i] = of[q];
                    continue;
                }
                pull_dense(&g, d->storage, fin, fl, x, y, z, i, f);
                collide_any(d->storage, d->compute, f, omega, gg, forced, o);
                for (int q = 0; q < 19; q++) st(fout, d->storage, q * g.plane + i, o[q]);
            }
}

/* ---- Zou-He boundary cells ------------------------------------------------
 * Unknown directions per face (face 1..6: unknowns coming from -x, +x, -y,
 * +y, -z, +z) and the 4 middle pairs per axis: LbmContainers.h:633-646. */
static const int ZH_UNK[6][5] = {{10, 13, 14, 15, 16}, {0, 3, 4, 5, 6},   {11, 4, 13, 17, 18},
                                 {1, 3, 7, 8, 14},     {12, 6, 8, 15, 17}, {2, 5, 7, 16, 18}};
static const int ZH_MID[3][4] = {{1, 2, 7, 8}, {0, 2, 5, 6}, {0, 1, 3, 4}};

static int zh_opp(int i) { return i < 9 ? i + 10 : i - 10; }

/* binary64: known sum K (LbmContainers.h:203-215), middle sum M (:217-227),
 * rho / normal velocity (:229-239), usqr and eq (:242-272, with eqopp =
 * eq + rho w 6 c.u), unknowns f_i = f_opp + eq_i - eq_opp (:274-286), then
 * BGK with that rho and u (:351-443). */
static void zouhe_d(double f[19], int face, int velocity, double value, double omega, double out[19])
{
    const int* unk = ZH_UNK[face - 1];
    const int axis = (face - 1) / 2;
    double known = 0., middle = 0.;
    for (int k = 0; k < 5; k++) known += f[zh_opp(unk[k])];
    for (int k = 0; k < 4; k++) {
        middle += f[ZH_MID[axis][k]];
        middle += f[ZH_MID[axis][k] + 10];
    }
    middle += f[9];
    const double sgn = unk[0] < 9 ? 1. : -1.;
    double rho, ua;
    if (velocity) {
        ua = -sgn * value;
        rho = (middle + known * 2.) / (1. - value);
    } else {
        rho = value;
        ua = (((middle + known * 2.) / rho) - 1.) * sgn;
    }
    double u[3] = {0., 0., 0.};
    u[axis] = ua;
    const double usqr = 1.5 * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    const double ck[9] = {u[0], u[1], u[2], u[0] + u[1], u[0] - u[1], u[0] + u[2], u[0] - u[2],
                          u[1] + u[2], u[1] - u[2]};
    double eq[19];
    for (int i = 0; i < 9; i++) {
        const double wi = i < 3 ? 1. / 18. : 1. / 36.;
        eq[i] = rho * wi * (1. - 3. * ck[i] + 4.5 * ck[i] * ck[i] - usqr);
        eq[i + 10] = eq[i] + rho * wi * 6. * ck[i];
    }
    eq[9] = rho * (1. / 3.) * (1. - usqr);
    for (int k = 0; k < 5; k++) {
        const int iu = unk[k], ik = zh_opp(iu);
        f[iu] = f[ik] + eq[iu] - eq[ik];
    }
    for (int q = 0; q < 19; q++) out[q] = (1. - omega) * f[q] + omega * eq[q];
}

/* the same in binary32 (the NX_COMPUTE_F32 collision of binary32 storage) */
static void zouhe_f(float f[19], int face, int velocity, float value, float omega, float out[19])
{
    const int* unk = ZH_UNK[face - 1];
    const int axis = (face - 1) / 2;
    float known = 0.f, middle = 0.f;
    for (int k = 0; k < 5; k++) known += f[zh_opp(unk[k])];
    for (int k = 0; k < 4; k++) {
        middle += f[ZH_MID[axis][k]];
        middle += f[ZH_MID[axis][k] + 10];
    }
    middle += f[9];
    const float sgn = unk[0] < 9 ? 1.f : -1.f;
    float rho, ua;
    if (velocity) {
        ua = -sgn * value;
        rho = (middle + known * 2.f) / (1.f - value);
    } else {
        rho = value;
        ua = (((middle + known * 2.f) / rho) - 1.f) * sgn;
    }
    float u[3] = {0.f, 0.f, 0.f};
    u[axis] = ua;
    const float usqr = 1.5f * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    const float ck[9] = {u[0], u[1], u[2], u[0] + u[1], u[0] - u[1], u[0] + u[2], u[0] - u[2],
                         u[1] + u[2], u[1] - u[2]};
    float eq[19];
    for (int i = 0; i < 9; i++) {
        const float wi = i < 3 ? (float)(1. / 18.) : (float)(1. / 36.);
        eq[i] = rho * wi * (1.f - 3.f * ck[i] + 4.5f * ck[i] * ck[i] - usqr);
        eq[i + 10] = eq[i] + rho * wi * 6.f * ck[i];
    }
    eq[9] = rho * (float)(1. / 3.) * (1.f - usqr);
    for (int k = 0; k < 5; k++) {
        const int iu = unk[k], ik = zh_opp(iu);
        f[iu] = f[ik] + eq[iu] - eq[ik];
    }
    for (int q = 0; q < 19; q++) out[q] = (1.f - omega) * f[q] + omega * eq[q];
}

void or_step_dense_zouhe(const or_dense* d, const void* fin, void* fout, const uint32_t* flags,
                         double omega, const double value[6])
{
    const geo_t g = geo_of(d);
    const double zero[3] = {0, 0, 0};
    const int f64 = d->storage == OR_F64;
#pragma omp parallel for schedule(static) num_threads(or_get_threads())
    for (int z = 0; z < g.Z; z++)
        for (int y = 0; y < g.Y; y++)
            for (int x = 0; x < g.X; x++) {
                const size_t i = (size_t)x + (size_t)g.X * ((size_t)y + (size_t)g.Y * z);
                const uint32_t fl = flags[i];
                if (cls_of(fl) != OR_BULK) continue;
                double f[19], o[19];
                pull_dense(&g, d->storage, fin, fl, x, y, z, i, f);
                const int face = (int)((fl >> OR_ZH_SHIFT) & 7u);
                const int vel = (fl & OR_ZH_VELOCITY) != 0;
                if (face >= 1 && face <= 6) {
                    if (f64) {
                        zouhe_d(f, face, vel, value[face - 1], omega, o);
                    } else {
                        float ff[19], of[19];
                        for (int q = 0; q < 19; q++) ff[q] = (float)f[q];
                        zouhe_f(ff, face, vel, (float)value[face - 1], (float)omega, of);
                        for (int q = 0; q < 19; q++) o[q] = (double)of[q];
                    }
                } else {
                    collide_any(d->storage, d->compute, f, omega, zero, 0, o);
                }
                for (int q = 0; q < 19; q++) st(fout, d->storage, q * g.plane + i, o[q]);
            }
}

void or_run_dense(const or_dense* d, void* pop0, void* pop1, const uint32_t* flags, double omega,
                  const double g[3], int nsteps)
{
    void* p[2] = {pop0, pop1};
    for (int t = 0; t < nsteps; t++) or_step_dense(d, p[t & 1], p[(t + 1) & 1], flags, omega, g);
}

/* moments of one cell (LbmTools.h:401-433) as double */
static void cell_moments(int s, int cmode, uint32_t fl, const double f[19], const double own[3],
                         double* rho, double u[3])
{
    const int cls = cls_of(fl);
    if (cls == OR_BULK) {
        if (is_dev(s)) {
            float ff[19], dr, r, uf[3];
            for (int q = 0; q < 19; q++) ff[q] = (float)f[q];
            macro_dev(ff, &dr, &r, uf);
            *rho = r;
            for (int a = 0; a < 3; a++) u[a] = uf[a];
        } else if (cmode == OR_COMPUTE_F64 || s == OR_F64) {
            macro_d(f, rho, u);
        } else {
            float ff[19], r, uf[3];
            for (int q = 0; q < 19; q++) ff[q] = (float)f[q];
            macro_f(ff, &r, uf);
            *rho = r;
            for (int a = 0; a < 3; a++) u[a] = uf[a];
        }
    } else if (cls == OR_MOVINGWALL) {
        *rho = 1.0;
        if (cmode == OR_COMPUTE_F64 || s == OR_F64) {
            for (int a = 0; a < 3; a++) u[a] = own[a] / (6. * 1. / 18.);
        } else {
            for (int a = 0; a < 3; a++) u[a] = (double)((float)own[a] / (float)(6. * 1. / 18.));
        }
    } else {
        *rho = 0;
        u[0] = u[1] = u[2] = 0;
    }
}

static inline void st_moment(void* p, int s, size_t i, double v)
{
    if (s == OR_F64)
        ((double*)p)[i] = v;
    else
        ((float*)p)[i] = (float)v;
}

void or_moments_dense(const or_dense* d, const void* fin, const uint32_t* flags, void* rho,
                      void* u)
{
    const geo_t g = geo_of(d);
#pragma omp parallel for schedule(static) num_threads(or_get_threads())
    for (int z = 0; z < g.Z; z++)
        for (int y = 0; y < g.Y; y++)
            for (int x = 0; x < g.X; x++) {
                const size_t i = (size_t)x + (size_t)g.X * ((size_t)y + (size_t)g.Y * z);
                const uint32_t fl = flags[i];
                double f[19] = {0}, own[3] = {0}, r, uu[3];
                if (cls_of(fl) == OR_BULK) pull_dense(&g, d->storage, fin, fl, x, y, z, i, f);
                if (cls_of(fl) == OR_MOVINGWALL)
                    for (int a = 0; a < 3; a++) own[a] = ld(fin, d->storage, a * g.plane + i);
                cell_moments(d->storage, d->compute, fl, f, own, &r, uu);
                st_moment(rho, d->storage, i, r);
                for (int a = 0; a < 3; a++) st_moment(u, d->storage, a * g.plane + i, uu[a]);
            }
}

double or_mass_dense(const or_dense* d, const void* fin, const uint32_t* flags)
{
    const geo_t g = geo_of(d);
    double m = 0;
    for (size_t i = 0; i < g.plane; i++) {
        if (cls_of(flags[i]) != OR_BULK) continue;
        double r = 0;
        for (int q = 0; q < 19; q++) r += dec_full(fin, d->storage, q * g.plane + i, q);
        m += r;
    }
    return m;
}

void or_residual_dense(const or_dense* d, const void* fin, const uint32_t* flags, float* u_prev,
                       double out[2])
{
    const geo_t g = geo_of(d);
    double num = 0, den = 0;
    for (int z = 0; z < g.Z; z++)
        for (int y = 0; y < g.Y; y++)
            for (int x = 0; x < g.X; x++) {
                const size_t i = (size_t)x + (size_t)g.X * ((size_t)y + (size_t)g.Y * z);
                const uint32_t fl = flags[i];
                if (cls_of(fl) != OR_BULK) continue;
                double f[19], own[3] = {0}, r, uu[3];
                pull_dense(&g, d->storage, fin, fl, x, y, z, i, f);
                cell_moments(d->storage, d->compute, fl, f, own, &r, uu);
                for (int a = 0; a < 3; a++) {
                    const float un = (float)uu[a];
                    const double du = (double)un - (double)u_prev[a * g.plane + i];
                    num += du * du;
                    den += (double)un * (double)un;
                    u_prev[a * g.plane + i] = un;
                }
            }
    out[0] = num;
    out[1] = den;
}

/* ===================== block-sparse grid (bGrid 8^3) ===================== */
static inline int in_dom(const int32_t dim[3], int x, int y, int z)
{
    return x < dim[0] && y < dim[1] && z < dim[2];
}

static inline int voxel_active(const int32_t dim[3], const uint8_t* act, int x, int y, int z)
{
    return in_dom(dim, x, y, z) &&
           act[(size_t)x + (size_t)dim[0] * ((size_t)y + (size_t)dim[1] * z)] != 0;
}

/* blocks holding >= 1 active voxel (SpanClassifier.h:170-191) */
uint32_t or_bgrid_count(const int32_t dim[3], const uint8_t* active)
{
    const int nbx = (dim[0] + 7) / 8, nby = (dim[1] + 7) / 8, nbz = (dim[2] + 7) / 8;
    uint32_t n = 0;
    for (int bz = 0; bz < nbz; bz++)
        for (int by = 0; by < nby; by++)
            for (int bx = 0; bx < nbx; bx++) {
                int any = 0;
                for (int v = 0; v < 512 && !any; v++)
                    any = voxel_active(dim, active, bx * 8 + (v & 7), by * 8 + ((v >> 3) & 7),
                                       bz * 8 + (v >> 6));
                n += (uint32_t)any;
            }
    return n;
}

/* Block list in bz, by, bx order; origin 8*(bx,by,bz); mask bit x+8y+64z
 * (StaticBlock.h:79-96, bGrid_imp.h:80-136); connectivity i+3j+9k for
 * offset (i-1,j-1,k-1) (bGrid_imp.h:139-172).  Returns the active voxel
 * count.  periodic (project-defined) wraps the block lattice. */
uint64_t or_bgrid_build(const int32_t dim[3], const uint8_t* active, int32_t periodic,
                        int32_t* origins, uint32_t* masks, uint32_t* conn)
{
    const int nbx = (dim[0] + 7) / 8, nby = (dim[1] + 7) / 8, nbz = (dim[2] + 7) / 8;
    const size_t nb3 = (size_t)nbx * nby * nbz;
    uint32_t* rank = (uint32_t*)malloc(nb3 * sizeof(uint32_t));
    uint32_t n = 0;
    uint64_t nact = 0;
    for (int bz = 0; bz < nbz; bz++)
        for (int by = 0; by < nby; by++)
            for (int bx = 0; bx < nbx; bx++) {
                uint32_t bits[16] = {0};
                int cnt = 0;
                for (int v = 0; v < 512; v++)
                    if (voxel_active(dim, active, bx * 8 + (v & 7), by * 8 + ((v >> 3) & 7),
                                     bz * 8 + (v >> 6))) {
                        bits[v >> 5] |= 1u << (v & 31);
                        cnt++;
                    }
                const size_t l = (size_t)bx + (size_t)nbx * ((size_t)by + (size_t)nby * bz);
                if (!cnt) {
                    rank[l] = OR_INVALID_BLOCK;
                    continue;
                }
                rank[l] = n;
                origins[3 * n + 0] = 8 * bx;
                origins[3 * n + 1] = 8 * by;
                origins[3 * n + 2] = 8 * bz;
                memcpy(masks + 16 * (size_t)n, bits, sizeof(bits));
                nact += (uint64_t)cnt;
                n++;
            }
    for (uint32_t b = 0; b < n; b++) {
        const int bx = origins[3 * b] / 8, by = origins[3 * b + 1] / 8, bz = origins[3 * b + 2] / 8;
        for (int k = 0; k < 3; k++)
            for (int j = 0; j < 3; j++)
                for (int i = 0; i < 3; i++) {
                    int x = bx + i - 1, y = by + j - 1, z = bz + k - 1;
                    if (periodic & 1) x = (x + nbx) % nbx;
                    if (periodic & 2) y = (y + nby) % nby;
                    if (periodic & 4) z = (z + nbz) % nbz;
                    uint32_t r = OR_INVALID_BLOCK;
                    if (x >= 0 && y >= 0 && z >= 0 && x < nbx && y < nby && z < nbz)
                        r = rank[(size_t)x + (size_t)nbx * ((size_t)y + (size_t)nby * z)];
                    conn[27 * (size_t)b + i + 3 * j + 9 * k] = r;
                }
    }
    free(rank);
    return nact;
}

static inline int blk_active(const or_block* d, uint32_t b, int v)
{
    return (d->masks[16 * (size_t)b + (v >> 5)] >> (v & 31)) & 1u;
}

/* neighbour voxel of (b, x,y,z) at (dx,dy,dz): bPartition_imp.h:218-309 then
 * the validity test of helpNghPitch (:109-124); returns 0 when invalid */
static inline int bnbr(const or_block* d, uint32_t b, int x, int y, int z, int dx, int dy, int dz,
                       uint32_t* nb, int* nv)
{
    int a = x + dx, bb = y + dy, c = z + dz;
    const int xf = a < 0 ? -1 : (a >= 8 ? 1 : 0);
    const int yf = bb < 0 ? -1 : (bb >= 8 ? 1 : 0);
    const int zf = c < 0 ? -1 : (c >= 8 ? 1 : 0);
    uint32_t t = b;
    if (xf | yf | zf) t = d->conn[27 * (size_t)b + (xf + 1) + 3 * (yf + 1) + 9 * (zf + 1)];
    if (t == OR_INVALID_BLOCK) return 0;
    a -= 8 * xf;
    bb -= 8 * yf;
    c -= 8 * zf;
    const int v = a + 8 * bb + 64 * c;
    if (!blk_active(d, t, v)) return 0;
    *nb = t;
    *nv = v;
    return 1;
}

static inline size_t bidx(uint32_t b, int q, int v) { return ((size_t)b * 19 + q) * 512 + v; }

void or_cavity_flags_block(const or_block* d, uint32_t* flags)
{
    const int X = d->dim[0], Y = d->dim[1], Z = d->dim[2];
    for (uint32_t b = 0; b < d->n_blocks; b++)
        for (int v = 0; v < 512; v++) {
            const int x = d->origins[3 * b] + (v & 7), y = d->origins[3 * b + 1] + ((v >> 3) & 7),
                      z = d->origins[3 * b + 2] + (v >> 6);
            int cls = OR_UNDEFINED;
            if (blk_active(d, b, v)) {
                cls = OR_BULK;
                if (x == 0 || x == X - 1 || y == 0 || y == Y - 1 || z == 0 || z == Z - 1) {
                    cls = OR_BOUNCEBACK;
                    if (y == Y - 1) cls = OR_MOVINGWALL;
                }
            }
            flags[(size_t)b * 512 + v] = (uint32_t)cls << OR_CLASS_SHIFT;
        }
}

void or_fill_block(const or_block* d, void* pop, const uint32_t* flags, double ulb,
                   const double ulid[3])
{
    for (uint32_t b = 0; b < d->n_blocks; b++)
        for (int v = 0; v < 512; v++) {
            const int cls = cls_of(flags[(size_t)b * 512 + v]);
            for (int q = 0; q < 19; q++) {
                if (cls == OR_BULK)
                    st_full(pop, d->storage, bidx(b, q, v), q, W[q]);
                else
                    st(pop, d->storage, bidx(b, q, v),
                       cls == OR_UNDEFINED ? 0. : fill_value(cls, q, ulb, ulid));
            }
        }
}

void or_wall_mask_block(const or_block* d, const uint32_t* in, uint32_t* out)
{
    const size_t n = (size_t)d->n_blocks * 512;
    uint32_t* tmp = (uint32_t*)malloc(n * sizeof(uint32_t) + 1);
    for (uint32_t b = 0; b < d->n_blocks; b++)
        for (int v = 0; v < 512; v++) {
            const size_t i = (size_t)b * 512 + v;
            const uint32_t fl = in[i];
            if (!blk_active(d, b, v) || cls_of(fl) != OR_BULK) {
                tmp[i] = fl;
                continue;
            }
            uint32_t m = 0, open = 0, closed = 0;
            for (int q = 0; q < 19; q++) {
                if (q == 9) continue;
                uint32_t nb;
                int nv;
                if (!bnbr(d, b, v & 7, (v >> 3) & 7, v >> 6, -C[q][0], -C[q][1], -C[q][2], &nb, &nv)) {
                    m |= 1u << q;
                    open = OR_FLAG_OPEN;
                } else if (cls_of(in[(size_t)nb * 512 + nv]) != OR_BULK) {
                    m |= 1u << q;
                    closed = 1;
                }
            }
            tmp[i] = ((uint32_t)OR_BULK << OR_CLASS_SHIFT) | m | open |
                     (open && closed ? OR_FLAG_MIXED : 0u);
        }
    memcpy(out, tmp, n * sizeof(uint32_t));
    free(tmp);
}

static void pull_block(const or_block* d, const void* fin, uint32_t fl, uint32_t b, int v,
                       double f[19])
{
    const int s = d->storage, x = v & 7, y = (v >> 3) & 7, z = v >> 6;
    for (int q = 0; q < 19; q++) {
        if (q == 9) {
            f[9] = dec_bulk(fin, s, bidx(b, 9, v), 9);
            continue;
        }
        uint32_t nb = 0;
        int nv = 0;
        const int valid = bnbr(d, b, x, y, z, -C[q][0], -C[q][1], -C[q][2], &nb, &nv);
        if (fl & (1u << q)) {
            const int o = OPP[q];
            const double own = dec_bulk(fin, s, bidx(b, o, v), o);
            const double nbv = valid ? ld(fin, s, bidx(nb, o, nv)) : 0.;
            f[q] = add_p(s, own, nbv);
        } else {
            f[q] = dec_bulk(fin, s, bidx(nb, q, nv), q);
        }
    }
}

void or_step_block(const or_block* d, const void* fin, void* fout, const uint32_t* flags,
                   double omega, const double g3[3])
{
    const double zero[3] = {0, 0, 0};
    const double* gg = g3 ? g3 : zero;
    const int forced = gg[0] != 0. || gg[1] != 0. || gg[2] != 0.;
#pragma omp parallel for schedule(static) num_threads(or_get_threads())
    for (long b = 0; b < (long)d->n_blocks; b++)
        for (int v = 0; v < 512; v++) {
            if (!blk_active(d, (uint32_t)b, v)) continue;
            const uint32_t fl = flags[(size_t)b * 512 + v];
            if (cls_of(fl) != OR_BULK) continue;
            double f[19], o[19];
            pull_block(d, fin, fl, (uint32_t)b, v, f);
            collide_any(d->storage, d->compute, f, omega, gg, forced, o);
            for (int q = 0; q < 19; q++) st(fout, d->storage, bidx((uint32_t)b, q, v), o[q]);
        }
}

void or_run_block(const or_block* d, void* pop0, void* pop1, const uint32_t* flags, double omega,
                  const double g[3], int nsteps)
{
    void* p[2] = {pop0, pop1};
    for (int t = 0; t < nsteps; t++) or_step_block(d, p[t & 1], p[(t + 1) & 1], flags, omega, g);
}

/* rho [b*512+v], u [(b*3+a)*512+v]  (bField AoSoA with cardinality 3) */
void or_moments_block(const or_block* d, const void* fin, const uint32_t* flags, void* rho,
                      void* u)
{
    for (uint32_t b = 0; b < d->n_blocks; b++)
        for (int v = 0; v < 512; v++) {
            const size_t i = (size_t)b * 512 + v;
            const uint32_t fl = blk_active(d, b, v) ? flags[i] : (uint32_t)OR_UNDEFINED << 24;
            double f[19] = {0}, own[3] = {0}, r, uu[3];
            if (cls_of(fl) == OR_BULK) pull_block(d, fin, fl, b, v, f);
            if (cls_of(fl) == OR_MOVINGWALL)
                for (int a = 0; a < 3; a++) own[a] = ld(fin, d->storage, bidx(b, a, v));
            cell_moments(d->storage, d->compute, fl, f, own, &r, uu);
            st_moment(rho, d->storage, i, r);
            for (int a = 0; a < 3; a++) st_moment(u, d->storage, ((size_t)b * 3 + a) * 512 + v, uu[a]);
        }
}

double or_mass_block(const or_block* d, const void* fin, const uint32_t* flags)
{
    double m = 0;
    for (uint32_t b = 0; b < d->n_blocks; b++)
        for (int v = 0; v < 512; v++) {
            if (!blk_active(d, b, v) || cls_of(flags[(size_t)b * 512 + v]) != OR_BULK) continue;
            double r = 0;
            for (int q = 0; q < 19; q++) r += dec_full(fin, d->storage, bidx(b, q, v), q);
            m += r;
        }
    return m;
}
