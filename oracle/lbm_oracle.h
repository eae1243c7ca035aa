/*
 * lbm_oracle.h -- CPU oracle for the D3Q19 lattice-Boltzmann step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the
 * reference's algorithm, used by tests/ as the checker and by bench.py as
 * the timed CPU baseline.  No product code path may call it.
 *
 * Constants (storage / compute / classification / flag bits / layouts)
 * are the ones of include/neon_amd.h; they are repeated here so the oracle
 * builds with gcc alone.
 */
#ifndef LBM_ORACLE_H
#define LBM_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#define OR_F64 0
#define OR_F32 1
#define OR_F32DEV 2
#define OR_F16DEV 3
#define OR_COMPUTE_F64 0
#define OR_COMPUTE_F32 1
#define OR_COMPUTE_REF 2
#define OR_BOUNCEBACK 0
#define OR_MOVINGWALL 1
#define OR_BULK 2
#define OR_UNDEFINED 3
#define OR_CLASS_SHIFT 24
#define OR_MASK_BITS 0x7FDFFu
#define OR_FLAG_OPEN (1u << 28)
#define OR_FLAG_MIXED (1u << 29)
#define OR_ZH_SHIFT 19
#define OR_ZH_VELOCITY (1u << 22)
#define OR_ZH_BITS (0xFu << 19)
#define OR_INVALID_BLOCK 0xFFFFFFFFu

typedef struct {
    int32_t dim[3];
    int32_t storage;
    int32_t compute;
    int32_t periodic;
} or_dense;

typedef struct {
    int32_t dim[3];
    int32_t storage;
    int32_t compute;
    int32_t periodic;
    uint32_t n_blocks;
    const int32_t* origins; /* 3 x n */
    const uint32_t* masks;  /* 16 x n */
    const uint32_t* conn;   /* 27 x n */
} or_block;

/* lattice tables (D3Q19.h:23-132) */
void or_lattice(int32_t c[19][3], double w[19], int32_t opp[19]);

/* IEEE binary16 <-> binary32, round to nearest even */
uint16_t or_f32_to_f16(float f);
float or_f16_to_f32(uint16_t h);

void or_set_threads(int n);
int or_get_threads(void);

/* dense grid */
void or_cavity_flags_dense(const or_dense* d, uint32_t* flags);
void or_fill_dense(const or_dense* d, void* pop, const uint32_t* flags, double ulb,
                   const double ulid[3]);
void or_wall_mask_dense(const or_dense* d, const uint32_t* in, uint32_t* out);
void or_step_dense(const or_dense* d, const void* fin, void* fout, const uint32_t* flags,
                   double omega, const double g[3]);
/* Step with Zou-He boundary cells (bits 19..22 of a bulk cell's flag word;
 * no body force; storage F64 + compute F64, or F32 + compute F32). */
void or_step_dense_zouhe(const or_dense* d, const void* fin, void* fout, const uint32_t* flags,
                         double omega, const double value[6]);
void or_moments_dense(const or_dense* d, const void* fin, const uint32_t* flags, void* rho,
                      void* u);
double or_mass_dense(const or_dense* d, const void* fin, const uint32_t* flags);
void or_residual_dense(const or_dense* d, const void* fin, const uint32_t* flags,
                       float* u_prev, double out[2]);
/* n full steps of the two-population scheme; result in pop[n % 2] */
void or_run_dense(const or_dense* d, void* pop0, void* pop1, const uint32_t* flags,
                  double omega, const double g[3], int nsteps);

/* block-sparse grid */
uint32_t or_bgrid_count(const int32_t dim[3], const uint8_t* active);
uint64_t or_bgrid_build(const int32_t dim[3], const uint8_t* active, int32_t periodic,
                        int32_t* origins, uint32_t* masks, uint32_t* conn);
void or_cavity_flags_block(const or_block* d, uint32_t* flags);
void or_fill_block(const or_block* d, void* pop, const uint32_t* flags, double ulb,
                   const double ulid[3]);
void or_wall_mask_block(const or_block* d, const uint32_t* in, uint32_t* out);
void or_step_block(const or_block* d, const void* fin, void* fout, const uint32_t* flags,
                   double omega, const double g[3]);
void or_moments_block(const or_block* d, const void* fin, const uint32_t* flags, void* rho,
                      void* u);
double or_mass_block(const or_block* d, const void* fin, const uint32_t* flags);
void or_run_block(const or_block* d, void* pop0, void* pop1, const uint32_t* flags,
                  double omega, const double g[3], int nsteps);

/* grid refinement over `depth` levels (mres_oracle.c); per-level arrays are
 * indexed by level, dims[3 * l + axis] */
#define OR_MRES_MAX_DEPTH 5
#define OR_MRES_BGK19 0 /* lattice and collision: BGK D3Q19 */
#define OR_MRES_KBC27 1 /* KBC D3Q27 */
void or_mres_dims(const int32_t dim[3], int depth, int32_t* dims);
int or_mres_build(const int32_t dim[3], int depth, const uint8_t* const* lam, uint8_t* const* st);
void or_mres_setup(const int32_t dim[3], int depth, int lat, int storage, double ulb, uint8_t* const* st,
                   void* const* fin, void* const* fout, float* const* ss);
void or_mres_setup_cells(const int32_t dim[3], int depth, int lat, int storage, const uint8_t* const* cls,
                         const double uw[3], uint8_t* const* st, void* const* fin, void* const* fout,
                         float* const* ss);
double or_mres_omega(double omega0, int level, int num_levels);
void or_mres_step(const int32_t dim[3], int depth, int lat, int s, const uint8_t* const* st, void* const* fin,
                  void* const* fout, const float* const* ss, double omega0);
void or_kbc27_collide(int storage, const void* fin, void* fout, double omega);

#endif
