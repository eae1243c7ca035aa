/*
 * neon_amd.h -- C ABI of the MI355X-native D3Q19 lattice-Boltzmann step.
 *
 * Every entry point returns NX_OK (0) or an NX_ERR_* code; the text of the
 * last error of the calling thread is returned by nx_last_error().  No C++
 * exception crosses this boundary.  Device pointers are plain pointers
 * obtained from nx_alloc (or any HIP allocation on the current device);
 * the library never frees memory it did not allocate.  Launches are
 * asynchronous on the given stream (NULL = the default stream).
 *
 * Cell flag word (one uint32 per cell, the counterpart of the reference's
 * CellType {wallNghBitflag, classification}, CellType.h:3-35):
 *   bits  0..18  wall-neighbour mask: bit q set <=> the upstream cell x-c_q of
 *                population q is not a bulk cell (LbmTools.h:327-377)
 *   bits 24..25  classification (NX_CELL_*)
 *   bit  28      NX_FLAG_OPEN: at least one masked link points outside the
 *                grid or at an inactive voxel (its bounce-back term is 0)
 *   bit  29      NX_FLAG_MIXED: masked links of both kinds (some upstream
 *                cells exist, some do not)
 *   bits 19..21  Zou-He face of a bulk boundary cell (NX_ZH_*; 0: none),
 *   bit  22      NX_ZH_VELOCITY: prescribed normal velocity (else density);
 *                kept by the wall-mask kernels, read by
 *                nx_d3q19_step_dense_zouhe only
 *
 * Population layouts:
 *   dense (dGrid SoA, dField_imp.h:67-87):  f[q][z][y][x], index
 *          x + X*y + X*Y*z + X*Y*Z*q
 *   block (bGrid AoSoA, bPartition_imp.h:95-107): index
 *          (b*19 + q)*512 + x + 8*y + 64*z   (8^3 blocks)
 * Storage types: NX_F64, NX_F32 hold f; NX_F32DEV / NX_F16DEV hold f - w_q
 * for bulk cells and f itself for non-bulk (wall) cells, and collide in
 * binary32 on the deviations (rho = 1 + sum d, feq - w_q formed term by term)
 * so that rounding is relative to f - w_q rather than to w_q.
 */
#ifndef NEON_AMD_H
#define NEON_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------- */
#define NX_OK 0
#define NX_ERR_INVALID 1     /* bad argument (shape, dtype, null pointer) */
#define NX_ERR_HIP 2         /* HIP runtime error                         */
#define NX_ERR_OOM 3         /* device allocation failed                  */
#define NX_ERR_UNSUPPORTED 4 /* combination not implemented               */
#define NX_ERR_NOT_INIT 5    /* nx_init was not called                    */

/* ---- storage / compute types ----------------------------------------- */
#define NX_F64 0
#define NX_F32 1
#define NX_F32DEV 2 /* f - w_q in binary32 */
#define NX_F16DEV 3 /* f - w_q in binary16, round to nearest even */

#define NX_COMPUTE_F64 0 /* every operation in binary64                 */
#define NX_COMPUTE_F32 1 /* every operation in binary32                 */
#define NX_COMPUTE_REF 2 /* float storage with the reference's literal  */
                         /* promotions (LbmTools.h:216-279)             */

/* ---- cell classification (CellType.h:5-11) --------------------------- */
#define NX_CELL_BOUNCEBACK 0
#define NX_CELL_MOVINGWALL 1
#define NX_CELL_BULK 2
#define NX_CELL_UNDEFINED 3
#define NX_CLASS_SHIFT 24
#define NX_MASK_BITS 0x7FDFFu
#define NX_FLAG_OPEN (1u << 28)
#define NX_FLAG_MIXED (1u << 29)
/* Zou-He boundary cells (the pressure / velocity cells of
 * benchmarks/lbm-flow-over-sphere/src/CellType.h:7-15): the face names the
 * side the unknown populations come from (NX_ZH_XMIN: the cell's -x
 * neighbour is outside the fluid, the unknowns are the 5 populations with
 * c_x = +1), value[face - 1] of nx_d3q19_step_dense_zouhe the prescribed
 * density, or with NX_ZH_VELOCITY the velocity along the inward normal. */
#define NX_ZH_SHIFT 19
#define NX_ZH_XMIN 1
#define NX_ZH_XMAX 2
#define NX_ZH_YMIN 3
#define NX_ZH_YMAX 4
#define NX_ZH_ZMIN 5
#define NX_ZH_ZMAX 6
#define NX_ZH_VELOCITY (1u << 22)
#define NX_ZH_BITS (0xFu << 19)

#define NX_PERIODIC_X 1
#define NX_PERIODIC_Y 2
#define NX_PERIODIC_Z 4

typedef struct nx_stream_s* nx_stream;
typedef struct nx_event_s* nx_event;
typedef struct nx_bgrid_s* nx_bgrid;
typedef struct nx_block_tiles_s* nx_block_tiles;
typedef struct nx_container_s* nx_container;
typedef struct nx_batch_s* nx_batch;
typedef struct nx_mgrid_s* nx_mgrid;

/* Dense grid: one partition of a dGrid (dGrid_imp.h:7-185). */
typedef struct {
    int32_t dim[3];   /* X, Y, Z cells, wall layer included          */
    int32_t storage;  /* NX_F64 / NX_F32 / NX_F32DEV / NX_F16DEV     */
    int32_t compute;  /* NX_COMPUTE_*                                 */
    int32_t periodic; /* NX_PERIODIC_* bits                           */
} nx_dense_desc;

/* Block-sparse grid: one partition of a bGrid<8,8,8> (bGrid_imp.h:19-246). */
typedef struct {
    int32_t dim[3];
    int32_t storage;
    int32_t compute;
    int32_t periodic;
    uint32_t n_blocks;
    uint32_t reserved;
    const int32_t* origins; /* device, 3 x n_blocks: block origin (x,y,z)   */
    const uint32_t* masks;  /* device, 16 x n_blocks: active bits, x+8y+64z */
    const uint32_t* conn;   /* device, 27 x n_blocks: neighbour block or    */
                            /* 0xFFFFFFFF, direction i+3j+9k               */
} nx_block_desc;

typedef struct {
    double omega;    /* BGK relaxation frequency 1/tau                  */
    double force[3]; /* body acceleration g (Guo forcing); 0 = none     */
} nx_lbm_params;

/* Launch configuration of the container's interior kernel. */
typedef struct {
    int32_t wg;          /* threads per workgroup: 256 or 512             */
    int32_t cpt;         /* cells per thread: 1 or 2                      */
    int32_t load_policy; /* 0: cached loads, 1: non-temporal loads        */
    int32_t xcd_remap;   /* 1: contiguous slab of the grid per XCD        */
    int32_t vx;          /* cells per lane along x: 1, or 16 bytes of     */
                         /* cells (2 f64, 4 f32); not binary16            */
    int32_t boundary;    /* 0: interior bitmap, other lanes run the full  */
                         /* pull; 1: every lane reads its flags, masked   */
                         /* links are fixed up, one collision for all;    */
                         /* 2: as 1, interior lanes read the bitmap only  */
                         /* 3: as 2, and lanes in the two outer layers of */
                         /* the box fetch flags and wall-link values with */
                         /* the streaming loads (no dependent round trip) */
} nx_tuning;

/* ---- runtime (replaces libNeonSys GpuDevice / Backend, GpuDevice.h:151-190,
 *      src/set/Backend.cpp:46-55,404-419) -------------------------------- */
int nx_init(int device);
const char* nx_last_error(void);
int nx_device_count(int* count);
int nx_device_name(char* buf, int len);
int nx_sync(void); /* Backend::syncAll */
int nx_stream_create(nx_stream* s);
int nx_stream_sync(nx_stream s);
int nx_stream_destroy(nx_stream s);
int nx_event_create(nx_event* e);
int nx_event_record(nx_event e, nx_stream s);
int nx_event_sync(nx_event e);
int nx_event_elapsed_ms(nx_event start, nx_event stop, float* ms);
int nx_event_destroy(nx_event e);

/* ---- memory (replaces MemDevice / MemMirror updateDeviceData / updateHostData) */
int nx_alloc(size_t bytes, void** dev);
int nx_free(void* dev);
int nx_memset(void* dev, int value, size_t bytes, nx_stream s);
int nx_h2d(void* dst, const void* src, size_t bytes, nx_stream s);
int nx_d2h(void* dst, const void* src, size_t bytes, nx_stream s);
int nx_d2d(void* dst, const void* src, size_t bytes, nx_stream s);
int nx_host_alloc(size_t bytes, void** host); /* pinned */
/* Plain device-to-device streaming copy kernel (16-byte aligned); no
 * counterpart in the reference: the achievable-bandwidth yardstick. */
int nx_copy_kernel(void* dst, const void* src, size_t bytes, nx_stream s);
/* The dense step's access pattern without arithmetic (no counterpart in the
 * reference): 19 SoA planes of dim[0]*dim[1]*dim[2] elements of elem_bytes
 * (2, 4 or 8); every cell reads plane q at its upstream cell i - c_q
 * (clamped to the grid) and writes plane q at i.  Its bandwidth is the
 * ceiling of the pull-streaming step on this layout. */
int nx_shifted_copy(void* dst, const void* src, const int32_t dim[3], int32_t elem_bytes, nx_stream s);
int nx_host_free(void* host);

/* Bytes of one population field (19 cardinalities) for a storage type. */
size_t nx_pop_bytes_dense(const nx_dense_desc* d);
size_t nx_pop_bytes_block(const nx_block_desc* d);

/* ---- dense D3Q19 (dGrid) -----------------------------------------------
 * nx_cavity_flags_dense: lid-driven cavity classification of
 *   RunCavityTwoPop.cu:208-224 (outer layer bounceBack, y = Y-1 movingWall,
 *   mask bits cleared).
 * nx_d3q19_fill_dense: populations of RunCavityTwoPop.cu:168-206 by class:
 *   bulk f=w_q, bounceBack f=0, movingWall f=-6 w_q ulb (c_q . ulid).
 * nx_d3q19_wall_mask_dense: LbmContainers::computeWallNghMask
 *   (LbmTools.h:327-377); may run in place (flags_in == flags_out).
 * nx_d3q19_step_dense: LbmContainers::iteration (LbmTools.h:284-325):
 *   pull stream + bounce-back + BGK for every bulk cell; fin != fout.
 * nx_d3q19_moments_dense: LbmContainers::computeRhoAndU
 *   (LbmTools.h:379-437); rho [n], u [3][n]; double for NX_F64 storage,
 *   float otherwise.
 */
int nx_cavity_flags_dense(const nx_dense_desc* d, uint32_t* flags, nx_stream s);
int nx_d3q19_fill_dense(const nx_dense_desc* d, void* pop, const uint32_t* flags,
                        double ulb, const double ulid[3], nx_stream s);
int nx_d3q19_wall_mask_dense(const nx_dense_desc* d, const uint32_t* flags_in,
                             uint32_t* flags_out, nx_stream s);
int nx_d3q19_step_dense(const nx_dense_desc* d, const void* fin, void* fout,
                        const uint32_t* flags, const nx_lbm_params* p, nx_stream s);
/* As nx_d3q19_step_dense, and bulk cells with a Zou-He face bit run the
 * Zou-He completion of streamAndZouhe
 * (benchmarks/lbm-flow-over-sphere/src/LbmContainers.h:145-287; tables of
 * unknown and middle directions :633-646): pull,
 * known sum K over the 5 outgoing populations, middle sum M over the 9 with
 * no normal component; density face: rho = value, normal velocity
 * u.n = 1 - (M + 2K) / rho; velocity face: u.n = value,
 * rho = (M + 2K) / (1 - u.n) (the reference's velocity branch reads its
 * velocity as a density, :128-133 and :230; this is the textbook form);
 * tangential velocity 0; each unknown f_i = f_opp(i) + feq_i - feq_opp(i);
 * BGK with that rho and u.  value[6] is indexed by face - 1.  Storage
 * NX_F64 (compute F64) or NX_F32 (compute F32); force must be zero.
 * Containers reject flag fields holding Zou-He cells, and
 * nx_d3q19_step_dense treats them as plain bulk cells. */
int nx_d3q19_step_dense_zouhe(const nx_dense_desc* d, const void* fin, void* fout,
                              const uint32_t* flags, const nx_lbm_params* p,
                              const double value[6], nx_stream s);
int nx_d3q19_moments_dense(const nx_dense_desc* d, const void* fin,
                           const uint32_t* flags, void* rho, void* u, nx_stream s);

/* Container: LbmContainers::iteration (LbmTools.h:284-325) as a launch plan
 * bound to (fin, fout, flags, params), run once per time step like
 * Neon::set::Container::run (Containter.h:25-35).  Creation classifies the
 * bulk cells once (interior bitmap, held by the container); the flags must
 * not change while the container is alive.  A run is one kernel on the stream
 * and produces exactly the populations of nx_d3q19_step_dense.  When both
 * fields hold identical values in every non-bulk cell (as the reference's
 * set-up leaves them, RunCavityTwoPop.cu:168-206) the kernel also writes
 * those unchanged wall values into fout, so every cache line is written
 * whole; info reports this as copies_walls.  Creation also notes whether
 * the walls at both x ends (rows at least two cells from the y and z faces)
 * hold one common value per population, which lets the edge-speculating
 * kernel (boundary 3) store that value without loading it.  Wall values are
 * thus read when the container is made: after changing them, make the
 * container again.  The two fields of a ping-pong pair take one container
 * each (LbmIteration.h:41-42). */
int nx_d3q19_iteration_dense(const nx_dense_desc* d, const void* fin, void* fout,
                             const uint32_t* flags, const nx_lbm_params* p, nx_container* c);
int nx_container_run(nx_container c, nx_stream s);
/* n_steps of the ping-pong pair a, b, a, b, ... (b steps b's input, which is
 * a's output, back into a's input; LbmIterationD3Q19::run n times,
 * LbmIteration.h:56-69).  On a created stream, every 64 steps replay one
 * hipGraph captured at the first call (rebuilt when the launch
 * configuration changes); on the default stream each step is launched.
 * Not for concurrent use of the same containers from several threads. */
int nx_container_run_pair(nx_container a, nx_container b, int32_t n_steps, nx_stream s);
int nx_container_info(nx_container c, uint64_t* n_interior, uint64_t* n_boundary,
                      int* copies_walls);
int nx_container_destroy(nx_container c);

/* Launch configuration of the container's interior kernel, per (storage,
 * compute) pair, process-wide.  No counterpart in the reference (its launch
 * shape is fixed: dGrid_imp.h:16,27).  autotune times every configuration
 * on the container's buffers (fout is overwritten, fin only read), keeps the
 * fastest and returns it with its time per step. */
int nx_d3q19_get_tuning_dense(const nx_dense_desc* d, nx_tuning* t);
int nx_d3q19_set_tuning_dense(const nx_dense_desc* d, const nx_tuning* t);
int nx_container_autotune(nx_container c, nx_stream s, nx_tuning* best, float* best_ms);

/* FP64 reductions on the GPU (wave64 DPP + two-level tree; deterministic).
 * The per-workgroup partials live in one library-owned scratch buffer per
 * device, shared by every mass and residual entry point (dense and block)
 * and by nx_conv3d_bwd_data_bf16 / nx_conv3d_bwd_weight_bf16: calls of these
 * on one device must be ordered (one stream, or synchronised between
 * streams); two of them in flight at once on different streams race on the
 * scratch.  The buffer grows on demand (hipMalloc after a device-wide
 * synchronisation).
 * mass: out[0] = sum over bulk cells of sum_q f_q  (device double*).
 * residual: u from the post-stream moments of fin; out[0] = sum |u-u_prev|^2,
 *   out[1] = sum |u|^2 over bulk cells, then u_prev <- u
 *   (u_prev: float [3][n], caller-owned). */
int nx_d3q19_mass_dense(const nx_dense_desc* d, const void* fin,
                        const uint32_t* flags, double* out, nx_stream s);
int nx_d3q19_residual_dense(const nx_dense_desc* d, const void* fin,
                            const uint32_t* flags, float* u_prev, double* out,
                            nx_stream s);
/* Allocation-free variant of nx_d3q19_residual_dense: same arguments, same
 * kernels and bit-identical results, but its workgroup partials live in a
 * static device array of the library (at most 2048 partials of two doubles,
 * 32 KiB per GPU, held until the process ends) instead of the growing
 * scratch above, so it calls no hipMalloc, hipFree or allocating library
 * routine.  No counterpart in the reference: it is for callers that cannot
 * allocate (code under HIP graph capture, tight serving or convergence
 * loops, harnesses that forbid allocation).  The array's device address is
 * looked up by nx_init (or by the first call on a device nx_init did not
 * select).  Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (different streams, or host threads) race. */
int nx_d3q19_residual_dense_noalloc(const nx_dense_desc* d, const void* fin,
                                    const uint32_t* flags, float* u_prev, double* out,
                                    nx_stream s);

/* ---- block-sparse D3Q19 (bGrid, 8^3 blocks) -----------------------------
 * Host builder restating bGrid's constructor for one partition
 * (bGrid_imp.h:19-246, Partitioner1D.h:106-255, SpanClassifier.h:170-236):
 * blocks holding >= 1 active voxel, ordered bz, by, bx; bitmask bit
 * x+8y+64z; connectivity direction i+3j+9k for offset (i-1,j-1,k-1).
 * periodic != 0 wraps the connectivity (no counterpart in the reference;
 * periodic == 0 reproduces it).  `active` is X*Y*Z bytes, x fastest. */
int nx_bgrid_create(const int32_t dim[3], const uint8_t* active, int32_t periodic,
                    nx_bgrid* g);
int nx_bgrid_create_cb(const int32_t dim[3],
                       int (*active)(int32_t x, int32_t y, int32_t z, void* user),
                       void* user, int32_t periodic, nx_bgrid* g);
int nx_bgrid_info(nx_bgrid g, uint32_t* n_blocks, uint64_t* n_active);
int nx_bgrid_tables(nx_bgrid g, int32_t* origins, uint32_t* masks, uint32_t* conn);
int nx_bgrid_destroy(nx_bgrid g);

int nx_cavity_flags_block(const nx_block_desc* d, uint32_t* flags, nx_stream s);
int nx_d3q19_fill_block(const nx_block_desc* d, void* pop, const uint32_t* flags,
                        double ulb, const double ulid[3], nx_stream s);
int nx_d3q19_wall_mask_block(const nx_block_desc* d, const uint32_t* flags_in,
                             uint32_t* flags_out, nx_stream s);
int nx_d3q19_step_block(const nx_block_desc* d, const void* fin, void* fout,
                        const uint32_t* flags, const nx_lbm_params* p, nx_stream s);
/* Block-to-XCD assignment of nx_d3q19_step_block (1: each XCD walks one
 * contiguous range of blocks, the default; 0: hardware round-robin).  No
 * counterpart in the reference. */
int nx_d3q19_set_block_remap(int on);
/* Lanes of nx_d3q19_step_block at 2-byte storage (NX_F16DEV): 1 -- two
 * x-adjacent voxels per lane, one 4-byte load and store per population;
 * 0 -- one voxel per lane.  Same results bit for bit.  No counterpart in
 * the reference. */
int nx_d3q19_set_block_pairs(int on);
/* The same step with the block neighbourhood staged in LDS (the reference
 * reads neighbour blocks through bPartition::getNghData,
 * bPartition_imp.h:218-309): one workgroup per tile
 * of up to 2 x 2 x 1 blocks reads each member's population planes whole and
 * the tile's halo from the neighbour blocks once, and every voxel takes its
 * streamed values from LDS.  Bit for bit equal to nx_d3q19_step_block.
 * nx_block_tiles_create builds the tile table of a descriptor (blocking
 * copies of the origins); the table belongs to that descriptor's block
 * tables.  2- and 4-byte storage only (NX_ERR_UNSUPPORTED for NX_F64). */
int nx_block_tiles_create(const nx_block_desc* d, nx_block_tiles* t);
int nx_block_tiles_info(nx_block_tiles t, uint32_t* n_tiles);
int nx_block_tiles_destroy(nx_block_tiles t);
int nx_d3q19_step_block_tiled(const nx_block_desc* d, nx_block_tiles t, const void* fin,
                              void* fout, const uint32_t* flags, const nx_lbm_params* p,
                              nx_stream s);
int nx_d3q19_moments_block(const nx_block_desc* d, const void* fin,
                           const uint32_t* flags, void* rho, void* u, nx_stream s);
int nx_d3q19_mass_block(const nx_block_desc* d, const void* fin,
                        const uint32_t* flags, double* out, nx_stream s);
int nx_d3q19_residual_block(const nx_block_desc* d, const void* fin,
                            const uint32_t* flags, float* u_prev, double* out,
                            nx_stream s);

/* ---- batched independent cavities (no counterpart in the reference) ------
 * n members, each a dense grid dims[3k..3k+2] with classification classes
 * (uint8 NX_CELL_*, concatenated in member order, x fastest), relaxation
 * omega[k], lid ulb[k] * ulid[3k..3k+2] (fill of RunCavityTwoPop.cu:168-206).
 * One arena per population field; member k's populations are its dGrid SoA
 * layout.  Steps run as one persistent launch for all members; nx_batch_run
 * captures check_every steps plus the residual test in a hipGraph and
 * replays it until every member's residual
 *   r = sqrt(sum |u_t - u_{t-K}|^2 / sum |u_t|^2)   (bulk cells, FP64 sums)
 * is below tol (converged members stop stepping) or max_steps are taken.
 * A member whose residual sums are not finite (its flow diverged) stops
 * stepping too and reports a NaN residual; n_unconverged counts only the
 * members still stepping. 
 * nx_batch_init sets bulk populations to the equilibrium of (rho, u)
 * (device arrays in arena cell order; NULL: rest).  nx_batch_member returns a
 * member's dense view (arena 0 holds the latest populations after an even
 * number of steps). */
int nx_batch_create(uint32_t n, const int32_t* dims, const uint8_t* classes,
                    const double* omega, const double* ulb, const double* ulid,
                    int32_t storage, nx_batch* b);
int nx_batch_info(nx_batch b, uint32_t* n_members, uint64_t* n_cells, uint32_t* n_tiles);
int nx_batch_member(nx_batch b, uint32_t k, int which, nx_dense_desc* d, void** pop,
                    uint32_t** flags);
int nx_batch_init(nx_batch b, const float* rho, const float* u, nx_stream s);
/* After nx_batch_init from a predicted field: `iterations` times, two steps
 * then every bulk cell's equilibrium part is moved to the prescribed u,
 * f <- f - feq(rho, u(f)) + feq(rho, u) (consistent initial conditions after
 * Mei, Luo, Lallemand & d'Humieres 2006): density and non-equilibrium parts
 * adapt to the field before the free iteration starts.  Only the members
 * whose byte in `members` (host, n_members; NULL: all) is non-zero step and
 * are constrained.  Step counters are not advanced (the caller accounts for
 * the 2 * iterations steps); every member is stepping afterwards. */
int nx_batch_constrain(nx_batch b, const float* u, int32_t iterations, const uint8_t* members,
                       nx_stream s);
int nx_batch_step(nx_batch b, int32_t k_steps, nx_stream s);
/* nx_batch_run replays on the batch's own stream and first waits for the
 * whole device (hipDeviceSynchronize), so work the caller queued on any
 * stream (nx_batch_init, nx_batch_step, ...) is complete before it starts;
 * it returns after its last replay has finished. */
int nx_batch_run(nx_batch b, int32_t max_steps, int32_t check_every, double tol,
                 int32_t* steps, double* residual, uint32_t* n_unconverged);
/* Launch of the persistent step: tiles of wg (256 or 512) cells, wgs_per_cu
 * workgroups per CU (0: one workgroup per tile).  A persistent grid deals
 * the tiles to the 8 XCDs in chunks of 65536 consecutive cells, so that one
 * L2 serves a member's neighbouring planes; nx_batch_autotune also times
 * tile-by-tile dealing and keeps the faster. */
int nx_batch_set_launch(nx_batch b, int32_t wg, int32_t wgs_per_cu);
/* Step variant: xcd_chunks (1: deal tiles to XCDs in 65536-cell chunks, 0:
 * tile by tile) and edge (1: cells in the two outer layers of a cavity fetch
 * their wall-link values with the streaming loads; 0: after their flag word).
 * Both give identical results.  nx_batch_get_launch reports the current
 * launch (after nx_batch_autotune: the fastest one measured). */
int nx_batch_set_variant(nx_batch b, int32_t xcd_chunks, int32_t edge);
int nx_batch_get_launch(nx_batch b, int32_t* wg, int32_t* wgs_per_cu, int32_t* xcd_chunks,
                        int32_t* edge);
/* Workgroups of the persistent step for n_tiles tiles on n_cu CUs with
 * wgs_per_cu per CU (0: one per tile): min(n_tiles, n_cu * wgs_per_cu) in
 * 64-bit arithmetic, at least 1.  Host-only, no device work.
 * nx_batch_launch_grid reports the grid and how the batch's current launch
 * deals tiles to the XCDs: 0 round robin, 1 persistent grid in XCD chunks,
 * 2 one workgroup per tile with each XCD taking a contiguous range of tiles
 * (xcd_chunks = 1 with wgs_per_cu = 0). */
int nx_batch_grid_size(uint32_t n_tiles, int32_t n_cu, int32_t wgs_per_cu, uint32_t* grid);
int nx_batch_launch_grid(nx_batch b, uint32_t* grid, int32_t* xcd_chunked);
int nx_batch_autotune(nx_batch b, nx_stream s, int32_t* best_wg, int32_t* best_wgs_per_cu,
                      float* best_ms);
int nx_batch_destroy(nx_batch b);

/* ---- grid refinement (Neon's mGrid with depth 2 to 5, BGK D3Q19 or KBC
 * D3Q27) ----------------------------------------------------------------------
 * Level 0 is the finest level (the domain dim[3]); level l has voxels of
 * edge 2^l, ceil(dim / 2^l) of them per axis, and a level-(l+1) voxel v
 * covers the level-l voxels 2v + {0,1}^3.  Each level is stored as a bGrid of
 * 8^3 blocks (AoSoA populations, as nx_d3q19_step_block).
 * nx_mgrid_create_levels replaces the mGrid constructor
 *   (libNeonDomain/src/domain/details/mGrid/mGrid.cpp:89-345): lams[l] is
 *   the level-l activation lambda at every level-l voxel (evaluated at its
 *   level-0 origin v 2^l), uint8, x fastest, for l < depth (2..5).  Every
 *   2^3 group holding an active voxel becomes active (and so does its
 *   parent); from the coarsest level down, refined voxels whose in-domain
 *   neighbours are all refined are removed; the strong-balance pass then
 *   activates the voxels that keep every neighbour of an active voxel of
 *   level l covered on level l or l + 1.  storage NX_F64 or NX_F32.
 *   lattice NX_MGRID_BGK19 (collide.h:8-149 on the D3Q19 of
 *   lattice.h:108-173) or NX_MGRID_KBC27 (collide.h:148-284 on the D3Q27 of
 *   lattice.h:14-106); fields then hold 19 or 27 populations per voxel
 *   (nx_mgrid_populations).
 * nx_mgrid_create: the same for depth 2 (lam0, lam1) and BGK D3Q19.
 * nx_mgrid_depth: the depth and, if dims is not NULL, the voxels per axis of
 *   every level, dims[3 * l + axis].
 * nx_mgrid_state_level: the state byte of every voxel of one level (dense,
 *   x fastest): bit 0 active, bit 1 refined (has children), bits 2..3 the
 *   NX_CELL_* class once nx_mgrid_cavity has run.  nx_mgrid_state: levels 0
 *   and 1 (either pointer may be NULL).
 * nx_mgrid_cavity replaces initLidDrivenCavity + initSumStore
 *   (apps/lbmMultiRes/lidDrivenCavity.h:8-78, init.h:8-86): leaf voxels of
 *   level 0 on the domain faces are bounce-back, y = Y-1 the lid moving at
 *   (ulb, 0, 0); populations w_q / 0 / -6 w_q c_q.ulid in both fields; the
 *   coalescence weights 1 / (2 n) of every refined voxel.
 * nx_mgrid_setup_cells is the same set-up for any problem: cls[l] gives the
 *   class of every level-l voxel (NX_CELL_BULK, NX_CELL_BOUNCEBACK or
 *   NX_CELL_MOVINGWALL; read for the leaves, x fastest) and u_wall the
 *   velocity of the moving-wall cells -- the lid, or the inlet of the flow
 *   over a shape (flowOverShape.h:17-113), whose cells the stream treats
 *   alike.
 * nx_mgrid_step runs n steps of the coarsest level, each the reference's
 *   nonUniformTimestepRecursive (apps/lbmMultiRes/lbmMultiRes.h:141-263)
 *   with the baseline operation order and the fine store: per level l from
 *   the top, collide l (collide.h:8-149) and store l into l + 1
 *   (store.h:98-163), the step of level l - 1, stream l with explosion and
 *   coalescence (stream.h:170-225), and below the top level all of it once
 *   more (two level-l steps per level-(l+1) step).  omega0 is the coarsest
 *   level's relaxation; level l uses computeOmega (util.h:40-45).  The
 *   fine-to-coarse sums are added in the fine voxels' z, y, x order (the
 *   reference adds with atomics in no fixed order), so results are
 *   deterministic.
 * A level's stream runs fused with its next collide (one pass instead of
 * two, same results) when no bulk leaf of the level has a link without a
 * source; the last stream of each level in a call is not fused, so the
 * fields between calls are those of the unfused sequence.  nx_mgrid_info
 * reports the fused levels (fine_fused, bit l for level l, after
 * nx_mgrid_cavity) with the block, active and leaf voxel counts per level
 * (arrays of depth entries) and the number of fine sources of all stores.
 * nx_mgrid_fields copies fin (which 0) or fout (which 1) of one level into
 *   a host array [19][Z][Y][X] of that level's dimensions; entries of
 *   inactive voxels are not written.  Synchronises the device.
 * A handle is not for concurrent use from several host threads (a step
 * swaps level 0's two buffers when it is queued). */
#define NX_MGRID_BGK19 0 /* BGK collision on D3Q19 (the app's BGK build) */
#define NX_MGRID_KBC27 1 /* KBC collision on D3Q27 (the app's default build) */
int nx_mgrid_create_levels(const int32_t dim[3], int32_t depth, const uint8_t* const* lams, int32_t storage,
                           int32_t lattice, nx_mgrid* g);
int nx_mgrid_create(const int32_t dim[3], const uint8_t* lam0, const uint8_t* lam1, int32_t storage,
                    nx_mgrid* g);
int nx_mgrid_depth(nx_mgrid g, int32_t* depth, int32_t* dims);
int nx_mgrid_populations(nx_mgrid g, int32_t* q);
int nx_mgrid_info(nx_mgrid g, uint32_t* n_blocks, uint64_t* n_active, uint64_t* n_leaf, uint64_t* n_store,
                  int32_t* fine_fused);
int nx_mgrid_state_level(nx_mgrid g, int32_t level, uint8_t* st);
int nx_mgrid_state(nx_mgrid g, uint8_t* st0, uint8_t* st1);
int nx_mgrid_cavity(nx_mgrid g, double ulb);
int nx_mgrid_setup_cells(nx_mgrid g, const uint8_t* const* cls, const double u_wall[3]);
int nx_mgrid_step(nx_mgrid g, double omega0, int32_t n_steps, nx_stream s);
int nx_mgrid_fields(nx_mgrid g, int32_t level, int32_t which, void* dense);
int nx_mgrid_destroy(nx_mgrid g);

/* ---- geometry (benchmark workloads; no counterpart in the reference) ------
 * Union of n seeded spheres, centres uniform in the box, radii uniform in
 * [rmin, rmax), SplitMix64 stream (draws per sphere: cx, cy, cz, r); voxel
 * (x,y,z) is solid when |(x,y,z) - c| < r, periodic axes measured across
 * the wrap.  `solid` is X*Y*Z bytes, x fastest. */
int nx_geometry_spheres(const int32_t dim[3], uint32_t n_spheres, double rmin, double rmax,
                        uint64_t seed, int32_t periodic, uint8_t* solid);
/* Lid-driven cavity classes (RunCavityTwoPop.cu:208-224) plus n_boxes
 * bounce-back boxes, edges uniform in {emin..emax}, lower corner uniform in
 * {1 .. dim-1-edge} per axis (draws per box: ex, ey, ez, x0, y0, z0),
 * clipped to the interior.  `cls` is X*Y*Z bytes of NX_CELL_*. */
int nx_geometry_cavity(const int32_t dim[3], uint32_t n_boxes, int32_t emin, int32_t emax,
                       uint64_t seed, uint8_t* cls);

/* ---- BF16 3D convolution on the matrix cores (learned warm start; no
 * counterpart in the reference) --------------------------------------------
 * Channels-last volumes x[n][z][y][x][c] (bf16 bit patterns), weights
 * w[co][tap][ci] (bf16), tap = (dz*k + dy)*k + dx over offsets
 * -(k-1)/2..(k-1)/2, zero padding, stride 1; every product accumulated in
 * FP32.  shape = {n, depth, height, width} (powers of two); cin, cout
 * multiples of 16 in [16, 256]; ksize 1 or 3.  x, w, dy 16-byte aligned.
 *   fwd:        y = conv(x, w) + bias (bias may be NULL), ReLU with
 *               NX_CONV_RELU; y bf16, or float with NX_CONV_OUT_F32
 *   bwd_data:   dx = gradient of the convolution w.r.t. x (bf16 or float)
 *   bwd_weight: dw[co][tap][ci] (float), db[co] (float, may be NULL)
 * bwd_data and bwd_weight use the library's per-device workspace: calls
 * from several host threads on one device must not overlap. */
#define NX_CONV_RELU 1
#define NX_CONV_OUT_F32 2
int nx_conv3d_fwd_bf16(const uint16_t* x, const uint16_t* w, const float* bias, void* y,
                       const int32_t shape[4], int32_t cin, int32_t cout, int32_t ksize,
                       int32_t flags, nx_stream s);
int nx_conv3d_bwd_data_bf16(const uint16_t* dy, const uint16_t* w, void* dx,
                            const int32_t shape[4], int32_t cin, int32_t cout, int32_t ksize,
                            int32_t flags, nx_stream s);
int nx_conv3d_bwd_weight_bf16(const uint16_t* x, const uint16_t* dy, float* dw, float* db,
                              const int32_t shape[4], int32_t cin, int32_t cout,
                              int32_t ksize, nx_stream s);

/* ---- U-Net surrogate for the warm start (no counterpart in the reference) -
 * levels 2..4, widths multiples of 16 (<= 128), input [n][size^3][in_ch]
 * float (in_ch <= 16), output [n][size^3][out_ch] float (out_ch <= 16),
 * cubes of side `size` (power of two).  Encoder: two 3^3 convolutions + ReLU
 * per level and 2x2x2 average pooling; decoder: nearest upsampling, skip
 * concatenation, two 3^3 convolutions + ReLU; a 1^3 output convolution.  BF16
 * weights and activations, FP32 accumulation, FP32 master weights.  The
 * parameter vector (float) holds, per convolution in the order encoder
 * (level 0..L-1, 2 each), decoder (level 0..L-2, 2 each), output:
 * W[co][tap][ci] then b[co] (first convolution cin = 16, output cout = 16,
 * channels past in_ch / out_ch unused).
 * nx_unet_train_step: forward, loss = sum_v weight[v] sum_c (y - target)^2
 * (weight [n][size^3]), backward, one Adam step (beta 0.9/0.999, eps 1e-8);
 * *loss (host, may be NULL: then no synchronisation) receives the loss
 * before the step.  nx_unet_backward: parameter gradients (nx_unet_grads)
 * of sum(y * dy) for the input of the last nx_unet_forward. */
typedef struct nx_unet_s* nx_unet;
typedef struct {
    int32_t levels;
    int32_t width[4];
    int32_t size;
    int32_t max_batch;
    int32_t in_ch;
    int32_t out_ch;
    int32_t reserved;
    uint64_t seed;      /* He-normal initialisation (SplitMix64 stream) */
} nx_unet_desc;
int nx_unet_create(const nx_unet_desc* d, nx_unet* u);
int nx_unet_info(nx_unet u, uint64_t* n_params, uint32_t* n_convs);
int nx_unet_params(nx_unet u, float* host, int32_t set);
int nx_unet_grads(nx_unet u, float* host);
int nx_unet_forward(nx_unet u, const float* x, int32_t batch, float* y, nx_stream s);
int nx_unet_backward(nx_unet u, const float* dy, int32_t batch, nx_stream s);
int nx_unet_train_step(nx_unet u, const float* x, const float* target, const float* weight,
                       int32_t batch, float lr, double* loss, nx_stream s);
int nx_unet_destroy(nx_unet u);

#ifdef __cplusplus
}
#endif
#endif /* NEON_AMD_H */
