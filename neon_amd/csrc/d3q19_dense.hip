// d3q19_dense.hip -- D3Q19 kernels on a dense grid (the dGrid path).
//
// Layout: SoA populations f[q][z][y][x] (dField_imp.h:67-87,
// dPartition.h:90-99), one uint32 flag word per cell.  One thread per cell,
// 256 cells (4 wavefronts) per workgroup over the linear cell index, so every
// population load and store of a wavefront is one contiguous 64-cell run;
// the +-1 x shifts of the pull read a neighbouring cache line that the
// adjacent wavefront of the same workgroup fetches anyway.  Workgroups are
// remapped so each XCD walks a contiguous slab (L2 keeps the row-boundary
// lines).  Each population is read once and written once per step:
// 2*19*sizeof(storage) bytes per cell update.
#include <cmath>
#include <mutex>
#include <string>

#include "dense_common.hpp"

namespace nx {
namespace dense {
namespace {

// Generic step (any boundary mode, including periodic wrap).
template <int S, int CM, bool FORCE, bool PERIODIC>
__global__ void __launch_bounds__(kWG)
    k_step(const typename Stor<S>::T* __restrict__ fin, typename Stor<S>::T* __restrict__ fout,
           const uint32_t* __restrict__ flags, DGeom g, Par p)
{
    using R = typename Real<S, CM>::R;
    const uint32_t i = xcd_remap(blockIdx.x, gridDim.x) * kWG + threadIdx.x;
    if (i >= g.n) return;
    const uint32_t fl = __builtin_nontemporal_load(&flags[i]);
    if (cls_of(fl) != NX_CELL_BULK) return;
    typename Stor<S>::P f[Q];
    pull<S, PERIODIC>(fin, g, i, fl, f);
    R out[Q];
    collide<S, CM, FORCE>(f, p.omega, p.g, out);
#pragma unroll
    for (int q = 0; q < Q; q++)
        __builtin_nontemporal_store(Stor<S>::template enc<R>(out[q], q), &fout[q * g.plane + i]);
}

// Step with Zou-He boundary cells (no body force): bulk cells whose flag
// carries a face in bits 19..21 complete their 5 unknown populations and
// collide with the prescribed density or velocity (zouhe_bgk); every other
// bulk cell is the plain step.
struct ZhValues {
    double v[6];
};

template <int S, int CM, bool PERIODIC>
__global__ void __launch_bounds__(kWG)
    k_step_zh(const typename Stor<S>::T* __restrict__ fin, typename Stor<S>::T* __restrict__ fout,
              const uint32_t* __restrict__ flags, DGeom g, Par p, ZhValues zh)
{
    using R = typename Real<S, CM>::R;
    const uint32_t i = xcd_remap(blockIdx.x, gridDim.x) * kWG + threadIdx.x;
    if (i >= g.n) return;
    const uint32_t fl = __builtin_nontemporal_load(&flags[i]);
    if (cls_of(fl) != NX_CELL_BULK) return;
    typename Stor<S>::P f[Q];
    pull<S, PERIODIC>(fin, g, i, fl, f);
    R out[Q];
    const int face = (int)((fl >> NX_ZH_SHIFT) & 7u);
    if (face >= 1 && face <= 6) {
        R fr[Q];
#pragma unroll
        for (int q = 0; q < Q; q++) fr[q] = (R)f[q];
        zouhe_bgk<R>(fr, face, (fl & NX_ZH_VELOCITY) != 0, (R)zh.v[face - 1], (R)p.omega, out);
    } else {
        collide<S, CM, false>(f, p.omega, p.g, out);
    }
#pragma unroll
    for (int q = 0; q < Q; q++)
        __builtin_nontemporal_store(Stor<S>::template enc<R>(out[q], q), &fout[q * g.plane + i]);
}

// Cavity classification, RunCavityTwoPop.cu:208-224.
__global__ void __launch_bounds__(kWG) k_cavity_flags(uint32_t* flags, DGeom g)
{
    const uint32_t i = blockIdx.x * kWG + threadIdx.x;
    if (i >= g.n) return;
    int x, y, z;
    coords(g, i, x, y, z);
    int cls = NX_CELL_BULK;
    if (x == 0 || x == g.X - 1 || y == 0 || y == g.Y - 1 || z == 0 || z == g.Z - 1) {
        cls = NX_CELL_BOUNCEBACK;
        if (y == g.Y - 1) cls = NX_CELL_MOVINGWALL;
    }
    flags[i] = (uint32_t)cls << NX_CLASS_SHIFT;
}

// Initial populations by class, RunCavityTwoPop.cu:168-206.
template <int S>
__global__ void __launch_bounds__(kWG)
    k_fill(typename Stor<S>::T* pop, const uint32_t* flags, DGeom g, double ulb, double l0,
           double l1, double l2)
{
    const uint32_t i = blockIdx.x * kWG + threadIdx.x;
    if (i >= g.n) return;
    const int cls = cls_of(flags[i]);
#pragma unroll
    for (int q = 0; q < Q; q++) {
        typename Stor<S>::T v;
        if (cls == NX_CELL_BULK) {
            v = Stor<S>::enc_full(wq(q), q);
        } else if (cls == NX_CELL_MOVINGWALL) {
            const double val = -6. * wq(q) * ulb * (CX[q] * l0 + CY[q] * l1 + CZ[q] * l2);
            // binary16 goes through binary32 (the oracle's rounding path)
            v = (typename Stor<S>::T)(typename Stor<S>::P)val;
        } else {
            v = (typename Stor<S>::T)0.;
        }
        pop[q * g.plane + i] = v;
    }
}

// computeWallNghMask, LbmTools.h:327-377: bit q <=> upstream cell of q is not
// bulk.  In place is allowed: only the mask bits of bulk cells change and the
// kernel reads classification bits only.
__global__ void __launch_bounds__(kWG)
    k_wall_mask(const uint32_t* fin, uint32_t* fout, DGeom g)
{
    const uint32_t i = blockIdx.x * kWG + threadIdx.x;
    if (i >= g.n) return;
    const uint32_t fl = fin[i];
    if (cls_of(fl) != NX_CELL_BULK) {
        fout[i] = fl;
        return;
    }
    int x, y, z;
    coords(g, i, x, y, z);
    uint32_t m = 0, open = 0, closed = 0;
#pragma unroll
    for (int q = 0; q < Q; q++) {
        if (q == 9) continue;
        uint32_t j;
        if (!upstream(g, x, y, z, q, j)) {
            m |= 1u << q;
            open = NX_FLAG_OPEN;
        } else if (cls_of(fin[j]) != NX_CELL_BULK) {
            m |= 1u << q;
            closed = 1;
        }
    }
    fout[i] = ((uint32_t)NX_CELL_BULK << NX_CLASS_SHIFT) | m | open |
              (open && closed ? NX_FLAG_MIXED : 0u) | (fl & NX_ZH_BITS);
}

template <typename M>
__device__ __forceinline__ void put_moments(M* rho, M* u, const DGeom& g, uint32_t i, double r,
                                            const double uu[3])
{
    rho[i] = (M)r;
    u[i] = (M)uu[0];
    u[g.plane + i] = (M)uu[1];
    u[2 * g.plane + i] = (M)uu[2];
}

// computeRhoAndU, LbmTools.h:379-437
template <int S, int CM, bool PERIODIC>
__global__ void __launch_bounds__(kWG)
    k_moments(const typename Stor<S>::T* __restrict__ fin, const uint32_t* __restrict__ flags,
              void* rho_, void* u_, DGeom g)
{
    using R = typename Real<S, CM>::R;
    using M = typename std::conditional<S == NX_F64, double, float>::type;
    const uint32_t i = blockIdx.x * kWG + threadIdx.x;
    if (i >= g.n) return;
    const uint32_t fl = flags[i];
    const int cls = cls_of(fl);
    typename Stor<S>::P f[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) f[q] = 0;
    if (cls == NX_CELL_BULK) pull<S, PERIODIC>(fin, g, i, fl, f);
    if (cls == NX_CELL_MOVINGWALL)
        for (int a = 0; a < 3; a++) f[a] = Stor<S>::raw(fin[a * g.plane + i]);
    R r, uu[3];
    cell_moments<S, CM>(cls, f, r, uu);
    const double ud[3] = {(double)uu[0], (double)uu[1], (double)uu[2]};
    put_moments<M>((M*)rho_, (M*)u_, g, i, (double)r, ud);
}

constexpr int kRedGrid = 2048;

// FP64 mass: sum over bulk cells of sum_q f_q (f decoded, summed in q order)
template <int S>
__global__ void __launch_bounds__(kWG)
    k_mass(const typename Stor<S>::T* __restrict__ fin, const uint32_t* __restrict__ flags,
           DGeom g, double* partials)
{
    __shared__ double lds[16];
    double acc[1] = {0.0};
    for (uint32_t i = blockIdx.x * kWG + threadIdx.x; i < g.n; i += gridDim.x * kWG) {
        if (cls_of(flags[i]) != NX_CELL_BULK) continue;
        double r = 0;
#pragma unroll
        for (int q = 0; q < Q; q++) r += Stor<S>::full(fin[q * g.plane + i], q);
        acc[0] += r;
    }
    block_sum_f64<1>(acc, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = acc[0];
}

// residual: sum |u - u_prev|^2 and sum |u|^2 over bulk cells; u_prev <- u
template <int S, int CM, bool PERIODIC>
__global__ void __launch_bounds__(kWG)
    k_residual(const typename Stor<S>::T* __restrict__ fin, const uint32_t* __restrict__ flags,
               float* __restrict__ u_prev, DGeom g, double* partials)
{
    using R = typename Real<S, CM>::R;
    __shared__ double lds[32];
    double acc[2] = {0.0, 0.0};
    for (uint32_t i = blockIdx.x * kWG + threadIdx.x; i < g.n; i += gridDim.x * kWG) {
        const uint32_t fl = flags[i];
        if (cls_of(fl) != NX_CELL_BULK) continue;
        typename Stor<S>::P f[Q];
        pull<S, PERIODIC>(fin, g, i, fl, f);
        R r, uu[3];
        cell_moments<S, CM>(NX_CELL_BULK, f, r, uu);
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const float un = (float)uu[a];
            const double du = (double)un - (double)u_prev[a * g.plane + i];
            acc[0] += du * du;
            acc[1] += (double)un * (double)un;
            u_prev[a * g.plane + i] = un;
        }
    }
    block_sum_f64<2>(acc, lds);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = acc[0];
        partials[2 * blockIdx.x + 1] = acc[1];
    }
}

template <int S, int CM>
void dispatch_step(bool force, bool per, const void* fin, void* fout, const uint32_t* flags,
                   const DGeom& g, const Par& p, hipStream_t s)
{
    using T = typename Stor<S>::T;
    if (force) {
        if (per)
            hipLaunchKernelGGL((k_step<S, CM, true, true>), dim3(nwg(g)), dim3(kWG), 0, s,
                               (const T*)fin, (T*)fout, flags, g, p);
        else
            hipLaunchKernelGGL((k_step<S, CM, true, false>), dim3(nwg(g)), dim3(kWG), 0, s,
                               (const T*)fin, (T*)fout, flags, g, p);
    } else {
        if (per)
            hipLaunchKernelGGL((k_step<S, CM, false, true>), dim3(nwg(g)), dim3(kWG), 0, s,
                               (const T*)fin, (T*)fout, flags, g, p);
        else
            hipLaunchKernelGGL((k_step<S, CM, false, false>), dim3(nwg(g)), dim3(kWG), 0, s,
                               (const T*)fin, (T*)fout, flags, g, p);
    }
}

}  // namespace
}  // namespace dense
}  // namespace nx


using namespace nx;
using namespace nx::dense;

extern "C" size_t nx_pop_bytes_dense(const nx_dense_desc* d)
{
    if (!d) return 0;
    const size_t n = (size_t)d->dim[0] * d->dim[1] * d->dim[2] * 19;
    switch (d->storage) {
        case NX_F64: return n * 8;
        case NX_F32:
        case NX_F32DEV: return n * 4;
        case NX_F16DEV: return n * 2;
    }
    return 0;
}

extern "C" int nx_cavity_flags_dense(const nx_dense_desc* d, uint32_t* flags, nx_stream s)
{
    if (int e = check(d)) return e;
    if (!flags) return set_error(NX_ERR_INVALID, "null flags");
    const DGeom g = make_geom(d);
    hipLaunchKernelGGL(k_cavity_flags, dim3(nwg(g)), dim3(kWG), 0, (hipStream_t)s, flags, g);
    NX_CHECK_LAUNCH("k_cavity_flags");
    return NX_OK;
}


extern "C" int nx_d3q19_fill_dense(const nx_dense_desc* d, void* pop, const uint32_t* flags,
                                   double ulb, const double ulid[3], nx_stream s)
{
    if (int e = check(d)) return e;
    if (!pop || !flags || !ulid) return set_error(NX_ERR_INVALID, "null pointer");
    const DGeom g = make_geom(d);
    const hipStream_t st = (hipStream_t)s;
    if (int e = by_type(d, [&](auto sc, auto) {
            constexpr int S = decltype(sc)::value;
            hipLaunchKernelGGL((k_fill<S>), dim3(nwg(g)), dim3(kWG), 0, st,
                               (typename Stor<S>::T*)pop, flags, g, ulb, ulid[0], ulid[1], ulid[2]);
        }))
        return e;
    NX_CHECK_LAUNCH("k_fill");
    return NX_OK;
}

extern "C" int nx_d3q19_wall_mask_dense(const nx_dense_desc* d, const uint32_t* flags_in,
                                        uint32_t* flags_out, nx_stream s)
{
    if (int e = check(d)) return e;
    if (!flags_in || !flags_out) return set_error(NX_ERR_INVALID, "null flags");
    const DGeom g = make_geom(d);
    hipLaunchKernelGGL(k_wall_mask, dim3(nwg(g)), dim3(kWG), 0, (hipStream_t)s, flags_in,
                       flags_out, g);
    NX_CHECK_LAUNCH("k_wall_mask");
    return NX_OK;
}

extern "C" int nx_d3q19_step_dense(const nx_dense_desc* d, const void* fin, void* fout,
                                   const uint32_t* flags, const nx_lbm_params* p, nx_stream s)
{
    if (int e = check(d)) return e;
    if (!fin || !fout || !flags || !p) return set_error(NX_ERR_INVALID, "null pointer");
    if (fin == fout)
        return set_error(NX_ERR_INVALID, "fin and fout must differ (two-population scheme)");
    const DGeom g = make_geom(d);
    const Par par{p->omega, {p->force[0], p->force[1], p->force[2]}};
    const bool force = p->force[0] != 0. || p->force[1] != 0. || p->force[2] != 0.;
    if (force && d->compute == NX_COMPUTE_REF)
        return set_error(NX_ERR_UNSUPPORTED,
                         "body force requires NX_COMPUTE_F32 or NX_COMPUTE_F64");
    const bool per = d->periodic != 0;
    const hipStream_t st = (hipStream_t)s;
    if (int e = by_type(d, [&](auto sc, auto cc) {
            constexpr int S = decltype(sc)::value;
            constexpr int CM = decltype(cc)::value;
            dispatch_step<S, CM>(force, per, fin, fout, flags, g, par, st);
        }))
        return e;
    NX_CHECK_LAUNCH("k_step");
    return NX_OK;
}

extern "C" int nx_d3q19_step_dense_zouhe(const nx_dense_desc* d, const void* fin, void* fout,
                                         const uint32_t* flags, const nx_lbm_params* p,
                                         const double value[6], nx_stream s)
{
    if (int e = check(d)) return e;
    if (!fin || !fout || !flags || !p || !value) return set_error(NX_ERR_INVALID, "null pointer");
    if (fin == fout)
        return set_error(NX_ERR_INVALID, "fin and fout must differ (two-population scheme)");
    if (p->force[0] != 0. || p->force[1] != 0. || p->force[2] != 0.)
        return set_error(NX_ERR_UNSUPPORTED, "Zou-He step: body force must be zero");
    const bool f64 = d->storage == NX_F64 && d->compute == NX_COMPUTE_F64;
    const bool f32 = d->storage == NX_F32 && d->compute == NX_COMPUTE_F32;
    if (!f64 && !f32)
        return set_error(NX_ERR_UNSUPPORTED,
                         "Zou-He step: storage/compute NX_F64/NX_COMPUTE_F64 or NX_F32/NX_COMPUTE_F32");
    for (int k = 0; k < 6; k++)
        if (!std::isfinite(value[k])) return set_error(NX_ERR_INVALID, "Zou-He step: non-finite value");
    const DGeom g = make_geom(d);
    const Par par{p->omega, {0., 0., 0.}};
    ZhValues zh;
    for (int k = 0; k < 6; k++) zh.v[k] = value[k];
    const hipStream_t st = (hipStream_t)s;
    auto go = [&](auto sc, auto cc, auto pc) {
        constexpr int S = decltype(sc)::value, CM = decltype(cc)::value;
        constexpr bool PER = decltype(pc)::value;
        using T = typename Stor<S>::T;
        hipLaunchKernelGGL((k_step_zh<S, CM, PER>), dim3(nwg(g)), dim3(kWG), 0, st, (const T*)fin,
                           (T*)fout, flags, g, par, zh);
    };
    if (f64) {
        if (g.periodic) go(std::integral_constant<int, NX_F64>{}, std::integral_constant<int, NX_COMPUTE_F64>{}, std::true_type{});
        else go(std::integral_constant<int, NX_F64>{}, std::integral_constant<int, NX_COMPUTE_F64>{}, std::false_type{});
    } else {
        if (g.periodic) go(std::integral_constant<int, NX_F32>{}, std::integral_constant<int, NX_COMPUTE_F32>{}, std::true_type{});
        else go(std::integral_constant<int, NX_F32>{}, std::integral_constant<int, NX_COMPUTE_F32>{}, std::false_type{});
    }
    NX_CHECK_LAUNCH("k_step_zh");
    return NX_OK;
}

extern "C" int nx_d3q19_moments_dense(const nx_dense_desc* d, const void* fin,
                                      const uint32_t* flags, void* rho, void* u, nx_stream s)
{
    if (int e = check(d)) return e;
    if (!fin || !flags || !rho || !u) return set_error(NX_ERR_INVALID, "null pointer");
    const DGeom g = make_geom(d);
    const hipStream_t st = (hipStream_t)s;
    if (int e = by_type(d, [&](auto sc, auto cc) {
            constexpr int S = decltype(sc)::value;
            constexpr int CM = decltype(cc)::value;
            using T = typename Stor<S>::T;
            if (g.periodic)
                hipLaunchKernelGGL((k_moments<S, CM, true>), dim3(nwg(g)), dim3(kWG), 0, st,
                                   (const T*)fin, flags, rho, u, g);
            else
                hipLaunchKernelGGL((k_moments<S, CM, false>), dim3(nwg(g)), dim3(kWG), 0, st,
                                   (const T*)fin, flags, rho, u, g);
        }))
        return e;
    NX_CHECK_LAUNCH("k_moments");
    return NX_OK;
}

extern "C" int nx_d3q19_mass_dense(const nx_dense_desc* d, const void* fin, const uint32_t* flags,
                                   double* out, nx_stream s)
{
    if (int e = check(d)) return e;
    if (!fin || !flags || !out) return set_error(NX_ERR_INVALID, "null pointer");
    const DGeom g = make_geom(d);
    const int nb = (int)std::min<uint32_t>(nwg(g), kRedGrid);
    void* part = nullptr;
    if (int e = scratch(sizeof(double) * nb, &part)) return e;
    const hipStream_t st = (hipStream_t)s;
    if (int e = by_type(d, [&](auto sc, auto) {
            constexpr int S = decltype(sc)::value;
            hipLaunchKernelGGL((k_mass<S>), dim3(nb), dim3(kWG), 0, st,
                               (const typename Stor<S>::T*)fin, flags, g, (double*)part);
        }))
        return e;
    NX_CHECK_LAUNCH("k_mass");
    return finish_reduce((const double*)part, nb, 1, out, st);
}

namespace {

// Residual kernels + the fixed-order second stage, with the nb workgroup
// partials (two doubles each) in `part`.
int residual_launch(const nx_dense_desc* d, const DGeom& g, const void* fin, const uint32_t* flags,
                    float* u_prev, double* out, void* part, int nb, hipStream_t st)
{
    if (int e = by_type(d, [&](auto sc, auto cc) {
            constexpr int S = decltype(sc)::value;
            constexpr int CM = decltype(cc)::value;
            using T = typename Stor<S>::T;
            if (g.periodic)
                hipLaunchKernelGGL((k_residual<S, CM, true>), dim3(nb), dim3(kWG), 0, st,
                                   (const T*)fin, flags, u_prev, g, (double*)part);
            else
                hipLaunchKernelGGL((k_residual<S, CM, false>), dim3(nb), dim3(kWG), 0, st,
                                   (const T*)fin, flags, u_prev, g, (double*)part);
        }))
        return e;
    NX_CHECK_LAUNCH("k_residual");
    return finish_reduce((const double*)part, nb, 2, out, st);
}

}  // namespace

extern "C" int nx_d3q19_residual_dense(const nx_dense_desc* d, const void* fin,
                                       const uint32_t* flags, float* u_prev, double* out,
                                       nx_stream s)
{
    if (int e = check(d)) return e;
    if (!fin || !flags || !u_prev || !out) return set_error(NX_ERR_INVALID, "null pointer");
    const DGeom g = make_geom(d);
    const int nb = (int)std::min<uint32_t>(nwg(g), kRedGrid);
    void* part = nullptr;
    if (int e = scratch(2 * sizeof(double) * nb, &part)) return e;
    return residual_launch(d, g, fin, flags, u_prev, out, part, nb, (hipStream_t)s);
}

// ---- allocation-free residual ---------------------------------------------
// The workgroup partials of nx_d3q19_residual_dense_noalloc live in this
// file-scope device array: at most kRedGrid reduction workgroups of two
// doubles each (32 KiB), held by every GPU the library runs on until the
// process ends.  Every partial is written before the second stage reads it,
// so nothing needs clearing between calls.  In host code the array's name is
// only a stand-in: kernels receive its device address, looked up once per
// device with hipGetSymbolAddress.
constexpr int kNoAllocDoubles = 2 * kRedGrid;
static __device__ __align__(256) double scratch[kNoAllocDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
double* g_noalloc_ptr[kMaxDevices] = {};
}  // namespace

// Looks up this device's copy of the array (the current device); called by
// nx_init and, for a device nx_init did not select, by the entry point's
// first call there.  hipGetSymbolAddress allocates nothing.
int nx::dense_noalloc_setup()
{
    int dev = 0;
    NX_CHECK_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return set_error(NX_ERR_INVALID, "device index out of range");
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (g_noalloc_ptr[dev]) return NX_OK;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(::scratch));
    if (e != hipSuccess) return hip_error(e, "hipGetSymbolAddress(scratch)");
    if (!ptr) return set_error(NX_ERR_HIP, "hipGetSymbolAddress(scratch): null address");
    g_noalloc_ptr[dev] = (double*)ptr;
    return NX_OK;
}

extern "C" int nx_d3q19_residual_dense_noalloc(const nx_dense_desc* d, const void* fin,
                                               const uint32_t* flags, float* u_prev, double* out,
                                               nx_stream s)
{
    if (int e = check(d)) return e;
    if (!fin || !flags || !u_prev || !out) return set_error(NX_ERR_INVALID, "null pointer");
    const DGeom g = make_geom(d);
    const int nb = (int)std::min<uint32_t>(nwg(g), kRedGrid);
    if (2 * nb > kNoAllocDoubles)
        return set_error(NX_ERR_INVALID, "nx_d3q19_residual_dense_noalloc: " + std::to_string(nb) +
                                             " reduction partials exceed the static scratch of " +
                                             std::to_string(kNoAllocDoubles / 2));
    int dev = 0;
    NX_CHECK_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return set_error(NX_ERR_INVALID, "device index out of range");
    double* part = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_noalloc_mu);
        part = g_noalloc_ptr[dev];
    }
    if (!part) {
        if (int e = dense_noalloc_setup()) return e;
        std::lock_guard<std::mutex> lk(g_noalloc_mu);
        part = g_noalloc_ptr[dev];
    }
    return residual_launch(d, g, fin, flags, u_prev, out, part, nb, (hipStream_t)s);
}

