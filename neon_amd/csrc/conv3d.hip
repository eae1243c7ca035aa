// BF16 3D convolution forward, backward-data and backward-weight on the gfx950
// matrix cores, FP32 accumulation (the learned warm start's network).  No
// counterpart in the reference.
//
// All three are implicit GEMMs.  On grids whose rows hold >= 32 voxels the
// operands are staged in LDS (the input block with its halo for the forward,
// a row chunk's 27 shifted tiles for the weight gradient); smaller grids
// read their fragments straight from global memory (the 27-fold reuse of
// every input voxel served by L1/L2):
//   forward      M = voxels, N = cout, K = taps*cin      (A: 8 channels of one
//                tap per lane = one 16-byte load; B: 8 weights, 16 bytes);
//                for rows of >= 32 voxels the input block with its halo is
//                staged in LDS, 16 channels at a time (k_fwd_tile)
//   bwd-data     the forward on the flipped, transposed weights
//   bwd-weight   M = cout, N = cin, K = voxels (split over workgroups,
//                FP32 partials, fixed-order reduction): all 27 taps of a
//                chunk from one LDS staging when rows hold >= 32 voxels
//                (k_bwd_w27), else one tap per workgroup (k_bwd_w)
// MFMA 16x16x32 bf16 operand layout: lane l holds row (A) / column (B) l%16
// and the 8 K values of group l/16; D register r of lane l is row 4*(l/16)+r,
// column l%16.
#include "conv3d.hpp"
#include "nx_common.hpp"

#include <algorithm>

namespace nx::conv {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

struct Geo {
    uint32_t V;          // voxels = n*d*h*w
    int D, H, W;         // powers of two
    int lw, lh;          // log2 W, log2 H
};

__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);  // quiet NaN
    u += 0x7fffu + ((u >> 16) & 1u);                                            // nearest even
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ bf16x8 ld8(const uint16_t* p)
{
    return __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(p));
}

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

template <int KS>
__device__ __forceinline__ void tap_offset(int tap, int& dz, int& dy, int& dx)
{
    if constexpr (KS == 3) {
        dz = tap / 9 - 1;
        dy = (tap / 3) % 3 - 1;
        dx = tap % 3 - 1;
    } else {
        dz = dy = dx = 0;
    }
}

// One wave: 32 voxels x (16*NT) output channels; 4 waves per workgroup.
// grid.y selects the block of 16*NT output channels.
template <int NT, int KS>
__global__ void __launch_bounds__(256)
    k_fwd(const uint16_t* __restrict__ x, const uint16_t* __restrict__ w,
          const float* __restrict__ bias, void* __restrict__ y, Geo g, int cin, int cout, int flags)
{
    constexpr int T = KS * KS * KS;
    const int K = T * cin, KC = (K + 31) / 32;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r16 = lane & 15, grp = lane >> 4;
    const uint32_t vbase = (blockIdx.x * 4u + (uint32_t)wave) * 32u;
    if (vbase >= g.V) return;  // whole wave
    const int co0 = (int)blockIdx.y * NT * 16;

    int px[2], py[2], pz[2];
    uint32_t pv[2];
    bool pok[2];
#pragma unroll
    for (int t = 0; t < 2; t++) {
        const uint32_t v = vbase + 16u * t + (uint32_t)r16;
        pok[t] = v < g.V;
        pv[t] = v;
        px[t] = (int)(v & (uint32_t)(g.W - 1));
        py[t] = (int)((v >> g.lw) & (uint32_t)(g.H - 1));
        pz[t] = (int)((v >> (g.lw + g.lh)) & (uint32_t)(g.D - 1));
    }
    const uint16_t* wrow[NT];
#pragma unroll
    for (int n = 0; n < NT; n++) wrow[n] = w + (size_t)(co0 + n * 16 + r16) * K;

    f32x4 acc[2][NT];
#pragma unroll
    for (int t = 0; t < 2; t++)
#pragma unroll
        for (int n = 0; n < NT; n++) acc[t][n] = f32x4{0.f, 0.f, 0.f, 0.f};

    // K position of this lane group: kk = 32*kc + 8*grp = tap*cin + ci.
    // Fragments are double-buffered: the loads of chunk kc+1 are issued
    // before the MFMAs of chunk kc.
    int tap = 0, ci = 8 * grp;
    while (ci >= cin) { ci -= cin; ++tap; }
    auto fetch = [&](int kc, bf16x8* a, bf16x8* b) {
        const int kk = kc * 32 + 8 * grp;
        if (kk < K) {
            int dz, dy, dx;
            tap_offset<KS>(tap, dz, dy, dx);
            const int off = (dz * g.H + dy) * g.W + dx;
#pragma unroll
            for (int t = 0; t < 2; t++) {
                const bool ok = pok[t] && (unsigned)(px[t] + dx) < (unsigned)g.W &&
                                (unsigned)(py[t] + dy) < (unsigned)g.H &&
                                (unsigned)(pz[t] + dz) < (unsigned)g.D;
                a[t] = ok ? ld8(x + (size_t)(int64_t)((int64_t)pv[t] + off) * cin + ci) : bf16x8{};
            }
#pragma unroll
            for (int n = 0; n < NT; n++) b[n] = ld8(wrow[n] + kk);
        } else {
#pragma unroll
            for (int t = 0; t < 2; t++) a[t] = bf16x8{};
#pragma unroll
            for (int n = 0; n < NT; n++) b[n] = bf16x8{};
        }
        ci += 32;
        while (ci >= cin) { ci -= cin; ++tap; }
    };
    auto mma = [&](const bf16x8* a, const bf16x8* b) {
#pragma unroll
        for (int t = 0; t < 2; t++)
#pragma unroll
            for (int n = 0; n < NT; n++) acc[t][n] = mfma(a[t], b[n], acc[t][n]);
    };
    bf16x8 a0[2], b0[NT], a1[2], b1[NT];
    fetch(0, a0, b0);
    for (int kc = 0; kc < KC; kc += 2) {
        if (kc + 1 < KC) fetch(kc + 1, a1, b1);
        mma(a0, b0);
        if (kc + 1 < KC) {
            if (kc + 2 < KC) fetch(kc + 2, a0, b0);
            mma(a1, b1);
        }
    }

    const bool relu = flags & kRelu, f32out = flags & kOutF32;
#pragma unroll
    for (int n = 0; n < NT; n++) {
        const int co = co0 + n * 16 + r16;
        const float bv = bias ? bias[co] : 0.f;
#pragma unroll
        for (int t = 0; t < 2; t++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const uint32_t v = vbase + 16u * t + 4u * grp + (uint32_t)r;
                if (v >= g.V) continue;
                float s = acc[t][n][r] + bv;
                if (relu) s = s > 0.f ? s : 0.f;
                if (f32out)
                    static_cast<float*>(y)[(size_t)v * cout + co] = s;
                else
                    static_cast<uint16_t*>(y)[(size_t)v * cout + co] = f2bf(s);
            }
    }
}

// 3^3 forward (and backward-data) with the input tile in LDS, for rows of
// >= 32 voxels.  A workgroup computes a 32 x 4 x 4 block of output voxels
// (x, y, z).  For each slice of 16 input channels it stages the block's input
// with its one-voxel halo, 34 x 6 x 6 voxels x 16 channels (39 KB), in
// channels-last order (one 16-byte LDS write per load); each wave then
// takes 4 rows of 32 voxels, reads every A fragment (8 channels of one tap
// for one voxel) with one ds_read_b128, and uses each weight fragment for
// all of its rows.  Each input voxel is read from memory ~2x instead of 27x.
// CINC = 16 fixes the channel count (one slice, fewer registers); 0 reads it
// from the argument.
template <int NT, int CINC>
__global__ void __launch_bounds__(256)
    k_fwd_tile(const uint16_t* __restrict__ x, const uint16_t* __restrict__ w,
               const float* __restrict__ bias, void* __restrict__ y, Geo g, int cin_arg, int cout, int flags)
{
    const int cin = CINC ? CINC : cin_arg;           // CINC: channel count fixed at compile time
    constexpr int TX = 32, TY = 4, TZ = 4, HX = TX + 2, HY = TY + 2, HZ = TZ + 2, CS = 16;
    constexpr int RW = TY * TZ / 4;                  // rows per wave
    constexpr int KS_ = 27 * CS, KC = (KS_ + 31) / 32;
    constexpr int ITEMS = HX * HY * HZ * (CS / 8);   // 16-byte pieces of one slice's halo tile
    __shared__ __attribute__((aligned(16))) uint16_t tile[HX * HY * HZ * CS];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r16 = lane & 15, grp = lane >> 4;
    const int co0 = (int)blockIdx.y * NT * 16;
    const int K = 27 * cin;
    // tile origin: blocks x fastest, then y, then z (z spans batch * depth)
    const int nbx = g.W / TX, nby = g.H / TY, nbz = g.D / TZ;
    const int bx = (int)(blockIdx.x % (uint32_t)nbx);
    const int by = (int)((blockIdx.x / (uint32_t)nbx) % (uint32_t)nby);
    const int bzall = (int)(blockIdx.x / (uint32_t)(nbx * nby));   // (sample, z block)
    const int sample = bzall / nbz, z0 = (bzall % nbz) * TZ, y0 = by * TY, x0 = bx * TX;
    const size_t sample_base = (size_t)sample * g.D * g.H * g.W;

    const uint16_t* wrow[NT];
#pragma unroll
    for (int n = 0; n < NT; n++) wrow[n] = w + (size_t)(co0 + n * 16 + r16) * K;
    f32x4 acc[RW][2][NT];
#pragma unroll
    for (int j = 0; j < RW; j++)
#pragma unroll
        for (int t = 0; t < 2; t++)
#pragma unroll
            for (int n = 0; n < NT; n++) acc[j][t][n] = f32x4{0.f, 0.f, 0.f, 0.f};

    for (int c0 = 0; c0 < cin; c0 += CS) {
        if (c0) __syncthreads();                     // the previous slice's reads are done
        for (int it = tid; it < ITEMS; it += 256) {
            const int piece = it & 1, hv = it >> 1;
            const int hx = hv % HX, hy = (hv / HX) % HY, hz = hv / (HX * HY);
            const int gx = x0 - 1 + hx, gy = y0 - 1 + hy, gz = z0 - 1 + hz;
            uint4 v = uint4{0u, 0u, 0u, 0u};
            if ((unsigned)gx < (unsigned)g.W && (unsigned)gy < (unsigned)g.H && (unsigned)gz < (unsigned)g.D)
                v = *reinterpret_cast<const uint4*>(
                    x + (sample_base + ((size_t)gz * g.H + gy) * g.W + gx) * cin + c0 + 8 * piece);
            *reinterpret_cast<uint4*>(tile + (size_t)hv * CS + 8 * piece) = v;
        }
        __syncthreads();

        // K step kc of the slice: lane group grp holds tap (32 kc + 8 grp) / 16,
        // channels c0 + (8 grp) % 16 ..+7; past the last tap it contributes
        // zeros -- the MFMAs run for every lane (a lane left out of one would
        // not receive its accumulator rows)
        auto fetch_b = [&](int kc, bf16x8* b) {
            const int kk = kc * 32 + 8 * grp;
            const int tap = kk / CS, ci = kk % CS;
#pragma unroll
            for (int n = 0; n < NT; n++)
                b[n] = kk < KS_ ? ld8(wrow[n] + tap * cin + c0 + ci) : bf16x8{};
        };
        bf16x8 b0[NT], b1[NT];
        fetch_b(0, b0);
        for (int kc = 0; kc < KC; kc++) {
            if (kc + 1 < KC) fetch_b(kc + 1, b1);
            const int kk = kc * 32 + 8 * grp;
            const bool kok = kk < KS_;
            const int tap = kok ? kk / CS : 26, ci = kok ? kk % CS : 0;
            int dz, dyo, dx;
            tap_offset<3>(tap, dz, dyo, dx);
#pragma unroll
            for (int j = 0; j < RW; j++) {
                const int row = wave * RW + j;
                const int ly = row % TY, lz = row / TY;
#pragma unroll
                for (int t = 0; t < 2; t++) {
                    const int hv = ((lz + dz + 1) * HY + (ly + dyo + 1)) * HX + (16 * t + r16 + dx + 1);
                    const bf16x8 a = kok ? ld8(tile + (size_t)hv * CS + ci) : bf16x8{};
#pragma unroll
                    for (int n = 0; n < NT; n++) acc[j][t][n] = mfma(a, b0[n], acc[j][t][n]);
                }
            }
#pragma unroll
            for (int n = 0; n < NT; n++) b0[n] = b1[n];
        }
    }

    const bool relu = flags & kRelu, f32out = flags & kOutF32;
#pragma unroll
    for (int j = 0; j < RW; j++) {
        const int row = wave * RW + j;
        const int ly = row % TY, lz = row / TY;
        const size_t vrow = sample_base + ((size_t)(z0 + lz) * g.H + (y0 + ly)) * g.W + x0;
#pragma unroll
        for (int n = 0; n < NT; n++) {
            const int co = co0 + n * 16 + r16;
            const float bv = bias ? bias[co] : 0.f;
#pragma unroll
            for (int t = 0; t < 2; t++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const size_t v = vrow + 16 * t + 4 * grp + r;
                    float sv = acc[j][t][n][r] + bv;
                    if (relu) sv = sv > 0.f ? sv : 0.f;
                    if (f32out)
                        static_cast<float*>(y)[v * cout + co] = sv;
                    else
                        static_cast<uint16_t*>(y)[v * cout + co] = f2bf(sv);
                }
        }
    }
}

// Weight gradient of one tap for a (16*NM) x (16*NN) block of (co, ci).
// grid.x = T * splits (tap fastest), grid.y = channel blocks.  Each wave
// takes every 4th 32-voxel chunk of its split and writes its own partial:
// part[(split*4 + wave)][tap][co][ci].
// The GEMM's K index is the voxel, but channels are contiguous in memory:
// each wave loads its chunk of dy (32 voxels x 16*NM channels) and of the
// shifted input (32 voxels x 16*NN channels) with 16-byte loads and writes
// it transposed into its own LDS tile [channel][voxel], from which every
// MFMA fragment (8 consecutive voxels of one channel) is one ds_read_b128.
// The next chunk's global loads are issued before the current chunk's MFMAs.
template <int NM, int NN, int KS>
__global__ void __launch_bounds__(256)
    k_bwd_w(const uint16_t* __restrict__ x, const uint16_t* __restrict__ dy,
            float* __restrict__ part, Geo g, int cin, int cout, uint32_t chunks_per_split)
{
    constexpr int T = KS * KS * KS;
    constexpr int CM = 16 * NM, CN = 16 * NN;       // channels of the two operands
    constexpr int PITCH = 40;                        // halves per LDS row (32 voxels + pad)
    __shared__ __attribute__((aligned(16))) uint16_t lds[4][(CM + CN) * PITCH];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r16 = lane & 15, grp = lane >> 4;
    const int tap = (int)(blockIdx.x % T);
    const uint32_t split = blockIdx.x / T;
    const int mblocks = cout / CM;
    const int co0 = (int)(blockIdx.y % mblocks) * CM;
    const int ci0 = (int)(blockIdx.y / mblocks) * CN;
    int dz, dyo, dx;
    tap_offset<KS>(tap, dz, dyo, dx);
    const int off = (dz * g.H + dyo) * g.W + dx;
    uint16_t* la = lds[wave];
    uint16_t* lb = lds[wave] + CM * PITCH;

    f32x4 acc[NM][NN];
#pragma unroll
    for (int m = 0; m < NM; m++)
#pragma unroll
        for (int n = 0; n < NN; n++) acc[m][n] = f32x4{0.f, 0.f, 0.f, 0.f};

    // element e = lane + 64 k of a chunk: voxel quad e / (C/8) (4 consecutive
    // voxels), channel group e % (C/8) (8 channels): four 16-byte loads, then
    // eight 8-byte LDS writes (4 voxels of one channel each)
    constexpr int GA = CM / 8, GB = CN / 8;
    constexpr int EA = 8 * GA, EB = 8 * GB;
    constexpr int NA = (EA + 63) / 64, NB = (EB + 63) / 64;
    auto fetch = [&](uint32_t c, uint4 (*va)[4], uint4 (*vb)[4]) {
#pragma unroll
        for (int k = 0; k < NA; k++) {
            const int e = lane + 64 * k;
            const uint32_t v0 = c * 32u + 4u * (uint32_t)(e / GA);
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const uint32_t v = v0 + (uint32_t)t;
                const bool ok = e < EA && v < g.V;
                va[k][t] = ok ? *reinterpret_cast<const uint4*>(dy + (size_t)v * cout + co0 + 8 * (e % GA))
                              : uint4{0u, 0u, 0u, 0u};
            }
        }
#pragma unroll
        for (int k = 0; k < NB; k++) {
            const int e = lane + 64 * k;
            const uint32_t v0 = c * 32u + 4u * (uint32_t)(e / GB);
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const uint32_t v = v0 + (uint32_t)t;
                const int px = (int)(v & (uint32_t)(g.W - 1));
                const int py = (int)((v >> g.lw) & (uint32_t)(g.H - 1));
                const int pz = (int)((v >> (g.lw + g.lh)) & (uint32_t)(g.D - 1));
                const bool ok = e < EB && v < g.V && (unsigned)(px + dx) < (unsigned)g.W &&
                                (unsigned)(py + dyo) < (unsigned)g.H && (unsigned)(pz + dz) < (unsigned)g.D;
                vb[k][t] = ok ? *reinterpret_cast<const uint4*>(x + (size_t)(int64_t)((int64_t)v + off) * cin +
                                                                ci0 + 8 * (e % GB))
                              : uint4{0u, 0u, 0u, 0u};
            }
        }
    };
    // channel j of two voxels' 8-channel words -> one 32-bit word (voxel a low)
    auto pair = [](const uint4& a, const uint4& b, int j) {
        const uint32_t wa = (j >> 1) == 0 ? a.x : (j >> 1) == 1 ? a.y : (j >> 1) == 2 ? a.z : a.w;
        const uint32_t wb = (j >> 1) == 0 ? b.x : (j >> 1) == 1 ? b.y : (j >> 1) == 2 ? b.z : b.w;
        return (j & 1) ? ((wa >> 16) | (wb & 0xffff0000u)) : ((wa & 0xffffu) | (wb << 16));
    };
    auto put = [&](uint16_t* tile, int e, int G, const uint4* q4) {
        const int quad = e / G, ch = 8 * (e % G);
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const uint32_t w2[2] = {pair(q4[0], q4[1], j), pair(q4[2], q4[3], j)};
            __builtin_memcpy(__builtin_assume_aligned(tile + (ch + j) * PITCH + 4 * quad, 8), w2, 8);
        }
    };
    auto stage = [&](uint4 (*va)[4], uint4 (*vb)[4]) {
#pragma unroll
        for (int k = 0; k < NA; k++)
            if (lane + 64 * k < EA) put(la, lane + 64 * k, GA, va[k]);
#pragma unroll
        for (int k = 0; k < NB; k++)
            if (lane + 64 * k < EB) put(lb, lane + 64 * k, GB, vb[k]);
    };
    // 16-byte fragment from the LDS tile (byte copy: the tile is written as
    // 16-bit values, so the read must not be reordered by type-based aliasing)
    auto frag = [&](const uint16_t* row) {
        uint4 t;
        __builtin_memcpy(&t, __builtin_assume_aligned(row, 16), 16);
        return __builtin_bit_cast(bf16x8, t);
    };
    auto mma = [&]() {
        bf16x8 a[NM], b[NN];
#pragma unroll
        for (int m = 0; m < NM; m++) a[m] = frag(la + (m * 16 + r16) * PITCH + 8 * grp);
#pragma unroll
        for (int n = 0; n < NN; n++) b[n] = frag(lb + (n * 16 + r16) * PITCH + 8 * grp);
#pragma unroll
        for (int m = 0; m < NM; m++)
#pragma unroll
            for (int n = 0; n < NN; n++) acc[m][n] = mfma(a[m], b[n], acc[m][n]);
    };

    const uint32_t n_chunks = (g.V + 31u) / 32u;
    const uint32_t c0 = split * chunks_per_split;
    const uint32_t c1 = min(n_chunks, c0 + chunks_per_split);
    uint4 va[NA][4], vb[NB][4], wa[NA][4], wb[NB][4];
    uint32_t c = c0 + (uint32_t)wave;
    if (c < c1) fetch(c, va, vb);
    for (; c < c1; c += 8u) {
        const uint32_t cn = c + 4u;
        if (cn < c1) fetch(cn, wa, wb);
        stage(va, vb);
        __builtin_amdgcn_wave_barrier();
        mma();
        __builtin_amdgcn_wave_barrier();
        if (cn < c1) {
            if (cn + 4u < c1) fetch(cn + 4u, va, vb);
            stage(wa, wb);
            __builtin_amdgcn_wave_barrier();
            mma();
            __builtin_amdgcn_wave_barrier();
        }
    }

    float* out = part + ((size_t)(split * 4u + (uint32_t)wave) * T + tap) * (size_t)cout * cin;
#pragma unroll
    for (int m = 0; m < NM; m++)
#pragma unroll
        for (int n = 0; n < NN; n++)
#pragma unroll
            for (int r = 0; r < 4; r++)
                out[(size_t)(co0 + m * 16 + 4 * grp + r) * cin + ci0 + n * 16 + r16] = acc[m][n][r];
}

// Weight gradient of all 27 taps of a 3^3 convolution for a (16*NM) x 16
// block of (co, ci), rows of at least 32 voxels (W >= 32).  grid.x = splits,
// grid.y = channel blocks.  The workgroup walks 32-voxel chunks (one row
// segment each): it stages the chunk's dy (32 voxels x 16*NM channels) and,
// for every tap, the input the tap reads (the 9 neighbour row segments, each
// in the 3 x shifts: 27 tiles of 16 channels x 32 voxels) transposed into
// LDS [channel][voxel], so every MFMA fragment is one 16-byte LDS read and
// each input element is fetched from memory once per chunk instead of once
// per tap.  Wave w accumulates taps w, w+4, ..., w+24; one partial per split:
// part[split][tap][co][ci].
// With part_b, the workgroups of the first input-channel block also sum the
// staged dy over their voxels (bias gradient): part_b[split][co].
template <int NM>
__global__ void __launch_bounds__(256)
    k_bwd_w27(const uint16_t* __restrict__ x, const uint16_t* __restrict__ dy,
              float* __restrict__ part, float* __restrict__ part_b, Geo g, int cin, int cout,
              uint32_t chunks_per_split)
{
    constexpr int CM = 16 * NM, CN = 16, T = 27, TW = 7;
    constexpr int PITCH = 40;                            // halves per LDS row (32 voxels + pad)
    constexpr int XITEMS = T * 8 * 2;                    // (tile, voxel quad, 8-channel group)
    constexpr int DITEMS = 8 * (CM / 8);                 // (voxel quad, 8-channel group)
    __shared__ __attribute__((aligned(16))) uint16_t lx[T * CN * PITCH];
    __shared__ __attribute__((aligned(16))) uint16_t ld[CM * PITCH];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r16 = lane & 15, grp = lane >> 4;
    const int mblocks = cout / CM;
    const int co0 = (int)(blockIdx.y % mblocks) * CM;
    const int ci0 = (int)(blockIdx.y / mblocks) * CN;

    f32x4 acc[TW][NM];
#pragma unroll
    for (int t = 0; t < TW; t++)
#pragma unroll
        for (int m = 0; m < NM; m++) acc[t][m] = f32x4{0.f, 0.f, 0.f, 0.f};

    // this thread's staging items: x items tid, tid + 256 (< XITEMS), dy item tid (< DITEMS)
    uint4 qx[2][4], qd[4];
    auto fetch = [&](uint32_t c) {
        const uint32_t v0 = c * 32u;
        const int x0 = (int)(v0 & (uint32_t)(g.W - 1));
        const int y0 = (int)((v0 >> g.lw) & (uint32_t)(g.H - 1));
        const int z0 = (int)((v0 >> (g.lw + g.lh)) & (uint32_t)(g.D - 1));
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const int it = tid + 256 * k;
            const int tile = it >> 4, quad = (it >> 1) & 7, cg = it & 1;
            int dz, dyo, dx;
            tap_offset<3>(tile, dz, dyo, dx);
            const bool row = it < XITEMS && (unsigned)(z0 + dz) < (unsigned)g.D &&
                             (unsigned)(y0 + dyo) < (unsigned)g.H;
            const int64_t base = (int64_t)v0 + (int64_t)(dz * g.H + dyo) * g.W + dx;
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int pos = 4 * quad + t;
                const bool ok = row && (unsigned)(x0 + pos + dx) < (unsigned)g.W;
                qx[k][t] = ok ? *reinterpret_cast<const uint4*>(x + (size_t)(base + pos) * cin + ci0 + 8 * cg)
                              : uint4{0u, 0u, 0u, 0u};
            }
        }
        {
            const int quad = tid / (CM / 8), cg = tid % (CM / 8);
#pragma unroll
            for (int t = 0; t < 4; t++)
                qd[t] = tid < DITEMS ? *reinterpret_cast<const uint4*>(dy + (size_t)(v0 + 4u * quad + t) * cout +
                                                                      co0 + 8 * cg)
                                     : uint4{0u, 0u, 0u, 0u};
        }
    };
    // channel j of two voxels' 8-channel words -> one 32-bit word (voxel a low)
    auto pair = [](const uint4& a, const uint4& b, int j) {
        const uint32_t wa = (j >> 1) == 0 ? a.x : (j >> 1) == 1 ? a.y : (j >> 1) == 2 ? a.z : a.w;
        const uint32_t wb = (j >> 1) == 0 ? b.x : (j >> 1) == 1 ? b.y : (j >> 1) == 2 ? b.z : b.w;
        return (j & 1) ? ((wa >> 16) | (wb & 0xffff0000u)) : ((wa & 0xffffu) | (wb << 16));
    };
    // 4 voxels x 8 channels -> eight 8-byte LDS writes (4 voxels of one channel each)
    auto put = [&](uint16_t* tile, int ch, int quad, const uint4* q4) {
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const uint32_t w2[2] = {pair(q4[0], q4[1], j), pair(q4[2], q4[3], j)};
            __builtin_memcpy(__builtin_assume_aligned(tile + (ch + j) * PITCH + 4 * quad, 8), w2, 8);
        }
    };
    auto stage = [&]() {
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const int it = tid + 256 * k;
            if (it < XITEMS) put(lx + (it >> 4) * CN * PITCH, 8 * (it & 1), (it >> 1) & 7, qx[k]);
        }
        if (tid < DITEMS) put(ld, 8 * (tid % (CM / 8)), tid / (CM / 8), qd);
    };
    auto frag = [&](const uint16_t* row) {
        uint4 t;
        __builtin_memcpy(&t, __builtin_assume_aligned(row, 16), 16);
        return __builtin_bit_cast(bf16x8, t);
    };

    const uint32_t n_chunks = g.V / 32u;
    const uint32_t c0 = blockIdx.x * chunks_per_split;
    const uint32_t c1 = min(n_chunks, c0 + chunks_per_split);
    const bool bias = part_b != nullptr && ci0 == 0 && tid < CM;
    float sb = 0.f;
    if (c0 < c1) fetch(c0);
    for (uint32_t c = c0; c < c1; c++) {
        stage();
        __syncthreads();
        if (c + 1 < c1) fetch(c + 1);
        if (bias)
            for (int v = 0; v < 32; v++) sb += __builtin_bit_cast(float, (uint32_t)ld[tid * PITCH + v] << 16);
        bf16x8 a[NM];
#pragma unroll
        for (int m = 0; m < NM; m++) a[m] = frag(ld + (m * 16 + r16) * PITCH + 8 * grp);
#pragma unroll
        for (int t = 0; t < TW; t++) {
            const int tap = wave + 4 * t;
            if (tap < T) {
                const bf16x8 b = frag(lx + (tap * CN + r16) * PITCH + 8 * grp);
#pragma unroll
                for (int m = 0; m < NM; m++) acc[t][m] = mfma(a[m], b, acc[t][m]);
            }
        }
        __syncthreads();
    }

    if (bias) part_b[(size_t)blockIdx.x * cout + co0 + tid] = sb;
#pragma unroll
    for (int t = 0; t < TW; t++) {
        const int tap = wave + 4 * t;
        if (tap >= T) continue;
        float* out = part + ((size_t)blockIdx.x * T + tap) * (size_t)cout * cin;
#pragma unroll
        for (int m = 0; m < NM; m++)
#pragma unroll
            for (int r = 0; r < 4; r++)
                out[(size_t)(co0 + m * 16 + 4 * grp + r) * cin + ci0 + r16] = acc[t][m][r];
    }
}

// dw[co][tap][ci] = sum_p part[p][tap][co][ci] in two fixed-order levels
// (deterministic, coalesced over the outputs, 64 outputs per workgroup with
// 4 threads per output adding every 4th partial, the four sums combined in
// a fixed order):
//   level 1 (grid.y = G slices of S partials): the sum of slice y replaces
//            its first partial, part[y*S] (only this workgroup touches it)
//   level 2: dw = sum over y of part[y*S]
__global__ void __launch_bounds__(256)
    k_reduce_w(float* __restrict__ part, int P, int S, int T, int cin, int cout, float* __restrict__ dw,
               int level)
{
    __shared__ float lds[4][64];
    const size_t n = (size_t)cout * T * cin;
    const int e = threadIdx.x & 63, k0 = threadIdx.x >> 6;
    const size_t i = (size_t)blockIdx.x * 64 + e;
    const size_t stride = (size_t)T * cout * cin;
    size_t at = 0;
    float s = 0.f;
    if (i < n) {
        const int ci = (int)(i % cin);
        const int tap = (int)((i / cin) % T);
        const int co = (int)(i / ((size_t)cin * T));
        at = ((size_t)tap * cout + co) * cin + ci;
        if (level == 1) {
            const int k1 = min(P, (int)(blockIdx.y + 1) * S);
            for (int k = (int)blockIdx.y * S + k0; k < k1; k += 4) s += part[(size_t)k * stride + at];
        } else {
            for (int y = k0; y * S < P; y += 4) s += part[(size_t)y * S * stride + at];
        }
    }
    lds[k0][e] = s;
    __syncthreads();
    if (k0 == 0 && i < n) {
        const float t = ((lds[0][e] + lds[1][e]) + lds[2][e]) + lds[3][e];
        if (level == 1)
            part[(size_t)blockIdx.y * S * stride + at] = t;
        else
            dw[i] = t;
    }
}

// Per-channel sums of dy over voxel ranges: part[blockIdx.x][c].
__global__ void __launch_bounds__(256)
    k_colsum(const uint16_t* __restrict__ dy, uint32_t V, int C, uint32_t per_block,
             float* __restrict__ part)
{
    __shared__ float lds[256];
    const int rows = 256 / C;  // C <= 256; threads past rows*C idle
    const int c = threadIdx.x % C, j = threadIdx.x / C;
    const uint32_t v0 = blockIdx.x * per_block, v1 = min(V, v0 + per_block);
    float s = 0.f;
    if (j < rows)
        for (uint32_t v = v0 + (uint32_t)j; v < v1; v += (uint32_t)rows)
            s += __builtin_bit_cast(float, (uint32_t)dy[(size_t)v * C + c] << 16);
    lds[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x < C) {
        float t = 0.f;
        for (int k = 0; k < rows; k++) t += lds[k * C + threadIdx.x];
        part[(size_t)blockIdx.x * C + threadIdx.x] = t;
    }
}

// out[c] = sum_k part[k][c]: one workgroup per channel, fixed-order tree
__global__ void __launch_bounds__(256)
    k_reduce_cols(const float* __restrict__ part, int P, int C, float* __restrict__ out)
{
    __shared__ float lds[256];
    const int c = (int)blockIdx.x;
    float s = 0.f;
    for (int k = threadIdx.x; k < P; k += 256) s += part[(size_t)k * C + c];
    lds[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) lds[threadIdx.x] += lds[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = lds[0];
}

__global__ void __launch_bounds__(256)
    k_flip(const uint16_t* __restrict__ w, uint16_t* __restrict__ wt, int T, int cin, int cout)
{
    const size_t n = (size_t)cout * T * cin;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        // i indexes wt[ci][tap][co]
        const int co = (int)(i % cout);
        const int tap = (int)((i / cout) % T);
        const int ci = (int)(i / ((size_t)cout * T));
        wt[i] = w[((size_t)co * T + (T - 1 - tap)) * cin + ci];
    }
}

Geo geo_of(const Shape& s)
{
    Geo g;
    g.V = (uint32_t)((uint64_t)s.n * s.d * s.h * s.w);
    g.D = s.d;
    g.H = s.h;
    g.W = s.w;
    g.lw = __builtin_ctz((unsigned)s.w);
    g.lh = __builtin_ctz((unsigned)s.h);
    return g;
}

bool pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

template <typename F>
int by_ks(int ks, F&& f)
{
    if (ks == 3) { f(std::integral_constant<int, 3>{}); return NX_OK; }
    if (ks == 1) { f(std::integral_constant<int, 1>{}); return NX_OK; }
    return set_error(NX_ERR_INVALID, "conv3d: kernel size must be 1 or 3");
}

// largest supported tile count (8, 6, 4, 3, 2, 1) dividing c16 = channels/16,
// at most `cap`
int tiles_of(int c16, int cap)
{
    for (int t : {8, 6, 4, 3, 2})
        if (t <= cap && c16 % t == 0) return t;
    return 1;
}

int launch_fwd(const Shape& s, const uint16_t* x, const uint16_t* w, const float* bias, void* y,
               int flags, hipStream_t st)
{
    const Geo g = geo_of(s);
    if (g.V == 0) return NX_OK;
    // LDS-tiled kernel: 3^3, rows of >= 32 voxels, blocks of 32 x 4 x 4
    if (s.ks == 3 && s.w >= 32 && s.h >= 4 && s.d >= 4) {
        const int nt = tiles_of(s.cout / 16, 4);
        const dim3 grid((unsigned)((size_t)s.n * (s.d / 4) * (s.h / 4) * (s.w / 32)),
                        (unsigned)(s.cout / (16 * nt)));
        auto go = [&](auto NTc) {
            constexpr int NT = decltype(NTc)::value;
            if (s.cin == 16)
                hipLaunchKernelGGL((k_fwd_tile<NT, 16>), grid, dim3(256), 0, st, x, w, bias, y, g, s.cin, s.cout,
                                   flags);
            else
                hipLaunchKernelGGL((k_fwd_tile<NT, 0>), grid, dim3(256), 0, st, x, w, bias, y, g, s.cin, s.cout,
                                   flags);
        };
        switch (nt) {
            case 4: go(std::integral_constant<int, 4>{}); break;
            case 3: go(std::integral_constant<int, 3>{}); break;
            case 2: go(std::integral_constant<int, 2>{}); break;
            default: go(std::integral_constant<int, 1>{}); break;
        }
        return NX_OK;
    }
    const int nt = tiles_of(s.cout / 16, 8);
    const dim3 grid((g.V + 127u) / 128u, (unsigned)(s.cout / (16 * nt)));
    return by_ks(s.ks, [&](auto KS) {
        switch (nt) {
            case 8: hipLaunchKernelGGL((k_fwd<8, KS>), grid, dim3(256), 0, st, x, w, bias, y, g, s.cin, s.cout, flags); break;
            case 6: hipLaunchKernelGGL((k_fwd<6, KS>), grid, dim3(256), 0, st, x, w, bias, y, g, s.cin, s.cout, flags); break;
            case 4: hipLaunchKernelGGL((k_fwd<4, KS>), grid, dim3(256), 0, st, x, w, bias, y, g, s.cin, s.cout, flags); break;
            case 3: hipLaunchKernelGGL((k_fwd<3, KS>), grid, dim3(256), 0, st, x, w, bias, y, g, s.cin, s.cout, flags); break;
            case 2: hipLaunchKernelGGL((k_fwd<2, KS>), grid, dim3(256), 0, st, x, w, bias, y, g, s.cin, s.cout, flags); break;
            default: hipLaunchKernelGGL((k_fwd<1, KS>), grid, dim3(256), 0, st, x, w, bias, y, g, s.cin, s.cout, flags); break;
        }
    });
}

struct WPlan {
    int nm, nn, T;
    bool all_taps;  // k_bwd_w27: 3^3 kernel, rows of >= 32 voxels
    uint32_t splits, chunks_per_split;
    size_t part_floats;
    int partials;   // partials per output element (reduced by k_reduce_w)
};

WPlan plan_w(const Shape& s)
{
    WPlan p;
    p.T = s.ks * s.ks * s.ks;
    p.all_taps = s.ks == 3 && s.w >= 32;
    if (p.all_taps) {
        p.nm = tiles_of(s.cout / 16, 4);
        p.nn = 1;
        const uint64_t V = (uint64_t)s.n * s.d * s.h * s.w;
        const uint32_t n_chunks = (uint32_t)(V / 32);
        const uint32_t blocks = (uint32_t)((s.cout / (16 * p.nm)) * (s.cin / 16));
        // about three workgroups per CU (the kernel's register use allows
        // three), each walking a contiguous run of chunks
        uint32_t splits = (768u + blocks - 1) / blocks;
        // at most 16 M partial floats (64 MB of workspace)
        const uint32_t cap = (uint32_t)std::max<uint64_t>(1, (16ull << 20) / ((uint64_t)27 * s.cout * s.cin));
        splits = std::max(1u, std::min({splits, cap, n_chunks}));
        p.chunks_per_split = (n_chunks + splits - 1) / splits;
        p.splits = (n_chunks + p.chunks_per_split - 1) / p.chunks_per_split;
        p.part_floats = (size_t)p.splits * 27 * s.cout * s.cin;
        p.partials = (int)p.splits;
        return p;
    }
    p.nm = tiles_of(s.cout / 16, 4);
    p.nn = tiles_of(s.cin / 16, 4);
    const uint64_t V = (uint64_t)s.n * s.d * s.h * s.w;
    const uint32_t n_chunks = (uint32_t)((V + 31) / 32);
    const uint32_t blocks = (uint32_t)(p.T * (s.cout / (16 * p.nm)) * (s.cin / (16 * p.nn)));
    uint32_t splits = (2048u + blocks - 1) / blocks;
    splits = std::max(1u, std::min({splits, 512u, (n_chunks + 3u) / 4u}));
    p.chunks_per_split = (n_chunks + splits - 1) / splits;
    p.splits = (n_chunks + p.chunks_per_split - 1) / p.chunks_per_split;
    if (p.splits == 0) p.splits = 1;
    p.part_floats = (size_t)p.splits * 4 * p.T * s.cout * s.cin;
    p.partials = (int)(p.splits * 4);
    return p;
}

constexpr uint32_t kColBlocks = 512;

}  // namespace

int check(const Shape& s)
{
    if (s.n < 1 || !pow2(s.d) || !pow2(s.h) || !pow2(s.w))
        return set_error(NX_ERR_INVALID, "conv3d: batch >= 1 and power-of-two depth/height/width required");
    if (s.cin < 16 || s.cin % 16 || s.cout < 16 || s.cout % 16 || s.cin > 256 || s.cout > 256)
        return set_error(NX_ERR_INVALID, "conv3d: channel counts must be multiples of 16 in [16, 256]");
    if (s.ks != 1 && s.ks != 3) return set_error(NX_ERR_INVALID, "conv3d: kernel size must be 1 or 3");
    if ((uint64_t)s.n * s.d * s.h * s.w * std::max(s.cin, s.cout) >= (1ull << 31))
        return set_error(NX_ERR_INVALID, "conv3d: tensor too large (2^31 elements)");
    return NX_OK;
}

int fwd(const Shape& s, const uint16_t* x, const uint16_t* w, const float* bias, void* y, int flags,
        hipStream_t st)
{
    if (int e = launch_fwd(s, x, w, bias, y, flags, st)) return e;
    NX_CHECK_LAUNCH("k_fwd");
    return NX_OK;
}

int flip(const Shape& s, const uint16_t* w, uint16_t* wt, hipStream_t st)
{
    const int T = s.ks * s.ks * s.ks;
    const size_t n = (size_t)s.cout * T * s.cin;
    hipLaunchKernelGGL(k_flip, dim3((unsigned)std::min<size_t>((n + 255) / 256, 4096)), dim3(256), 0,
                       st, w, wt, T, s.cin, s.cout);
    NX_CHECK_LAUNCH("k_flip");
    return NX_OK;
}

int bwd_data(const Shape& s, const uint16_t* dy, const uint16_t* wt, void* dx, int flags,
             hipStream_t st)
{
    Shape t = s;
    t.cin = s.cout;
    t.cout = s.cin;
    if (int e = launch_fwd(t, dy, wt, nullptr, dx, flags & kOutF32, st)) return e;
    NX_CHECK_LAUNCH("k_fwd (bwd-data)");
    return NX_OK;
}

size_t bwd_weight_ws(const Shape& s)
{
    const WPlan p = plan_w(s);
    const size_t cols = (size_t)kColBlocks * s.cout;
    // all-taps plan: weight partials, then the bias partials (splits x cout)
    if (p.all_taps) return (p.part_floats + (size_t)p.splits * s.cout) * sizeof(float);
    return (std::max(p.part_floats, cols)) * sizeof(float);
}

int bwd_weight(const Shape& s, const uint16_t* x, const uint16_t* dy, float* dw, float* db, void* ws,
               hipStream_t st)
{
    const Geo g = geo_of(s);
    const WPlan p = plan_w(s);
    float* part = static_cast<float*>(ws);
    const dim3 grid(p.all_taps ? p.splits : p.T * p.splits,
                    (unsigned)((s.cout / (16 * p.nm)) * (s.cin / (16 * p.nn))));
    if (p.all_taps) {
        auto go27 = [&](auto NM) {
            hipLaunchKernelGGL((k_bwd_w27<decltype(NM)::value>), grid, dim3(256), 0, st, x, dy, part,
                               db ? part + p.part_floats : nullptr, g, s.cin, s.cout, p.chunks_per_split);
        };
        switch (p.nm) {
            case 4: go27(std::integral_constant<int, 4>{}); break;
            case 3: go27(std::integral_constant<int, 3>{}); break;
            case 2: go27(std::integral_constant<int, 2>{}); break;
            default: go27(std::integral_constant<int, 1>{}); break;
        }
        NX_CHECK_LAUNCH("k_bwd_w27");
    }
    int err = p.all_taps ? 0 : by_ks(s.ks, [&](auto KS) {
        auto go = [&](auto NM, auto NN) {
            hipLaunchKernelGGL((k_bwd_w<decltype(NM)::value, decltype(NN)::value, KS>), grid,
                               dim3(256), 0, st, x, dy, part, g, s.cin, s.cout, p.chunks_per_split);
        };
        auto by_nn = [&](auto NM) {
            switch (p.nn) {
                case 4: go(NM, std::integral_constant<int, 4>{}); break;
                case 3: go(NM, std::integral_constant<int, 3>{}); break;
                case 2: go(NM, std::integral_constant<int, 2>{}); break;
                default: go(NM, std::integral_constant<int, 1>{}); break;
            }
        };
        switch (p.nm) {
            case 4: by_nn(std::integral_constant<int, 4>{}); break;
            case 3: by_nn(std::integral_constant<int, 3>{}); break;
            case 2: by_nn(std::integral_constant<int, 2>{}); break;
            default: by_nn(std::integral_constant<int, 1>{}); break;
        }
    });
    if (err) return err;
    NX_CHECK_LAUNCH("k_bwd_w");
    const size_t nw = (size_t)s.cout * p.T * s.cin;
    // level 1 in slices of S partials, then the G slice sums
    const int G = std::max(1, std::min(64, p.partials / 32));
    const int S = (p.partials + G - 1) / G;
    const unsigned gx = (unsigned)((nw + 63) / 64);
    hipLaunchKernelGGL(k_reduce_w, dim3(gx, (unsigned)((p.partials + S - 1) / S)), dim3(256), 0, st, part,
                       p.partials, S, p.T, s.cin, s.cout, dw, 1);
    hipLaunchKernelGGL(k_reduce_w, dim3(gx), dim3(256), 0, st, part, p.partials, S, p.T, s.cin, s.cout,
                       dw, 2);
    NX_CHECK_LAUNCH("k_reduce_w");
    if (db && p.all_taps) {
        hipLaunchKernelGGL(k_reduce_cols, dim3(s.cout), dim3(256), 0, st, part + p.part_floats,
                           (int)p.splits, s.cout, db);
        NX_CHECK_LAUNCH("k_reduce_cols");
    } else if (db) {
        // the column partials reuse the workspace after the weight reduction
        const uint32_t per = (g.V + kColBlocks - 1) / kColBlocks;
        hipLaunchKernelGGL(k_colsum, dim3(kColBlocks), dim3(256), 0, st, dy, g.V, s.cout, per, part);
        NX_CHECK_LAUNCH("k_colsum");
        hipLaunchKernelGGL(k_reduce_cols, dim3(s.cout), dim3(256), 0, st, part, (int)kColBlocks,
                           s.cout, db);
        NX_CHECK_LAUNCH("k_reduce_cols");
    }
    return NX_OK;
}

}  // namespace nx::conv

using namespace nx;

namespace {

int shape_of(const int32_t shape[4], int cin, int cout, int ks, conv::Shape& s)
{
    if (!shape) return set_error(NX_ERR_INVALID, "conv3d: null shape");
    s = conv::Shape{shape[0], shape[1], shape[2], shape[3], cin, cout, ks};
    return conv::check(s);
}

bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

}  // namespace

extern "C" {

int nx_conv3d_fwd_bf16(const uint16_t* x, const uint16_t* w, const float* bias, void* y,
                       const int32_t shape[4], int32_t cin, int32_t cout, int32_t ksize, int32_t flags,
                       nx_stream s)
{
    conv::Shape sh;
    if (int e = shape_of(shape, cin, cout, ksize, sh)) return e;
    if (!x || !w || !y) return set_error(NX_ERR_INVALID, "nx_conv3d_fwd_bf16: null pointer");
    if (!aligned16(x) || !aligned16(w))
        return set_error(NX_ERR_INVALID, "nx_conv3d_fwd_bf16: x and w must be 16-byte aligned");
    return conv::fwd(sh, x, w, bias, y, flags & (conv::kRelu | conv::kOutF32), (hipStream_t)s);
}

int nx_conv3d_bwd_data_bf16(const uint16_t* dy, const uint16_t* w, void* dx, const int32_t shape[4],
                            int32_t cin, int32_t cout, int32_t ksize, int32_t flags, nx_stream s)
{
    conv::Shape sh;
    if (int e = shape_of(shape, cin, cout, ksize, sh)) return e;
    if (!dy || !w || !dx) return set_error(NX_ERR_INVALID, "nx_conv3d_bwd_data_bf16: null pointer");
    if (!aligned16(dy)) return set_error(NX_ERR_INVALID, "nx_conv3d_bwd_data_bf16: dy must be 16-byte aligned");
    void* ws = nullptr;
    const size_t nw = (size_t)cout * ksize * ksize * ksize * cin;
    if (int e = scratch(nw * sizeof(uint16_t), &ws)) return e;
    if (int e = conv::flip(sh, w, static_cast<uint16_t*>(ws), (hipStream_t)s)) return e;
    return conv::bwd_data(sh, dy, static_cast<uint16_t*>(ws), dx, flags & conv::kOutF32, (hipStream_t)s);
}

int nx_conv3d_bwd_weight_bf16(const uint16_t* x, const uint16_t* dy, float* dw, float* db,
                              const int32_t shape[4], int32_t cin, int32_t cout, int32_t ksize,
                              nx_stream s)
{
    conv::Shape sh;
    if (int e = shape_of(shape, cin, cout, ksize, sh)) return e;
    if (!x || !dy || !dw) return set_error(NX_ERR_INVALID, "nx_conv3d_bwd_weight_bf16: null pointer");
    void* ws = nullptr;
    if (int e = scratch(conv::bwd_weight_ws(sh), &ws)) return e;
    return conv::bwd_weight(sh, x, dy, dw, db, ws, (hipStream_t)s);
}

}  // extern "C"
