// The learned warm start's network: a 3D U-Net on channels-last BF16
// volumes, trained on the device with the conv3d kernels of conv3d.hip
// (FP32 master weights, Adam) -- no counterpart in the reference.
//
// Architecture for L levels (2..4) and widths w[0..L-1] (multiples of 16):
//   encoder l:  a_l = relu(conv3(in_l)),  e_l = relu(conv3(a_l)),
//               in_{l+1} = avgpool2(e_l)            (in_0 = input, 16 ch)
//   decoder l = L-2..0:  cat_l = [upsample2(low), e_l]  (low = e_{L-1}, then d_{l+1})
//               b_l = relu(conv3(cat_l)),  d_l = relu(conv3(b_l))
//   output:     y = conv1(d_0) (16 channels, the first out_ch used), FP32.
// Parameter order (FP32 vector): for every conv in the order enc(0,0),
// enc(0,1), ..., enc(L-1,1), dec(0,0), dec(0,1), ..., dec(L-2,1), out:
// W[co][tap][ci] then b[co].  BF16 is used for weights and activations,
// every product is accumulated in FP32.
#include "conv3d.hpp"
#include "nx_common.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

using namespace nx;

namespace {

constexpr int kC0 = 16;  // padded input / output channel count

__device__ __forceinline__ uint16_t f2bf(float f)
{
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float bf2f(uint16_t h) { return __builtin_bit_cast(float, (uint32_t)h << 16); }

#define GRID_STRIDE(i, n) \
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < (n); i += (size_t)gridDim.x * 256)

__global__ void __launch_bounds__(256)
    k_pack_input(const float* __restrict__ x, uint16_t* __restrict__ out, size_t V, int in_ch)
{
    GRID_STRIDE(i, V * kC0)
    {
        const size_t v = i / kC0;
        const int c = (int)(i % kC0);
        out[i] = c < in_ch ? f2bf(x[v * in_ch + c]) : (uint16_t)0;
    }
}

// voxel index of (b, z, y, x) on a side-s cube and its parent on side s/2
__device__ __forceinline__ size_t parent_of(size_t v, int ls)
{
    const size_t s = (size_t)1 << ls, h = s >> 1;
    const size_t x = v & (s - 1), y = (v >> ls) & (s - 1), z = (v >> (2 * ls)) & (s - 1);
    const size_t b = v >> (3 * ls);
    return ((b * h + (z >> 1)) * h + (y >> 1)) * h + (x >> 1);
}

// out[b][z][y][x][c] = mean of the 2x2x2 children on the side-2h cube (ls = log2 of 2h)
__global__ void __launch_bounds__(256)
    k_pool(const uint16_t* __restrict__ in, uint16_t* __restrict__ out, size_t Vh, int ls, int C)
{
    const size_t s = (size_t)1 << ls, h = s >> 1;
    GRID_STRIDE(i, Vh * C)
    {
        const size_t vh = i / C;
        const int c = (int)(i % C);
        const size_t x = vh % h, y = (vh / h) % h, z = (vh / (h * h)) % h, b = vh / (h * h * h);
        float acc = 0.f;
        for (int k = 0; k < 8; k++) {
            const size_t v = ((b * s + 2 * z + (k >> 2)) * s + 2 * y + ((k >> 1) & 1)) * s + 2 * x + (k & 1);
            acc += bf2f(in[v * C + c]);
        }
        out[i] = f2bf(acc * 0.125f);
    }
}

// out[v] = [low[parent(v)] (Cu channels), skip[v] (Cs channels)]
__global__ void __launch_bounds__(256)
    k_upcat(const uint16_t* __restrict__ low, const uint16_t* __restrict__ skip,
            uint16_t* __restrict__ out, size_t V, int ls, int Cu, int Cs)
{
    const int Cc = Cu + Cs;
    GRID_STRIDE(i, V * Cc)
    {
        const size_t v = i / Cc;
        const int c = (int)(i % Cc);
        out[i] = c < Cu ? low[parent_of(v, ls) * Cu + c] : skip[v * Cs + (c - Cu)];
    }
}

// gradient through upsampling: out[vh][c] = sum over the 8 children of g[child][c], c < Cu
__global__ void __launch_bounds__(256)
    k_up_bwd(const uint16_t* __restrict__ g, uint16_t* __restrict__ out, size_t Vh, int ls, int Cc,
             int Cu)
{
    const size_t s = (size_t)1 << ls, h = s >> 1;
    GRID_STRIDE(i, Vh * Cu)
    {
        const size_t vh = i / Cu;
        const int c = (int)(i % Cu);
        const size_t x = vh % h, y = (vh / h) % h, z = (vh / (h * h)) % h, b = vh / (h * h * h);
        float acc = 0.f;
        for (int k = 0; k < 8; k++) {
            const size_t v = ((b * s + 2 * z + (k >> 2)) * s + 2 * y + ((k >> 1) & 1)) * s + 2 * x + (k & 1);
            acc += bf2f(g[v * Cc + c]);
        }
        out[i] = f2bf(acc);
    }
}

// gradient of a skip tensor e_l: its slice of the concatenation gradient plus
// the pooling gradient gin_next[parent]/8
__global__ void __launch_bounds__(256)
    k_skip_grad(const uint16_t* __restrict__ gcat, int Cc, int off, const uint16_t* __restrict__ gnext,
                uint16_t* __restrict__ out, size_t V, int ls, int C)
{
    GRID_STRIDE(i, V * C)
    {
        const size_t v = i / C;
        const int c = (int)(i % C);
        const float a = bf2f(gcat[v * Cc + off + c]);
        const float b = bf2f(gnext[parent_of(v, ls) * C + c]);
        out[i] = f2bf(a + b * 0.125f);
    }
}

__global__ void __launch_bounds__(256)
    k_relu_bwd(const uint16_t* __restrict__ y, const uint16_t* __restrict__ g, uint16_t* __restrict__ out,
               size_t n)
{
    GRID_STRIDE(i, n)
    {
        const uint16_t h = y[i];
        out[i] = ((h & 0x8000u) == 0 && (h & 0x7fffu) != 0) ? g[i] : (uint16_t)0;
    }
}

__global__ void __launch_bounds__(256)
    k_unpack_output(const float* __restrict__ o, float* __restrict__ y, size_t V, int oc)
{
    GRID_STRIDE(i, V * oc)
    {
        y[i] = o[(i / oc) * kC0 + i % oc];
    }
}

__global__ void __launch_bounds__(256)
    k_pack_grad(const float* __restrict__ dy, uint16_t* __restrict__ g, size_t V, int oc)
{
    GRID_STRIDE(i, V * kC0)
    {
        const int c = (int)(i % kC0);
        g[i] = c < oc ? f2bf(dy[(i / kC0) * oc + c]) : (uint16_t)0;
    }
}

// loss = sum_v wgt[v] sum_{c<oc} (o - t)^2, gradient 2 wgt (o - t) in bf16
__global__ void __launch_bounds__(256)
    k_mse(const float* __restrict__ o, const float* __restrict__ t, const float* __restrict__ wgt,
          uint16_t* __restrict__ g, double* __restrict__ part, size_t V, int oc)
{
    __shared__ double lds[16];
    double acc[1] = {0.0};
    for (size_t v = (size_t)blockIdx.x * 256 + threadIdx.x; v < V; v += (size_t)gridDim.x * 256) {
        const float w = wgt[v];
        for (int c = 0; c < kC0; c++) {
            float gv = 0.f;
            if (c < oc) {
                const float d = o[v * kC0 + c] - t[v * oc + c];
                acc[0] += (double)w * (double)d * (double)d;
                gv = 2.f * w * d;
            }
            g[v * kC0 + c] = f2bf(gv);
        }
    }
    block_sum_f64<1>(acc, lds);
    if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
}

__global__ void __launch_bounds__(256)
    k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
           uint16_t* __restrict__ wb, size_t n, float lr, float b1, float b2, float eps, float c1, float c2)
{
    GRID_STRIDE(i, n)
    {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float wi = w[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
        w[i] = wi;
        wb[i] = f2bf(wi);
    }
}

__global__ void __launch_bounds__(256) k_to_bf16(const float* __restrict__ w, uint16_t* __restrict__ wb, size_t n)
{
    GRID_STRIDE(i, n) { wb[i] = f2bf(w[i]); }
}

unsigned blocks_for(size_t n) { return (unsigned)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 8192)); }

struct ConvSpec {
    int cin, cout, ks, level;
    size_t w_off, b_off;  // into the parameter vector (floats)
};

}  // namespace

struct nx_unet_s {
    nx_unet_desc d{};
    int L = 0, S = 0, B = 0, ls0 = 0;
    int wd[4] = {};
    std::vector<ConvSpec> convs;
    size_t n_params = 0;
    float *w32 = nullptr, *g32 = nullptr, *m1 = nullptr, *m2 = nullptr;
    uint16_t *wbf = nullptr, *wt = nullptr;
    std::vector<void*> allocs;
    uint16_t *in[4] = {}, *a[4] = {}, *e[4] = {}, *cat[4] = {}, *bb[4] = {}, *dd[4] = {};
    uint16_t *g1[4] = {}, *g2[4] = {}, *gcat[4] = {}, *gin[4] = {};
    float* out32 = nullptr;
    uint16_t* gout = nullptr;
    void* ws = nullptr;
    double *loss_part = nullptr, *loss_dev = nullptr;
    int adam_t = 0;
    int fwd_batch = 0;  // batch of the last forward (backward needs it)

    int enc(int l, int k) const { return 2 * l + k; }
    int dec(int l, int k) const { return 2 * L + 2 * l + k; }
    int outc() const { return 4 * L - 2; }
    int side(int l) const { return S >> l; }
    size_t vox(int l, int b) const { return (size_t)b * side(l) * side(l) * side(l); }
    conv::Shape shape(int ci, int b) const
    {
        const ConvSpec& c = convs[ci];
        const int s = side(c.level);
        return conv::Shape{b, s, s, s, c.cin, c.cout, c.ks};
    }
};

namespace {

int dalloc(nx_unet_s* u, size_t bytes, void** p)
{
    bytes = (bytes + 255) & ~(size_t)255;
    NX_CHECK_HIP(hipMalloc(p, std::max<size_t>(bytes, 256)));
    u->allocs.push_back(*p);
    return NX_OK;
}

template <typename T>
int dalloc_t(nx_unet_s* u, size_t count, T** p)
{
    return dalloc(u, count * sizeof(T), reinterpret_cast<void**>(p));
}

void release(nx_unet_s* u)
{
    for (void* p : u->allocs) (void)hipFree(p);
    u->allocs.clear();
}

// SplitMix64 (Steele, Lea & Flood 2014) and Box-Muller normals for He init
struct Rng {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
    double normal()
    {
        const double u1 = uniform(), u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

int conv_fwd(nx_unet_s* u, int ci, int b, const uint16_t* x, void* y, int flags, hipStream_t st)
{
    const ConvSpec& c = u->convs[ci];
    return conv::fwd(u->shape(ci, b), x, u->wbf + c.w_off, u->w32 + c.b_off, y, flags, st);
}

int conv_bwd(nx_unet_s* u, int ci, int b, const uint16_t* x, const uint16_t* gy, uint16_t* gx,
             hipStream_t st)
{
    const ConvSpec& c = u->convs[ci];
    const conv::Shape s = u->shape(ci, b);
    if (int e = conv::bwd_weight(s, x, gy, u->g32 + c.w_off, u->g32 + c.b_off, u->ws, st)) return e;
    if (gx) return conv::bwd_data(s, gy, u->wt + c.w_off, gx, 0, st);
    return NX_OK;
}

int relu_bwd(const uint16_t* y, const uint16_t* g, uint16_t* out, size_t n, hipStream_t st)
{
    hipLaunchKernelGGL(k_relu_bwd, dim3(blocks_for(n)), dim3(256), 0, st, y, g, out, n);
    NX_CHECK_LAUNCH("k_relu_bwd");
    return NX_OK;
}

int forward(nx_unet_s* u, const float* x, int b, hipStream_t st)
{
    const int L = u->L;
    const size_t V0 = u->vox(0, b);
    hipLaunchKernelGGL(k_pack_input, dim3(blocks_for(V0 * kC0)), dim3(256), 0, st, x, u->in[0], V0,
                       u->d.in_ch);
    NX_CHECK_LAUNCH("k_pack_input");
    for (int l = 0; l < L; l++) {
        if (int e = conv_fwd(u, u->enc(l, 0), b, u->in[l], u->a[l], conv::kRelu, st)) return e;
        if (int e = conv_fwd(u, u->enc(l, 1), b, u->a[l], u->e[l], conv::kRelu, st)) return e;
        if (l < L - 1) {
            const size_t Vh = u->vox(l + 1, b);
            hipLaunchKernelGGL(k_pool, dim3(blocks_for(Vh * u->wd[l])), dim3(256), 0, st, u->e[l],
                               u->in[l + 1], Vh, u->ls0 - l, u->wd[l]);
            NX_CHECK_LAUNCH("k_pool");
        }
    }
    const uint16_t* low = u->e[L - 1];
    for (int l = L - 2; l >= 0; l--) {
        const size_t V = u->vox(l, b);
        const int Cu = u->wd[l + 1], Cs = u->wd[l];
        hipLaunchKernelGGL(k_upcat, dim3(blocks_for(V * (Cu + Cs))), dim3(256), 0, st, low, u->e[l],
                           u->cat[l], V, u->ls0 - l, Cu, Cs);
        NX_CHECK_LAUNCH("k_upcat");
        if (int e = conv_fwd(u, u->dec(l, 0), b, u->cat[l], u->bb[l], conv::kRelu, st)) return e;
        if (int e = conv_fwd(u, u->dec(l, 1), b, u->bb[l], u->dd[l], conv::kRelu, st)) return e;
        low = u->dd[l];
    }
    if (int e = conv_fwd(u, u->outc(), b, low, u->out32, conv::kOutF32, st)) return e;
    u->fwd_batch = b;
    return NX_OK;
}

// parameter gradients from the output gradient in u->gout (bf16, 16 channels)
int backward(nx_unet_s* u, int b, hipStream_t st)
{
    const int L = u->L;
    for (int ci = 0; ci < (int)u->convs.size(); ci++)
        if (int e = conv::flip(u->shape(ci, b), u->wbf + u->convs[ci].w_off, u->wt + u->convs[ci].w_off, st))
            return e;
    const uint16_t* top = L > 1 ? u->dd[0] : u->e[0];
    if (int e = conv_bwd(u, u->outc(), b, top, u->gout, u->g1[0], st)) return e;
    for (int l = 0; l < L - 1; l++) {
        const size_t V = u->vox(l, b);
        const size_t n = V * u->wd[l];
        if (int e = relu_bwd(u->dd[l], u->g1[l], u->g2[l], n, st)) return e;
        if (int e = conv_bwd(u, u->dec(l, 1), b, u->bb[l], u->g2[l], u->g1[l], st)) return e;
        if (int e = relu_bwd(u->bb[l], u->g1[l], u->g2[l], n, st)) return e;
        if (int e = conv_bwd(u, u->dec(l, 0), b, u->cat[l], u->g2[l], u->gcat[l], st)) return e;
        const size_t Vh = u->vox(l + 1, b);
        const int Cu = u->wd[l + 1], Cc = u->wd[l + 1] + u->wd[l];
        hipLaunchKernelGGL(k_up_bwd, dim3(blocks_for(Vh * Cu)), dim3(256), 0, st, u->gcat[l], u->g1[l + 1],
                           Vh, u->ls0 - l, Cc, Cu);
        NX_CHECK_LAUNCH("k_up_bwd");
    }
    for (int l = L - 1; l >= 0; l--) {
        const size_t V = u->vox(l, b);
        const size_t n = V * u->wd[l];
        if (l < L - 1) {
            hipLaunchKernelGGL(k_skip_grad, dim3(blocks_for(n)), dim3(256), 0, st, u->gcat[l],
                               u->wd[l + 1] + u->wd[l], u->wd[l + 1], u->gin[l + 1], u->g1[l], V,
                               u->ls0 - l, u->wd[l]);
            NX_CHECK_LAUNCH("k_skip_grad");
        }
        if (int e = relu_bwd(u->e[l], u->g1[l], u->g2[l], n, st)) return e;
        if (int e = conv_bwd(u, u->enc(l, 1), b, u->a[l], u->g2[l], u->g1[l], st)) return e;
        if (int e = relu_bwd(u->a[l], u->g1[l], u->g2[l], n, st)) return e;
        if (int e = conv_bwd(u, u->enc(l, 0), b, u->in[l], u->g2[l], l > 0 ? u->gin[l] : nullptr, st))
            return e;
    }
    return NX_OK;
}

int check_batch(nx_unet_s* u, int b)
{
    if (!u) return set_error(NX_ERR_INVALID, "nx_unet: null handle");
    if (b < 1 || b > u->B) return set_error(NX_ERR_INVALID, "nx_unet: batch out of range [1, max_batch]");
    return NX_OK;
}

}  // namespace

extern "C" {

int nx_unet_create(const nx_unet_desc* d, nx_unet* out)
{
    if (!d || !out) return set_error(NX_ERR_INVALID, "nx_unet_create: null pointer");
    const int L = d->levels, S = d->size;
    if (L < 2 || L > 4) return set_error(NX_ERR_INVALID, "nx_unet_create: levels must be 2..4");
    if (S < (1 << (L - 1)) || (S & (S - 1)))
        return set_error(NX_ERR_INVALID, "nx_unet_create: size must be a power of two >= 2^(levels-1)");
    if (d->in_ch < 1 || d->in_ch > kC0 || d->out_ch < 1 || d->out_ch > kC0 || d->max_batch < 1)
        return set_error(NX_ERR_INVALID, "nx_unet_create: in_ch/out_ch must be 1..16, max_batch >= 1");
    for (int l = 0; l < L; l++)
        if (d->width[l] < 16 || d->width[l] % 16 || d->width[l] > 128)
            return set_error(NX_ERR_INVALID, "nx_unet_create: widths must be multiples of 16 in [16, 128]");
    if ((uint64_t)d->max_batch * S * S * S * 256 >= (1ull << 31))
        return set_error(NX_ERR_INVALID, "nx_unet_create: max_batch * size^3 too large");

    auto* u = new nx_unet_s;
    u->d = *d;
    u->L = L;
    u->S = S;
    u->B = d->max_batch;
    u->ls0 = __builtin_ctz((unsigned)S);
    for (int l = 0; l < L; l++) u->wd[l] = d->width[l];
    size_t off = 0;
    auto add = [&](int cin, int cout, int ks, int level) {
        ConvSpec c{cin, cout, ks, level, off, 0};
        off += (size_t)cout * ks * ks * ks * cin;
        c.b_off = off;
        off += (size_t)cout;
        u->convs.push_back(c);
    };
    for (int l = 0; l < L; l++) {
        add(l == 0 ? kC0 : u->wd[l - 1], u->wd[l], 3, l);
        add(u->wd[l], u->wd[l], 3, l);
    }
    for (int l = 0; l < L - 1; l++) {
        add(u->wd[l + 1] + u->wd[l], u->wd[l], 3, l);
        add(u->wd[l], u->wd[l], 3, l);
    }
    add(u->wd[0], kC0, 1, 0);
    u->n_params = off;

    auto fail = [&](int e) {
        release(u);
        delete u;
        return e;
    };
    const size_t P = u->n_params;
    int e = NX_OK;
    if ((e = dalloc_t(u, P, &u->w32)) || (e = dalloc_t(u, P, &u->g32)) || (e = dalloc_t(u, P, &u->m1)) ||
        (e = dalloc_t(u, P, &u->m2)) || (e = dalloc_t(u, P, &u->wbf)) || (e = dalloc_t(u, P, &u->wt)))
        return fail(e);
    for (int l = 0; l < L; l++) {
        const size_t V = u->vox(l, u->B);
        const int cin = l == 0 ? kC0 : u->wd[l - 1];
        if ((e = dalloc_t(u, V * cin, &u->in[l])) || (e = dalloc_t(u, V * u->wd[l], &u->a[l])) ||
            (e = dalloc_t(u, V * u->wd[l], &u->e[l])) || (e = dalloc_t(u, V * u->wd[l], &u->g1[l])) ||
            (e = dalloc_t(u, V * u->wd[l], &u->g2[l])))
            return fail(e);
        if (l > 0 && (e = dalloc_t(u, V * cin, &u->gin[l]))) return fail(e);
        if (l < L - 1) {
            const size_t cc = (size_t)u->wd[l + 1] + u->wd[l];
            if ((e = dalloc_t(u, V * cc, &u->cat[l])) || (e = dalloc_t(u, V * cc, &u->gcat[l])) ||
                (e = dalloc_t(u, V * u->wd[l], &u->bb[l])) || (e = dalloc_t(u, V * u->wd[l], &u->dd[l])))
                return fail(e);
        }
    }
    const size_t V0 = u->vox(0, u->B);
    size_t ws = 0;
    for (int ci = 0; ci < (int)u->convs.size(); ci++) ws = std::max(ws, conv::bwd_weight_ws(u->shape(ci, u->B)));
    if ((e = dalloc_t(u, V0 * kC0, &u->out32)) || (e = dalloc_t(u, V0 * kC0, &u->gout)) ||
        (e = dalloc(u, ws, &u->ws)) || (e = dalloc_t(u, 8192, &u->loss_part)) ||
        (e = dalloc_t(u, 4, &u->loss_dev)))
        return fail(e);

    // He-normal weights over the real fan-in, zero biases and zero padding
    std::vector<float> h(P, 0.f);
    Rng rng{d->seed};
    for (size_t ci = 0; ci < u->convs.size(); ci++) {
        const ConvSpec& c = u->convs[ci];
        const int T = c.ks * c.ks * c.ks;
        const int real_in = ci == 0 ? d->in_ch : c.cin;
        const int real_out = (int)ci == u->outc() ? d->out_ch : c.cout;
        const double sd = std::sqrt(2.0 / (double)(T * real_in));
        for (int co = 0; co < c.cout; co++)
            for (int t = 0; t < T; t++)
                for (int i = 0; i < c.cin; i++) {
                    const double r = rng.normal();
                    if (co < real_out && i < real_in)
                        h[c.w_off + ((size_t)co * T + t) * c.cin + i] = (float)(sd * r);
                }
    }
    if (hipMemcpy(u->w32, h.data(), P * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(u->m1, 0, P * sizeof(float)) != hipSuccess ||
        hipMemset(u->m2, 0, P * sizeof(float)) != hipSuccess ||
        hipMemset(u->g32, 0, P * sizeof(float)) != hipSuccess)
        return fail(hip_error(hipGetLastError(), "nx_unet_create: upload"));
    hipLaunchKernelGGL(k_to_bf16, dim3(blocks_for(P)), dim3(256), 0, 0, u->w32, u->wbf, P);
    if (hipDeviceSynchronize() != hipSuccess) return fail(hip_error(hipGetLastError(), "nx_unet_create"));
    *out = u;
    return NX_OK;
}

int nx_unet_info(nx_unet u, uint64_t* n_params, uint32_t* n_convs)
{
    if (!u) return set_error(NX_ERR_INVALID, "nx_unet_info: null handle");
    if (n_params) *n_params = u->n_params;
    if (n_convs) *n_convs = (uint32_t)u->convs.size();
    return NX_OK;
}

int nx_unet_params(nx_unet u, float* host, int set)
{
    if (!u || !host) return set_error(NX_ERR_INVALID, "nx_unet_params: null pointer");
    const size_t bytes = u->n_params * sizeof(float);
    NX_CHECK_HIP(hipDeviceSynchronize());
    if (set) {
        NX_CHECK_HIP(hipMemcpy(u->w32, host, bytes, hipMemcpyHostToDevice));
        NX_CHECK_HIP(hipMemset(u->m1, 0, bytes));
        NX_CHECK_HIP(hipMemset(u->m2, 0, bytes));
        u->adam_t = 0;
        hipLaunchKernelGGL(k_to_bf16, dim3(blocks_for(u->n_params)), dim3(256), 0, 0, u->w32, u->wbf,
                           u->n_params);
        NX_CHECK_LAUNCH("k_to_bf16");
        NX_CHECK_HIP(hipDeviceSynchronize());
    } else {
        NX_CHECK_HIP(hipMemcpy(host, u->w32, bytes, hipMemcpyDeviceToHost));
    }
    return NX_OK;
}

int nx_unet_grads(nx_unet u, float* host)
{
    if (!u || !host) return set_error(NX_ERR_INVALID, "nx_unet_grads: null pointer");
    NX_CHECK_HIP(hipDeviceSynchronize());
    NX_CHECK_HIP(hipMemcpy(host, u->g32, u->n_params * sizeof(float), hipMemcpyDeviceToHost));
    return NX_OK;
}

int nx_unet_forward(nx_unet u, const float* x, int32_t batch, float* y, nx_stream s)
{
    if (int e = check_batch(u, batch)) return e;
    if (!x || !y) return set_error(NX_ERR_INVALID, "nx_unet_forward: null pointer");
    const hipStream_t st = (hipStream_t)s;
    if (int e = forward(u, x, batch, st)) return e;
    const size_t V0 = u->vox(0, batch);
    hipLaunchKernelGGL(k_unpack_output, dim3(blocks_for(V0 * u->d.out_ch)), dim3(256), 0, st, u->out32, y,
                       V0, u->d.out_ch);
    NX_CHECK_LAUNCH("k_unpack_output");
    return NX_OK;
}

int nx_unet_backward(nx_unet u, const float* dy, int32_t batch, nx_stream s)
{
    if (int e = check_batch(u, batch)) return e;
    if (!dy) return set_error(NX_ERR_INVALID, "nx_unet_backward: null pointer");
    if (batch != u->fwd_batch)
        return set_error(NX_ERR_INVALID, "nx_unet_backward: batch differs from the last forward");
    const hipStream_t st = (hipStream_t)s;
    const size_t V0 = u->vox(0, batch);
    hipLaunchKernelGGL(k_pack_grad, dim3(blocks_for(V0 * kC0)), dim3(256), 0, st, dy, u->gout, V0,
                       u->d.out_ch);
    NX_CHECK_LAUNCH("k_pack_grad");
    return backward(u, batch, st);
}

int nx_unet_train_step(nx_unet u, const float* x, const float* target, const float* weight, int32_t batch,
                       float lr, double* loss, nx_stream s)
{
    if (int e = check_batch(u, batch)) return e;
    if (!x || !target || !weight) return set_error(NX_ERR_INVALID, "nx_unet_train_step: null pointer");
    const hipStream_t st = (hipStream_t)s;
    if (int e = forward(u, x, batch, st)) return e;
    const size_t V0 = u->vox(0, batch);
    const unsigned nb = std::min<unsigned>(blocks_for(V0), 8192u);
    hipLaunchKernelGGL(k_mse, dim3(nb), dim3(256), 0, st, u->out32, target, weight, u->gout, u->loss_part,
                       V0, u->d.out_ch);
    NX_CHECK_LAUNCH("k_mse");
    if (int e = finish_reduce(u->loss_part, (int)nb, 1, u->loss_dev, st)) return e;
    if (int e = backward(u, batch, st)) return e;
    u->adam_t++;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = (float)(1.0 / (1.0 - std::pow((double)b1, u->adam_t)));
    const float c2 = (float)(1.0 / (1.0 - std::pow((double)b2, u->adam_t)));
    hipLaunchKernelGGL(k_adam, dim3(blocks_for(u->n_params)), dim3(256), 0, st, u->w32, u->g32, u->m1, u->m2,
                       u->wbf, u->n_params, lr, b1, b2, eps, c1, c2);
    NX_CHECK_LAUNCH("k_adam");
    if (loss) {
        NX_CHECK_HIP(hipMemcpyAsync(loss, u->loss_dev, sizeof(double), hipMemcpyDeviceToHost, st));
        NX_CHECK_HIP(hipStreamSynchronize(st));
    }
    return NX_OK;
}

int nx_unet_destroy(nx_unet u)
{
    if (!u) return NX_OK;
    (void)hipDeviceSynchronize();
    release(u);
    delete u;
    return NX_OK;
}

}  // extern "C"
