// nx_runtime.hip -- device selection, streams, events, memory, error state
// and the second reduction stage.  Replaces the single-device subset of
// libNeonSys (GpuDevice.h:151-190 launch/error handling,
// gpuDeviceMemory.cpp) and libNeonSet's Backend (src/set/Backend.cpp:46-55,
// 404-419): one device, launches asynchronous on a caller stream.
#include <algorithm>
#include <mutex>
#include <string>

#include "nx_common.hpp"

namespace nx {

namespace {
thread_local std::string g_err;
std::mutex g_mu;
constexpr int kMaxDev = 64;
void* g_scratch[kMaxDev] = {};
size_t g_scratch_bytes[kMaxDev] = {};
}  // namespace

int set_error(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

int hip_error(hipError_t e, const char* what)
{
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? NX_ERR_OOM : NX_ERR_HIP;
}

int scratch(size_t bytes, void** ptr)
{
    int dev = 0;
    NX_CHECK_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDev) return set_error(NX_ERR_INVALID, "device index out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch_bytes[dev] < bytes) {
        if (g_scratch[dev]) {
            // the old buffer may still be read by queued reductions
            NX_CHECK_HIP(hipDeviceSynchronize());
            NX_CHECK_HIP(hipFree(g_scratch[dev]));
            g_scratch[dev] = nullptr;
            g_scratch_bytes[dev] = 0;
        }
        size_t want = bytes < (1u << 20) ? (1u << 20) : bytes;
        NX_CHECK_HIP(hipMalloc(&g_scratch[dev], want));
        g_scratch_bytes[dev] = want;
    }
    *ptr = g_scratch[dev];
    return NX_OK;
}

// Second stage: one workgroup sums nv interleaved partials in a fixed order.
__global__ void __launch_bounds__(256) k_finish(const double* part, int n, int nv, double* out)
{
    __shared__ double lds[16 * 4];
    double acc[4] = {0, 0, 0, 0};
    for (int i = threadIdx.x; i < n; i += 256)
        for (int k = 0; k < nv; k++) acc[k] += part[i * nv + k];
    block_sum_f64<4>(acc, lds);
    if (threadIdx.x == 0)
        for (int k = 0; k < nv; k++) out[k] = acc[k];
}

// Plain streaming copy, 16 B per lane per iteration: the measure of
// achievable HBM bandwidth (and of counter scaling) on the device.
__global__ void __launch_bounds__(256) k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                              size_t n)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        dst[i] = src[i];
}

// The D3Q19 step's memory access pattern with no arithmetic: every cell of an
// X*Y*Z grid reads population plane q at its upstream cell (i - c_q, the
// index clamped to the grid) and writes plane q at i, one thread per cell,
// 256 cells per workgroup, each XCD walking one contiguous range of cells
// (as the dense step does).  What this copy streams is the ceiling of the
// pull-streaming step on SoA planes.
template <typename E>
__global__ void __launch_bounds__(256)
    k_shifted_copy(const E* __restrict__ src, E* __restrict__ dst, int32_t X, int32_t Y, uint32_t n)
{
    const uint32_t i = xcd_remap(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    if (i >= n) return;
    const int64_t XY = (int64_t)X * Y;
    E v[19];
#pragma unroll
    for (int q = 0; q < 19; q++) {
        const int64_t off = (int64_t)CX[q] + (int64_t)X * CY[q] + XY * CZ[q];
        int64_t j = (int64_t)i - off;
        j = j < 0 ? 0 : (j >= (int64_t)n ? (int64_t)n - 1 : j);
        v[q] = src[(uint64_t)q * n + (uint64_t)j];
    }
#pragma unroll
    for (int q = 0; q < 19; q++) __builtin_nontemporal_store(v[q], &dst[(uint64_t)q * n + i]);
}

int finish_reduce(const double* partials, int n, int nv, double* out, hipStream_t s)
{
    if (nv < 1 || nv > 4) return set_error(NX_ERR_INVALID, "finish_reduce: nv out of range");
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(256), 0, s, partials, n, nv, out);
    NX_CHECK_LAUNCH("k_finish");
    return NX_OK;
}

}  // namespace nx

using namespace nx;

extern "C" {

const char* nx_last_error(void) { return g_err.c_str(); }

int nx_init(int device)
{
    int n = 0;
    NX_CHECK_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n)
        return set_error(NX_ERR_INVALID, "nx_init: device " + std::to_string(device) +
                                             " not present (" + std::to_string(n) + " devices)");
    NX_CHECK_HIP(hipSetDevice(device));
    NX_CHECK_HIP(hipFree(nullptr));  // create the context now
    return dense_noalloc_setup();
}

int nx_device_count(int* count)
{
    if (!count) return set_error(NX_ERR_INVALID, "null pointer");
    NX_CHECK_HIP(hipGetDeviceCount(count));
    return NX_OK;
}

int nx_device_name(char* buf, int len)
{
    if (!buf || len < 1) return set_error(NX_ERR_INVALID, "bad buffer");
    int dev = 0;
    NX_CHECK_HIP(hipGetDevice(&dev));
    hipDeviceProp_t p;
    NX_CHECK_HIP(hipGetDeviceProperties(&p, dev));
    // the marketing name can be empty (no access to the product table)
    const std::string name = p.name[0] ? std::string(p.name) : std::string("AMD GPU");
    std::string s = name + " (" + p.gcnArchName + ", " + std::to_string(p.multiProcessorCount) +
                    " CUs, " + std::to_string(p.totalGlobalMem >> 30) + " GiB)";
    snprintf(buf, (size_t)len, "%s", s.c_str());
    return NX_OK;
}

int nx_sync(void)
{
    NX_CHECK_HIP(hipDeviceSynchronize());
    return NX_OK;
}

int nx_stream_create(nx_stream* s)
{
    if (!s) return set_error(NX_ERR_INVALID, "null pointer");
    hipStream_t h;
    NX_CHECK_HIP(hipStreamCreateWithFlags(&h, hipStreamNonBlocking));
    *s = (nx_stream)h;
    return NX_OK;
}

int nx_stream_sync(nx_stream s)
{
    NX_CHECK_HIP(hipStreamSynchronize((hipStream_t)s));
    return NX_OK;
}

int nx_stream_destroy(nx_stream s)
{
    if (s) NX_CHECK_HIP(hipStreamDestroy((hipStream_t)s));
    return NX_OK;
}

int nx_event_create(nx_event* e)
{
    if (!e) return set_error(NX_ERR_INVALID, "null pointer");
    hipEvent_t h;
    NX_CHECK_HIP(hipEventCreate(&h));
    *e = (nx_event)h;
    return NX_OK;
}

int nx_event_record(nx_event e, nx_stream s)
{
    NX_CHECK_HIP(hipEventRecord((hipEvent_t)e, (hipStream_t)s));
    return NX_OK;
}

int nx_event_sync(nx_event e)
{
    NX_CHECK_HIP(hipEventSynchronize((hipEvent_t)e));
    return NX_OK;
}

int nx_event_elapsed_ms(nx_event a, nx_event b, float* ms)
{
    if (!ms) return set_error(NX_ERR_INVALID, "null pointer");
    NX_CHECK_HIP(hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b));
    return NX_OK;
}

int nx_event_destroy(nx_event e)
{
    if (e) NX_CHECK_HIP(hipEventDestroy((hipEvent_t)e));
    return NX_OK;
}

int nx_alloc(size_t bytes, void** dev)
{
    if (!dev) return set_error(NX_ERR_INVALID, "null pointer");
    *dev = nullptr;
    if (bytes == 0) return NX_OK;
    NX_CHECK_HIP(hipMalloc(dev, bytes));
    return NX_OK;
}

int nx_free(void* dev)
{
    if (dev) NX_CHECK_HIP(hipFree(dev));
    return NX_OK;
}

int nx_memset(void* dev, int value, size_t bytes, nx_stream s)
{
    if (!bytes) return NX_OK;
    NX_CHECK_HIP(hipMemsetAsync(dev, value, bytes, (hipStream_t)s));
    return NX_OK;
}

int nx_h2d(void* dst, const void* src, size_t bytes, nx_stream s)
{
    if (!bytes) return NX_OK;
    NX_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)s));
    return NX_OK;
}

int nx_d2h(void* dst, const void* src, size_t bytes, nx_stream s)
{
    if (!bytes) return NX_OK;
    NX_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)s));
    return NX_OK;
}

int nx_d2d(void* dst, const void* src, size_t bytes, nx_stream s)
{
    if (!bytes) return NX_OK;
    NX_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)s));
    return NX_OK;
}

int nx_shifted_copy(void* dst, const void* src, const int32_t dim[3], int32_t elem_bytes, nx_stream s)
{
    if (!dst || !src || !dim) return set_error(NX_ERR_INVALID, "null pointer");
    if (dim[0] < 1 || dim[1] < 1 || dim[2] < 1)
        return set_error(NX_ERR_INVALID, "nx_shifted_copy: dimensions must be >= 1");
    const uint64_t n = (uint64_t)dim[0] * dim[1] * dim[2];
    if (n >= (1ull << 32) / 19) return set_error(NX_ERR_INVALID, "nx_shifted_copy: grid too large");
    if (dst == src) return set_error(NX_ERR_INVALID, "nx_shifted_copy: src and dst must differ");
    const unsigned nb = (unsigned)((n + 255) / 256);
    const hipStream_t st = (hipStream_t)s;
    switch (elem_bytes) {
        case 2: hipLaunchKernelGGL(k_shifted_copy<uint16_t>, dim3(nb), dim3(256), 0, st, (const uint16_t*)src,
                                   (uint16_t*)dst, dim[0], dim[1], (uint32_t)n); break;
        case 4: hipLaunchKernelGGL(k_shifted_copy<uint32_t>, dim3(nb), dim3(256), 0, st, (const uint32_t*)src,
                                   (uint32_t*)dst, dim[0], dim[1], (uint32_t)n); break;
        case 8: hipLaunchKernelGGL(k_shifted_copy<uint64_t>, dim3(nb), dim3(256), 0, st, (const uint64_t*)src,
                                   (uint64_t*)dst, dim[0], dim[1], (uint32_t)n); break;
        default: return set_error(NX_ERR_INVALID, "nx_shifted_copy: elem_bytes 2, 4 or 8");
    }
    NX_CHECK_LAUNCH("k_shifted_copy");
    return NX_OK;
}

int nx_copy_kernel(void* dst, const void* src, size_t bytes, nx_stream s)
{
    if (!dst || !src) return set_error(NX_ERR_INVALID, "null pointer");
    if (bytes % 16 || ((uintptr_t)dst | (uintptr_t)src) % 16)
        return set_error(NX_ERR_INVALID, "nx_copy_kernel: 16-byte aligned sizes and pointers only");
    const size_t n = bytes / 16;
    const unsigned nb = (unsigned)std::min<size_t>((n + 255) / 256, 256 * 64);
    hipLaunchKernelGGL(k_copy, dim3(nb ? nb : 1), dim3(256), 0, (hipStream_t)s, (const uint4*)src,
                       (uint4*)dst, n);
    NX_CHECK_LAUNCH("k_copy");
    return NX_OK;
}

int nx_host_alloc(size_t bytes, void** host)
{
    if (!host) return set_error(NX_ERR_INVALID, "null pointer");
    NX_CHECK_HIP(hipHostMalloc(host, bytes, hipHostMallocDefault));
    return NX_OK;
}

int nx_host_free(void* host)
{
    if (host) NX_CHECK_HIP(hipHostFree(host));
    return NX_OK;
}

}  // extern "C"
