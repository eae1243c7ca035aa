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
