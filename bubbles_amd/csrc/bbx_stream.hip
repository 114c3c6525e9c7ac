// bbx_stream.hip -- STREAM-style copy kernel: measures achievable HBM
// bandwidth on the device and calibrates the FETCH_SIZE / WRITE_SIZE
// counters against a kernel whose traffic is known exactly (2 * bytes).
#include "bbx_common.h"

namespace bbx {

// Eight 16-byte loads per lane are issued before any store, so a wave keeps
// 8 KiB of reads in flight; streaming (non-temporal) accesses keep the copy
// from thrashing the caches.  The tail of n not divisible by the grid's
// chunk is copied one element per lane.
constexpr int kCopyUnroll = 8;
typedef float f32x4_t __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(256) k_stream_copy(const float4 *__restrict__ a,
                                                     float4 *__restrict__ b, size_t n) {
    const size_t chunk = (size_t)blockDim.x * kCopyUnroll;
    const size_t full = n / chunk;
    for (size_t c = blockIdx.x; c < full; c += gridDim.x) {
        const size_t base = c * chunk + threadIdx.x;
        const f32x4_t *src = reinterpret_cast<const f32x4_t *>(a) + base;
        f32x4_t *dst = reinterpret_cast<f32x4_t *>(b) + base;
        f32x4_t v[kCopyUnroll];
#pragma unroll
        for (int u = 0; u < kCopyUnroll; u++) v[u] = __builtin_nontemporal_load(src + (size_t)u * blockDim.x);
#pragma unroll
        for (int u = 0; u < kCopyUnroll; u++) __builtin_nontemporal_store(v[u], dst + (size_t)u * blockDim.x);
    }
    for (size_t i = full * chunk + (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

}  // namespace bbx

using namespace bbx;

extern "C" int bbx_stream_copy(double bytes, int iters, double *gbps, double *avg_ms) {
    if (!(bytes >= 1024) || iters < 1) return fail(BBX_E_INVALID, "bytes >= 1 KiB and iters >= 1");
    const size_t n = (size_t)(bytes / sizeof(float4));
    float4 *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = BBX_OK;
    if (hipMalloc((void **)&a, n * sizeof(float4)) != hipSuccess ||
        hipMalloc((void **)&b, n * sizeof(float4)) != hipSuccess) {
        rc = fail(BBX_E_NOMEM, "stream copy buffers");
    }
    if (!rc && (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
                hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        rc = fail(BBX_E_HIP, "stream copy setup");
    if (!rc && hipMemsetAsync(a, 0, n * sizeof(float4), st) != hipSuccess) rc = fail(BBX_E_HIP, "memset");
    const unsigned blocks = 256 * 8;  // eight workgroups per CU, each looping over 32 KiB chunks
    if (!rc) {
        k_stream_copy<<<blocks, 256, 0, st>>>(a, b, n);  // warm-up
        (void)hipEventRecord(e0, st);
        for (int k = 0; k < iters; k++) k_stream_copy<<<blocks, 256, 0, st>>>(a, b, n);
        (void)hipEventRecord(e1, st);
        if (hipEventSynchronize(e1) != hipSuccess) rc = fail(BBX_E_HIP, "stream copy run");
    }
    if (!rc) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        double per = ms / iters;
        if (avg_ms) *avg_ms = per;
        if (gbps) *gbps = 2.0 * (double)n * sizeof(float4) / (per * 1e-3) / 1e9;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}
