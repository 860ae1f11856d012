// peaks.hip -- measured device rates that the roofline figures are set
// against, beside the vendor peaks: a streaming copy, the FP64 and FP32
// MFMA shapes the Hex20 integration uses, and FP64 atomic adds to distinct
// addresses (the rate a scatter into a cached pattern is bound by).
#include "alls_internal.h"
#include "host_internal.h"

#include <algorithm>

namespace {

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef int i32x4 __attribute__((ext_vector_type(4)));

// 16-byte copy, U loads in flight per thread before their stores;
// non-temporal (streaming) accesses when NT.
template <int U, bool NT>
__global__ __launch_bounds__(256) void k_peak_copy(const i32x4* __restrict__ src, i32x4* __restrict__ dst, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + (U - 1) * stride < n; i += U * stride) {
        i32x4 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = NT ? __builtin_nontemporal_load(src + i + u * stride) : src[i + u * stride];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (NT) __builtin_nontemporal_store(v[u], dst + i + u * stride);
            else dst[i + u * stride] = v[u];
        }
    }
    for (; i < n; i += stride) dst[i] = src[i];
}

// CH independent MFMA chains per wave on register operands; one lane
// stores a result that depends on all of them so that nothing is dropped.
template <int CH>
__global__ __launch_bounds__(256) void k_peak_mfma_f64(int iters, double* sink) {
    const double a = 1.0 + 1e-9 * threadIdx.x, b = 1.0 - 1e-9 * threadIdx.x;
    f64x4 c[CH];
#pragma unroll
    for (int k = 0; k < CH; ++k) c[k] = f64x4{0, 0, 0, 0};
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int k = 0; k < CH; ++k) c[k] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c[k], 0, 0, 0);
    }
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < CH; ++k) s += c[k][k & 3];
    if (threadIdx.x == 0) sink[blockIdx.x] = s;
}

template <int CH>
__global__ __launch_bounds__(256) void k_peak_mfma_f32(int iters, double* sink) {
    const float a = 1.0f + 1e-6f * threadIdx.x, b = 1.0f - 1e-6f * threadIdx.x;
    f32x4 c[CH];
#pragma unroll
    for (int k = 0; k < CH; ++k) c[k] = f32x4{0, 0, 0, 0};
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int k = 0; k < CH; ++k) c[k] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c[k], 0, 0, 0);
    }
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < CH; ++k) s += c[k][k & 3];
    if (threadIdx.x == 0) sink[blockIdx.x] = (double)s;
}

// FP64 atomic adds (no return) spread over `slots` doubles: thread t adds
// to slot (t * 9973 + i * 65537) mod slots, so lanes of a wave hit distinct
// lines and every slot is hit about equally.
__global__ __launch_bounds__(256) void k_peak_atomic_f64(double* buf, int64_t slots, int iters) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (int i = 0; i < iters; ++i) atomicAdd(&buf[(t * 9973 + (int64_t)i * 65537) % slots], 1.0);
}

}  // namespace

namespace {

struct PeakBufs {
    void* a = nullptr;
    void* b = nullptr;
    ~PeakBufs() {
        if (a) (void)hipFree(a);
        if (b) (void)hipFree(b);
    }
};

// Average milliseconds of `reps` launches of f after one untimed launch.
template <typename F>
hipError_t time_ms(hipStream_t s, int reps, F f, float* ms) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) {
        f();
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e0, s);
    for (int r = 0; r < reps && e == hipSuccess; ++r) {
        f();
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e1, s);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, e0, e1);
    if (e == hipSuccess) *ms /= (float)reps;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    return e;
}

}  // namespace

extern "C" {

alls_status alls_device_peaks(alls_ctx* ctx, alls_peaks* out) {
    if (!ctx || !out) return ALLS_E_INVALID;
    ctx->err.clear();
    (void)hipSetDevice(ctx->device);
    *out = alls_peaks{};
    hipStream_t s = ctx->stream;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess || cus <= 0)
        cus = 256;
    const int64_t copy_bytes = int64_t(1) << 30;
    const int64_t slots = int64_t(1) << 24;   // 128 MB of doubles for the atomics
    PeakBufs buf;
    hipError_t e = hipMalloc(&buf.a, (size_t)copy_bytes);
    if (e == hipSuccess) e = hipMalloc(&buf.b, (size_t)copy_bytes);
    if (e == hipSuccess) e = hipMemsetAsync(buf.a, 0, (size_t)copy_bytes, s);
    if (e == hipSuccess) e = hipMemsetAsync(buf.b, 0, (size_t)copy_bytes, s);
    if (e != hipSuccess) return alls::hip_fail(e, "device peaks: buffers", &ctx->err);
    float ms = 0.f;
    // Streaming copy: 1 GiB read + 1 GiB written; the best of a few shapes
    // (loads in flight per thread, streaming or cached accesses, blocks per CU).
    {
        const int64_t n16 = copy_bytes / 16;
        const i32x4* src = static_cast<const i32x4*>(buf.a);
        i32x4* dst = static_cast<i32x4*>(buf.b);
        double best = 0.0;
        for (int variant = 0; variant < 6; ++variant) {
            const unsigned grid = (unsigned)(cus * (variant % 3 == 0 ? 8 : variant % 3 == 1 ? 16 : 32));
            e = time_ms(s, 5, [&] {
                if (variant < 3)
                    hipLaunchKernelGGL((k_peak_copy<8, true>), dim3(grid), dim3(256), 0, s, src, dst, n16);
                else
                    hipLaunchKernelGGL((k_peak_copy<4, false>), dim3(grid), dim3(256), 0, s, src, dst, n16);
            }, &ms);
            if (e != hipSuccess) return alls::hip_fail(e, "device peaks: copy", &ctx->err);
            best = std::max(best, 2.0 * (double)copy_bytes / (ms * 1e-3) / 1e9);
        }
        out->copy_GBs = best;
    }
    // MFMA: 16x16x4 shapes, 2048 flops each; 8 blocks of 4 waves per CU,
    // 4 or 8 independent chains per wave (the better of the two).
    {
        const int iters = 8192;
        const unsigned grid = (unsigned)(cus * 8);
        double* sink = static_cast<double*>(buf.b);
        auto rate = [&](int ch) { return (double)grid * 4 * iters * ch * 2048.0 / (ms * 1e-3) / 1e12; };
        e = time_ms(s, 3, [&] { hipLaunchKernelGGL((k_peak_mfma_f64<4>), dim3(grid), dim3(256), 0, s, iters, sink); }, &ms);
        if (e != hipSuccess) return alls::hip_fail(e, "device peaks: mfma f64", &ctx->err);
        out->mfma_f64_TFLOPs = rate(4);
        e = time_ms(s, 3, [&] { hipLaunchKernelGGL((k_peak_mfma_f64<8>), dim3(grid), dim3(256), 0, s, iters, sink); }, &ms);
        if (e != hipSuccess) return alls::hip_fail(e, "device peaks: mfma f64", &ctx->err);
        out->mfma_f64_TFLOPs = std::max(out->mfma_f64_TFLOPs, rate(8));
        e = time_ms(s, 3, [&] { hipLaunchKernelGGL((k_peak_mfma_f32<4>), dim3(grid), dim3(256), 0, s, iters, sink); }, &ms);
        if (e != hipSuccess) return alls::hip_fail(e, "device peaks: mfma f32", &ctx->err);
        out->mfma_f32_TFLOPs = rate(4);
        e = time_ms(s, 3, [&] { hipLaunchKernelGGL((k_peak_mfma_f32<8>), dim3(grid), dim3(256), 0, s, iters, sink); }, &ms);
        if (e != hipSuccess) return alls::hip_fail(e, "device peaks: mfma f32", &ctx->err);
        out->mfma_f32_TFLOPs = std::max(out->mfma_f32_TFLOPs, rate(8));
    }
    // FP64 atomic adds to 16M distinct doubles.
    {
        const int iters = 64;
        const unsigned grid = (unsigned)(cus * 16);
        e = time_ms(s, 3, [&] {
            hipLaunchKernelGGL(k_peak_atomic_f64, dim3(grid), dim3(256), 0, s, static_cast<double*>(buf.a), slots,
                               iters);
        }, &ms);
        if (e != hipSuccess) return alls::hip_fail(e, "device peaks: atomics", &ctx->err);
        out->atomic_f64_Gops = (double)grid * 256 * iters / (ms * 1e-3) / 1e9;
    }
    return ALLS_OK;
}

}  // extern "C"
