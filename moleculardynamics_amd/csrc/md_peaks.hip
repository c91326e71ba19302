// md_device_peaks: the attainable peaks of the current device, measured with
// three kernels, for quoting roofline fractions against what this code can
// reach rather than against data-sheet figures:
//   * HBM: a copy of 16-byte elements, one per thread, between two 1 GiB
//     buffers (bytes = read + write);
//   * FP64 and FP32 vector FMA: every lane runs 8 independent FMA chains
//     (a = a * b + c) of 4,096 links, enough waves to fill every SIMD; flops
//     = 2 per FMA. FP32 is built without the packed-FP32 forms (as all of this
//     library is), so its figure is the non-packed rate this code runs at.
// Each kernel is launched twice untimed, then timed over 5 launches with HIP
// events on a stream of its own.
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/moldyn_hip.h"

namespace {

constexpr int kChains = 8;
constexpr int kLinks = 4096;
constexpr int kThreads = 256;

// One element per thread: on gfx950 this beats grid-stride and unrolled
// copies (6.2 against 4.6-5.4 TB/s in a sweep of 4-64 workgroups per CU and
// 1-8 elements per pass).
__global__ __launch_bounds__(kThreads) void k_copy(const double2* __restrict__ src,
                                                   double2* __restrict__ dst, long n) {
    const long i = static_cast<long>(blockIdx.x) * kThreads + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

template <class T>
__global__ __launch_bounds__(kThreads) void k_fma(T b, T c, T* __restrict__ sink) {
    T a[kChains];
#pragma unroll
    for (int k = 0; k < kChains; ++k) a[k] = static_cast<T>(threadIdx.x + k);
    for (int l = 0; l < kLinks; ++l) {
#pragma unroll
        for (int k = 0; k < kChains; ++k) a[k] = fma(a[k], b, c);
    }
    T s = 0;
#pragma unroll
    for (int k = 0; k < kChains; ++k) s += a[k];
    // b = 1 and c = 0 keep every value finite; the store never happens for
    // those arguments but the compiler cannot know it
    if (s == static_cast<T>(-1.0)) sink[threadIdx.x] = s;
}

template <class F>
hipError_t time_ms(hipStream_t s, F&& launch, float* ms) {
    hipEvent_t e0, e1;
    hipError_t e = hipEventCreate(&e0);
    if (e != hipSuccess) return e;
    e = hipEventCreate(&e1);
    if (e != hipSuccess) {
        (void)hipEventDestroy(e0);
        return e;
    }
    launch();
    launch();
    e = hipEventRecord(e0, s);
    for (int r = 0; r < 5 && e == hipSuccess; ++r) launch();
    if (e == hipSuccess) e = hipEventRecord(e1, s);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipEventElapsedTime(ms, e0, e1);
    *ms /= 5.0f;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return e;
}

}  // namespace

md_status md_device_peaks(double* hbm_bytes_per_s, double* fp64_flops, double* fp32_flops) {
    hipDeviceProp_t prop{};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
        return MD_ERR_HIP;
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return MD_ERR_HIP;
    md_status st = MD_OK;
    const long bytes = 1L << 30;
    const long n = bytes / static_cast<long>(sizeof(double2));
    double2 *src = nullptr, *dst = nullptr;
    double* sink64 = nullptr;
    float* sink32 = nullptr;
    float ms = 0.0f;
    const unsigned copy_blocks = static_cast<unsigned>((n + kThreads - 1) / kThreads);
    const unsigned fma_blocks = static_cast<unsigned>(prop.multiProcessorCount) * 32;
    if (hipMalloc(reinterpret_cast<void**>(&src), bytes) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&dst), bytes) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&sink64), kThreads * sizeof(double)) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&sink32), kThreads * sizeof(float)) != hipSuccess ||
        hipMemsetAsync(src, 0, bytes, s) != hipSuccess) {
        st = MD_ERR_HIP;
    }
    if (st == MD_OK && hbm_bytes_per_s != nullptr) {
        if (time_ms(s, [&] { hipLaunchKernelGGL(k_copy, dim3(copy_blocks), dim3(kThreads), 0, s, src, dst, n); },
                    &ms) != hipSuccess)
            st = MD_ERR_HIP;
        else
            *hbm_bytes_per_s = 2.0 * static_cast<double>(bytes) / (ms * 1e-3);
    }
    const double flops = 2.0 * kChains * kLinks * static_cast<double>(kThreads) * fma_blocks;
    if (st == MD_OK && fp64_flops != nullptr) {
        if (time_ms(s, [&] {
                hipLaunchKernelGGL(k_fma<double>, dim3(fma_blocks), dim3(kThreads), 0, s, 1.0, 0.0,
                                   sink64);
            }, &ms) != hipSuccess)
            st = MD_ERR_HIP;
        else
            *fp64_flops = flops / (ms * 1e-3);
    }
    if (st == MD_OK && fp32_flops != nullptr) {
        if (time_ms(s, [&] {
                hipLaunchKernelGGL(k_fma<float>, dim3(fma_blocks), dim3(kThreads), 0, s, 1.0f, 0.0f,
                                   sink32);
            }, &ms) != hipSuccess)
            st = MD_ERR_HIP;
        else
            *fp32_flops = flops / (ms * 1e-3);
    }
    (void)hipStreamSynchronize(s);
    for (void* p : {static_cast<void*>(src), static_cast<void*>(dst), static_cast<void*>(sink64),
                    static_cast<void*>(sink32)})
        if (p) (void)hipFree(p);
    (void)hipStreamDestroy(s);
    return st;
}
