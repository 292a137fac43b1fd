// kwgpu_peak -- measures, on the device it runs on, the two ceilings the FD pricer is
// measured against: FP64 FMA throughput of the vector ALUs (independent FMA chains in
// registers, no memory traffic) and device-to-device copy bandwidth (a streaming copy
// kernel over 1 GiB). Prints one JSON line. Diagnostic only; not part of the library.
#include <hip/hip_runtime.h>

#include <cstdio>

namespace {

constexpr int kChains = 8;     // independent FMA chains per thread
constexpr int kIters = 4096;   // iterations of the chain loop

__global__ __launch_bounds__(256) void fma_loop(double* out, double a, double b) {
    double acc[kChains];
#pragma unroll
    for (int c = 0; c < kChains; ++c) acc[c] = threadIdx.x * 1e-9 + c;
    for (int it = 0; it < kIters; ++it) {
#pragma unroll
        for (int c = 0; c < kChains; ++c) acc[c] = fma(acc[c], a, b);
    }
    double s = 0;
#pragma unroll
    for (int c = 0; c < kChains; ++c) s += acc[c];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void copy_kernel(const double2* __restrict__ src,
                                                   double2* __restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

#define CHECK(x)                                                                   \
    do {                                                                           \
        hipError_t e_ = (x);                                                       \
        if (e_ != hipSuccess) {                                                    \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));          \
            return 1;                                                              \
        }                                                                          \
    } while (0)

}  // namespace

int main() {
    hipDeviceProp_t prop;
    CHECK(hipGetDeviceProperties(&prop, 0));
    const int blocks = prop.multiProcessorCount * 32;  // 8 waves per SIMD
    double* out = nullptr;
    CHECK(hipMalloc(&out, (size_t)blocks * 256 * sizeof(double)));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));

    // FP64 FMA: warm-up, then the best of 5 launches
    hipLaunchKernelGGL(fma_loop, dim3(blocks), dim3(256), 0, 0, out, 0.999999, 1e-7);
    CHECK(hipGetLastError());
    CHECK(hipDeviceSynchronize());
    float best = 1e30f;
    for (int r = 0; r < 5; ++r) {
        CHECK(hipEventRecord(e0, 0));
        hipLaunchKernelGGL(fma_loop, dim3(blocks), dim3(256), 0, 0, out, 0.999999, 1e-7);
        CHECK(hipEventRecord(e1, 0));
        CHECK(hipEventSynchronize(e1));
        float ms;
        CHECK(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) best = ms;
    }
    const double flops = 2.0 * kChains * kIters * (double)blocks * 256;
    const double fp64_tflops = flops / (best * 1e-3) / 1e12;

    // copy: 1 GiB read + 1 GiB written per launch
    const size_t bytes = (size_t)1 << 30;
    const size_t n = bytes / sizeof(double2);
    double2 *a = nullptr, *b = nullptr;
    CHECK(hipMalloc(&a, bytes));
    CHECK(hipMalloc(&b, bytes));
    CHECK(hipMemset(a, 0, bytes));
    const int cblocks = prop.multiProcessorCount * 16;
    hipLaunchKernelGGL(copy_kernel, dim3(cblocks), dim3(256), 0, 0, a, b, n);
    CHECK(hipGetLastError());
    CHECK(hipDeviceSynchronize());
    float cbest = 1e30f;
    for (int r = 0; r < 5; ++r) {
        CHECK(hipEventRecord(e0, 0));
        hipLaunchKernelGGL(copy_kernel, dim3(cblocks), dim3(256), 0, 0, a, b, n);
        CHECK(hipEventRecord(e1, 0));
        CHECK(hipEventSynchronize(e1));
        float ms;
        CHECK(hipEventElapsedTime(&ms, e0, e1));
        if (ms < cbest) cbest = ms;
    }
    const double copy_gbs = 2.0 * bytes / (cbest * 1e-3) / 1e9;
    std::printf("{\"device\": \"%s\", \"cus\": %d, \"fp64_fma_tflops\": %.2f, "
                "\"fp64_fma_ms\": %.4f, \"copy_gb_per_s\": %.1f, \"copy_ms\": %.4f}\n",
                prop.gcnArchName, prop.multiProcessorCount, fp64_tflops, best, copy_gbs, cbest);
    (void)hipFree(out);
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}
