// bs.hip -- closed-form European price and the FD1D-BS control-variate combine.
#include <math.h>

#include "kwgpu_internal.h"

namespace kwgpu {


namespace {

__global__ __launch_bounds__(256) void bs_kernel(const kwgpu_option* opts, size_t n,
                                                 double* prices) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) prices[i] = bs_one(opts[i]);
}

// Fd1d_BlackScholes_Pricer::price, src/Pricer/kwFd1d_BlackScholes.cpp:38-40:
// price += BS(euro) - FD(euro).
__global__ __launch_bounds__(256) void fd1d_bs_combine_kernel(const kwgpu_option* opts, size_t n,
                                                              double* amer, const double* euro) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        kwgpu_option a = opts[i];
        a.e = 0;
        amer[i] += bs_one(a) - euro[i];
    }
}

}  // namespace

hipError_t launch_bs(const kwgpu_option* opts, size_t n, double* prices, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const unsigned blocks = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(bs_kernel, dim3(blocks), dim3(256), 0, s, opts, n, prices);
    return hipGetLastError();
}

hipError_t launch_fd1d_bs_combine(const kwgpu_option* opts, size_t n, double* amer,
                                  const double* euro, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const unsigned blocks = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(fd1d_bs_combine_kernel, dim3(blocks), dim3(256), 0, s, opts, n, amer,
                       euro);
    return hipGetLastError();
}

}  // namespace kwgpu
