// kwgpu_internal.h -- shared device/host definitions of the gfx950 FD pricer.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kwgpu.h"

namespace kwgpu {

// One batch of PDE solves as the kernels see it. All pointers are device pointers.
//
// PDE p (0 <= p < m) is solved on the grid of option rep[p] (identity when rep is
// null) and its solution is interpolated for the options list[off[p] .. off[p+1])
// (just option rep[p] when off is null). `order`, when given, is the processing
// order of PDEs (slot -> p), used to group PDEs of similar cost into one wave.
struct FdBatch {
    const kwgpu_option* opts;
    const uint32_t* rep;
    const uint32_t* off;
    const uint32_t* list;
    const uint32_t* order;
    const int32_t* x_dims;  // per option (indexed by rep[p]); null = x_dim
    const int32_t* t_dims;  // per option; null = t_dim
    int m;
    int x_dim, t_dim;
    double density, scale, theta;
    double* prices;
    int32_t* status;
    double* ws;             // lane layout workspace (null for the wave layout)
    long long ws_stride;    // slots per workspace row (>= m)
    int ws_rows;            // rows per workspace array (>= every x_dim in the batch)
};

// Limits of the FP32-mixed march. Both depend only on the PDE (T, sigma, r, q) and the
// mesh scale, so every option of a PDE falls on the same side.
// Overflow: a call's far payoff e^(scale sigma sqrt T) must stay within FP32's range
// (~e^88.7); above e^80 its FP32 increments can overflow.
__host__ __device__ inline bool mixed_overflow_risk(const kwgpu_option& o, double scale) {
    return !(scale * o.z * sqrt(o.t) <= 80.0);
}
// Drift: over the option's life the drift |r - q - sigma^2/2| T moves the solution by
// more than the mesh's half-width scale sigma sqrt T, into the boundary rows. There the
// mixed march's FP32 substitutions miss the 1e-4 bar. A CPU emulation of the mixed march
// (2880 options: vol 1-5 %, T 1-30 years, r -3 %-20 %, q 0-30 %, on grids from 256 x 64
// to 2048 x 1095) measured at most 1.2e-5 with the drift within the half-width and up to
// 2e-1 beyond it.
__host__ __device__ inline bool mixed_drift_risk(const kwgpu_option& o, double scale) {
    return !(fabs(o.r - o.q - 0.5 * o.z * o.z) * o.t <= scale * o.z * sqrt(o.t));
}

// Per-PDE parameters derived from an option (src/Pricer/kwFd1d.cpp:71-86).
struct PdeParams {
    double t, a0, ax, axx;
    double yMin, yMax, dy;  // sinh mesh: x_j = density * sinh(yMin (1 - j dy) + yMax j dy)
    double dt;              // uniform time step T / (tDim - 1)
    int x_dim, t_dim;
    bool american;
    bool call;
    bool valid;
};

__device__ __forceinline__ bool finite_pos(double v) { return v > 0.0 && v < 1.0e300; }

// Grid and coefficient setup mirror the reference's expression order without FMA
// contraction so the grid matches the CPU pricer to within libm ulps.
#pragma clang fp contract(off)
__device__ __forceinline__ PdeParams pde_params(const kwgpu_option& o, int x_dim, int t_dim,
                                                double density, double scale) {
    PdeParams p;
    p.t = o.t;
    p.a0 = -o.r;
    p.ax = o.r - o.q - o.z * o.z / 2;
    p.axx = o.z * o.z / 2;
    p.x_dim = x_dim;
    p.t_dim = t_dim;
    p.american = o.e != 0;
    p.call = !(o.w < 0);
    p.valid = finite_pos(o.t) && finite_pos(o.z) && finite_pos(o.s) && finite_pos(o.k) &&
              (o.r == o.r) && (o.q == o.q) && fabs(o.r) < 1e6 && fabs(o.q) < 1e6 && x_dim >= 3 &&
              t_dim >= 2;
    // t-grid, src/Pricer/kwFd1d.cpp:95-103
    p.dt = (o.t - 0.0) / (double)(t_dim - 1);
    // x-grid, src/Pricer/kwFd1d.cpp:105-125
    const double xMin = 0.0 - scale * o.z * sqrt(o.t);
    const double xMax = 0.0 + scale * o.z * sqrt(o.t);
    p.yMin = asinh((xMin - 0.0) / density);
    p.yMax = asinh((xMax - 0.0) / density);
    p.dy = 1. / (double)(x_dim - 1);
    return p;
}

__device__ __forceinline__ double grid_x(const PdeParams& p, double density, int j) {
    const double yj = j * p.dy;
    return 0.0 + density * sinh(p.yMin * (1.0 - yj) + p.yMax * yj);
}

// payoff in units of K, src/Pricer/kwFd1d.cpp:127-139
__device__ __forceinline__ double payoff(const PdeParams& p, double x) {
    return p.call ? fmax(0.0, exp(x) - 1.) : fmax(0.0, 1. - exp(x));
}

// Row j of B = I - theta dt A (src/Math/kwFd1d.cpp:73-114) from x_{j-1}, x_j, x_{j+1}.
// Rows j >= N are identity padding rows.
__device__ __forceinline__ void b_row(const PdeParams& p, double th, int j, double xm, double x0,
                                      double xp, double& bl, double& b, double& bu) {
    const int N = p.x_dim;
    const double dt = p.dt, a0 = p.a0, ax = p.ax, axx = p.axx;
    if (j >= N) {
        bl = 0.0;
        b = 1.0;
        bu = 0.0;
    } else if (j == 0) {
        const double inv_dx = 1. / (xp - x0);
        bl = 0;
        b = 1 - th * dt * (a0 - inv_dx * ax);
        bu = -th * dt * (inv_dx * ax);
    } else if (j == N - 1) {
        const double inv_dx = 1. / (x0 - xm);
        bl = -th * dt * (-inv_dx * ax);
        b = 1 - th * dt * (a0 + inv_dx * ax);
        bu = 0;
    } else {
        const double inv_dxu = 1. / (xp - x0);
        const double inv_dxm = 1. / (xp - xm);
        const double inv_dxd = 1. / (x0 - xm);
        const double inv_dx2u = 2. * inv_dxu * inv_dxm;
        const double inv_dx2m = 2. * inv_dxd * inv_dxu;
        const double inv_dx2l = 2. * inv_dxd * inv_dxm;
        bl = -th * dt * (-inv_dxm * ax + inv_dx2l * axx);
        b = 1 - th * dt * (a0 - inv_dx2m * axx);
        bu = -th * dt * (inv_dxm * ax + inv_dx2u * axx);
    }
}

// Fd1d::value (src/Math/kwFd1d.cpp:139-158): index of the first node with x_j >= x*,
// found from the inverse of the sinh mesh and corrected against the exact nodes.
// Returns 0 or N when x* is outside (x_0, x_{N-1}].
__device__ __forceinline__ int bracket(const PdeParams& p, double density, double xs) {
    const int N = p.x_dim;
    const double u = (asinh(xs / density) - p.yMin) / (p.yMax - p.yMin);
    const double jf = ceil(u * (double)(N - 1));
    int j = 0;
    if (jf >= (double)N)
        j = N;
    else if (jf > 0.0)
        j = (int)jf;
    while (j > 0 && grid_x(p, density, j - 1) >= xs) --j;
    while (j < N && grid_x(p, density, j) < xs) ++j;
    return j;
}
// BlackScholes_Pricer::priceOne, src/Pricer/kwBlackScholes.cpp:27-50:
// NaN for American options; cdfNormal(x) = (1 + erf(x / sqrt 2)) / 2 (src/Math/kwMath.cpp:9-13).
__device__ __forceinline__ double cdf_normal(double x) {
    return 0.5 * (1 + erf(x / 1.4142135623730951));
}

__device__ __forceinline__ double bs_one(const kwgpu_option& a) {
    if (a.e) return NAN;
    const double k = a.k, q = a.q, r = a.r, s = a.s, t = a.t, z = a.z;
    const double w = (double)a.w;
    const double zt = z * sqrt(t);
    const double d1 = 1 / zt * (log(s / k) + (r - q + 0.5 * z * z) * t);
    const double d2 = d1 - zt;
    return w * (s * cdf_normal(w * d1) * exp(-q * t) - k * cdf_normal(w * d2) * exp(-r * t));
}
#pragma clang fp contract(on)

}  // namespace kwgpu

#include <string>

// Kernel launchers (defined in the .hip files, called by the C-ABI layer).
namespace kwgpu {
// Context internals shared by the C-ABI translation units (kwgpu_api.hip).
int ctx_fail(kwgpu_ctx* c, int code, const std::string& msg);
hipStream_t ctx_stream(kwgpu_ctx* c);
// kwgpu_fd1d_price_device with an optional per-option precision (host array of n).
int fd_price_device_prec(kwgpu_ctx* c, const kwgpu_option* d_opts, size_t n,
                         const kwgpu_fd1d_params* p, double* d_prices, int32_t* d_status,
                         void* stream, const uint8_t* host_prec);
int ctx_device(kwgpu_ctx* c);

// Wave layout, FP64 / FP32-mixed. Returns hipErrorInvalidValue when x_dim does not fit.
// queue: null for one workgroup per PDE group, or a zeroed device counter for a
// persistent grid that takes groups from it.
hipError_t launch_fd1d_wave(const FdBatch& b, int max_x_dim, int precision, hipStream_t s,
                            int* queue = nullptr);
hipError_t launch_zero_counters(int* c, int n, hipStream_t s);
// Wave-layout kernel variants: M rows per lane, L lanes per option, W waves per option;
// one option per segment holds up to M * L * W grid points.
struct WaveVariant {
    int M, L, W;
    int capacity() const { return M * L * W; }
};
int wave_variant_count();
WaveVariant wave_variant(int v);
int default_wave_variant(int max_x_dim);  // -1 when the grid does not fit
hipError_t launch_fd1d_wave_variant(const FdBatch& b, int variant, int precision, hipStream_t s,
                                    int* queue = nullptr);
bool wave_layout_fits(int x_dim, int precision);
// Lane layout (any grid size).
hipError_t launch_fd1d_lane(const FdBatch& b, hipStream_t s);
size_t lane_workspace_doubles(int m, int max_x_dim);
hipError_t launch_bs(const kwgpu_option* opts, size_t n, double* prices, hipStream_t s);
hipError_t launch_fd1d_bs_combine(const kwgpu_option* opts, size_t n, double* amer,
                                  const double* euro, hipStream_t s);
}  // namespace kwgpu
