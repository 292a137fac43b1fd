// fd1d_lane.hip -- lane layout of the batched Crank-Nicolson American FD solver.
//
// One PDE per lane; the factored operator and the state live in HBM as
// option-fastest structure-of-arrays (element (j, slot) at [j * stride + slot]), so a
// wavefront's 64 lanes touch 64 consecutive doubles per row. Each lane runs the
// reference's serial Thomas recurrences over all rows (src/Math/kwMath.cpp:16-49)
// with the same time-invariant factorisation and step identity as the wave layout
// (fd1d_wave.hip). Handles any grid size; its HBM traffic per point and step
// (read ap, bt, V; write f; read f, gm, V, payoff; write V) makes it bandwidth-bound.
#include <math.h>

#include "kwgpu_internal.h"

namespace kwgpu {
namespace {

enum { kAp = 0, kBt, kGm, kPay, kV, kF, kArrays };

__global__ __launch_bounds__(256) void fd1d_lane_f64(FdBatch b, int rows) {
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= b.m) return;
    const int p = b.order ? (int)b.order[slot] : slot;
    const uint32_t oi = b.rep ? b.rep[p] : (uint32_t)p;
    const kwgpu_option opt = b.opts[oi];
    const int xd = b.x_dims ? b.x_dims[oi] : b.x_dim;
    const int td = b.t_dims ? b.t_dims[oi] : b.t_dim;
    const PdeParams P = pde_params(opt, xd, td, b.density, b.scale);
    const bool solve = P.valid && xd <= rows;
    const long long S = b.ws_stride;
    double* ws = b.ws;
    auto at = [&](int arr, int j) -> double& {
        return ws[((long long)arr * rows + j) * S + slot];
    };
    const int N = solve ? xd : 0;
    const double th = b.theta, inv_th = 1.0 / th, c_th = (1.0 - th) / th;

    // set-up: rolling window over the sinh mesh, Thomas pivots
    {
        double xm = 0.0, x0 = N > 0 ? grid_x(P, b.density, 0) : 0.0;
        double xp = N > 1 ? grid_x(P, b.density, 1) : 0.0;
        double dprev = 1.0, bup = 0.0;
        for (int j = 0; j < N; ++j) {
            double bl, bb, bu;
            b_row(P, th, j, xm, x0, xp, bl, bb, bu);
            const double gam = bup / dprev;
            const double d = bb - bl * gam;
            at(kAp, j) = inv_th / d;
            at(kBt, j) = bl / d;
            at(kGm, j) = bu / d;
            const double pay = payoff(P, x0);
            at(kV, j) = pay;
            at(kPay, j) = (P.american && j < N - 1) ? pay : -INFINITY;
            dprev = d;
            bup = bu;
            xm = x0;
            x0 = xp;
            xp = j + 2 < N ? grid_x(P, b.density, j + 2) : 0.0;
        }
    }
    const int steps = solve ? td - 1 : 0;
    for (int step = 0; step < steps; ++step) {
        double f = 0.0;
        for (int j = 0; j < N; ++j) {
            f = fma(-at(kBt, j), f, at(kAp, j) * at(kV, j));
            at(kF, j) = f;
        }
        double z = 0.0;
        for (int j = N - 1; j >= 0; --j) {
            z = fma(-at(kGm, j), z, at(kF, j));
            const double v = at(kV, j);
            at(kV, j) = fmax(fma(-c_th, v, z), at(kPay, j));
        }
    }

    // a NaN or inf anywhere in the march ends non-finite in the unprojected last row
    // (maxNum hides it in the projected rows): report NaN, as the reference's march would
    bool bad = false;
    for (int j = 0; j < (solve ? N : 0); ++j) bad |= !isfinite(at(kV, j));

    const uint32_t beg = b.off ? b.off[p] : 0u;
    const uint32_t end = b.off ? b.off[p + 1] : 1u;
    for (uint32_t k = beg; k < end; ++k) {
        const uint32_t o2 = b.off ? b.list[k] : oi;
        const kwgpu_option q = b.off ? b.opts[o2] : opt;
        double price = NAN;
        int st = KWGPU_STATUS_BAD_INPUT;
        if (solve && finite_pos(q.s) && finite_pos(q.k)) {
            const double xs = log(q.s / q.k);
            const int j = bracket(P, b.density, xs);
            if (j == 0 || j == N) {
                st = KWGPU_STATUS_OUT_OF_GRID;
            } else if (bad) {
                st = KWGPU_STATUS_NONFINITE;
            } else {
                const double x1 = grid_x(P, b.density, j), x0 = grid_x(P, b.density, j - 1);
                const double v = ((x1 - xs) * at(kV, j - 1) + (xs - x0) * at(kV, j)) / (x1 - x0);
                price = q.k * v;
                st = isfinite(price) ? KWGPU_STATUS_OK : KWGPU_STATUS_NONFINITE;
            }
        }
        b.prices[o2] = price;
        if (b.status) b.status[o2] = st;
    }
}

}  // namespace

size_t lane_workspace_doubles(int m, int max_x_dim) {
    return (size_t)kArrays * (size_t)max_x_dim * (size_t)m;
}

hipError_t launch_fd1d_lane(const FdBatch& b, hipStream_t s) {
    if (b.m <= 0) return hipSuccess;
    if (!b.ws || b.ws_stride < b.m) return hipErrorInvalidValue;
    // rows per array = workspace size / (arrays * stride), fixed by the caller's sizing
    const int rows = (int)(b.ws_rows);
    const int blocks = (b.m + 255) / 256;
    hipLaunchKernelGGL(fd1d_lane_f64, dim3(blocks), dim3(256), 0, s, b, rows);
    return hipGetLastError();
}

}  // namespace kwgpu
