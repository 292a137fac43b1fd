// fd1d_wave.hip -- wavefront layout of the batched Crank-Nicolson American FD solver.
//
// Replaces Fd1d::solve / Fd1d::solveOne / kw::solveTridiagonal / Fd1d::value
// (src/Math/kwFd1d.cpp:11-158, src/Math/kwMath.cpp:16-49) for a batch of PDEs.
//
// Layout: one PDE per L-lane segment of a wave64 (L = 16, 32 or 64, so 4, 2 or 1
// options per wave); lane l of a segment owns grid rows [l*M, l*M + M). Padding rows
// beyond x_dim are identity rows and stay zero.
//
// The operator is time-invariant (constant coefficients, uniform t-grid), so B = I -
// theta dt A is LU-factored once in FP64. Since E = I + (1-theta) dt A =
// (1/theta) I - ((1-theta)/theta) B, one step of the reference scheme  B V' = E V  is
//
//     V' = (1/theta) B^{-1} V - ((1-theta)/theta) V        (= 2 B^{-1} V - V at theta 1/2)
//
// followed by the reference's post-solve projection V' = max(V', payoff) on nodes
// 0..N-2 (src/Math/kwFd1d.cpp:127-132). B^{-1} V is the reference's Thomas algorithm:
// a forward recurrence f_j = (V_j - bl_j f_{j-1}) / d_j and a backward one
// z_j = f_j - (bu_j / d_j) z_{j+1}. Each lane runs both recurrences over its own rows
// with zero carry-in; the carries between lanes are first-order affine recurrences
// over lanes, solved as Kogge-Stone scans with DPP moves (row_shr/row_bcast forward,
// row_shl and readlane backward) whose multipliers depend only on B and so are
// precomputed; per-row fix-up coefficients fold the carries back in.
//
// Per point and step that is 7 instructions (2 forward, 1 backward, 3 fix-up and
// correction, 1 projection) plus the scans amortised over M rows. The state, the
// factors and the forward fix-up live in VGPRs; the backward fix-up and the payoff
// live in LDS (one ds_read_b128 per row and step) so that a wave stays within 256
// registers and two waves share each SIMD.
#include <math.h>

#include "kwgpu_internal.h"

namespace kwgpu {
namespace {

// 64-bit DPP move built from two 32-bit v_mov_b32_dpp. Lanes whose source is outside
// the row/wave read 0 (bound_ctrl); lanes disabled by ROWMASK read 0.
template <int CTRL, int ROWMASK = 0xF>
__device__ __forceinline__ double dpp(double v) {
    int2 a = __builtin_bit_cast(int2, v);
    a.x = __builtin_amdgcn_update_dpp(0, a.x, CTRL, ROWMASK, 0xF, true);
    a.y = __builtin_amdgcn_update_dpp(0, a.y, CTRL, ROWMASK, 0xF, true);
    return __builtin_bit_cast(double, a);
}
template <int LANE>
__device__ __forceinline__ double readlane(double v) {
    int2 a = __builtin_bit_cast(int2, v);
    a.x = __builtin_amdgcn_readlane(a.x, LANE);
    a.y = __builtin_amdgcn_readlane(a.y, LANE);
    return __builtin_bit_cast(double, a);
}

constexpr int kRowShr = 0x110;   // + n, n in 1..15
constexpr int kRowShl = 0x100;   // + n
constexpr int kRowBcast15 = 0x142;
constexpr int kRowBcast31 = 0x143;
constexpr int kWaveShr1 = 0x138;
constexpr int kWaveShl1 = 0x130;

// max(a, b) as one v_max_f64: fmax() would also re-canonicalise its operands every
// step (LLVM maxnum semantics), doubling the projection's cost.
__device__ __forceinline__ double vmax(double a, double b) {
    double r;
    asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

// Product of v[from..to] (inclusive); 1 when from > to.
__device__ __forceinline__ double lds_prod(const double* v, int from, int to) {
    double p = 1.0;
    for (int k = from; k <= to; ++k) p *= v[k];
    return p;
}

template <int M, int L>
__global__ __attribute__((amdgpu_flat_work_group_size(64, 64), amdgpu_waves_per_eu(2, 2)))
void fd1d_wave_f64(FdBatch b) {
    static_assert(L == 16 || L == 32 || L == 64, "segment width");
    static_assert(M >= 4, "rows per lane");
    constexpr int G = 64 / L;
    // One LDS region per wave, used three ways in turn:
    //   set-up scratch (256 doubles), then during the march the pairs
    //   (Q_i, payoff_i) of every lane at qp[(i * 64 + lane) * 2], then the solution
    //   for the interpolation at qp[seg * L * M + j].
    __shared__ double qp[2 * 64 * M];

    const int lane = threadIdx.x;
    const int seg = lane / L;
    const int ls = lane - seg * L;
    const int row16 = lane >> 4;
    const int pr = lane & 15;
    const int slot = blockIdx.x * G + seg;
    const bool live = slot < b.m;
    const int p = live ? (b.order ? (int)b.order[slot] : slot) : 0;
    const uint32_t oi = live ? (b.rep ? b.rep[p] : (uint32_t)p) : 0u;
    const kwgpu_option opt = b.opts[oi];
    const int xd = b.x_dims ? b.x_dims[oi] : b.x_dim;
    const int td = b.t_dims ? b.t_dims[oi] : b.t_dim;
    const PdeParams P = pde_params(opt, xd, td, b.density, b.scale);
    const bool solve = live && P.valid && xd <= L * M;
    const int N = solve ? xd : 0;  // N == 0: every row is an identity row
    const int steps = solve ? td - 1 : 0;
    const double th = b.theta;
    const double inv_th = 1.0 / th;
    const double c_th = (1.0 - th) / th;
    const int j0 = ls * M;

    // ---- set-up: grid, operator rows, payoff, LU factors (FP64) --------------------
    double xv[M], ap[M], bt[M], gm[M], R[M];
    double Q[M];  // only during set-up; the march reads it from LDS
    {
        double xg[M + 3];  // x_{j0-2} .. x_{j0+M}
#pragma unroll
        for (int i = 0; i < M + 3; ++i) {
            const int j = j0 - 2 + i;
            xg[i] = (j >= 0 && j < N) ? grid_x(P, b.density, j) : 0.0;
        }
        double blv[M], bv[M], buv[M];
        double bu_prev = 0.0;  // bu of row j0-1
        if (j0 >= 1 && j0 - 1 < N) {
            double t0, t1;
            b_row(P, th, j0 - 1, xg[0], xg[1], xg[2], t0, t1, bu_prev);
        }
#pragma unroll
        for (int i = 0; i < M; ++i) {
            const int j = j0 + i;  // x_{j-1}, x_j, x_{j+1} = xg[i+1], xg[i+2], xg[i+3]
            b_row(P, th, j, xg[i + 1], xg[i + 2], xg[i + 3], blv[i], bv[i], buv[i]);
            xv[i] = j < N ? payoff(P, xg[i + 2]) : 0.0;
        }

        // LU pivots d_j = b_j - bl_j bu_{j-1} / d_{j-1} are a serial recurrence; its
        // value entering each lane comes from a prefix product of the lanes' projective
        // maps (u, v) -> (b u - bl bu_prev v, u), normalised to avoid overflow.
        double m00 = 1, m01 = 0, m10 = 0, m11 = 1;
        {
            double bup = bu_prev;
#pragma unroll
            for (int i = 0; i < M; ++i) {
                const double e = blv[i] * bup;
                const double n00 = bv[i] * m00 - e * m10, n01 = bv[i] * m01 - e * m11;
                m10 = m00;
                m11 = m01;
                m00 = n00;
                m01 = n01;
                const double s = fmax(fmax(fabs(m00), fabs(m01)), fmax(fabs(m10), fabs(m11)));
                const double is = s > 0.0 ? 1.0 / s : 1.0;
                m00 *= is; m01 *= is; m10 *= is; m11 *= is;
                bup = buv[i];
            }
        }
        qp[4 * lane + 0] = m00;
        qp[4 * lane + 1] = m01;
        qp[4 * lane + 2] = m10;
        qp[4 * lane + 3] = m11;
        __syncthreads();
        double s00 = 1, s01 = 0, s10 = 0, s11 = 1;
        for (int k = seg * L; k < lane; ++k) {
            const double a00 = qp[4 * k], a01 = qp[4 * k + 1], a10 = qp[4 * k + 2],
                         a11 = qp[4 * k + 3];
            const double n00 = a00 * s00 + a01 * s10, n01 = a00 * s01 + a01 * s11;
            const double n10 = a10 * s00 + a11 * s10, n11 = a10 * s01 + a11 * s11;
            const double s = fmax(fmax(fabs(n00), fabs(n01)), fmax(fabs(n10), fabs(n11)));
            const double is = s > 0.0 ? 1.0 / s : 1.0;
            s00 = n00 * is; s01 = n01 * is; s10 = n10 * is; s11 = n11 * is;
        }
        double dprev = ls == 0 ? 1.0 : (s00 + s01) / (s10 + s11);

        // Thomas pivots of the reference (gam_j = bu_{j-1}/bet, bet = b_j - bl_j gam_j)
        double bup = bu_prev;
#pragma unroll
        for (int i = 0; i < M; ++i) {
            const double gam = bup / dprev;
            const double d = bv[i] - blv[i] * gam;
            ap[i] = inv_th / d;
            bt[i] = blv[i] / d;
            gm[i] = buv[i] / d;
            dprev = d;
            bup = buv[i];
        }
    }

    // Carry propagation coefficients: f_j = fhat_j + Pc_j f_in,
    // z_j = zhat_j + R_j f_in + Q_j z_in.
    double Pl, Ql;
    {
        double Pc[M];
        Pc[0] = -bt[0];
#pragma unroll
        for (int i = 1; i < M; ++i) Pc[i] = -bt[i] * Pc[i - 1];
        R[M - 1] = Pc[M - 1];
#pragma unroll
        for (int i = M - 2; i >= 0; --i) R[i] = Pc[i] - gm[i] * R[i + 1];
        Q[M - 1] = -gm[M - 1];
#pragma unroll
        for (int i = M - 2; i >= 0; --i) Q[i] = -gm[i] * Q[i + 1];
        Pl = Pc[M - 1];
        Ql = Q[0];
    }
    __syncthreads();
    qp[lane] = Pl;
    qp[64 + lane] = Ql;
    __syncthreads();
    // Forward scan multipliers (window products of Pl over lanes).
    const double F1 = pr >= 1 ? lds_prod(qp, lane, lane) : 0.0;
    const double F2 = pr >= 2 ? lds_prod(qp, lane - 1, lane) : 0.0;
    const double F4 = pr >= 4 ? lds_prod(qp, lane - 3, lane) : 0.0;
    const double F8 = pr >= 8 ? lds_prod(qp, lane - 7, lane) : 0.0;
    const double FB15 = (L >= 32 && (row16 & 1)) ? lds_prod(qp, 16 * row16, lane) : 0.0;
    const double FB31 = (L == 64 && row16 >= 2) ? lds_prod(qp, 32, lane) : 0.0;
    // Backward scan multipliers (window products of Ql).
    const double* lq = qp + 64;
    const double G1 = pr + 1 <= 15 ? lds_prod(lq, lane, lane) : 0.0;
    const double G2 = pr + 2 <= 15 ? lds_prod(lq, lane, lane + 1) : 0.0;
    const double G4 = pr + 4 <= 15 ? lds_prod(lq, lane, lane + 3) : 0.0;
    const double G8 = pr + 8 <= 15 ? lds_prod(lq, lane, lane + 7) : 0.0;
    const double XB = lds_prod(lq, lane, 16 * row16 + 15);
    // Segment edges: selects, not multiplies by 0, so one segment's values never
    // reach another's.
    const bool first_lane = ls == 0;
    const bool last_lane = ls == L - 1;

    // All segments of the wave iterate to the wave's longest march; lanes whose own
    // march is done keep their values.
    int wsteps = steps;
    if (L < 64) {
        __syncthreads();
        qp[lane] = (double)steps;
        __syncthreads();
        double mx = 0.0;
        for (int k = 0; k < 64; k += L) mx = fmax(mx, qp[k]);
        wsteps = (int)mx;
    }
    __syncthreads();
    double2* qpv = reinterpret_cast<double2*>(qp);
#pragma unroll
    for (int i = 0; i < M; ++i) {
        const int j = j0 + i;
        qpv[i * 64 + lane] = make_double2(Q[i], (P.american && j < N - 1) ? xv[i] : -INFINITY);
    }
    __syncthreads();

    // ---- march: tDim-1 backward steps (src/Math/kwFd1d.cpp:67-133) ----------------
    for (int step = 0; step < wsteps; ++step) {
        double f[M];
        f[0] = ap[0] * xv[0];
#pragma unroll
        for (int i = 1; i < M; ++i) f[i] = fma(-bt[i], f[i - 1], ap[i] * xv[i]);

        double c = f[M - 1];
        c = fma(F1, dpp<kRowShr + 1>(c), c);
        c = fma(F2, dpp<kRowShr + 2>(c), c);
        c = fma(F4, dpp<kRowShr + 4>(c), c);
        c = fma(F8, dpp<kRowShr + 8>(c), c);
        if constexpr (L >= 32) c = fma(FB15, dpp<kRowBcast15, 0xA>(c), c);
        if constexpr (L == 64) c = fma(FB31, dpp<kRowBcast31, 0xC>(c), c);
        double fin = dpp<kWaveShr1>(c);
        if constexpr (L < 64) fin = first_lane ? 0.0 : fin;

#pragma unroll
        for (int i = M - 2; i >= 0; --i) f[i] = fma(-gm[i], f[i + 1], f[i]);

        double cb = fma(R[0], fin, f[0]);
        cb = fma(G1, dpp<kRowShl + 1>(cb), cb);
        cb = fma(G2, dpp<kRowShl + 2>(cb), cb);
        cb = fma(G4, dpp<kRowShl + 4>(cb), cb);
        cb = fma(G8, dpp<kRowShl + 8>(cb), cb);
        if constexpr (L == 64) {
            double t = fma(XB, readlane<48>(cb), cb);
            cb = row16 == 2 ? t : cb;
            t = fma(XB, readlane<32>(cb), cb);
            cb = row16 == 1 ? t : cb;
            t = fma(XB, readlane<16>(cb), cb);
            cb = row16 == 0 ? t : cb;
        } else if constexpr (L == 32) {
            const double t0 = fma(XB, readlane<16>(cb), cb);
            const double t2 = fma(XB, readlane<48>(cb), cb);
            cb = row16 == 0 ? t0 : (row16 == 2 ? t2 : cb);
        }
        double zin = dpp<kWaveShl1>(cb);
        if constexpr (L < 64) zin = last_lane ? 0.0 : zin;

        if (step < steps) {
#pragma unroll
            for (int i = 0; i < M; ++i) {
                const double2 qv = qpv[i * 64 + lane];
                double t = fma(R[i], fin, f[i]);
                t = fma(qv.x, zin, t);
                xv[i] = vmax(fma(-c_th, xv[i], t), qv.y);
            }
        }
    }

    // ---- Fd1d::value: interpolation at log(S/K), price = K v -----------------------
    __syncthreads();
#pragma unroll
    for (int i = 0; i < M; ++i) qp[seg * L * M + j0 + i] = xv[i];
    __syncthreads();
    if (live) {
        const double* V = qp + seg * L * M;
        const uint32_t beg = b.off ? b.off[p] : 0u;
        const uint32_t end = b.off ? b.off[p + 1] : 1u;
        for (uint32_t k = beg + ls; k < end; k += L) {
            const uint32_t o2 = b.off ? b.list[k] : oi;
            const kwgpu_option q = b.off ? b.opts[o2] : opt;
            double price = NAN;
            int st = KWGPU_STATUS_BAD_INPUT;
            if (solve && finite_pos(q.s) && finite_pos(q.k)) {
                const double xs = log(q.s / q.k);
                const int j = bracket(P, b.density, xs);
                if (j == 0 || j == N) {
                    st = KWGPU_STATUS_OUT_OF_GRID;
                } else {
                    const double x1 = grid_x(P, b.density, j), x0 = grid_x(P, b.density, j - 1);
                    const double v = ((x1 - xs) * V[j - 1] + (xs - x0) * V[j]) / (x1 - x0);
                    price = q.k * v;
                    st = isfinite(price) ? KWGPU_STATUS_OK : KWGPU_STATUS_NONFINITE;
                }
            }
            b.prices[o2] = price;
            if (b.status) b.status[o2] = st;
        }
    }
}

template <int M, int L>
hipError_t launch_wave(const FdBatch& b, hipStream_t s) {
    constexpr int G = 64 / L;
    const int blocks = (b.m + G - 1) / G;
    hipLaunchKernelGGL((fd1d_wave_f64<M, L>), dim3(blocks), dim3(64), 0, s, b);
    return hipGetLastError();
}

}  // namespace

bool wave_layout_fits(int x_dim, int precision) {
    (void)precision;
    return x_dim >= 3 && x_dim <= 1024;
}

hipError_t launch_fd1d_wave(const FdBatch& b, int max_x_dim, int precision, hipStream_t s) {
    if (b.m <= 0) return hipSuccess;
    if (!wave_layout_fits(max_x_dim, precision)) return hipErrorInvalidValue;
    if (max_x_dim <= 64) return launch_wave<4, 16>(b, s);
    if (max_x_dim <= 128) return launch_wave<8, 16>(b, s);
    if (max_x_dim <= 256) return launch_wave<16, 16>(b, s);
    if (max_x_dim <= 512) return launch_wave<16, 32>(b, s);
    return launch_wave<16, 64>(b, s);
}

}  // namespace kwgpu
