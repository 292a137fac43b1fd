// fd1d_wave.hip -- wavefront layout of the batched Crank-Nicolson American FD solver.
//
// Replaces Fd1d::solve / Fd1d::solveOne / kw::solveTridiagonal / Fd1d::value
// (src/Math/kwFd1d.cpp:11-158, src/Math/kwMath.cpp:16-49) for a batch of PDEs.
//
// Layout: one PDE per L-lane segment of a wave64 (L = 16, 32 or 64, so 4, 2 or 1
// options per wave), or per workgroup of W = 2 whole waves for 2048-point grids; lane
// l of a segment owns grid rows [l*M, l*M + M). Padding rows beyond x_dim are identity
// rows and stay zero.
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
// The factors are kept in LDU form: unit-lower l_j = bl_j / d_{j-1}, pivots d_j and
// the scaled upper u_j = bu_j / d_{j+1}. With g = L^{-1} V (g_j = V_j - l_j g_{j-1}) and
// w_j = g_j - u_j w_{j+1}, B^{-1} V = D^{-1} w, so with a_j = 1 / (theta d_j)
//     V'_j = a_j w_j - ((1-theta)/theta) V_j              (= a_j w_j - V_j at theta 1/2).
//
// FP64 march: 6 instructions per point and step at theta 1/2 (1 forward, 1 backward,
// 2 carry fix-ups, 1 correction, 1 projection) plus the scans amortised over M rows.
// The state, a, l, u and the forward fix-up live in VGPRs; the backward fix-up and the
// payoff live in LDS (one ds_read_b128 per row and step) so that a wave stays within
// 256 registers and two waves share each SIMD.
//
// FP32-mixed march: the same factors applied to the step's increment,
//     V' - V = (1/theta) B^{-1} (I - B) V = a (L D U)... = a_j w_j with g, w taken of (I - B) V,
// with the residual (I - B) V assembled in FP64 from FP64 differences of V (so its
// rounding error is relative to the increment, not to V), the substitutions, carries
// and fix-ups in FP32 on FP32 copies of the factors, and V accumulated, projected and
// interpolated in FP64.
#include <math.h>

#include <algorithm>

#include "kwgpu_internal.h"

namespace kwgpu {
namespace {

#ifndef KWGPU_PF1
#define KWGPU_PF1 3  // one option per wave
#endif
#ifndef KWGPU_PF2
#define KWGPU_PF2 0  // two-wave groups: read at the point of use (any buffer costs spills)
#endif
#ifndef KWGPU_PFG
#define KWGPU_PFG 0  // several options per wave
#endif

constexpr int kRowShr = 0x110;   // + n, n in 1..15
constexpr int kRowShl = 0x100;   // + n
constexpr int kRowBcast15 = 0x142;
constexpr int kRowBcast31 = 0x143;
constexpr int kWaveShr1 = 0x138;
constexpr int kWaveShl1 = 0x130;

// DPP move of a float or a double (two 32-bit v_mov_b32_dpp). Lanes whose source is
// outside the row/wave read 0 (bound_ctrl); lanes in rows disabled by ROWMASK keep
// `old`, so a caller that multiplies the result by 0 there can pass any finite value
// already in a register and save the zeroing moves.
template <int CTRL, int ROWMASK = 0xF, class T>
__device__ __forceinline__ T dpp(T v, T old = T(0)) {
    if constexpr (sizeof(T) == 8) {
        int2 a = __builtin_bit_cast(int2, v);
        const int2 o = __builtin_bit_cast(int2, old);
        a.x = __builtin_amdgcn_update_dpp(o.x, a.x, CTRL, ROWMASK, 0xF, true);
        a.y = __builtin_amdgcn_update_dpp(o.y, a.y, CTRL, ROWMASK, 0xF, true);
        return __builtin_bit_cast(T, a);
    } else {
        return __builtin_bit_cast(
            T, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, old), __builtin_bit_cast(int, v),
                                           CTRL, ROWMASK, 0xF, true));
    }
}
template <int LANE, class T>
__device__ __forceinline__ T readlane(T v) {
    if constexpr (sizeof(T) == 8) {
        int2 a = __builtin_bit_cast(int2, v);
        a.x = __builtin_amdgcn_readlane(a.x, LANE);
        a.y = __builtin_amdgcn_readlane(a.y, LANE);
        return __builtin_bit_cast(T, a);
    } else {
        return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), LANE));
    }
}

// max(a, b) as one v_max_f64: fmax() would also re-canonicalise its operands every
// step (LLVM maxnum semantics), doubling the projection's cost.
__device__ __forceinline__ double vmax(double a, double b) {
    double r;
    asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

template <int M, int L, int W>
struct Geo {
    static_assert(L == 16 || L == 32 || L == 64, "segment width");
    static_assert(W == 1 || (W == 2 && L == 64), "two-wave groups use whole waves");
    static_assert(M >= 4, "rows per lane");
    static constexpr int NT = 64 * W;           // threads per workgroup
    static constexpr int LT = W == 1 ? L : NT;  // lanes per option
    static constexpr int G = NT / LT;           // options per workgroup
};

struct Lanes {
    int gl, wv, lane, seg, ls, row16, wrow, pr;
};

template <class T>
struct ScanMul {
    T F1, F2, F4, F8, FB15, FB31;  // forward, window products of the lanes' P
    T G1, G2, G4, G8;              // backward, window products of the lanes' Q
    T X16, X32, X48;               // backward across rows: L = 64, nonzero only in the
                                   // row that takes lane 16 / 32 / 48; L = 32, X16 is
                                   // nonzero in rows 0 and 2 (each segment's first row)
    T FW, ZW;                      // two-wave groups: carries between the waves
    // two-wave groups, FP64 march with one barrier per step (both waves scan locally,
    // then exchange one value u): gin += AW u, win += BW u, where u is wave 0's last
    // forward carry (in wave 1) or wave 1's first backward carry (in wave 0, formed as
    // z1_local + TW0 c0). AW and BW are per lane and kept in LDS; TW0 is uniform.
    T TW0;
};

// Per-PDE state after set-up.
struct Pde {
    PdeParams P;
    int p;          // PDE index
    uint32_t oi;    // representative option
    kwgpu_option opt;
    bool live, solve;
    int N, steps, j0;
};

template <int M, int L, int W>
__device__ __forceinline__ Lanes lanes_of(int gl) {
    using Gm = Geo<M, L, W>;
    Lanes t;
    t.gl = gl;
    t.wv = gl >> 6;
    t.lane = gl & 63;
    t.seg = gl / Gm::LT;
    t.ls = gl - t.seg * Gm::LT;
    t.row16 = gl >> 4;
    t.wrow = t.lane >> 4;
    t.pr = gl & 15;
    return t;
}

template <int M, int L, int W>
__device__ __forceinline__ Pde pde_of(const FdBatch& b, int group, const Lanes& t) {
    using Gm = Geo<M, L, W>;
    Pde d;
    const int slot = group * Gm::G + t.seg;
    d.live = slot < b.m;
    d.p = d.live ? (b.order ? (int)b.order[slot] : slot) : 0;
    d.oi = d.live ? (b.rep ? b.rep[d.p] : (uint32_t)d.p) : 0u;
    d.opt = b.opts[d.oi];
    const int xd = b.x_dims ? b.x_dims[d.oi] : b.x_dim;
    const int td = b.t_dims ? b.t_dims[d.oi] : b.t_dim;
    d.P = pde_params(d.opt, xd, td, b.density, b.scale);
    d.solve = d.live && d.P.valid && xd <= Gm::LT * M;
    d.N = d.solve ? xd : 0;  // N == 0: every row is an identity row
    d.steps = d.solve ? td - 1 : 0;
    d.j0 = t.ls * M;
    return d;
}

// ---- set-up: grid, operator rows, payoff, LU factors, carry coefficients (FP64) ----
// Reciprocal from v_rcp_f64 and two Newton steps (within an ulp or two of 1/x for the
// normal, finite x used here).
__device__ __forceinline__ double rcp(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    return fma(r, fma(-x, r, 1.0), r);
}

// The sinh mesh x_j = density sinh(y_j), y_j = y_min (1 - j dy) + y_max j dy
// (src/Pricer/kwFd1d.cpp:105-125), for K consecutive nodes from j: sinh as
// (e^y - e^-y) / 2 with e^y_j a geometric progression (one exp for the first node, one
// for the ratio). Its error, a few 1e-16 relative in e^y, moves no price by more than
// ~1e-13.
template <int K>
__device__ __forceinline__ void mesh_run(const PdeParams& p, double density, int j,
                                         double (&x)[K]) {
#pragma clang fp contract(off)
    const double yj = j * p.dy;
    double e = exp(p.yMin * (1.0 - yj) + p.yMax * yj);
    double f = rcp(e);
    const double step = (p.yMax - p.yMin) * p.dy;
    const double up = exp(step), down = rcp(up);
#pragma unroll
    for (int i = 0; i < K; ++i) {
        x[i] = density * (0.5 * (e - f));
        e *= up;
        f *= down;
    }
}

// Row j of B (src/Math/kwFd1d.cpp:73-114) from the reciprocals of x_j - x_{j-1} (rl),
// x_{j+1} - x_j (ru) and x_{j+1} - x_{j-1} (rm); rows j >= N are identity padding.
__device__ __forceinline__ void b_row_rcp(const PdeParams& p, double th, int j, double rl,
                                          double ru, double rm, double& bl, double& b,
                                          double& bu) {
#pragma clang fp contract(off)
    const int N = p.x_dim;
    const double dt = p.dt, a0 = p.a0, ax = p.ax, axx = p.axx;
    if (j >= N) {
        bl = 0.0;
        b = 1.0;
        bu = 0.0;
    } else if (j == 0) {
        bl = 0;
        b = 1 - th * dt * (a0 - ru * ax);
        bu = -th * dt * (ru * ax);
    } else if (j == N - 1) {
        bl = -th * dt * (-rl * ax);
        b = 1 - th * dt * (a0 + rl * ax);
        bu = 0;
    } else {
        const double inv_dx2u = 2. * ru * rm;
        const double inv_dx2m = 2. * rl * ru;
        const double inv_dx2l = 2. * rl * rm;
        bl = -th * dt * (-rm * ax + inv_dx2l * axx);
        b = 1 - th * dt * (a0 - inv_dx2m * axx);
        bu = -th * dt * (rm * ax + inv_dx2u * axx);
    }
}

// 2x2 projective map [[a, b], [c, d]] of the pivot recurrence.
struct Map2 {
    double a, b, c, d;
};

// x . y, rescaled by a power of two (exact) so that its largest entry lies in [0.5, 1):
// the pivot a map yields, (a + b) / (c + d), does not depend on the scale.
__device__ __forceinline__ Map2 map_mul(const Map2& x, const Map2& y) {
    Map2 m{x.a * y.a + x.b * y.c, x.a * y.b + x.b * y.d, x.c * y.a + x.d * y.c,
           x.c * y.b + x.d * y.d};
    const double s = fmax(fmax(fabs(m.a), fabs(m.b)), fmax(fabs(m.c), fabs(m.d)));
    const int e = (s > 0.0 && s < INFINITY) ? -__builtin_amdgcn_frexp_exp(s) : 0;
    return {ldexp(m.a, e), ldexp(m.b, e), ldexp(m.c, e), ldexp(m.d, e)};
}

template <int CTRL, int ROWMASK = 0xF>
__device__ __forceinline__ Map2 dpp_map(const Map2& m) {
    return {dpp<CTRL, ROWMASK>(m.a), dpp<CTRL, ROWMASK>(m.b), dpp<CTRL, ROWMASK>(m.c),
            dpp<CTRL, ROWMASK>(m.d)};
}

__device__ __forceinline__ Map2 select_map(bool p, const Map2& x, const Map2& y) {
    return {p ? x.a : y.a, p ? x.b : y.b, p ? x.c : y.c, p ? x.d : y.d};
}

// qp: LDS scratch of at least 4 * NT doubles. On return xv holds the payoff, ap the
// scaled inverse pivots 1/(theta d), lo/up the unit-lower and scaled upper factors
// l_j = bl_j / d_{j-1}, u_j = bu_j / d_{j+1}, R/Q the carry fix-ups, blv/buv the
// off-diagonals of B, and sm the scan multipliers. Every cross-lane step below reads
// only lanes of the same option, or discards what it read with a select.
// TF is the type the march takes the factors, fix-ups and scan multipliers in: double
// for the FP64 march; float for the FP32-mixed one, which gets each as it is produced
// (so no FP64 copy of the factor arrays stays live beside the FP32 one).
template <int M, int L, int W, class TF>
__device__ __forceinline__ void setup(const FdBatch& b, const Pde& d, const Lanes& t, double* qp,
                                      double (&xv)[M], TF (&ap)[M], TF (&lo)[M],
                                      TF (&up)[M], TF (&R)[M], TF (&Q)[M],
                                      double (&blv)[M], double (&buv)[M], ScanMul<TF>& sm) {
    using Gm = Geo<M, L, W>;
    constexpr int LT = Gm::LT;
    const PdeParams& P = d.P;
    const int N = d.N, j0 = d.j0, gl = t.gl;
    const int pr = t.pr, wrow = t.wrow, wv = t.wv;
    const double th = b.theta;
    const double inv_th = 1.0 / th;
    double bv[M];
    double bu_prev = 0.0;  // bu of row j0-1
    {
        // nodes x_{j0-2} .. x_{j0+M} as a rolling window; nodes outside the mesh only
        // feed padding rows and get far-away values that keep every spacing nonzero
#pragma clang fp contract(off)
        const double yj = (j0 - 2) * P.dy;
        double e = exp(P.yMin * (1.0 - yj) + P.yMax * yj), f = rcp(e);
        const double grow = exp((P.yMax - P.yMin) * P.dy), shrink = rcp(grow);
        auto next_x = [&](int j) {
            const double x = (j >= 0 && j < N) ? b.density * (0.5 * (e - f))
                                               : (j < 0 ? -1e6 : 1e6) + j;
            e *= grow;
            f *= shrink;
            return x;
        };
        double xa = next_x(j0 - 2), xb = next_x(j0 - 1), xc = next_x(j0);
        double rl = rcp(xb - xa), ru = rcp(xc - xb);  // 1/(x_{j-1} - x_{j-2}), 1/(x_j - x_{j-1})
        double rm = rcp(xc - xa);
        if (j0 >= 1 && j0 - 1 < N) {
            double t0, t1;
            b_row_rcp(P, th, j0 - 1, rl, ru, rm, t0, t1, bu_prev);
        }
#pragma unroll
        for (int i = 0; i < M; ++i) {
            const int j = j0 + i;  // window (x_{j-1}, x_j, x_{j+1}) = (xb, xc, xd)
            const double xd = next_x(j + 1);
            rl = ru;
            ru = rcp(xd - xc);
            rm = rcp(xd - xb);
            b_row_rcp(P, th, j, rl, ru, rm, blv[i], bv[i], buv[i]);
            xv[i] = j < N ? payoff(P, xc) : 0.0;
            xb = xc;
            xc = xd;
        }
    }

    // LU pivots d_j = b_j - bl_j bu_{j-1} / d_{j-1} are a serial recurrence; its value
    // entering each lane is the projective image of the prefix product of the lanes'
    // maps (u, v) -> (b u - bl bu_prev v, u), taken by a Kogge-Stone scan over lanes.
    Map2 m{1, 0, 0, 1};
    {
        double bup = bu_prev;
#pragma unroll
        for (int i = 0; i < M; ++i) {
            const double e = blv[i] * bup;
            m = map_mul(Map2{bv[i], -e, 1.0, 0.0}, m);
            bup = buv[i];
        }
    }
    // inclusive prefix within the row, then across rows of the option
    Map2 s = m;
    s = select_map(pr >= 1, map_mul(s, dpp_map<kRowShr + 1>(s)), s);
    s = select_map(pr >= 2, map_mul(s, dpp_map<kRowShr + 2>(s)), s);
    s = select_map(pr >= 4, map_mul(s, dpp_map<kRowShr + 4>(s)), s);
    s = select_map(pr >= 8, map_mul(s, dpp_map<kRowShr + 8>(s)), s);
    if constexpr (L >= 32) s = select_map(wrow & 1, map_mul(s, dpp_map<kRowBcast15, 0xA>(s)), s);
    if constexpr (L == 64) s = select_map(wrow >= 2, map_mul(s, dpp_map<kRowBcast31, 0xC>(s)), s);
    // the map entering this lane: the previous lane's inclusive product
    Map2 e = dpp_map<kWaveShr1>(s);
    if constexpr (W == 2) {
        __syncthreads();  // the previous group's interpolation may still read qp
        if (wv == 0 && t.lane == 63) {
            qp[0] = s.a;
            qp[1] = s.b;
            qp[2] = s.c;
            qp[3] = s.d;
        }
        __syncthreads();
        if (wv == 1) {
            const Map2 w0{qp[0], qp[1], qp[2], qp[3]};
            e = t.lane == 0 ? w0 : map_mul(e, w0);
        }
    }
    double d0;  // this lane's first pivot
    {
        double dprev = t.ls == 0 ? 1.0 : (e.a + e.b) / (e.c + e.d);
        double bup = bu_prev;
        double rprev = rcp(dprev);
#pragma unroll
        for (int i = 0; i < M; ++i) {
            // Thomas pivots of the reference (gam_j = bu_{j-1}/bet, bet = b_j - bl_j gam_j)
            const double dd = bv[i] - blv[i] * (bup * rprev);
            const double rd = rcp(dd);
            lo[i] = (TF)(blv[i] * rprev);  // bl = 0 on row 0, whatever the entering pivot
            if (i == 0) d0 = dd;
            if (i > 0) up[i - 1] = (TF)(buv[i - 1] * rd);  // u_j = bu_j / d_{j+1}
            ap[i] = (TF)(inv_th * rd);
            rprev = rd;
            bup = buv[i];
        }
    }
    // the last row of a lane needs the next lane's first pivot (bu = 0 on the last real
    // row and on padding rows, so the segment end needs none)
    __syncthreads();
    qp[gl] = d0;
    __syncthreads();
    up[M - 1] = (TF)(buv[M - 1] * (t.ls + 1 < LT ? rcp(qp[gl + 1]) : 1.0));

    // Carry propagation coefficients: g_j = ghat_j + Pc_j g_in,
    // w_j = what_j + R_j g_in + Q_j w_in.
    double Pl, Ql;
    {
        double Pc[M];
        Pc[0] = -(double)lo[0];
#pragma unroll
        for (int i = 1; i < M; ++i) Pc[i] = -(double)lo[i] * Pc[i - 1];
        double r = Pc[M - 1], q = -(double)up[M - 1];
        R[M - 1] = (TF)r;
        Q[M - 1] = (TF)q;
#pragma unroll
        for (int i = M - 2; i >= 0; --i) {
            r = Pc[i] - (double)up[i] * r;
            q = -(double)up[i] * q;
            R[i] = (TF)r;
            Q[i] = (TF)q;
        }
        Pl = Pc[M - 1];
        Ql = q;
    }
    // Scan multipliers: products of the lanes' Pl (forward) and Ql (backward) over the
    // windows each scan stage spans. Lanes whose window leaves the row take 0 by select.
    {
        const double w2 = Pl * dpp<kRowShr + 1>(Pl);  // Pl over lanes l-1 .. l
        const double w4 = w2 * dpp<kRowShr + 2>(w2);  // l-3 .. l
        const double w8 = w4 * dpp<kRowShr + 4>(w4);  // l-7 .. l
        sm.F1 = (TF)(pr >= 1 ? Pl : 0.0);
        sm.F2 = (TF)(pr >= 2 ? w2 : 0.0);
        sm.F4 = (TF)(pr >= 4 ? w4 : 0.0);
        sm.F8 = (TF)(pr >= 8 ? w8 : 0.0);
        const double v2 = Ql * dpp<kRowShl + 1>(Ql);  // Ql over lanes l .. l+1
        const double v4 = v2 * dpp<kRowShl + 2>(v2);  // l .. l+3
        const double v8 = v4 * dpp<kRowShl + 4>(v4);  // l .. l+7
        sm.G1 = (TF)(pr + 1 <= 15 ? Ql : 0.0);
        sm.G2 = (TF)(pr + 2 <= 15 ? v2 : 0.0);
        sm.G4 = (TF)(pr + 4 <= 15 ? v4 : 0.0);
        sm.G8 = (TF)(pr + 8 <= 15 ? v8 : 0.0);
        // row-inclusive prefix of Pl and suffix of Ql
        double ip = Pl, sq = Ql;
        {
            double x;
            x = dpp<kRowShr + 1>(ip); ip = pr >= 1 ? ip * x : ip;
            x = dpp<kRowShr + 2>(ip); ip = pr >= 2 ? ip * x : ip;
            x = dpp<kRowShr + 4>(ip); ip = pr >= 4 ? ip * x : ip;
            x = dpp<kRowShr + 8>(ip); ip = pr >= 8 ? ip * x : ip;
            x = dpp<kRowShl + 1>(sq); sq = pr + 1 <= 15 ? sq * x : sq;
            x = dpp<kRowShl + 2>(sq); sq = pr + 2 <= 15 ? sq * x : sq;
            x = dpp<kRowShl + 4>(sq); sq = pr + 4 <= 15 ? sq * x : sq;
            x = dpp<kRowShl + 8>(sq); sq = pr + 8 <= 15 ? sq * x : sq;
        }
        // row products (uniform): Pl over rows 0..2 from the prefix at each row's end,
        // Ql over rows 1..3 from the suffix at each row's start
        const double rp0 = readlane<15>(ip), rp1 = readlane<31>(ip), rp2 = readlane<47>(ip);
        const double rs1 = readlane<16>(sq), rs2 = readlane<32>(sq), rs3 = readlane<48>(sq);
        sm.FB15 = (TF)((L >= 32 && (wrow & 1)) ? ip : 0.0);
        sm.FB31 = (TF)((L == 64 && wrow >= 2) ? (wrow == 3 ? ip * rp2 : ip) : 0.0);
        const bool w64 = L == 64, w32 = L == 32;
        sm.X16 = (TF)((w64 && wrow == 0) || (w32 && (wrow & 1) == 0) ? sq : 0.0);
        sm.X32 = (TF)(w64 && wrow == 1 ? sq : 0.0);
        sm.X48 = (TF)(w64 && wrow == 2 ? sq : 0.0);
        // two-wave groups: Pl over wave 1's lanes 0..l (carry of wave 0's last lane into
        // wave 1) and Ql over wave 0's lanes l..63 (carry of wave 1's first lane back)
        if constexpr (W == 2) {
            const double before = wrow == 0 ? 1.0 : wrow == 1 ? rp0 : wrow == 2 ? rp0 * rp1 : rp0 * rp1 * rp2;
            const double after = wrow == 3 ? 1.0 : wrow == 2 ? rs3 : wrow == 1 ? rs2 * rs3 : rs1 * rs2 * rs3;
            sm.FW = (TF)(wv == 1 ? ip * before : 0.0);
            sm.ZW = (TF)(wv == 0 ? sq * after : 0.0);
        } else {
            sm.FW = TF(0);
            sm.ZW = TF(0);
        }
    }
}

// All segments of a workgroup iterate to its longest march; lanes whose own march is
// done keep their values. *wmin (optional) receives the shortest march among segments
// that march at all: up to there every segment steps, so no lane needs its own test
// (segments with no march are identity rows, which stepping leaves unchanged).
template <int M, int L, int W>
__device__ __forceinline__ int group_steps(const Pde& d, const Lanes& t, double* qp,
                                           int* wmin = nullptr) {
    using Gm = Geo<M, L, W>;
    if constexpr (Gm::G == 1) {
        if (wmin) *wmin = d.steps;
        return d.steps;
    } else {
        __syncthreads();
        qp[t.gl] = (double)d.steps;
        __syncthreads();
        double mx = 0.0, mn = 1e300;
        for (int k = 0; k < Gm::NT; k += Gm::LT) {
            mx = fmax(mx, qp[k]);
            if (qp[k] > 0.0) mn = fmin(mn, qp[k]);
        }
        if (wmin) *wmin = mn < 1e300 ? (int)mn : 0;
        return (int)mx;
    }
}

// Wave-local inclusive forward scan of the lanes' local last values.
template <int L, class T>
__device__ __forceinline__ T fwd_scan(T c, const ScanMul<T>& sm) {
    c = fma(sm.F1, dpp<kRowShr + 1>(c), c);
    c = fma(sm.F2, dpp<kRowShr + 2>(c), c);
    c = fma(sm.F4, dpp<kRowShr + 4>(c), c);
    T s = dpp<kRowShr + 8>(c);
    c = fma(sm.F8, s, c);
    // rows outside the broadcast's mask keep s (finite) and have a zero multiplier
    if constexpr (L >= 32) {
        s = dpp<kRowBcast15, 0xA>(c, s);
        c = fma(sm.FB15, s, c);
    }
    if constexpr (L == 64) {
        s = dpp<kRowBcast31, 0xC>(c, s);
        c = fma(sm.FB31, s, c);
    }
    return c;
}

// Wave-local inclusive backward scan of the lanes' local first values.
template <int L, class T>
__device__ __forceinline__ T bwd_scan(T cb, const ScanMul<T>& sm, const Lanes& t) {
    cb = fma(sm.G1, dpp<kRowShl + 1>(cb), cb);
    cb = fma(sm.G2, dpp<kRowShl + 2>(cb), cb);
    cb = fma(sm.G4, dpp<kRowShl + 4>(cb), cb);
    cb = fma(sm.G8, dpp<kRowShl + 8>(cb), cb);
    // Across rows. A zero multiplier stands in for a select only inside one option:
    // 0 * x is NaN when x is not finite, so no lane may read another option's value.
    if constexpr (L == 64) {
        cb = fma(sm.X48, readlane<48>(cb), cb);
        cb = fma(sm.X32, readlane<32>(cb), cb);
        cb = fma(sm.X16, readlane<16>(cb), cb);
    } else if constexpr (L == 32) {
        // two options per wave: each takes the first lane of its own second row
        const T src = t.wrow < 2 ? readlane<16>(cb) : readlane<48>(cb);
        cb = fma(sm.X16, src, cb);
    }
    return cb;
}

// Carries of one step: forward scan of `c` (the lanes' local last values) giving each
// lane's f_in; `xchg` links the two waves of a W = 2 group.
template <int L, int W, class T>
__device__ __forceinline__ T fwd_carry(T c, const ScanMul<T>& sm, const Lanes& t, bool first_lane,
                                       T* xchg) {
    c = fwd_scan<L>(c, sm);
    T fin;
    if constexpr (W == 2) {
        if (t.wv == 0 && t.lane == 63) xchg[0] = c;
        __syncthreads();
        if (t.wv == 1) {
            const T c0 = xchg[0];
            c = fma(sm.FW, c0, c);
            fin = dpp<kWaveShr1>(c);
            fin = t.lane == 0 ? c0 : fin;
        } else {
            fin = dpp<kWaveShr1>(c);
        }
    } else {
        fin = dpp<kWaveShr1>(c);
        // a select, not a multiply by 0, so one segment's values never reach another's
        if constexpr (L < 64) fin = first_lane ? T(0) : fin;
    }
    return fin;
}

// Backward scan of `cb` (the lanes' local first values with the forward carry folded
// in) giving each lane's z_in.
template <int L, int W, class T>
__device__ __forceinline__ T bwd_carry(T cb, const ScanMul<T>& sm, const Lanes& t, bool last_lane,
                                       T* xchg) {
    cb = bwd_scan<L>(cb, sm, t);
    T zin;
    if constexpr (W == 2) {
        if (t.wv == 1 && t.lane == 0) xchg[1] = cb;
        __syncthreads();
        if (t.wv == 0) {
            const T z1 = xchg[1];
            cb = fma(sm.ZW, z1, cb);
            zin = dpp<kWaveShl1>(cb);
            zin = t.lane == 63 ? z1 : zin;
        } else {
            zin = dpp<kWaveShl1>(cb);
        }
    } else {
        zin = dpp<kWaveShl1>(cb);
        if constexpr (L < 64) zin = last_lane ? T(0) : zin;
    }
    return zin;
}

// ---- Fd1d::value: interpolation at log(S/K), price = K v ---------------------------
// MIXED: the state comes from the FP32-mixed march, whose prices beyond its drift limit
// (mixed_drift_risk) miss the 1e-4 bar: such options get NaN and KWGPU_STATUS_PRECISION.
template <int M, int L, int W, bool MIXED = false>
__device__ __forceinline__ void interpolate(const FdBatch& b, const Pde& d, const Lanes& t,
                                            const double (&xv)[M], double* qp) {
    using Gm = Geo<M, L, W>;
    // A NaN or inf anywhere in the march reaches the unprojected last row (the
    // projection's maxNum turns NaN into the payoff elsewhere, -inf there), where the
    // reference's own march would leave NaN: such an option is reported NaN / NONFINITE
    // instead of as the payoff its projected rows show.
    bool bad = false;
#pragma unroll
    for (int i = 0; i < M; ++i) bad |= d.j0 + i < d.N && !isfinite(xv[i]);
    bool seg_bad;
    if constexpr (W == 2) {
        seg_bad = __syncthreads_or(bad);  // one option per workgroup
    } else {
        const uint64_t votes = __ballot(bad);
        const int first = t.lane - t.ls;  // this option's first lane in the wave
        const uint64_t mine = L == 64 ? ~0ull : (((1ull << L) - 1) << first);
        seg_bad = (votes & mine) != 0;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < M; ++i) qp[t.seg * Gm::LT * M + d.j0 + i] = xv[i];
    __syncthreads();
    if (!d.live) return;
    const double* V = qp + t.seg * Gm::LT * M;
    const uint32_t beg = b.off ? b.off[d.p] : 0u;
    const uint32_t end = b.off ? b.off[d.p + 1] : 1u;
    for (uint32_t k = beg + t.ls; k < end; k += Gm::LT) {
        const uint32_t o2 = b.off ? b.list[k] : d.oi;
        const kwgpu_option q = b.off ? b.opts[o2] : d.opt;
        double price = NAN;
        int st = KWGPU_STATUS_BAD_INPUT;
        if (d.solve && finite_pos(q.s) && finite_pos(q.k)) {
            const double xs = log(q.s / q.k);
            const int j = bracket(d.P, b.density, xs);
            if (j == 0 || j == d.N) {
                st = KWGPU_STATUS_OUT_OF_GRID;
            } else if (MIXED && mixed_drift_risk(q, b.scale)) {
                st = KWGPU_STATUS_PRECISION;
            } else if (seg_bad) {
                st = KWGPU_STATUS_NONFINITE;
            } else {
                const double x1 = grid_x(d.P, b.density, j), x0 = grid_x(d.P, b.density, j - 1);
                const double v = ((x1 - xs) * V[j - 1] + (xs - x0) * V[j]) / (x1 - x0);
                price = q.k * v;
                st = isfinite(price) ? KWGPU_STATUS_OK : KWGPU_STATUS_NONFINITE;
            }
        }
        b.prices[o2] = price;
        if (b.status) b.status[o2] = st;
    }
}

// ---- FP64 march (src/Math/kwFd1d.cpp:67-133) -----------------------------------------
// One step of the scheme on the state xv; HALF: theta = 1/2, where the correction is
// a single FMA a_j w_j - V_j. Steps [0, wmin) are taken by every segment of the
// workgroup, steps [wmin, wsteps) only by segments whose own march (`steps`) lasts that
// long (several options per workgroup).
template <int M, int L, int W, bool HALF>
__device__ __forceinline__ void march_f64(double (&xv)[M], const double (&ap)[M],
                                          const double (&lo)[M], const double (&up)[M],
                                          const double (&R)[M], const double2* qpv,
                                          const ScanMul<double>& sm, const Lanes& t,
                                          bool first_lane, bool last_lane, double* xchg,
                                          int wmin, int wsteps, int steps, double c_th) {
    constexpr int NT = Geo<M, L, W>::NT;
    constexpr bool ONE = Geo<M, L, W>::G == 1;  // one option per workgroup
    // LDS reads of the (Q, payoff) pairs in flight ahead of the row updates. Measured on
    // one box: the 1024-point one-wave kernel 9.81 ms at 3, 10.03 at 4, 10.06 at 1 (and
    // 10.3 at 0); the 2048-point two-wave kernel 13.2 ms at 0, 15.6 at 1, 22.3 at 3,
    // where the extra registers turn into spill reloads inside the step loop; the
    // 256- and 512-point kernels (4 and 2 options per wave) 2.69 and 5.34 ms at 0, 2.88
    // and 6.33 at 3.
    constexpr int PF0 = W == 2 ? KWGPU_PF2 : ONE ? KWGPU_PF1 : KWGPU_PFG;
    constexpr int PF = PF0 < M ? PF0 : M - 1;  // 0: each pair read where it is used
    const int wv = __builtin_amdgcn_readfirstlane(t.wv);  // wave of a W = 2 group
    double g[M];
    // backward substitution of g in place and the step's carries: g_in (forward) and
    // w_in (backward) of this lane
    auto carries = [&](int step, double& gin, double& win) {
        const double c = g[M - 1];
#pragma unroll
        for (int i = M - 2; i >= 0; --i) g[i] = fma(-up[i], g[i + 1], g[i]);
        if constexpr (W == 2) {
            // both waves scan their own rows, then meet once: wave 0 publishes its last
            // forward carry c0, wave 1 its first backward carry with c0 left out; the
            // missing c0 terms are linear, with coefficients fixed at set-up. Slots
            // alternate with the step's parity, so one barrier orders every reuse.
            const double cw = fwd_scan<L>(c, sm);
            gin = dpp<kWaveShr1>(cw);
            const double cb = bwd_scan<L>(fma(R[0], gin, g[0]), sm, t);
            win = dpp<kWaveShl1>(cb);
            const int slot = 2 * (step & 1);
            if (wv == 0 && t.lane == 63) xchg[slot] = cw;
            if (wv == 1 && t.lane == 0) xchg[slot + 1] = cb;
            __syncthreads();
            const double other = xchg[slot + 1 - wv];
            const double u = wv == 0 ? fma(sm.TW0, readlane<63>(cw), other) : other;
            const double2 ab = qpv[M * NT + t.gl];
            gin = fma(ab.x, u, gin);
            win = fma(ab.y, u, win);
        } else {
            gin = fwd_carry<L, W>(c, sm, t, first_lane, xchg);
            win = bwd_carry<L, W>(fma(R[0], gin, g[0]), sm, t, last_lane, xchg);
        }
    };
    // Steps every segment takes: no per-lane test, so the loop is rotated. Each row's
    // update is followed at once by that row's forward substitution for the next step:
    // the substitution is a dependent chain of M FMAs, the row updates are independent
    // of each other, so interleaved they fill the chain's latency.
    if (wmin > 0) {
        g[0] = xv[0];
#pragma unroll
        for (int i = 1; i < M; ++i) g[i] = fma(-lo[i], g[i - 1], xv[i]);
    }
    for (int step = 0; step < wmin; ++step) {
        double gin, win;
        carries(step, gin, win);
        // no branch, so the LDS reads and fix-ups can also move into the backward scan;
        // the (Q, payoff) pairs are read PF rows ahead of their use
        double2 qbuf[PF > 0 ? PF : 1];
#pragma unroll
        for (int k = 0; k < PF; ++k) qbuf[k] = qpv[k * NT + t.gl];
#pragma unroll
        for (int i = 0; i < M; ++i) {
            double2 qv;
            if constexpr (PF == 0) {
                qv = qpv[i * NT + t.gl];
            } else {
                qv = qbuf[0];
#pragma unroll
                for (int k = 0; k + 1 < PF; ++k) qbuf[k] = qbuf[k + 1];
                if (i + PF < M) qbuf[PF - 1] = qpv[(i + PF) * NT + t.gl];
            }
            double w = fma(R[i], gin, g[i]);
            w = fma(qv.x, win, w);
            const double v = HALF ? fma(ap[i], w, -xv[i]) : fma(ap[i], w, -c_th * xv[i]);
            xv[i] = vmax(v, qv.y);
            g[i] = i == 0 ? xv[0] : fma(-lo[i], g[i - 1], xv[i]);  // next step's
        }
    }
    // Steps only the longer segments take (several options per workgroup): lanes whose
    // own march is done keep their values.
    if constexpr (!ONE) {
        for (int step = wmin; step < wsteps; ++step) {
            g[0] = xv[0];
#pragma unroll
            for (int i = 1; i < M; ++i) g[i] = fma(-lo[i], g[i - 1], xv[i]);
            double gin, win;
            carries(step, gin, win);
            if (step < steps) {
#pragma unroll
                for (int i = 0; i < M; ++i) {
                    const double2 qv = qpv[i * NT + t.gl];
                    double w = fma(R[i], gin, g[i]);
                    w = fma(qv.x, win, w);
                    const double v = HALF ? fma(ap[i], w, -xv[i]) : fma(ap[i], w, -c_th * xv[i]);
                    xv[i] = vmax(v, qv.y);
                }
            }
        }
    }
}

// GTH: general theta. The theta = 1/2 kernel holds only its own march, so set-up values
// that only the general march needs are not kept (and spilled) on its path.
template <int M, int L, int W, bool GTH>
__device__ __forceinline__ void solve_group_f64(const FdBatch& b, int group, double* qp,
                                                double* xchg) {
    using Gm = Geo<M, L, W>;
    constexpr int NT = Gm::NT;
    const Lanes t = lanes_of<M, L, W>(threadIdx.x);
    const Pde d = pde_of<M, L, W>(b, group, t);
    const double c_th = (1.0 - b.theta) / b.theta;
    double xv[M], ap[M], lo[M], up[M], R[M], Q[M], blv[M], buv[M];
    ScanMul<double> sm;
    setup<M, L, W>(b, d, t, qp, xv, ap, lo, up, R, Q, blv, buv, sm);
    const bool first_lane = t.ls == 0;
    const bool last_lane = t.ls == Gm::LT - 1;
    // the same for every lane of the workgroup: keep the loop counter scalar
    int wmin = 0;
    const int wsteps = __builtin_amdgcn_readfirstlane(group_steps<M, L, W>(d, t, qp, &wmin));
    wmin = __builtin_amdgcn_readfirstlane(wmin);
    double aw = 0.0, bw = 0.0;
    sm.TW0 = 0.0;
    if constexpr (W == 2) {
        // wave 1: c0 enters lane l's forward carry with FW of lane l-1 (1 in lane 0) and
        // its backward carry through the backward scan of R_0 times that; wave 0: wave
        // 1's first backward carry enters lane l's with ZW of lane l+1 (1 in lane 63)
        // Every shift runs on the whole wave before any select: a DPP move inside a
        // branch reads 0 from the lanes the branch turns off.
        const double fw_prev = dpp<kWaveShr1>(sm.FW);
        const double zw_next = dpp<kWaveShl1>(sm.ZW);
        const double fws = t.wv == 1 ? (t.lane == 0 ? 1.0 : fw_prev) : 0.0;
        const double tw = bwd_scan<L>(R[0] * fws, sm, t);
        const double tw_next = dpp<kWaveShl1>(tw);
        const double zws = t.wv == 0 ? (t.lane == 63 ? 1.0 : zw_next) : 0.0;
        aw = fws;
        bw = t.wv == 0 ? zws : tw_next;
        __syncthreads();
        if (t.wv == 1 && t.lane == 0) xchg[0] = tw;
        __syncthreads();
        sm.TW0 = readlane<0>(xchg[0]);  // uniform: keep it in scalar registers
    }

    // (Q_i, projection target_i) pairs of every thread at qpv[i * NT + gl]; two-wave
    // groups add (AW, BW) at qpv[M * NT + gl]
    __syncthreads();
    double2* qpv = reinterpret_cast<double2*>(qp);
#pragma unroll
    for (int i = 0; i < M; ++i) {
        const bool proj = d.P.american && d.j0 + i < d.N - 1;
        qpv[i * NT + t.gl] = make_double2(Q[i], proj ? xv[i] : -INFINITY);
    }
    if constexpr (W == 2) qpv[M * NT + t.gl] = make_double2(aw, bw);
    __syncthreads();
    const int steps = d.steps;
    // (the general form gives the same values at theta = 1/2: -c_th V is exactly -V)
    march_f64<M, L, W, !GTH>(xv, ap, lo, up, R, qpv, sm, t, first_lane, last_lane, xchg, wmin,
                             wsteps, steps, c_th);
    // the option's record and lane indices are derived again rather than kept in
    // registers over the march: the group and thread index go through an empty asm so
    // that the compiler cannot reuse the set-up's values (which it would otherwise keep,
    // and spill, across the loop)
    int g2 = group, tid = threadIdx.x;
    asm volatile("" : "+v"(g2), "+v"(tid));
    const Lanes t2 = lanes_of<M, L, W>(tid);
    interpolate<M, L, W>(b, pde_of<M, L, W>(b, g2, t2), t2, xv, qp);
}

// ---- FP32-mixed march ---------------------------------------------------------------
template <int M, int L, int W>
__device__ __forceinline__ void solve_group_mixed(const FdBatch& b, int group, double* qp,
                                                  double* xchg) {
    using Gm = Geo<M, L, W>;
    constexpr int NT = Gm::NT;
    const Lanes t = lanes_of<M, L, W>(threadIdx.x);
    const Pde d = pde_of<M, L, W>(b, group, t);
    double xv[M], blv[M], buv[M];
    float ap[M], lo[M], up[M], R[M], Q[M];
    ScanMul<float> sm;
    setup<M, L, W>(b, d, t, qp, xv, ap, lo, up, R, Q, blv, buv, sm);
    // diagonal of (I - B) is theta dt a0 on every real row (the stencils' weights sum
    // to zero); padding rows hold V = 0
    const double sdiag = b.theta * d.P.dt * d.P.a0;
    const bool first_lane = t.ls == 0;
    const bool last_lane = t.ls == Gm::LT - 1;
    // the same for every lane of the workgroup: keep the loop counter scalar
    const int wsteps = __builtin_amdgcn_readfirstlane(group_steps<M, L, W>(d, t, qp));
    float* xf = reinterpret_cast<float*>(xchg + 2);  // carries; xchg[0..1] exchange V

    // projection targets of every thread at qp[i * NT + gl]
    __syncthreads();
#pragma unroll
    for (int i = 0; i < M; ++i) {
        const bool proj = d.P.american && d.j0 + i < d.N - 1;
        qp[i * NT + t.gl] = proj ? xv[i] : -INFINITY;
    }
    __syncthreads();

    for (int step = 0; step < wsteps; ++step) {
        // residual (I - B) V in FP64: -bl (V_{j-1} - V_j) - bu (V_{j+1} - V_j) + sdiag V_j
        // (bl = 0 on the first row and bu = 0 on the last, so their outer neighbour
        // only has to be finite)
        double vprev = dpp<kWaveShr1>(xv[M - 1]);
        double vnext = dpp<kWaveShl1>(xv[0]);
        vprev = first_lane ? xv[0] : vprev;
        vnext = last_lane ? xv[M - 1] : vnext;
        if constexpr (W == 2) {
            // the waves meet between rows 1023 and 1024
            if (t.wv == 0 && t.lane == 63) xchg[0] = xv[M - 1];
            if (t.wv == 1 && t.lane == 0) xchg[1] = xv[0];
            __syncthreads();
            if (t.wv == 1 && t.lane == 0) vprev = xchg[0];
            if (t.wv == 0 && t.lane == 63) vnext = xchg[1];
            __syncthreads();
        }
        float g[M];
        double dm = xv[0] - vprev;
        float gprev = 0.0f;
#pragma unroll
        for (int i = 0; i < M; ++i) {
            const double dp = (i + 1 < M ? xv[i + 1] : vnext) - xv[i];
            double r = fma(blv[i], dm, sdiag * xv[i]);
            r = fma(-buv[i], dp, r);
            dm = dp;
            gprev = fmaf(-lo[i], gprev, (float)r);
            g[i] = gprev;
        }
        const float c = g[M - 1];
#pragma unroll
        for (int i = M - 2; i >= 0; --i) g[i] = fmaf(-up[i], g[i + 1], g[i]);
        const float gin = fwd_carry<L, W>(c, sm, t, first_lane, xf);
        const float win = bwd_carry<L, W>(fmaf(R[0], gin, g[0]), sm, t, last_lane, xf);
        if (Geo<M, L, W>::G == 1 || step < d.steps) {
#pragma unroll
            for (int i = 0; i < M; ++i) {
                float w = fmaf(R[i], gin, g[i]);
                w = fmaf(Q[i], win, w);
                xv[i] = vmax(fma((double)ap[i], (double)w, xv[i]), qp[i * NT + t.gl]);
            }
        }
    }
    // the option's record is read again rather than kept in registers over the march
    interpolate<M, L, W, true>(b, pde_of<M, L, W>(b, group, t), t, xv, qp);
}

// M rows per lane, L lanes per option (W = 1) or W whole waves per option; WPE is the
// number of waves per SIMD the register budget is cut for (2: <= 256 registers).
// With PERSIST the grid is persistent: workgroups take group indices from `queue`
// until all `ngroups` are done (every workgroup reaches the exit test).
template <int M, int L, int W, int WPE, bool PERSIST, bool MIXED, bool GTH>
__global__ __attribute__((amdgpu_flat_work_group_size(64 * W, 64 * W),
                          amdgpu_waves_per_eu(WPE, WPE)))
void fd1d_wave(FdBatch b, int* queue, int ngroups) {
    // (Q, payoff) pairs per row and thread, and (AW, BW) per thread for two-wave groups
    __shared__ double qp[2 * 64 * W * (M + (W == 2 ? 1 : 0))];
    __shared__ double xchg[4];
    auto run = [&](int g) {
        if constexpr (MIXED)
            solve_group_mixed<M, L, W>(b, g, qp, xchg);
        else
            solve_group_f64<M, L, W, GTH>(b, g, qp, xchg);
    };
    if constexpr (!PERSIST) {
        run(blockIdx.x);
    } else {
        __shared__ int next;
        for (;;) {
            __syncthreads();
            if (threadIdx.x == 0) next = atomicAdd(queue, 1);
            __syncthreads();
            const int g = next;
            if (g >= ngroups) break;
            run(g);
        }
    }
}

__global__ void zero_counters(int* c, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) c[i] = 0;
}

template <int M, int L, int W, bool MIXED, bool GTH>
hipError_t launch_wave(const FdBatch& b, hipStream_t s, int* queue) {
    constexpr int G = Geo<M, L, W>::G;
    constexpr int WPE = 2;
    const int ngroups = (b.m + G - 1) / G;
    int grid = ngroups;
    if (queue) {
        // enough workgroups to fill every SIMD at the register-limited occupancy
        int dev = 0, cus = 256;
        if (hipGetDevice(&dev) == hipSuccess)
            (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        grid = std::min(ngroups, cus * 4 * WPE / W);
        hipLaunchKernelGGL((fd1d_wave<M, L, W, WPE, true, MIXED, GTH>), dim3(grid), dim3(64 * W),
                           0, s, b, queue, ngroups);
    } else {
        hipLaunchKernelGGL((fd1d_wave<M, L, W, WPE, false, MIXED, GTH>), dim3(grid), dim3(64 * W),
                           0, s, b, queue, ngroups);
    }
    return hipGetLastError();
}

// The FP64 march has a theta = 1/2 kernel and a general one; the FP32-mixed march
// handles any theta in one form.
template <int M, int L, int W, bool MIXED>
hipError_t launch_wave(const FdBatch& b, hipStream_t s, int* queue) {
    if (!MIXED && b.theta != 0.5) return launch_wave<M, L, W, MIXED, !MIXED>(b, s, queue);
    return launch_wave<M, L, W, MIXED, false>(b, s, queue);
}

template <bool MIXED>
hipError_t launch_variant(const FdBatch& b, int v, hipStream_t s, int* queue) {
    switch (v) {
        case 0: return launch_wave<4, 16, 1, MIXED>(b, s, queue);
        case 1: return launch_wave<8, 16, 1, MIXED>(b, s, queue);
        case 2: return launch_wave<16, 16, 1, MIXED>(b, s, queue);
        case 3: return launch_wave<4, 32, 1, MIXED>(b, s, queue);
        case 4: return launch_wave<8, 32, 1, MIXED>(b, s, queue);
        case 5: return launch_wave<16, 32, 1, MIXED>(b, s, queue);
        case 6: return launch_wave<4, 64, 1, MIXED>(b, s, queue);
        case 7: return launch_wave<8, 64, 1, MIXED>(b, s, queue);
        case 8: return launch_wave<16, 64, 1, MIXED>(b, s, queue);
        case 9: return launch_wave<16, 64, 2, MIXED>(b, s, queue);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

// rows per lane M, lanes per option L, waves per option W (index = variant id)
static const WaveVariant kVariants[] = {{4, 16, 1}, {8, 16, 1}, {16, 16, 1}, {4, 32, 1},
                                        {8, 32, 1}, {16, 32, 1}, {4, 64, 1}, {8, 64, 1},
                                        {16, 64, 1}, {16, 64, 2}};

int wave_variant_count() { return (int)(sizeof(kVariants) / sizeof(kVariants[0])); }
WaveVariant wave_variant(int v) { return kVariants[v]; }

// Fewest lanes per option at 16 rows per lane (4 or 8 for the smallest grids).
int default_wave_variant(int max_x_dim) {
    if (max_x_dim <= 64) return 0;
    if (max_x_dim <= 128) return 1;
    if (max_x_dim <= 256) return 2;
    if (max_x_dim <= 512) return 5;
    if (max_x_dim <= 1024) return 8;
    if (max_x_dim <= 2048) return 9;
    return -1;
}

bool wave_layout_fits(int x_dim, int precision) {
    (void)precision;
    return x_dim >= 3 && x_dim <= 2048;
}

hipError_t launch_fd1d_wave_variant(const FdBatch& b, int variant, int precision, hipStream_t s,
                                    int* queue) {
    if (b.m <= 0) return hipSuccess;
    if (variant < 0 || variant >= wave_variant_count()) return hipErrorInvalidValue;
    if (precision == KWGPU_FP32_MIXED) return launch_variant<true>(b, variant, s, queue);
    return launch_variant<false>(b, variant, s, queue);
}

hipError_t launch_fd1d_wave(const FdBatch& b, int max_x_dim, int precision, hipStream_t s,
                            int* queue) {
    if (!wave_layout_fits(max_x_dim, precision)) return hipErrorInvalidValue;
    return launch_fd1d_wave_variant(b, default_wave_variant(max_x_dim), precision, s, queue);
}

hipError_t launch_zero_counters(int* c, int n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(zero_counters, dim3((n + 63) / 64), dim3(64), 0, s, c, n);
    return hipGetLastError();
}

}  // namespace kwgpu
