// surrogate.hip -- learned early-exercise premium in front of the FD pricer.
//
// The surrogate emulates one FD pricer configuration (density, scale, theta and either
// one x/t grid or per-option grids). An ensemble of MLPs 8 -> 256 -> 256 -> 256 -> 256
// -> 1 (SiLU) predicts, at a mesh node x_j of an option's own FD grid, the FD solution
// there minus the Black-Scholes price, as y = (premium - base) / (base + 1e-4 K) against
// a control variate base: the premium from the option's exercise boundary, solved by a
// fixed point (boundary_setup). A price is then formed exactly as the FD pricer forms its own
// (Fd1d::value, src/Math/kwFd1d.cpp:139-158): the network is
// evaluated at the two nodes that bracket log(S/K), BS is added at each node, and the
// two node values are interpolated linearly with the reference's weights. The
// interpolation error between nodes -- up to ~1e-4 of price, and oscillating with the
// position of log(S/K) in its cell -- is then reproduced rather than learned.
// Labels are the FD pricer's own values at nodes: an option with spot K e^{x_j} is
// priced at node j (interpolation weight 1).
// The three 256x256 layers run forward and backward on BF16 MFMA
// (v_mfma_f32_16x16x32_bf16, FP32 accumulation) through one strided GEMM kernel with
// fused epilogues; the 8-wide input and 1-wide output layers are FP32; Adam keeps FP32
// master weights and refreshes the BF16 copies after every step. The reference has no
// counterpart: the labels come from this build's FD pricer, which is also the fallback.
#include <math.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "kwgpu_internal.h"

namespace kwgpu {
namespace sur {

typedef __bf16 bf16;
typedef bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int H = 256;  // hidden width
constexpr int F = 8;    // features: log(S/K), T, sigma, r, q, w, then the FD grid's x and t points
constexpr int F0 = 6;   // ... of which the option's own (the training box applies to these)
constexpr int kW1 = 0, kB1 = kW1 + H * F, kW2 = kB1 + H, kB2 = kW2 + H * H, kW3 = kB2 + H,
              kB3 = kW3 + H * H, kW4 = kB3 + H, kB4 = kW4 + H * H, kW5 = kB4 + H, kB5 = kW5 + H,
              kP = kB5 + 1;
static_assert(kP == KWGPU_SURROGATE_PARAMS, "parameter layout");
// partial-sum slots of the column reductions: at most kPartRows row blocks, each
// at most kPartWidth sums (the layer-1 block W1 | b1), per member
constexpr int kPartRows = 128, kPartWidth = H * F + H;
// ... and of the weight-gradient GEMMs' K splits (at most kMaxSplits slots of H x H)
constexpr int kMaxSplits = 32;
constexpr size_t kPartFloats = std::max<size_t>((size_t)kPartRows * kPartWidth,
                                                (size_t)kMaxSplits * H * H);
// Target: y = (premium - base) / (base + kEps K), the part of the FD pricer's premium
// that the control variate (base, the premium from the exercise boundary below) misses,
// relative to it (floored at kEps K). A network error dy then costs (base + kEps K) dy in
// price: small where the premium is small, which is where an absolute error bar can be
// met.
constexpr double kEps = 1e-4;
constexpr uint32_t kFileVersion = 6;  // saved parameters belong to this target and features

// The gate's price scale: Ju & Zhong's (1999) correction of the Barone-Adesi & Whaley
// (1987) quadratic approximation of the early-exercise premium (American minus European
// Black-Scholes), with a continuous dividend yield q. It is cheap (one critical spot per
// option), so the candidate gate takes it for every option of a batch. With
// alpha = 2r / sigma^2, beta = 2 (r - q) / sigma^2, h = 1 - e^{-rT}, m = alpha / h
// (-> 2 / (sigma^2 T) as r -> 0) and D = (beta - 1)^2 + 4m,
//     lambda = (-(beta - 1) + w sqrt(D)) / 2,
// the critical spot S* solves (Barone-Adesi & Whaley)
//     w (S* - K) = E(S*) + w (1 - e^{-qT} Phi(w d1(S*))) S* / lambda,
// and on the continuation side of S* the premium is
//     A (S / S*)^lambda / (1 - chi),   chi = b x^2 + c x,   x = log(S / S*),
// A = w (S* / lambda)(1 - e^{-qT} Phi(w d1(S*))) (the premium at S*), with Ju & Zhong's
//     b = -(1 - h) m^2 / (2 D),
//     c = -[ m (h / r) dE/dT(S*) / (w sqrt(D) A) + (1 - h) m / (w sqrt D)
//            - (1 - h) w m^2 / D^{3/2} ],
// dE/dT the European price's derivative in maturity at S*; (h / r) -> T as r -> 0.
// (These are Ju & Zhong's b = (1 - h) alpha lambda'(h) / (2 (2 lambda + beta - 1)) and
// c = -(1 - h) alpha / (2 lambda + beta - 1) [dE/dh / A_h + 1 / h + lambda'(h) /
// (2 lambda + beta - 1)], A_h = A / h, with 2 lambda + beta - 1 = w sqrt D and
// dE/dh = dE/dT / (r (1 - h)) written out so that r = 0 needs no limit.) Where chi > 0
// the factor 1 / (1 - chi) is replaced by 1 + chi, equal to first order: it stays
// positive where chi reaches 1 (options with r or q near 0, whose premium is near 0).
// Beyond S* the premium is w (S - K) - E(S). It is 0 where early exercise never pays
// (calls with q <= 0, puts with r <= 0). Against the FD price at 1024^2 on 3000 options
// of the training box, the correction leaves an rms premium error of 0.056 against
// BAW's 0.32. Its critical spot also starts the exercise-boundary fixed point below.
// S* and the coefficients do not depend on the spot, so they are set up once per option
// (baseline_setup) and the premium taken at any spot from them (baseline_at): the
// surrogate evaluates it at the option's spot and at the two mesh nodes around it.
struct Baseline {
    double w, z2, qw, sq, eq, er, ss;
    double a, cb, cc;  // premium at S*, and chi = cb x^2 + cc x
    bool none;         // no early-exercise premium
};

__host__ __device__ inline double baw_euro(const kwgpu_option& o, const Baseline& c, double S,
                                           double& nd1) {  // European price; nd1 = Phi(w d1)
    const double cr = 0.70710678118654752440;  // 1 / sqrt(2)
    const double d1 = (log(S / o.k) + (o.r - o.q + 0.5 * c.z2) * o.t) / c.sq;
    nd1 = 0.5 * erfc(-c.w * d1 * cr);
    return c.w * (S * c.eq * nd1 - o.k * c.er * 0.5 * erfc(-c.w * (d1 - c.sq) * cr));
}

// The S* equation as h(log S), positive below the root (for calls and puts alike).
__host__ __device__ inline double baw_h(const kwgpu_option& o, const Baseline& c, double x) {
    const double S = exp(x);
    double nd1;
    const double g = c.w * (S - o.k) - baw_euro(o, c, S, nd1) - c.w * (1 - c.eq * nd1) * S / c.qw;
    return c.w < 0 ? g : -g;
}

// S* by the Illinois method (regula falsi that halves the stale end's value) in log S
// over [K/1000, K] (puts) or [K, 10^4 K] (calls), until the bracket is 1e-14 wide; a root
// outside the bracket gives its nearer end. Against 64 halvings of the same bracket:
// premiums within 5e-15 K over 1.7M options (vol 2-122 %, T to 30 years, r -3-20 %,
// q 0-30 %, rates or yields of 0 included) in about 19 evaluations of h instead of 64.
// Then A and the coefficients of chi at S*.
__host__ __device__ inline Baseline baseline_setup(const kwgpu_option& o) {
    Baseline c{};
    const double t = o.t, z = o.z, r = o.r, q = o.q;
    c.w = o.w < 0 ? -1.0 : 1.0;
    c.none = !(o.k > 0 && t > 0 && z > 0) || (c.w > 0 && !(q > 0)) || (c.w < 0 && !(r > 0));
    if (c.none) return c;
    c.z2 = z * z;
    const double hh = r > 0 ? -expm1(-r * t) : 0.0;         // h
    const double m = r > 0 ? 2 * r / (c.z2 * hh) : 2 / (c.z2 * t);  // alpha / h
    const double nn = 2 * (r - q) / c.z2;                    // beta
    const double dd = (nn - 1) * (nn - 1) + 4 * m, sd = sqrt(dd);
    c.qw = 0.5 * (-(nn - 1) + c.w * sd);
    c.sq = z * sqrt(t);
    c.eq = exp(-q * t);
    c.er = exp(-r * t);
    double a = c.w < 0 ? log(o.k * 1e-3) : log(o.k), b = c.w < 0 ? log(o.k) : log(o.k * 1e4);
    double fa = baw_h(o, c, a), fb = baw_h(o, c, b);
    if (!(fa > 0)) {
        c.ss = exp(a);
    } else if (fb > 0) {
        c.ss = exp(b);
    } else {
        int side = 0;
        for (int it = 0; it < 64 && b - a > 1e-14; ++it) {
            double x = (a * fb - b * fa) / (fb - fa);
            if (!(x > a && x < b)) x = 0.5 * (a + b);
            const double fx = baw_h(o, c, x);
            if (fx > 0) {
                a = x;
                fa = fx;
                if (side == 1) fb *= 0.5;
                side = 1;
            } else {
                b = x;
                fb = fx;
                if (side == -1) fa *= 0.5;
                side = -1;
            }
        }
        c.ss = exp(0.5 * (a + b));
    }
    double nd1s;
    baw_euro(o, c, c.ss, nd1s);
    c.a = c.w * (c.ss / c.qw) * (1 - c.eq * nd1s);
    // dE/dT at S*: S e^{-qT} n(d1) sigma / (2 sqrt T) - w q S e^{-qT} Phi(w d1)
    //              + w r K e^{-rT} Phi(w d2)
    const double d1 = (log(c.ss / o.k) + (r - q + 0.5 * c.z2) * t) / c.sq;
    const double dens = 0.39894228040143267794 * exp(-0.5 * d1 * d1);
    const double nd2 = 0.5 * erfc(-c.w * (d1 - c.sq) * 0.70710678118654752440);
    const double det = c.ss * c.eq * dens * z / (2 * sqrt(t)) - c.w * q * c.ss * c.eq * nd1s +
                       c.w * r * o.k * c.er * nd2;
    const double hr = r > 0 ? hh / r : t;  // h / r
    const double cb = -(1 - hh) * m * m / (2 * dd);
    const double cc = -(m * hr * det / (c.w * sd * c.a) + (1 - hh) * m / (c.w * sd) -
                        (1 - hh) * c.w * m * m / (dd * sd));
    const bool ok = c.a > 0 && isfinite(cb) && isfinite(cc);
    c.cb = ok ? cb : 0.0;
    c.cc = ok ? cc : 0.0;
    return c;
}

// The premium at spot s of an option whose baseline is c (o's own spot is not used).
__host__ __device__ inline double baseline_at(const kwgpu_option& o, const Baseline& c, double s) {
    if (c.none || !(s > 0)) return 0.0;
    if (c.w < 0 ? s > c.ss : s < c.ss) {
        const double x = log(s / c.ss);
        const double chi = (c.cb * x + c.cc) * x;
        const double f = chi > 0 ? 1 + chi : 1 / (1 - chi);
        return c.a * pow(s / c.ss, c.qw) * f;
    }
    double nd1;
    return c.w * (s - o.k) - baw_euro(o, c, s, nd1);
}

__host__ __device__ inline double baseline_premium(const kwgpu_option& o) {
    return baseline_at(o, baseline_setup(o), o.s);
}

// ---- control variate: the premium from the exercise boundary ------------------------
// The early-exercise premium of the American put with unit strike, rate r, yield q and
// volatility sigma, from its exercise boundary b(tau) (tau: time to maturity). The
// boundary solves, for every tau (Andersen, Lake & Offengenden 2016, their fixed point
// "A"; it is the premium representation below taken at S = b(tau)),
//     b(tau) = e^{-(r-q) tau} N(tau) / D(tau),
//     N = Phi(d-(tau, b(tau))) + r int_0^tau e^{ru} Phi(d-(tau - u, b(tau) / b(u))) du,
//     D = Phi(d+(tau, b(tau))) + q int_0^tau e^{qu} Phi(d+(tau - u, b(tau) / b(u))) du,
// d+-(t, z) = (ln z + (r - q) t +- sigma^2 t / 2) / (sigma sqrt t), and the premium at a
// spot S above b(T) is
//     int_0^T [r e^{-r(T-u)} Phi(-d-(T-u, S / b(u))) - q S e^{-q(T-u)} Phi(-d+(T-u, S / b(u)))] du
// (at or below b(T): 1 - S minus the European put). b(0+) = X = min(1, r / q) (1 when
// q <= 0). b is held as H(xi) = ln(b / X)^2 on kBn + 1 Chebyshev-Lobatto nodes in
// xi = sqrt(tau) over [0, sqrt T], where H is smooth. The fixed point takes kBit Jacobi
// steps from b_i = X (b_T / X)^{sqrt(tau_i / T)}, b_T the Barone-Adesi & Whaley critical
// spot at T; the integrals take a 17-point tanh-sinh rule on u = tau a_k, which holds
// the sqrt and log behaviour of the integrands at both ends. Calls price as puts by
// put-call symmetry: C(S, K; r, q) = P(K, S; q, r), and P(S, K) = K p(S / K) for the
// unit-strike premium p. Against the FD price at 1024^2 on 3000 options of the training
// box, the remainder FD - premium has rms 1.7e-3 -- the FD solver's own discretisation
// error, since the same at 16 nodes and 16 steps changes it by under 1e-4 -- against
// 0.064 for Ju & Zhong. No early exercise (0): puts with r <= 0, calls with q <= 0.
constexpr int kBn = 6, kBit = 8, kTs = 17;
// tanh-sinh rule on (0, 1), h = 0.375, |k| <= 8: abscissae a_k = (1 + tanh(pi/2 sinh(kh))) / 2
// and weights w_k = (h pi/4) cosh(kh) / cosh^2(pi/2 sinh(kh)), so that
// int_0^tau f(u) du ~ tau sum_k w_k f(tau a_k) (the w_k sum to 1 + 8e-9). a_{16-k} = 1 - a_k
// to the last digit, so tau - u is taken from the mirror point.
__host__ __device__ inline double ts_a(int k) {
    constexpr double a[kTs] = {
        2.1470805279391156e-14, 4.2864745391083903e-10, 3.9763144264366805e-07,
        4.5307652428000167e-05, 0.0012425717713878076, 0.013016565902161277,
        0.070215470655051695,   0.23042664730601614,   0.5,
        0.76957335269398386,    0.92978452934494826,   0.98698343409783873,
        0.99875742822861224,    0.99995469234757206,   0.99999960236855734,
        0.99999999957135244,    0.99999999999997846};
    return a[k];
}
__host__ __device__ inline double ts_w(int k) {
    constexpr double w[kTs] = {
        2.5465845514760737e-13, 3.5038654253880788e-09, 2.2469440055945116e-06,
        0.00017811502580315067, 0.0034393438106114739,  0.025766645199968952,
        0.09957717664276014,    0.22377431734671693,    0.2945243112740431,
        0.22377431734671693,    0.09957717664276014,    0.025766645199968952,
        0.0034393438106114739,  0.00017811502580315067, 2.2469440055945116e-06,
        3.5038654253880788e-09, 2.5465845514760737e-13};
    return w[k];
}
// sqrt(a_k): sqrt(u) = sqrt(tau) sqrt(a_k) and sqrt(tau - u) = sqrt(tau) sqrt(a_{16-k})
__host__ __device__ inline double ts_sa(int k) {
    constexpr double s[kTs] = {
        1.4652919599653563e-07, 2.07038028852392e-05,  0.00063058024282692843,
        0.0067310959306787601,  0.035250131508801602,  0.11409016566804203,
        0.26498201949387379,    0.48002775680789139,   0.70710678118654757,
        0.87725330018985048,    0.96425335329722772,   0.99347039920565261,
        0.99937852099623004,    0.99997734591718235,   0.99999980118425891,
        0.99999999978567622,    0.99999999999998923};
    return s[k];
}
// cos(m pi / 6)
__host__ __device__ inline double cos6(int m) {
    constexpr double c[12] = {1.0, 0.86602540378443865, 0.5, 0.0, -0.5, -0.86602540378443865,
                              -1.0, -0.86602540378443865, -0.5, 0.0, 0.5, 0.86602540378443865};
    return c[m % 12];
}
static_assert(kBn == 6, "cos6 holds the Chebyshev-Lobatto angles of six intervals");

__host__ __device__ inline double phi_cdf(double x) {
    return 0.5 * erfc(-x * 0.70710678118654752440);
}

struct Boundary {
    double a[kBn + 1];  // Chebyshev coefficients of H over z = 1 - 2 xi / sqrt T
    double x0, r, q, z, t, st;  // X, the put's rate and yield, sigma, T, sqrt T
    bool call, none;
};

// ln(b(u) / X) = -sqrt(H) at zz = 1 - 2 sqrt(u) / sqrt(T) (H by Clenshaw)
__host__ __device__ inline double boundary_lnb(const Boundary& bd, double zz) {
    double b1 = 0.0, b2 = 0.0;
    for (int k = kBn; k >= 1; --k) {
        const double b0 = 2.0 * zz * b1 - b2 + bd.a[k];
        b2 = b1;
        b1 = b0;
    }
    const double h = zz * b1 - b2 + bd.a[0];
    return -sqrt(fmax(h, 0.0));
}
__host__ __device__ inline double boundary_bT(const Boundary& bd) {  // b(T)
    return bd.x0 * exp(boundary_lnb(bd, -1.0));
}

// Chebyshev-Lobatto coefficients of H = ln(b / X)^2 from ln(b_i / X) at the nodes
__host__ __device__ inline void boundary_fit_log(Boundary& bd, const double (&lb)[kBn + 1]) {
    double hv[kBn + 1];
    for (int i = 0; i <= kBn; ++i) hv[i] = lb[i] * lb[i];
    for (int k = 0; k <= kBn; ++k) {
        double s = 0.5 * (hv[0] + hv[kBn] * cos6(k * kBn));
        for (int i = 1; i < kBn; ++i) s += hv[i] * cos6(k * i);
        bd.a[k] = (2.0 / kBn) * s * (k == 0 || k == kBn ? 0.5 : 1.0);
    }
}
__host__ __device__ inline void boundary_fit(Boundary& bd, const double (&b)[kBn + 1]) {
    double lb[kBn + 1];
    for (int i = 0; i <= kBn; ++i) lb[i] = log(b[i] / bd.x0);
    boundary_fit_log(bd, lb);
}

// Parameters, node times tau_i and the starting boundary b_i.
__host__ __device__ inline void boundary_start(const kwgpu_option& o, Boundary& bd,
                                               double (&b)[kBn + 1], double (&tau)[kBn + 1]) {
    bd.call = !(o.w < 0);
    bd.r = bd.call ? o.q : o.r;
    bd.q = bd.call ? o.r : o.q;
    bd.z = o.z;
    bd.t = o.t;
    bd.none = !(o.k > 0 && o.t > 0 && o.z > 0) || !(bd.r > 0);
    if (bd.none) return;
    bd.st = sqrt(o.t);
    bd.x0 = bd.q > 0 ? fmin(1.0, bd.r / bd.q) : 1.0;
    kwgpu_option u{};  // the unit-strike put, for b_T
    u.t = o.t;
    u.k = 1.0;
    u.z = o.z;
    u.r = bd.r;
    u.q = bd.q;
    u.s = 1.0;
    u.e = 1;
    u.w = -1;
    const double bT = fmin(baseline_setup(u).ss, bd.x0);
    for (int i = 0; i <= kBn; ++i) {
        const double f = 0.5 * (1.0 - cos6(i));  // xi_i / sqrt T
        tau[i] = o.t * f * f;
        b[i] = bd.x0 * pow(bT / bd.x0, f);
    }
    b[0] = bd.x0;
}

// One fixed-point step at node tau_i (b_i its current value), with bd fitted to the
// current boundary: each quadrature point k adds its terms to the integrals' sums
// (boundary_point; er, eq = e^{r u}, e^{q u} at the point), then boundary_next forms the
// new b_i. In the log domain, ln(b_i / b(u)) = ln(b_i / X) + sqrt(H(u)), so no
// exponential or logarithm of b(u) is taken.
struct NodeStep {
    double ti, xi, f, lbi, c;  // tau_i, sqrt(tau_i), sqrt(tau_i / T), ln(b_i / X), r - q - s^2 / 2
};
__host__ __device__ inline NodeStep node_step(const Boundary& bd, double ti, double lbi) {
    NodeStep n;
    n.ti = ti;
    n.xi = sqrt(ti);
    n.f = n.xi / bd.st;
    n.lbi = lbi;
    n.c = bd.r - bd.q - 0.5 * bd.z * bd.z;
    return n;
}
__host__ __device__ inline void boundary_point(const Boundary& bd, const NodeStep& n, int k,
                                               double er, double eq, double& ni, double& di) {
    const double dt = n.ti * ts_a(kTs - 1 - k);                // tau - u
    const double sd = bd.z * n.xi * ts_sa(kTs - 1 - k);        // sigma sqrt(tau - u)
    const double lb = boundary_lnb(bd, 1.0 - 2.0 * n.f * ts_sa(k));
    const double dm = (n.lbi - lb + n.c * dt) / sd;
    ni += ts_w(k) * er * phi_cdf(dm);
    di += ts_w(k) * eq * phi_cdf(dm + sd);
}
__host__ __device__ inline void boundary_sums(const Boundary& bd, double ti, double bi, int k0,
                                              int kstep, double& ni, double& di) {
    const NodeStep n = node_step(bd, ti, log(bi / bd.x0));
    ni = 0.0;
    di = 0.0;
    for (int k = k0; k < kTs; k += kstep) {
        const double u1 = ti * ts_a(k);
        boundary_point(bd, n, k, exp(bd.r * u1), exp(bd.q * u1), ni, di);
    }
}
__host__ __device__ inline double boundary_next(const Boundary& bd, double ti, double bi,
                                                double ni, double di) {
    const double s = bd.z, r = bd.r, q = bd.q;
    const double sd0 = s * sqrt(ti);
    const double d0 = (log(bi) + (r - q - 0.5 * s * s) * ti) / sd0;
    const double v = exp(-(r - q) * ti) * (phi_cdf(d0) + r * ti * ni) /
                     (phi_cdf(d0 + sd0) + q * ti * di);
    return fmin(v, bd.x0);
}
__host__ __device__ inline double boundary_step(const Boundary& bd, double ti, double bi) {
    double ni, di;
    boundary_sums(bd, ti, bi, 0, 1, ni, di);
    return boundary_next(bd, ti, bi, ni, di);
}

__host__ __device__ inline Boundary boundary_setup(const kwgpu_option& o) {
    Boundary bd{};
    double b[kBn + 1], tau[kBn + 1];
    boundary_start(o, bd, b, tau);
    if (bd.none) return bd;
    for (int it = 0; it < kBit; ++it) {
        boundary_fit(bd, b);
        double nb[kBn + 1];
        for (int i = 1; i <= kBn; ++i) nb[i] = boundary_step(bd, tau[i], b[i]);
        for (int i = 1; i <= kBn; ++i) b[i] = nb[i];
    }
    boundary_fit(bd, b);
    return bd;
}

// boundary_setup by a group of kGroup lanes (lane j of the group; groups aligned to
// kGroup lanes, every lane of a group taking part): node i in 1..kBn is stepped by lanes
// 2i - 2 and 2i - 1, each summing every other quadrature point, and the new values are
// exchanged by shuffle before each fit. The same arithmetic as boundary_setup up to the
// order of the two half sums.
constexpr int kGroup = 16;
__device__ inline Boundary boundary_setup_group(const kwgpu_option& o, int j) {
    static_assert(2 * kBn <= kGroup, "two lanes per node");
    Boundary bd{};
    double b[kBn + 1], tau[kBn + 1];
    boundary_start(o, bd, b, tau);
    if (bd.none) return bd;  // the same for every lane of the group
    const int node = j < 2 * kBn ? 1 + j / 2 : 0;
    double ti = tau[0], bi = b[0];
#pragma unroll
    for (int i = 1; i <= kBn; ++i)
        if (node == i) {
            ti = tau[i];
            bi = b[i];
        }
    // this lane's quadrature points k = k0 + 2m and their e^{r u}, e^{q u}, which do
    // not change between steps
    constexpr int KP = (kTs + 1) / 2;
    const int k0 = j & 1;
    double er[KP], eq[KP];
#pragma unroll
    for (int m = 0; m < KP; ++m) {
        const double u1 = ti * ts_a(min(k0 + 2 * m, kTs - 1));
        er[m] = exp(bd.r * u1);
        eq[m] = exp(bd.q * u1);
    }
    // ln(b_i / X), taken by the lanes of node i and exchanged (ln(b_0 / X) = 0)
    double lb[kBn + 1];
    for (int i = 0; i <= kBn; ++i) lb[i] = log(b[i] / bd.x0);
    double lbi = log(bi / bd.x0);
    for (int it = 0; it < kBit; ++it) {
        boundary_fit_log(bd, lb);
        double ni = 0.0, di = 0.0;
        if (node) {
            const NodeStep n = node_step(bd, ti, lbi);
#pragma unroll
            for (int m = 0; m < KP; ++m)
                if (k0 + 2 * m < kTs) boundary_point(bd, n, k0 + 2 * m, er[m], eq[m], ni, di);
        }
        ni += __shfl_xor(ni, 1, kGroup);
        di += __shfl_xor(di, 1, kGroup);
        if (node) {
            bi = boundary_next(bd, ti, bi, ni, di);
            lbi = log(bi / bd.x0);
        }
#pragma unroll
        for (int i = 1; i <= kBn; ++i) lb[i] = __shfl(lbi, 2 * i - 2, kGroup);
    }
    boundary_fit_log(bd, lb);
    return bd;
}

// The premium at spot s of the option whose boundary is bd (o's own spot is not used):
// the put's strike kk and unit spot x; at or below b(T) the exercise value minus the
// European put (boundary_exercised), above it kk T times the integral's sum over the
// quadrature points (boundary_integral: the points k0, k0 + kstep, ...).
__host__ __device__ inline void boundary_spot(const kwgpu_option& o, const Boundary& bd, double s,
                                              double& kk, double& x) {
    kk = bd.call ? s : o.k;        // the put's strike
    x = (bd.call ? o.k : s) / kk;  // and its spot, per unit strike
}
__host__ __device__ inline double boundary_exercised(const Boundary& bd, double kk, double x) {
    const double r = bd.r, q = bd.q, sg = bd.z, T = bd.t;
    const double sd = sg * bd.st;
    const double dm = (log(x) + (r - q - 0.5 * sg * sg) * T) / sd;
    const double eu = exp(-r * T) * phi_cdf(-dm) - x * exp(-q * T) * phi_cdf(-dm - sd);
    return kk * ((1.0 - x) - eu);
}
__host__ __device__ inline double boundary_integral(const Boundary& bd, double x, int k0, int kstep) {
    const double r = bd.r, q = bd.q, sg = bd.z, T = bd.t;
    const double lx = log(x / bd.x0), c = r - q - 0.5 * sg * sg;
    double p = 0.0;
    for (int k = k0; k < kTs; k += kstep) {
        const double dt = T * ts_a(kTs - 1 - k);            // T - u
        const double sd = sg * bd.st * ts_sa(kTs - 1 - k);  // sigma sqrt(T - u)
        const double dm = (lx - boundary_lnb(bd, 1.0 - 2.0 * ts_sa(k)) + c * dt) / sd;
        p += ts_w(k) * (r * exp(-r * dt) * phi_cdf(-dm) - q * x * exp(-q * dt) * phi_cdf(-dm - sd));
    }
    return p;
}
__host__ __device__ inline double boundary_at(const kwgpu_option& o, const Boundary& bd, double s) {
    if (bd.none || !(s > 0)) return 0.0;
    double kk, x;
    boundary_spot(o, bd, s, kk, x);
    if (x <= boundary_bT(bd)) return boundary_exercised(bd, kk, x);
    return kk * bd.t * boundary_integral(bd, x, 0, 1);
}

__host__ __device__ inline double boundary_premium(const kwgpu_option& o) {
    return boundary_at(o, boundary_setup(o), o.s);
}

__host__ __device__ inline double premium_of(double y, double k, double base) {
    return base + (base + kEps * k) * y;
}

// Training box, as (centre, half-width) of each raw feature.
constexpr double kLogMin = -0.69314718055994530942, kLogMax = 0.91629073187415506518;  // K/S 2, 0.4
constexpr double kBox[F0][2] = {
    {(kLogMin + kLogMax) / 2, (kLogMax - kLogMin) / 2},  // log(S/K)
    {(1.0 / 52 + 5.0) / 2, (5.0 - 1.0 / 52) / 2},        // T
    {0.35, 0.25},                                          // sigma
    {0.05, 0.05},                                          // r
    {0.04, 0.04},                                          // q
    {0.0, 1.0},                                            // w
};

// Normalised features of an option priced on an x_dim x t_dim grid; true when the
// option lies in the training box (American, K/S, T, sigma, r, q inside their ranges).
// The grid enters as log2 of its point counts (x 256 .. 2048 and t 32 .. 2048 map to
// [-1, 1]); a surrogate of a single grid sees constants there.
__host__ __device__ inline bool option_features(const kwgpu_option& o, int x_dim, int t_dim,
                                                float* f) {
    const double raw[F0] = {log(o.s / o.k), o.t, o.z, o.r, o.q, o.w < 0 ? -1.0 : 1.0};
    bool in = o.e != 0 && o.s > 0 && o.k > 0;
    for (int i = 0; i < F0; ++i) {
        const double u = (raw[i] - kBox[i][0]) / kBox[i][1];
        f[i] = (float)u;
        in = in && (u >= -1.0 - 1e-12) && (u <= 1.0 + 1e-12);
    }
    f[F0] = (float)((log2((double)x_dim) - 9.5) / 1.5);
    f[F0 + 1] = (float)((log2((double)t_dim) - 8.0) / 3.0);
    return in;
}

// Crank-Nicolson does not smooth the payoff's kink at the strike: its highest-frequency
// modes there are multiplied by (1 - 2 mu) / (1 + 2 mu) ~ -exp(-1 / mu) per step, mu =
// (sigma^2 / 2) dt / dx^2 on the strike's cell, so after the n steps of the march they
// keep exp(-n / mu) of their size. The FD prices near the strike then carry a grid-scale
// oscillation that no smooth function of the option and grid follows. Such an option
// gets no labels and is never served: n / mu must be at least kDampMin (e-foldings).
// Every option of the training box has n / mu >= 10.9 on the 1024 x 1024 grid; per-option
// grids of many x-points and few steps (2048 points, 365 steps a year) go below it.
constexpr double kDampMin = 10.0;

__device__ inline double kink_damping(const PdeParams& P, double density) {
    const int jk = bracket(P, density, 0.0);
    if (jk < 1 || jk >= P.x_dim) return 0.0;
    const double dx = grid_x(P, density, jk) - grid_x(P, density, jk - 1);
    const double mu = P.axx * P.dt / (dx * dx);
    return (double)(P.t_dim - 1) / mu;
}

// The two mesh nodes that bracket log(S/K) on the option's grid (Fd1d::value's
// bracket: the first node x_j >= log(S/K)), as options at those spots, with the
// interpolation abscissae. ok is false where the network has no labels to stand on:
// an invalid option, a spot outside the grid, a bracket at the first cell (the FD
// pricer rejects a spot at x_0, so no label exists there), or an undamped kink (above).
struct NodePair {
    kwgpu_option a, b;  // spots K e^{x_{j-1}}, K e^{x_j}
    double x0, x1, xs;
    bool ok;
};

__device__ inline NodePair node_pair(const kwgpu_option& o, int x_dim, int t_dim, double density,
                                     double scale) {
    NodePair q;
    const PdeParams P = pde_params(o, x_dim, t_dim, density, scale);
    q.xs = log(o.s / o.k);
    const int j = P.valid && isfinite(q.xs) ? bracket(P, density, q.xs) : 0;
    q.ok = P.valid && j >= 2 && j < x_dim && kink_damping(P, density) >= kDampMin;
    const int jj = q.ok ? j : 2;
    q.x0 = grid_x(P, density, jj - 1);
    q.x1 = grid_x(P, density, jj);
    q.a = o;
    q.b = o;
    q.a.s = o.k * exp(q.x0);
    q.b.s = o.k * exp(q.x1);
    return q;
}

__device__ inline double euro_bs(kwgpu_option o) {
    o.e = 0;
    return bs_one(o);
}

__device__ __forceinline__ float sigm(float z) { return 1.0f / (1.0f + __expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigm(z); }
__device__ __forceinline__ float dsilu(float z) {
    const float s = sigm(z);
    return s * (1.0f + z * (1.0f - s));
}

// ---- data ------------------------------------------------------------------------
// Grid of row i: per-row sizes when given, else the uniform ones.
struct GridArg {
    const int32_t* x_dims;
    const int32_t* t_dims;
    int x_dim, t_dim;
    double density, scale;
    __device__ int x(size_t i) const { return x_dims ? x_dims[i] : x_dim; }
    __device__ int t(size_t i) const { return t_dims ? t_dims[i] : t_dim; }
};

__global__ void features_kernel(const kwgpu_option* opts, size_t n, GridArg g, const double* premium,
                                float* feat, float* target, uint8_t* in_box) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const kwgpu_option o = opts[i];
    float f[F];
    const bool in = option_features(o, g.x(i), g.t(i), f);
    for (int k = 0; k < F; ++k) feat[i * F + k] = f[k];
    if (target) {
        const double base = boundary_premium(o);
        target[i] = (float)((premium[i] - base) / (base + kEps * o.k));
    }
    if (in_box) in_box[i] = in ? 1 : 0;
}

// Member-batched kernels: the pointer arguments are the first member's; member k's
// data follows at the given element strides, k = blockIdx.y (blockIdx.z for 2-D grids).

// batch rows of member k: its permutation idx + k * sidx picks the samples
__global__ void gather_kernel(const float* feat, const float* target, const uint32_t* idx,
                              long sidx, int n, float* bfeat, float* btarget, long srow) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int k = blockIdx.y;
    const uint32_t s = idx[k * sidx + i];
    for (int f = 0; f < F; ++f) bfeat[k * srow * F + i * F + f] = feat[(size_t)s * F + f];
    btarget[k * srow + i] = target[s];
}

// Labels at nodes: option i gives node options a_i (row i) and b_i (row n + i) with the
// option's grid sizes; ok[i] = 0 where node_pair has no labelled bracket.
__global__ void node_options_kernel(const kwgpu_option* opts, size_t n, GridArg g,
                                    kwgpu_option* nodes, int32_t* nx, int32_t* nt, uint8_t* ok) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const NodePair q = node_pair(opts[i], g.x(i), g.t(i), g.density, g.scale);
    nodes[i] = q.a;
    nodes[n + i] = q.b;
    if (nx) nx[i] = nx[n + i] = g.x(i);
    if (nt) nt[i] = nt[n + i] = g.t(i);
    ok[i] = q.ok ? 1 : 0;
}

// premium = FD price - BS European price at each node of pair i; keep[i] when the pair
// has a labelled bracket and both nodes priced to a finite premium
__global__ void node_premium_kernel(const kwgpu_option* nodes, size_t n, const double* fd,
                                    const int32_t* st, const uint8_t* ok, double* premium,
                                    uint8_t* keep) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double pa = fd[i] - euro_bs(nodes[i]);
    const double pb = fd[n + i] - euro_bs(nodes[n + i]);
    premium[i] = pa;
    premium[n + i] = pb;
    keep[i] = ok[i] && st[i] == KWGPU_STATUS_OK && st[n + i] == KWGPU_STATUS_OK && isfinite(pa) &&
                      isfinite(pb)
                  ? 1
                  : 0;
}

// Prediction rows: node a of option i in row i, node b in row cn + i; in_box[i] when the
// option is in the training box and has a labelled bracket.
__global__ void node_features_kernel(const kwgpu_option* opts, size_t cn, GridArg g, float* feat,
                                     uint8_t* in_box) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cn) return;
    const kwgpu_option o = opts[i];
    const int xd = g.x(i), td = g.t(i);
    const NodePair q = node_pair(o, xd, td, g.density, g.scale);
    float f[F];
    const bool in = option_features(o, xd, td, f);
    option_features(q.a, xd, td, feat + i * F);
    option_features(q.b, xd, td, feat + (cn + i) * F);
    if (in_box) in_box[i] = in && q.ok ? 1 : 0;
}

// Ensemble combine in price units, as the FD pricer interpolates: member m's node values
// v = BS + premium at x_{j-1} and x_j, its price ((x_j - x) v_{j-1} + (x - x_{j-1}) v_j) /
// (x_j - x_{j-1}); premium[i] = mean price - BS at the spot, spread[i] = max - min price.
// Candidates' prices from their node values: kGroup lanes per candidate (launched with
// kGroup * cn threads in 64-thread blocks). The group solves the candidate's exercise
// boundary; lanes 0-7 then price the control variate at the first node and lanes 8-15
// at the second, each lane summing every eighth quadrature point; lane 0 combines.
__global__ __launch_bounds__(64) void node_combine_kernel(const kwgpu_option* opts, size_t cn,
                                                          GridArg g, const float* y, int members,
                                                          size_t ld, double* premium,
                                                          double* spread) {
    static_assert(kGroup == 16, "two nodes x 8 lanes");
    const size_t gi = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t i = gi / kGroup;
    const int j = (int)(gi % kGroup);
    if (i >= cn) return;  // whole groups: cn counts groups
    const kwgpu_option o = opts[i];
    const NodePair q = node_pair(o, g.x(i), g.t(i), g.density, g.scale);
    if (!q.ok) {  // the same for the whole group
        if (j == 0) {
            premium[i] = NAN;
            spread[i] = INFINITY;
        }
        return;
    }
    const Boundary bd = boundary_setup_group(o, j);  // the nodes differ from o only in spot
    double bw = 0.0;
    if (!bd.none) {
        double kk, x;
        boundary_spot(o, bd, j < 8 ? q.a.s : q.b.s, kk, x);
        const bool ex = x <= boundary_bT(bd);
        double part = ex ? 0.0 : boundary_integral(bd, x, j & 7, 8);
        part += __shfl_xor(part, 1, kGroup);
        part += __shfl_xor(part, 2, kGroup);
        part += __shfl_xor(part, 4, kGroup);
        bw = ex ? boundary_exercised(bd, kk, x) : kk * bd.t * part;
    }
    const double bwb = __shfl(bw, 8, kGroup);
    if (j != 0) return;
    const double bwa = bw;
    const double bsa = euro_bs(q.a), bsb = euro_bs(q.b);
    const double ua = q.x1 - q.xs, ub = q.xs - q.x0, h = q.x1 - q.x0;
    double sum = 0, lo = 1e300, hi = -1e300;
    for (int m = 0; m < members; ++m) {
        const double va = bsa + premium_of(y[m * ld + i], o.k, bwa);
        const double vb = bsb + premium_of(y[m * ld + cn + i], o.k, bwb);
        const double v = (ua * va + ub * vb) / h;
        sum += v;
        lo = fmin(lo, v);
        hi = fmax(hi, v);
    }
    premium[i] = sum / members - euro_bs(o);
    spread[i] = hi - lo;
}

__global__ void baseline_kernel(const kwgpu_option* opts, size_t n, double* out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = baseline_premium(opts[i]);
}

// ---- layers ----------------------------------------------------------------------
// input layer in FP32: Z1 = x W1^T + b1, H1 = silu(Z1) in BF16
__global__ void l1_fwd(const float* x, long sx, const float* p, int n, float* Z, bf16* Hb,
                       long sz) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * H) return;
    const int k = blockIdx.y;
    x += k * sx;
    p += (long)k * kP;
    Hb += k * sz;
    const int m = i / H, j = i - m * H;
    float z = p[kB1 + j];
#pragma unroll
    for (int f = 0; f < F; ++f) z = fmaf(p[kW1 + j * F + f], x[m * F + f], z);
    if (Z) Z[k * sz + i] = z;  // prediction keeps no pre-activations
    Hb[i] = (bf16)silu(z);
}

enum { EPI_FWD = 0, EPI_BWD = 1, EPI_WGRAD = 2 };

struct GemmArgs {
    const bf16* A;  // A(m, k) = A[m * sam + k * sak]
    long sam, sak;
    const bf16* B;  // B(k, n) = B[k * sbk + n * sbn]
    long sbk, sbn;
    int M, N, K;
    const float* bias;  // EPI_FWD: bias[n]
    float* Z;           // EPI_FWD: writes pre-activations (unless null); EPI_BWD: reads them
    bf16* Hout;         // EPI_FWD: silu(Z); EPI_BWD: dZ = C * silu'(Z)
    float* C;           // EPI_WGRAD: C[z][m * N + n] = sum over this block's K range, one
                        // M x N slot per blockIdx.z (member, split), summed afterwards
    int k_split;        // EPI_WGRAD: K range per split
    int splits;         // EPI_WGRAD: K splits; blockIdx.z = member * splits + split
    long mA, mB, mBias, mZ, mH;  // member strides (elements)
};

// C(M x N) = A(M x K) B(K x N), 64 x 64 tile per 256-thread workgroup, each wave a
// 32 x 32 quarter as 2 x 2 MFMA 16x16x32 tiles; K in steps of BK through LDS.
template <int EPI, int BK>
__global__ __launch_bounds__(256) void gemm_bf16(GemmArgs g) {
    constexpr int BM = 64, BN = 64, LD = BK + 8;
    static_assert(BK % 32 == 0 && (EPI != EPI_WGRAD || BK == 32),
                  "k staged in 32-wide chunks; the transposing loader stages 32");
    __shared__ __attribute__((aligned(16))) bf16 As[BM * LD];
    __shared__ __attribute__((aligned(16))) bf16 Bs[BN * LD];
    const int t = threadIdx.x, w = t >> 6, l = t & 63;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    int kb = 0, ke = g.K;
    const int mem = EPI == EPI_WGRAD ? blockIdx.z / g.splits : blockIdx.z;
    if constexpr (EPI == EPI_WGRAD) {
        kb = (blockIdx.z - mem * g.splits) * g.k_split;
        ke = min(g.K, kb + g.k_split);
    }
    g.A += mem * g.mA;
    g.B += mem * g.mB;
    if constexpr (EPI == EPI_FWD) g.bias += mem * g.mBias;
    if constexpr (EPI != EPI_WGRAD) {
        if (g.Z) g.Z += mem * g.mZ;
        g.Hout += mem * g.mH;
    } else {
        g.C += (long)blockIdx.z * g.M * g.N;
    }
    const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
    floatx4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
    const int lr = t >> 2, lc = (t & 3) * 8;  // tile row 0..63, k offset 0/8/16/24
    for (int k0 = kb; k0 < ke; k0 += BK) {
        if constexpr (EPI == EPI_WGRAD) {
            // both operands have K (the batch) across rows and M / N along them: load 16
            // bytes along the row and transpose into the K-contiguous LDS tiles
            const int kk = t >> 3, c8 = (t & 7) * 8;  // k row 0..31, column group 0..56
            const int k = k0 + kk;
            bf16x8 va, vb;
            if (g.sam == 1 && g.sbn == 1 && k < ke && m0 + c8 + 8 <= g.M && n0 + c8 + 8 <= g.N) {
                va = *reinterpret_cast<const bf16x8*>(g.A + (long)k * g.sak + m0 + c8);
                vb = *reinterpret_cast<const bf16x8*>(g.B + (long)k * g.sbk + n0 + c8);
            } else {
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const int m = m0 + c8 + i, n = n0 + c8 + i;
                    va[i] = (k < ke && m < g.M) ? g.A[(long)m * g.sam + (long)k * g.sak] : (bf16)0.0f;
                    vb[i] = (k < ke && n < g.N) ? g.B[(long)k * g.sbk + (long)n * g.sbn] : (bf16)0.0f;
                }
            }
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                As[(c8 + i) * LD + kk] = va[i];
                Bs[(c8 + i) * LD + kk] = vb[i];
            }
        } else {
            // BK / 32 chunks of 16 bytes per thread and operand, all loads issued before
            // the first LDS store
            constexpr int KC = BK / 32;
            bf16x8 va[KC], vb[KC];
#pragma unroll
            for (int c = 0; c < KC; ++c) {
                const int m = m0 + lr, n = n0 + lr, k = k0 + 32 * c + lc;
                if (g.sak == 1 && m < g.M && k + 8 <= ke) {
                    va[c] = *reinterpret_cast<const bf16x8*>(g.A + (long)m * g.sam + k);
                } else {
#pragma unroll
                    for (int i = 0; i < 8; ++i)
                        va[c][i] = (m < g.M && k + i < ke)
                                       ? g.A[(long)m * g.sam + (long)(k + i) * g.sak]
                                       : (bf16)0.0f;
                }
                if (g.sbk == 1 && n < g.N && k + 8 <= ke) {
                    vb[c] = *reinterpret_cast<const bf16x8*>(g.B + (long)n * g.sbn + k);
                } else {
#pragma unroll
                    for (int i = 0; i < 8; ++i)
                        vb[c][i] = (n < g.N && k + i < ke)
                                       ? g.B[(long)(k + i) * g.sbk + (long)n * g.sbn]
                                       : (bf16)0.0f;
                }
            }
#pragma unroll
            for (int c = 0; c < KC; ++c) {
                *reinterpret_cast<bf16x8*>(&As[lr * LD + 32 * c + lc]) = va[c];
                *reinterpret_cast<bf16x8*>(&Bs[lr * LD + 32 * c + lc]) = vb[c];
            }
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < BK; kk += 32)
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const bf16x8 a = *reinterpret_cast<const bf16x8*>(
                    &As[(wm + 16 * i + (l & 15)) * LD + kk + 8 * (l >> 4)]);
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    const bf16x8 b = *reinterpret_cast<const bf16x8*>(
                        &Bs[(wn + 16 * j + (l & 15)) * LD + kk + 8 * (l >> 4)]);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[i][j], 0, 0, 0);
                }
            }
        __syncthreads();
    }
    // D layout of v_mfma_f32_16x16x32: lane l, element e -> row 4 (l / 16) + e, col l % 16
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int m = m0 + wm + 16 * i + 4 * (l >> 4) + e;
                const int n = n0 + wn + 16 * j + (l & 15);
                if (m >= g.M || n >= g.N) continue;
                const float v = acc[i][j][e];
                const long o = (long)m * g.N + n;
                if constexpr (EPI == EPI_FWD) {
                    const float z = v + g.bias[n];
                    if (g.Z) g.Z[o] = z;
                    g.Hout[o] = (bf16)silu(z);
                } else if constexpr (EPI == EPI_BWD) {
                    g.Hout[o] = (bf16)(v * dsilu(g.Z[o]));
                } else {
                    g.C[o] = v;
                }
            }
}

// output layer: y = silu(Z4) W5 + b5; with targets also dy = 2 (y - t) / n and the loss
// sum. kOutSPW samples per wave; the loss is reduced in the block before one atomic.
constexpr int kOutSPW = 8;
__global__ __launch_bounds__(256) void out_fwd(const float* Z4, long sz, const float* p, int n,
                                               const float* target, float* y, long sy, float* dy,
                                               long srow, float* loss_sum) {
    __shared__ float part[4];
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
    {
        const int k = blockIdx.y;
        Z4 += k * sz;
        p += (long)k * kP;
        y += k * sy;
        if (target) {
            target += k * srow;
            dy += k * srow;
            loss_sum += k;
        }
    }
    const int m0 = (blockIdx.x * 4 + wave) * kOutSPW;
    float loss = 0.0f;
    for (int k = 0; k < kOutSPW; ++k) {
        const int m = m0 + k;
        if (m >= n) break;  // uniform per wave
        float s = 0.0f;
#pragma unroll
        for (int c = l; c < H; c += 64) s = fmaf(p[kW5 + c], silu(Z4[(long)m * H + c]), s);
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
        if (l == 0) {
            const float v = s + p[kB5];
            y[m] = v;
            if (target) {
                const float d = v - target[m];
                dy[m] = 2.0f * d / (float)n;
                loss += d * d;
            }
        }
    }
    if (l == 0) part[wave] = loss;
    __syncthreads();
    if (target && threadIdx.x == 0) atomicAdd(loss_sum, part[0] + part[1] + part[2] + part[3]);
}

// ---- prediction: the whole member MLP in one kernel ----------------------------------
// kPredRows node rows of one member (blockIdx.y) per 256-thread workgroup. The input
// layer (FP32) writes silu(Z1) in BF16 to LDS; each 256 x 256 layer then runs on MFMA
// 16x16x32, wave w computing output columns 64w .. 64w + 63 of every row of the tile
// with its weights W[n][k] read from L2; once every wave has read the layer's input,
// bias and SiLU are applied to the accumulators and overwrite it in LDS. The output
// layer is reduced from the last layer's accumulators. The arithmetic is forward()'s --
// the same FP32 input layer, BF16 activations and weights, the same MFMA sequence over
// k with FP32 accumulation, FP32 pre-activations of the last layer -- up to the order
// of the output layer's FP32 sum, but no activation goes to HBM: prediction moved
// 0.15 GB per 25k node rows of a 4-member ensemble through the separate kernels.
constexpr int kPredRows = 128;
constexpr int kPredLd = H + 8;  // LDS row pitch in BF16 elements: rows 16 bytes apart mod 128

__global__ __attribute__((amdgpu_flat_work_group_size(256, 256), amdgpu_waves_per_eu(2, 2)))
void mlp_predict(const float* __restrict__ x, const float* __restrict__ pall,
                 const bf16* __restrict__ wball, int n, float* __restrict__ yall, long sy) {
    static_assert(H == 256 && kPredRows % 16 == 0, "four waves of 64 output columns");
    constexpr int RT = kPredRows / 16;  // 16-row MFMA tiles
    __shared__ __attribute__((aligned(16))) bf16 act[kPredRows * kPredLd];
    __shared__ float xs[kPredRows * F];
    __shared__ float red[4][kPredRows];
    const int t = threadIdx.x, w = t >> 6, l = t & 63;
    const int mem = blockIdx.y;
    const float* p = pall + (long)mem * kP;
    const bf16* wb = wball + (long)mem * 3 * H * H;
    const int r0 = blockIdx.x * kPredRows;
    for (int i = t; i < kPredRows * F; i += 256)  // rows past n read as 0
        xs[i] = r0 + i / F < n ? x[(long)r0 * F + i] : 0.0f;
    __syncthreads();
    {  // input layer, as l1_fwd: thread t computes hidden unit t of every row
        float w1[F];
#pragma unroll
        for (int f = 0; f < F; ++f) w1[f] = p[kW1 + t * F + f];
        const float b1 = p[kB1 + t];
        for (int r = 0; r < kPredRows; ++r) {
            float z = b1;
#pragma unroll
            for (int f = 0; f < F; ++f) z = fmaf(w1[f], xs[r * F + f], z);
            act[r * kPredLd + t] = (bf16)silu(z);
        }
    }
    __syncthreads();
    const int col = 64 * w + (l & 15);  // + 16 j: this lane's output columns
    for (int layer = 0; layer < 3; ++layer) {
        const bf16* W = wb + (long)layer * H * H;
        floatx4 acc[RT][4];
#pragma unroll
        for (int i = 0; i < RT; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
        // k in the 32-wide steps of gemm_bf16, so each accumulator sees the same MFMAs
#pragma unroll 2
        for (int k0 = 0; k0 < H; k0 += 32) {
            bf16x8 b[4];
#pragma unroll
            for (int j = 0; j < 4; ++j)
                b[j] = *reinterpret_cast<const bf16x8*>(W + (long)(col + 16 * j) * H + k0 +
                                                        8 * (l >> 4));
#pragma unroll
            for (int i = 0; i < RT; ++i) {
                const bf16x8 a = *reinterpret_cast<const bf16x8*>(
                    &act[(16 * i + (l & 15)) * kPredLd + k0 + 8 * (l >> 4)]);
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[j], acc[i][j], 0, 0, 0);
            }
        }
        __syncthreads();  // every wave has read this layer's input
        const float* bias = p + kB2 + layer * (H * H + H);
        // D layout of v_mfma_f32_16x16x32: lane l, element e -> row 4 (l / 16) + e, col l % 16
        if (layer < 2) {
#pragma unroll
            for (int i = 0; i < RT; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const float bj = bias[col + 16 * j];
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        act[(16 * i + 4 * (l >> 4) + e) * kPredLd + col + 16 * j] =
                            (bf16)silu(acc[i][j][e] + bj);
                }
            __syncthreads();
        } else {
            // output layer: y = b5 + sum_c W5[c] silu(Z4[c]); this lane's four columns,
            // then the 16 lanes of a row group, then the four waves
            float w5[4], bj[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                w5[j] = p[kW5 + col + 16 * j];
                bj[j] = bias[col + 16 * j];
            }
#pragma unroll
            for (int i = 0; i < RT; ++i)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    float s = 0.0f;
#pragma unroll
                    for (int j = 0; j < 4; ++j) s = fmaf(w5[j], silu(acc[i][j][e] + bj[j]), s);
#pragma unroll
                    for (int o = 1; o < 16; o <<= 1) s += __shfl_xor(s, o);
                    if ((l & 15) == 0) red[w][16 * i + 4 * (l >> 4) + e] = s;
                }
            __syncthreads();
            if (t < kPredRows && r0 + t < n)
                yall[mem * sy + r0 + t] = p[kB5] + ((red[0][t] + red[1][t]) + (red[2][t] + red[3][t]));
        }
    }
}

// dZ4 = dy W5 silu'(Z4) (BF16 for the next GEMMs)
__global__ void out_bwd(const float* Z4, long sz, const float* p, const float* dy, long srow, int n,
                        bf16* dZ, long sdz) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * H) return;
    const int mem = blockIdx.y;
    Z4 += mem * sz;
    p += (long)mem * kP;
    dy += mem * srow;
    dZ += mem * sdz;
    const int m = i / H, c = i - m * H;
    dZ[i] = (bf16)(dy[m] * p[kW5 + c] * dsilu(Z4[i]));
}

// grad[kB? + n] += sum_m dZ[m][n]   and, for the output layer,
// grad[kW5 + n] += sum_m dy[m] silu(Z4[m][n]), grad[kB5] += sum_m dy[m]
// The column reductions below: 256 threads = 64 columns x 4 row phases; the four
// phases are summed in LDS, and each block stores one partial sum per column into its
// own slot of `part` (added up by sum_parts).
__device__ __forceinline__ float phase_sum(float v, float* lds) {
    lds[threadIdx.x] = v;
    __syncthreads();
    const int c = threadIdx.x & 63;
    return lds[c] + lds[64 + c] + lds[128 + c] + lds[192 + c];
}
__global__ __launch_bounds__(256) void colsum_bf16(const bf16* dZ, long sdz, int n, float* part) {
    __shared__ float lds[256];
    dZ += blockIdx.z * sdz;
    part += ((long)blockIdx.z * gridDim.y + blockIdx.y) * H;  // this block's column sums
    const int col = blockIdx.x * 64 + (threadIdx.x & 63);
    const int r0 = blockIdx.y * 4 + (threadIdx.x >> 6);
    float s = 0.0f;
#pragma unroll 4
    for (int m = r0; m < n; m += 4 * gridDim.y) s += (float)dZ[(long)m * H + col];
    s = phase_sum(s, lds);
    if (threadIdx.x < 64) part[col] = s;
}
__global__ __launch_bounds__(256) void out_grad(const float* Z4, long sz, const float* dy,
                                                long srow, int n, float* part) {
    __shared__ float lds[256];
    Z4 += blockIdx.z * sz;
    dy += blockIdx.z * srow;
    part += ((long)blockIdx.z * gridDim.y + blockIdx.y) * (H + 1);  // this block's W5 | b5 sums
    const int col = blockIdx.x * 64 + (threadIdx.x & 63);
    const int r0 = blockIdx.y * 4 + (threadIdx.x >> 6);
    float s = 0.0f, sb = 0.0f;
    for (int m = r0; m < n; m += 4 * gridDim.y) {
        s = fmaf(dy[m], silu(Z4[(long)m * H + col]), s);
        if (col == 0) sb += dy[m];
    }
    s = phase_sum(s, lds);
    __syncthreads();
    sb = phase_sum(sb, lds);
    if (threadIdx.x < 64) {
        part[col] = s;
        if (col == 0) part[H] = sb;
    }
}
// input layer gradients: grad W1[j][f] += sum_m dZ1[m][j] x[m][f], b1[j] += sum_m dZ1[m][j]
// Layer-1 gradient: seven sums per hidden column (six W1 weights and the bias). Each
// lane owns four adjacent columns, so a wave reads one whole 512-byte dZ1 row per
// iteration and the row's six inputs once (wave-uniform); the four waves' partial sums
// meet in LDS and each block stores its 1792 sums into its own slot of `part`.
typedef bf16 bf16x4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void l1_grad(const bf16* dZ1, long sdz, const float* x, long sx,
                                               int n, float* part) {
    static_assert(H == 256, "one wave covers a row as 64 lanes x 4 columns");
    constexpr int V = 4 * (F + 1);
    __shared__ float lds[V * 256];
    dZ1 += blockIdx.y * sdz;
    x += blockIdx.y * sx;
    part += ((long)blockIdx.y * gridDim.x + blockIdx.x) * (H * F + H);  // W1 | b1 order
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    float s[V] = {};
    // U rows per wave per pass, all loads issued before any use (one wave per SIMD
    // here, so the loads' latency is hidden only by having several in flight)
    constexpr int U = 8;
    const int stride = 4 * gridDim.x;
    for (int m0 = blockIdx.x * 4 + w; m0 < n; m0 += U * stride) {
        bf16x4 d4[U];
        float xf[U][F];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            // rows past the end load the last row (no branch around the loads) and
            // contribute d = 0
            const int m = m0 + u * stride;
            const int mc = min(m, n - 1);
            const bf16x4 v = *(const bf16x4*)(dZ1 + (long)mc * H + 4 * lane);
            d4[u] = m < n ? v : bf16x4{};
#pragma unroll
            for (int f = 0; f < F; ++f) xf[u][f] = x[mc * F + f];
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float d = (float)d4[u][j];
#pragma unroll
                for (int f = 0; f < F; ++f) s[j * (F + 1) + f] = fmaf(d, xf[u][f], s[j * (F + 1) + f]);
                s[j * (F + 1) + F] += d;
            }
    }
#pragma unroll
    for (int v = 0; v < V; ++v) lds[v * 256 + threadIdx.x] = s[v];
    __syncthreads();
    for (int o = threadIdx.x; o < V * 64; o += 256) {
        const int v = o >> 6, l = o & 63;
        const float t = lds[v * 256 + l] + lds[v * 256 + 64 + l] + lds[v * 256 + 128 + l] +
                        lds[v * 256 + 192 + l];
        const int col = 4 * l + v / (F + 1), f = v % (F + 1);
        part[f < F ? col * F + f : H * F + col] = t;
    }
}

// grad[member][offset + o] += sum over the nb blocks' partial sums part[member][b][o]: the
// second stage of the column reductions above. Many blocks adding into the same few
// hundred addresses with device-scope atomics serialise on them; this reads each
// partial once, coalesced.
// 64 outputs per 256-thread block: the four waves take every fourth partial, 4 loads in
// flight each, and meet in LDS.
__global__ __launch_bounds__(256) void sum_parts(const float* part, int nb, int width,
                                                 float* grad, int offset) {
    __shared__ float lds[256];
    const int o = blockIdx.x * 64 + (threadIdx.x & 63), ph = threadIdx.x >> 6;
    const int oc = min(o, width - 1);  // lanes past the end load a valid slot, write nothing
    part += (long)blockIdx.y * nb * width + oc;
    float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, s3 = 0.0f;
    int b = ph;
    for (; b + 12 < nb; b += 16) {
        s0 += part[(long)b * width];
        s1 += part[(long)(b + 4) * width];
        s2 += part[(long)(b + 8) * width];
        s3 += part[(long)(b + 12) * width];
    }
    for (; b < nb; b += 4) s0 += part[(long)b * width];
    const float t = phase_sum((s0 + s1) + (s2 + s3), lds);
    if (threadIdx.x < 64 && o < width) grad[(long)blockIdx.y * kP + offset + o] += t;
}

__global__ void adam_kernel(float* p, const float* g, float* m1, float* m2, int np, float lr,
                            float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const long o = (long)blockIdx.y * kP;  // member
    p += o;
    g += o;
    m1 += o;
    m2 += o;
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.0f - b1) * gi;
    const float v = b2 * m2[i] + (1.0f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = v;
    p[i] -= lr * (a / c1) / (sqrtf(v / c2) + eps);
}

// BF16 copies of W2..W4 and of their transposes: blockIdx.y = 3 * member + layer
// (one 16 x 16 tile per 256-thread block; the transpose goes through LDS so that both
// copies are written in 32-byte row segments)
__global__ __launch_bounds__(256) void to_bf16(const float* p, bf16* w, bf16* wT) {
    static_assert(H % 16 == 0, "16 x 16 tiles");
    __shared__ bf16 tile[16][17];
    const int t = threadIdx.x, tr = blockIdx.x / (H / 16), tc = blockIdx.x % (H / 16);
    const int k = blockIdx.y / 3, l = blockIdx.y - 3 * k;
    const long o = (long)k * 3 * H * H + (long)l * H * H;
    const int r = tr * 16 + (t >> 4), c = tc * 16 + (t & 15);
    const bf16 v = (bf16)p[(long)k * kP + kW2 + l * (H * H + H) + r * H + c];
    w[o + r * H + c] = v;
    tile[t >> 4][t & 15] = v;
    __syncthreads();
    wT[o + (long)(tc * 16 + (t >> 4)) * H + tr * 16 + (t & 15)] = tile[t & 15][t >> 4];  // WT[c][r]
}

__global__ void zero_f(float* p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

// ---- serving ---------------------------------------------------------------------
// The option's class in a gate: put/call, then its maturity band, then its band of
// log(S/K) (kwgpu_gate in kwgpu.h).
__device__ __forceinline__ int gate_class(const kwgpu_gate& g, const kwgpu_option& o) {
    int a = 0;
    while (a < g.n_t - 1 && o.t >= g.t_edges[a]) ++a;
    const double m = log(o.s / o.k);
    int b = 0;
    while (b < g.n_m - 1 && m >= g.m_edges[b]) ++b;
    return ((o.w < 0 ? 0 : 1) * g.n_t + a) * g.n_m + b;
}

// Route 0: surrogate (in the box, ensemble spread within its class's spread bound): price
// = BS + premium; premium == nullptr (no prediction made) routes every option to FD.
// Route 1: the FD fallback at its own precision. Route 2: FP64 FD, when the fallback is
// FP32-mixed, for options beyond the mixed march's limits (mixed_overflow_risk,
// mixed_drift_risk) at mesh scale mixed_scale (0: the fallback is FP64).
// Each block also counts its options of routes 1 and 2; the compaction below places
// them in index order, so a batch routed the same way gives the same FD subset on
// every call.
constexpr int kGateBlock = 256;

__global__ __launch_bounds__(kGateBlock) void gate_kernel(
    const kwgpu_option* opts, size_t n, const double* premium, const double* spread,
    const uint8_t* in_box, const kwgpu_gate gate, double mixed_scale, double* prices,
    int32_t* route, uint32_t* blk_counts) {
    __shared__ uint32_t wc[2][kGateBlock / 64];
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    int r = -1;
    if (i < n) {
        const kwgpu_option o = opts[i];
        double price = NAN;
        if (premium && in_box[i] && spread[i] <= gate.spread_max[gate_class(gate, o)] &&
            isfinite(premium[i])) {
            kwgpu_option e = o;
            e.e = 0;
            price = bs_one(e) + premium[i];
            r = isfinite(price) ? 0 : 1;
        } else {
            r = 1;
        }
        if (r != 0)
            r = mixed_scale > 0.0 &&
                        (mixed_overflow_risk(o, mixed_scale) || mixed_drift_risk(o, mixed_scale))
                    ? 2
                    : 1;
        prices[i] = price;
        route[i] = r;
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t v1 = __ballot(r == 1), v2 = __ballot(r == 2);
    if (lane == 0) {
        wc[0][w] = (uint32_t)__popcll(v1);
        wc[1][w] = (uint32_t)__popcll(v2);
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        uint32_t c = 0;
        for (int q = 0; q < kGateBlock / 64; ++q) c += wc[threadIdx.x][q];
        blk_counts[2 * blockIdx.x + threadIdx.x] = c;
    }
}

// Exclusive prefix sums of the per-block counts (in place, one workgroup) and the two
// totals.
__global__ __launch_bounds__(1024) void scan_block_counts(uint32_t* blk_counts, unsigned nb,
                                                          int* totals) {
    __shared__ uint32_t s1[1024], s2[1024];
    const unsigned t = threadIdx.x;
    const unsigned per = (nb + 1023) / 1024;
    const unsigned b0 = min(nb, t * per), b1 = min(nb, b0 + per);
    uint32_t a1 = 0, a2 = 0;
    for (unsigned b = b0; b < b1; ++b) {
        a1 += blk_counts[2 * b];
        a2 += blk_counts[2 * b + 1];
    }
    s1[t] = a1;
    s2[t] = a2;
    __syncthreads();
    for (unsigned d = 1; d < 1024; d <<= 1) {
        const uint32_t x1 = t >= d ? s1[t - d] : 0, x2 = t >= d ? s2[t - d] : 0;
        __syncthreads();
        s1[t] += x1;
        s2[t] += x2;
        __syncthreads();
    }
    uint32_t o1 = s1[t] - a1, o2 = s2[t] - a2;
    for (unsigned b = b0; b < b1; ++b) {
        const uint32_t c1 = blk_counts[2 * b], c2 = blk_counts[2 * b + 1];
        blk_counts[2 * b] = o1;
        blk_counts[2 * b + 1] = o2;
        o1 += c1;
        o2 += c2;
    }
    if (t == 1023) {
        totals[0] = (int)s1[1023];
        totals[1] = (int)s2[1023];
    }
}

// The indices of routes 1 and 2 in index order, from the scanned block offsets: route
// 1's first, then route 2's (from idx + totals[0]).
__global__ __launch_bounds__(kGateBlock) void compact_routes(const int32_t* route, size_t n,
                                                             const uint32_t* blk_off,
                                                             const int* totals, uint32_t* idx) {
    __shared__ uint32_t wc[2][kGateBlock / 64];
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int r = i < n ? route[i] : 0;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t v1 = __ballot(r == 1), v2 = __ballot(r == 2);
    if (lane == 0) {
        wc[0][w] = (uint32_t)__popcll(v1);
        wc[1][w] = (uint32_t)__popcll(v2);
    }
    __syncthreads();
    if (r == 1 || r == 2) {
        const int k = r - 1;
        uint32_t base = blk_off[2 * blockIdx.x + k];
        for (int q = 0; q < w; ++q) base += wc[k][q];
        const uint64_t below = (1ull << lane) - 1;
        const uint32_t at = (k == 0 ? 0u : (uint32_t)totals[0]) + base;
        idx[at + __popcll((k == 0 ? v1 : v2) & below)] = (uint32_t)i;
    }
}

// Serving candidates: in the training box, with a labelled bracket, in a class the gate
// leaves open, and with a baseline scale (the Ju & Zhong premium + 1e-4 K, about the
// price scale of the network's error) of at most that class's scale bound; flag[i] = 1
// for a candidate, else 0. Launched in blocks of kCandBlock: an option's work (the
// critical spot) is long and a batch can be a few ten thousand options, so small blocks
// spread it over every CU (in 256-thread blocks a 32768-option batch left half the CUs
// idle).
constexpr int kCandBlock = 64;

__global__ __launch_bounds__(kCandBlock) void candidate_kernel(const kwgpu_option* opts, size_t n,
                                                               GridArg g, const kwgpu_gate gate,
                                                               int32_t* flag) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const kwgpu_option o = opts[i];
    float f[F];
    const bool in = option_features(o, g.x(i), g.t(i), f);
    const int k = gate_class(gate, o);
    int c = 0;
    if (in && gate.spread_max[k] >= 0.0 && node_pair(o, g.x(i), g.t(i), g.density, g.scale).ok)
        c = baseline_premium(o) + kEps * o.k <= gate.scale_max[k] ? 1 : 0;
    flag[i] = c;
}

// Per-block counts of the candidates (blocks of kGateBlock options), as gate_kernel keeps
// its route counts, so that compact_routes lists the candidates in index order.
__global__ __launch_bounds__(kGateBlock) void count_candidates(const int32_t* flag, size_t n,
                                                               uint32_t* blk_counts) {
    __shared__ uint32_t wc[kGateBlock / 64];
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = i < n ? flag[i] : 0;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t v = __ballot(c == 1);
    if (lane == 0) wc[w] = (uint32_t)__popcll(v);
    __syncthreads();
    if (threadIdx.x < 2) {
        uint32_t t = 0;
        if (threadIdx.x == 0)
            for (int q = 0; q < kGateBlock / 64; ++q) t += wc[q];
        blk_counts[2 * blockIdx.x + threadIdx.x] = t;
    }
}

// No prediction (premium NaN, in_box 0) for every option; then the candidates' own.
__global__ void clear_prediction(size_t n, double* premium, double* spread, uint8_t* in_box) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    premium[i] = NAN;
    spread[i] = INFINITY;
    in_box[i] = 0;
}

__global__ void scatter_prediction(const uint32_t* idx, int nc, const double* pc, const double* sc,
                                   const uint8_t* ic, double* premium, double* spread,
                                   uint8_t* in_box) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nc) return;
    const uint32_t i = idx[k];
    premium[i] = pc[k];
    spread[i] = sc[k];
    in_box[i] = ic[k];
}

__global__ void gather_opts(const kwgpu_option* opts, const uint32_t* idx, int n, kwgpu_option* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = opts[idx[i]];
}

__global__ void gather_dims(const int32_t* dims, const uint32_t* idx, int n, int32_t* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = dims[idx[i]];
}

__global__ void scatter_prices(const double* sub, const int32_t* sub_status, const uint32_t* idx,
                               int n, double* prices) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) prices[idx[i]] = sub_status[i] == KWGPU_STATUS_OK ? sub[i] : NAN;
}

// Every option priced by FD in place (a gate that serves nothing): route 1, and NaN for
// an option whose FD status is not OK, as scatter_prices leaves it.
__global__ void all_to_fd(size_t n, const int32_t* status, double* prices, int32_t* route) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    route[i] = 1;
    if (status[i] != KWGPU_STATUS_OK) prices[i] = NAN;
}

}  // namespace sur
}  // namespace kwgpu

// ---------------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------------
using namespace kwgpu::sur;

struct kwgpu_surrogate {
    int members = 0;
    int device = 0;
    float* pall = nullptr;             // FP32 master weights, member-major (kP each)
    bf16* wball = nullptr;             // BF16 W2 | W3 | W4, member-major (3 H^2 each)
    bf16* wballT = nullptr;            // ... and their transposes (backward-data GEMMs)
    std::vector<float*> p;             // pall + m kP
    std::vector<bf16*> wb;             // wball + 3 m H^2
    std::vector<bf16*> wbT;            // wballT + 3 m H^2
    float *grad = nullptr, *m1 = nullptr, *m2 = nullptr, *loss = nullptr;  // per member
    // per-member scratch slots of `cap` rows each (grown on demand)
    size_t cap = 0;
    float *x = nullptr, *tg = nullptr, *y = nullptr, *dy = nullptr, *Z = nullptr;
    bf16 *Hb = nullptr, *dZa = nullptr, *dZb = nullptr;
    uint32_t* idx = nullptr;
    float* yall = nullptr;  // prediction: the members' outputs, 2 rows per option
    size_t yall_cap = 0;
    float* xp = nullptr;    // prediction: node features, 2 rows per option
    size_t xp_cap = 0;
    float* part = nullptr;  // per-block partial sums of the column reductions (kPartRows
                            // blocks x kPartWidth per member)
    void* serve = nullptr;  // kwgpu_surrogate_price's per-option buffers (grown on demand)
    size_t serve_cap = 0;   // options they hold
    // the FD configuration the network emulates (set by kwgpu_surrogate_fit; by default
    // the 1024 x 1024 reference defaults), and whether its grids are per option
    double density = 0.25, scale = 50.0, theta = 0.5;
    int x_dim = 1024, t_dim = 1024;
    bool per_option = false;
    // end of the last prediction and its stream: a prediction on another stream waits
    // for it before it reuses the scratch buffers
    hipEvent_t ev_done = nullptr;
    hipStream_t done_stream = nullptr;
};

namespace {

#define HIPCHK(expr)                                                              \
    do {                                                                          \
        hipError_t e_ = (expr);                                                   \
        if (e_ != hipSuccess)                                                     \
            return kwgpu::ctx_fail(ctx, KWGPU_ERR_HIP,                            \
                                   std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

unsigned nblk(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b); }

// Does `fd` describe the FD configuration the surrogate emulates? (Same density, scale
// and theta; the same single grid, or per-option grids for a per-option surrogate.)
bool grid_matches(const kwgpu_surrogate* s, const kwgpu_fd1d_params* fd) {
    if (!fd || fd->density != s->density || fd->scale != s->scale || fd->theta != s->theta)
        return false;
    if (s->per_option) return fd->x_dims && fd->t_dims;
    return !fd->x_dims && !fd->t_dims && fd->x_dim == s->x_dim && fd->t_dim == s->t_dim;
}

GridArg grid_arg(const kwgpu_surrogate* s, const kwgpu_fd1d_params* fd) {
    GridArg g{};
    g.density = s->density;
    g.scale = s->scale;
    g.x_dim = s->x_dim;
    g.t_dim = s->t_dim;
    if (fd && s->per_option) {
        g.x_dims = fd->x_dims;
        g.t_dims = fd->t_dims;
    }
    return g;
}

void free_scratch(kwgpu_surrogate* s) {
    for (void* q : {(void*)s->x, (void*)s->tg, (void*)s->y, (void*)s->dy, (void*)s->Z,
                    (void*)s->Hb, (void*)s->dZa, (void*)s->dZb, (void*)s->idx})
        if (q) (void)hipFree(q);
    s->x = s->tg = s->y = s->dy = s->Z = nullptr;
    s->Hb = s->dZa = s->dZb = nullptr;
    s->idx = nullptr;
    s->cap = 0;
}

// batch scratch, one slot of `rows` samples per member: x, targets, y, dy, Z and H for 4
// layers, two dZ buffers
hipError_t ensure_scratch(kwgpu_surrogate* s, size_t rows) {
    if (s->cap >= rows) return hipSuccess;
    free_scratch(s);
    const size_t M = (size_t)s->members;
    hipError_t e;
    if ((e = hipMalloc(&s->x, M * rows * F * sizeof(float))) ||
        (e = hipMalloc(&s->tg, M * rows * sizeof(float))) ||
        (e = hipMalloc(&s->y, M * rows * sizeof(float))) ||
        (e = hipMalloc(&s->dy, M * rows * sizeof(float))) ||
        (e = hipMalloc(&s->Z, M * 4 * rows * H * sizeof(float))) ||
        (e = hipMalloc(&s->Hb, M * 4 * rows * H * sizeof(bf16))) ||
        (e = hipMalloc(&s->dZa, M * rows * H * sizeof(bf16))) ||
        (e = hipMalloc(&s->dZb, M * rows * H * sizeof(bf16))) ||
        (e = hipMalloc(&s->idx, rows * sizeof(uint32_t))))
        return e;
    s->cap = rows;
    return hipSuccess;
}

GemmArgs gemm_fwd(const bf16* Hin, const bf16* W, int n, const float* bias, float* Z, bf16* Hout) {
    GemmArgs g{};
    g.A = Hin; g.sam = H; g.sak = 1;        // H_{l-1}[m][k]
    g.B = W; g.sbk = 1; g.sbn = H;          // W_l[n][k]
    g.M = n; g.N = H; g.K = H;
    g.bias = bias; g.Z = Z; g.Hout = Hout;
    return g;
}

hipError_t launch_gemm(int epi, const GemmArgs& g, hipStream_t st, int z) {
    dim3 grid(nblk(g.N, 64), nblk(g.M, 64), z);
    // k in 32-wide stages: 64- and 128-wide stages (more loads in flight and MFMAs per
    // barrier) raise the register count to 108 / 184 and measured slower (the layer
    // GEMMs are bound by their FP32 pre-activation traffic, not by the operand loads)
    if (epi == EPI_FWD)
        hipLaunchKernelGGL((gemm_bf16<EPI_FWD, 32>), grid, dim3(256), 0, st, g);
    else if (epi == EPI_BWD)
        hipLaunchKernelGGL((gemm_bf16<EPI_BWD, 32>), grid, dim3(256), 0, st, g);
    else
        hipLaunchKernelGGL((gemm_bf16<EPI_WGRAD, 32>), grid, dim3(256), 0, st, g);
    return hipGetLastError();
}

// Forward of members m0 .. m0+cnt-1 (parameters), scratch slots 0 .. cnt-1, n rows each:
// features from s->x (slot k, or slot 0 for every member when shared_x); y into
// yout + k sy; with targets (s->tg) also dy and the loss sums s->loss[k].
hipError_t forward(kwgpu_surrogate* s, int m0, int cnt, int n, bool shared_x, float* yout,
                   long sy, bool with_targets, hipStream_t st) {
    const float* p = s->p[m0];
    const bf16* wb = s->wb[m0];
    const long cap = (long)s->cap, sz = 4 * cap * H, sx = shared_x ? 0 : cap * F;
    float* Z = s->Z;
    bf16* Hb = s->Hb;
    const size_t L = (size_t)n * H;
    // without targets (prediction) no backward follows: only the last layer's FP32
    // pre-activations, which the output layer reads, are stored
    const bool keep = with_targets;
    hipLaunchKernelGGL(l1_fwd, dim3(nblk(L, 256), cnt), dim3(256), 0, st, s->x, sx, p, n,
                       keep ? Z : nullptr, Hb, sz);
    for (int l = 0; l < 3; ++l) {
        GemmArgs g = gemm_fwd(Hb + l * L, wb + l * H * H, n, p + kB2 + l * (H * H + H),
                              keep || l == 2 ? Z + (l + 1) * L : nullptr, Hb + (l + 1) * L);
        g.mA = sz;
        g.mB = 3 * H * H;
        g.mBias = kP;
        g.mZ = sz;
        g.mH = sz;
        if (hipError_t e = launch_gemm(EPI_FWD, g, st, cnt)) return e;
    }
    hipLaunchKernelGGL(out_fwd, dim3(nblk(n, 4 * kOutSPW), cnt), dim3(256), 0, st, Z + 3 * L, sz, p,
                       n, with_targets ? (const float*)s->tg : nullptr, yout, sy, s->dy, cap,
                       s->loss);
    return hipGetLastError();
}

// Backward after forward() for the same members; gradients into s->grad slots 0 .. cnt-1
// (zeroed here).
hipError_t backward(kwgpu_surrogate* s, int m0, int cnt, int n, hipStream_t st) {
    const float* p = s->p[m0];
    const bf16* wbT = s->wbT[m0];
    const long cap = (long)s->cap, sz = 4 * cap * H, sx = cap * F, sdz = cap * H, srow = cap;
    const size_t L = (size_t)n * H;
    float* g = s->grad;
    hipLaunchKernelGGL(zero_f, dim3(nblk((size_t)cnt * kP, 256)), dim3(256), 0, st, g,
                       (size_t)cnt * kP);
    // row splits of the column reductions: 16 rows per thread at the default batch
    const int rs = std::max(1, std::min(kPartRows, n / 64));
    hipLaunchKernelGGL(out_grad, dim3(H / 64, rs, cnt), dim3(256), 0, st, s->Z + 3 * L, sz, s->dy,
                       srow, n, s->part);
    hipLaunchKernelGGL(sum_parts, dim3(nblk(H + 1, 64), cnt), dim3(256), 0, st, s->part, rs, H + 1,
                       g, kW5);
    hipLaunchKernelGGL(out_bwd, dim3(nblk(L, 256), cnt), dim3(256), 0, st, s->Z + 3 * L, sz, p,
                       s->dy, srow, n, s->dZa, sdz);
    bf16* dz = s->dZa;
    bf16* dzn = s->dZb;
    for (int l = 2; l >= 0; --l) {  // hidden layer l+2: weights W_{l+2}, input H_{l+1}
        const int splits = std::max(1, std::min(kMaxSplits, n / 512));
        GemmArgs gw{};
        gw.A = dz; gw.sam = 1; gw.sak = H;               // A(o, m) = dZ[m][o]
        gw.B = s->Hb + l * L; gw.sbk = H; gw.sbn = 1;    // B(m, i) = H_{l+1}[m][i]
        gw.M = H; gw.N = H; gw.K = n;
        gw.C = s->part;  // one H x H slot per (member, split)
        gw.k_split = ((n + splits - 1) / splits + 31) / 32 * 32;
        gw.splits = splits;
        gw.mA = sdz;
        gw.mB = sz;
        if (hipError_t e = launch_gemm(EPI_WGRAD, gw, st, cnt * splits)) return e;
        hipLaunchKernelGGL(sum_parts, dim3(nblk(H * H, 64), cnt), dim3(256), 0, st, s->part,
                           splits, H * H, g, kW2 + l * (H * H + H));
        hipLaunchKernelGGL(colsum_bf16, dim3(H / 64, rs, cnt), dim3(256), 0, st, dz, sdz, n,
                           s->part);
        hipLaunchKernelGGL(sum_parts, dim3(nblk(H, 64), cnt), dim3(256), 0, st, s->part, rs, H, g,
                           kB2 + l * (H * H + H));
        GemmArgs gb{};
        gb.A = dz; gb.sam = H; gb.sak = 1;                     // A(m, o) = dZ[m][o]
        gb.B = wbT + l * H * H; gb.sbk = 1; gb.sbn = H;       // B(o, i) = W[o][i] = WT[i][o]
        gb.M = n; gb.N = H; gb.K = H;
        gb.Z = s->Z + l * L;                                   // pre-activations of H_{l+1}
        gb.Hout = dzn;
        gb.mA = sdz;
        gb.mB = 3 * H * H;
        gb.mZ = sz;
        gb.mH = sdz;
        if (hipError_t e = launch_gemm(EPI_BWD, gb, st, cnt)) return e;
        std::swap(dz, dzn);
    }
    const int rs1 = std::max(1, std::min(64, n / 128));
    static_assert(64 <= kPartRows, "partial slots");
    hipLaunchKernelGGL(l1_grad, dim3(rs1, cnt), dim3(256), 0, st, dz, sdz, s->x, sx, n, s->part);
    hipLaunchKernelGGL(sum_parts, dim3(nblk(kPartWidth, 64), cnt), dim3(256), 0, st, s->part, rs1,
                       kPartWidth, g, kW1);
    return hipGetLastError();
}

// BF16 copies of members m0 .. m0+cnt-1
hipError_t refresh_bf16(kwgpu_surrogate* s, int m0, int cnt, hipStream_t st) {
    hipLaunchKernelGGL(to_bf16, dim3((H / 16) * (H / 16), 3 * cnt), dim3(256), 0, st, s->p[m0],
                       s->wb[m0], s->wbT[m0]);
    return hipGetLastError();
}

void init_params(std::vector<float>& w, uint64_t seed) {
    std::mt19937_64 rng(seed);
    w.assign(kP, 0.0f);
    auto fill = [&](int off, int rows, int fan_in) {
        const float a = std::sqrt(6.0f / (float)fan_in);  // He-uniform for SiLU
        std::uniform_real_distribution<float> u(-a, a);
        for (int i = 0; i < rows * fan_in; ++i) w[off + i] = u(rng);
    };
    fill(kW1, H, F);
    fill(kW2, H, H);
    fill(kW3, H, H);
    fill(kW4, H, H);
    fill(kW5, 1, H);
    for (int i = 0; i < H; ++i) w[kW5 + i] *= 0.1f;
}

}  // namespace

extern "C" {

void kwgpu_surrogate_train_params_default(kwgpu_surrogate_train_params* p) {
    if (!p) return;
    p->epochs = 30;
    p->batch = 8192;
    p->lr = 2e-3f;
    p->lr_final = 2e-5f;
    p->beta1 = 0.9f;
    p->beta2 = 0.999f;
    p->eps = 1e-8f;
    p->seed = 7;
}

int kwgpu_surrogate_create(kwgpu_ctx* ctx, int32_t members, uint64_t seed, kwgpu_surrogate** out) {
    if (!ctx || !out || members < 1 || members > 16)
        return kwgpu::ctx_fail(ctx, KWGPU_ERR_INVALID, "kwgpu_surrogate_create: bad argument");
    *out = nullptr;
    HIPCHK(hipSetDevice(kwgpu::ctx_device(ctx)));
    auto* s = new kwgpu_surrogate();
    s->members = members;
    s->device = kwgpu::ctx_device(ctx);
    if (hipEventCreateWithFlags(&s->ev_done, hipEventDisableTiming)) {
        delete s;
        return kwgpu::ctx_fail(ctx, KWGPU_ERR_HIP, "kwgpu_surrogate_create: event");
    }
    std::vector<float> w;
    hipStream_t st = kwgpu::ctx_stream(ctx);
    const size_t M = (size_t)members;
    if (hipMalloc(&s->pall, M * kP * sizeof(float)) ||
        hipMalloc(&s->wball, M * 3 * H * H * sizeof(bf16)) ||
        hipMalloc(&s->wballT, M * 3 * H * H * sizeof(bf16)) ||
        hipMalloc(&s->part, M * kPartFloats * sizeof(float)) ||
        hipMalloc(&s->grad, M * kP * sizeof(float)) || hipMalloc(&s->m1, M * kP * sizeof(float)) ||
        hipMalloc(&s->m2, M * kP * sizeof(float)) || hipMalloc(&s->loss, M * sizeof(float))) {
        kwgpu_surrogate_destroy(s);
        return kwgpu::ctx_fail(ctx, KWGPU_ERR_NOMEM, "kwgpu_surrogate_create: allocation");
    }
    for (int m = 0; m < members; ++m) {
        s->p.push_back(s->pall + (size_t)m * kP);
        s->wb.push_back(s->wball + (size_t)m * 3 * H * H);
        s->wbT.push_back(s->wballT + (size_t)m * 3 * H * H);
        init_params(w, seed * 1000003ull + (uint64_t)m);
        if (hipMemcpy(s->p[m], w.data(), kP * sizeof(float), hipMemcpyHostToDevice)) {
            kwgpu_surrogate_destroy(s);
            return kwgpu::ctx_fail(ctx, KWGPU_ERR_HIP, "kwgpu_surrogate_create: upload");
        }
    }
    if (refresh_bf16(s, 0, members, st) || hipStreamSynchronize(st)) {
        kwgpu_surrogate_destroy(s);
        return kwgpu::ctx_fail(ctx, KWGPU_ERR_HIP, "kwgpu_surrogate_create: upload");
    }
    *out = s;
    return KWGPU_OK;
}

void kwgpu_surrogate_destroy(kwgpu_surrogate* s) {
    if (!s) return;
    (void)hipSetDevice(s->device);
    (void)hipDeviceSynchronize();
    free_scratch(s);
    for (void* q : {(void*)s->pall, (void*)s->wball, (void*)s->wballT, (void*)s->grad,
                    (void*)s->m1, (void*)s->m2,
                    (void*)s->loss, (void*)s->yall, (void*)s->xp, (void*)s->part, s->serve})
        if (q) (void)hipFree(q);
    if (s->ev_done) (void)hipEventDestroy(s->ev_done);
    delete s;
}

int kwgpu_surrogate_features(kwgpu_ctx* ctx, const kwgpu_option* opts, size_t n, int32_t x_dim,
                             int32_t t_dim, float* feat) {
    if (!opts || !feat || x_dim < 3 || t_dim < 2)
        return kwgpu::ctx_fail(ctx, KWGPU_ERR_INVALID, "kwgpu_surrogate_features: bad argument");
    for (size_t i = 0; i < n; ++i) (void)option_features(opts[i], x_dim, t_dim, feat + i * F);
    return KWGPU_OK;
}

int kwgpu_surrogate_baseline(kwgpu_ctx* ctx, const kwgpu_option* opts, size_t n, double* out) {
    if (!opts || !out) return kwgpu::ctx_fail(ctx, KWGPU_ERR_INVALID, "kwgpu_surrogate_baseline: bad argument");
    if (n == 0) return KWGPU_OK;
    hipStream_t st = kwgpu::ctx_stream(ctx);
    kwgpu_option* d_o = nullptr;
    double* d_v = nullptr;
    int rc = KWGPU_OK;
    if (hipMalloc(&d_o, n * sizeof(kwgpu_option)) || hipMalloc(&d_v, n * sizeof(double))) {
        rc = kwgpu::ctx_fail(ctx, KWGPU_ERR_NOMEM, "kwgpu_surrogate_baseline: allocation");
    } else {
        hipError_t e = hipMemcpyAsync(d_o, opts, n * sizeof(kwgpu_option), hipMemcpyHostToDevice, st);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(baseline_kernel, dim3(nblk(n, 256)), dim3(256), 0, st, d_o, n, d_v);
            e = hipGetLastError();
        }
        if (e == hipSuccess)
            e = hipMemcpyAsync(out, d_v, n * sizeof(double), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess)
            rc = kwgpu::ctx_fail(ctx, KWGPU_ERR_HIP, "kwgpu_surrogate_baseline: HIP error");
    }
    (void)hipFree(d_o);
    (void)hipFree(d_v);
    return rc;
}

int kwgpu_surrogate_get_params(kwgpu_ctx* ctx, const kwgpu_surrogate* s, int32_t member,
                               float* host) {
    if (!s || member < 0 || member >= s->members || !host)
        return kwgpu::ctx_fail(ctx, KWGPU_ERR_INVALID, "kwgpu_surrogate_get_params: bad argument");
    HIPCHK(hipMemcpy(host, s->p[member], kP * sizeof(float), hipMemcpyDeviceToHost));
    return KWGPU_OK;
}

int kwgpu_surrogate_set_params(kwgpu_ctx* ctx, kwgpu_surrogate* s, int32_t member,
                               const float* host) {
    if (!s || member < 0 || member >= s->members || !host)
        return kwgpu::ctx_fail(ctx, KWGPU_ERR_INVALID, "kwgpu_surrogate_set_params: bad argument");
    hipStream_t st = kwgpu::ctx_stream(ctx);
    HIPCHK(hipMemcpy(s->p[member], host, kP * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(refresh_bf16(s, member, 1, st));
    HIPCHK(hipStreamSynchronize(st));
    return KWGPU_OK;
}

int kwgpu_surrogate_loss_grad(kwgpu_ctx* ctx, kwgpu_surrogate* s, int32_t member,
                              const float* host_x, const float* host_t, size_t n, float* host_loss,
                              float* host_grad) {
    if (!s || member < 0 || member >= s->members || !host_x || !host_t || n == 0 ||
        n > (1u << 17))
        return kwgpu::ctx_fail(ctx, KWGPU_ERR_INVALID, "kwgpu_surrogate_loss_grad: bad argument");
    hipStream_t st = kwgpu::ctx_stream(ctx);
    HIPCHK(ensure_scratch(s, n));
    HIPCHK(hipMemcpy(s->x, host_x, n * F * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(s->tg, host_t, n * sizeof(float), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(zero_f, dim3(1), dim3(64), 0, st, s->loss, (size_t)1);
    HIPCHK(forward(s, member, 1, (int)n, false, s->y, (long)s->cap, true, st));
    HIPCHK(backward(s, member, 1, (int)n, st));
    HIPCHK(hipStreamSynchronize(st));
    float l = 0;
    HIPCHK(hipMemcpy(&l, s->loss, sizeof(float), hipMemcpyDeviceToHost));
    if (host_loss) *host_loss = l / (float)n;
    if (host_grad) HIPCHK(hipMemcpy(host_grad, s->grad, kP * sizeof(float), hipMemcpyDeviceToHost));
    return KWGPU_OK;
}

}  // extern "C"

namespace {

// Trains every member on rows (option, premium) held on the device, row i's features
// taken on grid g; `rows` (host, optional) lists the rows to use, all n otherwise. Member
// m shuffles with its own stream seeded from tp->seed; the next epoch's permutation is
// drawn on a host thread while the current epoch runs, into the other of two pinned /
// device buffer pairs.
int train_rows(kwgpu_ctx* ctx, kwgpu_surrogate* s, const kwgpu_option* d_opts,
               const double* d_premium, size_t n, const GridArg& g,
               const std::vector<uint32_t>* rows, const kwgpu_surrogate_train_params* tp,
               float* loss_out) {
    hipStream_t st = kwgpu::ctx_stream(ctx);
    const int B = tp->batch;
    const int M = s->members;
    const size_t nr = rows ? rows->size() : n;
    if (nr < (size_t)B)
        return kwgpu::ctx_fail(ctx, KWGPU_ERR_INVALID,
                               "kwgpu_surrogate_train: fewer usable rows (" + std::to_string(nr) +
                                   ") than a batch");
    float *feat = nullptr, *tgt = nullptr;
    uint32_t* dperm[2] = {nullptr, nullptr};
    uint32_t* hperm[2] = {nullptr, nullptr};
    int rc = KWGPU_OK;
    auto cleanup = [&]() {
        (void)hipStreamSynchronize(st);
        (void)hipFree(feat);
        (void)hipFree(tgt);
        for (int k = 0; k < 2; ++k) {
            (void)hipFree(dperm[k]);
            (void)hipHostFree(hperm[k]);
        }
    };
    const size_t pbytes = (size_t)M * nr * sizeof(uint32_t);
    if (hipMalloc(&feat, n * F * sizeof(float)) || hipMalloc(&tgt, n * sizeof(float)) ||
        hipMalloc(&dperm[0], pbytes) || hipMalloc(&dperm[1], pbytes) ||
        hipHostMalloc((void**)&hperm[0], pbytes, hipHostMallocDefault) ||
        hipHostMalloc((void**)&hperm[1], pbytes, hipHostMallocDefault) || ensure_scratch(s, B)) {
        cleanup();
        return kwgpu::ctx_fail(ctx, KWGPU_ERR_NOMEM, "kwgpu_surrogate_train: allocation");
    }
    hipLaunchKernelGGL(features_kernel, dim3(nblk(n, 256)), dim3(256), 0, st, d_opts, n, g,
                       d_premium, feat, tgt, (uint8_t*)nullptr);
    // every member steps at once (one launch per layer for the whole ensemble), each on
    // its own shuffle of the rows
    std::vector<std::mt19937_64> rng;
    for (int m = 0; m < M; ++m) rng.emplace_back(tp->seed + 7919ull * (uint64_t)m);
    auto draw = [&](uint32_t* dst) {
        for (int m = 0; m < M; ++m) {
            uint32_t* first = dst + (size_t)m * nr;
            if (rows)
                std::copy(rows->begin(), rows->end(), first);
            else
                std::iota(first, first + nr, 0u);
            std::shuffle(first, first + nr, rng[m]);
        }
    };
    const size_t steps_per_epoch = nr / B;  // drop the ragged tail
    const double total = (double)steps_per_epoch * tp->epochs;
    const long cap = (long)s->cap;
    hipLaunchKernelGGL(zero_f, dim3(nblk((size_t)M * kP, 256)), dim3(256), 0, st, s->m1,
                       (size_t)M * kP);
    hipLaunchKernelGGL(zero_f, dim3(nblk((size_t)M * kP, 256)), dim3(256), 0, st, s->m2,
                       (size_t)M * kP);
    draw(hperm[0]);
    if (hipError_t e = hipMemcpyAsync(dperm[0], hperm[0], pbytes, hipMemcpyHostToDevice, st)) {
        cleanup();
        return kwgpu::ctx_fail(ctx, KWGPU_ERR_HIP, std::string("permutation: ") + hipGetErrorString(e));
    }
    long step = 0;
    std::vector<float> lsum(M);
    for (int ep = 0; ep < tp->epochs && rc == KWGPU_OK; ++ep) {
        const uint32_t* didx = dperm[ep & 1];
        // the next epoch's order, drawn while this one runs: its host buffer was last
        // copied from before the previous epoch's end, which was waited for
        std::thread next;
        if (ep + 1 < tp->epochs) next = std::thread(draw, hperm[(ep + 1) & 1]);
        hipLaunchKernelGGL(zero_f, dim3(1), dim3(64), 0, st, s->loss, (size_t)M);
        for (size_t b = 0; b < steps_per_epoch; ++b, ++step) {
            hipLaunchKernelGGL(gather_kernel, dim3(nblk(B, 256), M), dim3(256), 0, st, feat, tgt,
                               didx + b * B, (long)nr, B, s->x, s->tg, cap);
            hipError_t e = forward(s, 0, M, B, false, s->y, cap, true, st);
            if (!e) e = backward(s, 0, M, B, st);
            const double frac = step / total;
            const float lr = (float)(tp->lr_final + 0.5 * (tp->lr - tp->lr_final) *
                                                         (1.0 + std::cos(M_PI * frac)));
            const float c1 = 1.0f - std::pow(tp->beta1, (float)(step + 1));
            const float c2 = 1.0f - std::pow(tp->beta2, (float)(step + 1));
            hipLaunchKernelGGL(adam_kernel, dim3(nblk(kP, 256), M), dim3(256), 0, st, s->pall,
                               s->grad, s->m1, s->m2, kP, lr, tp->beta1, tp->beta2, tp->eps, c1,
                               c2);
            if (!e) e = refresh_bf16(s, 0, M, st);
            if (e) {
                rc = kwgpu::ctx_fail(ctx, KWGPU_ERR_HIP,
                                     std::string("train step: ") + hipGetErrorString(e));
                break;
            }
        }
        if (next.joinable()) next.join();
        if (rc == KWGPU_OK && ep + 1 < tp->epochs) {
            // the other device buffer was last read by the previous epoch (stream order)
            if (hipError_t e = hipMemcpyAsync(dperm[(ep + 1) & 1], hperm[(ep + 1) & 1], pbytes,
                                              hipMemcpyHostToDevice, st))
                rc = kwgpu::ctx_fail(ctx, KWGPU_ERR_HIP, hipGetErrorString(e));
        }
        if (hipError_t e = hipStreamSynchronize(st)) {
            if (rc == KWGPU_OK) rc = kwgpu::ctx_fail(ctx, KWGPU_ERR_HIP, hipGetErrorString(e));
            break;
        }
        (void)hipMemcpy(lsum.data(), s->loss, M * sizeof(float), hipMemcpyDeviceToHost);
        if (loss_out)
            for (int m = 0; m < M; ++m)
                loss_out[(size_t)ep * M + m] = lsum[m] / (float)(steps_per_epoch * B);
    }
    cleanup();
    return rc;
}

}  // namespace

extern "C" {

int kwgpu_surrogate_train(kwgpu_ctx* ctx, kwgpu_surrogate* s, const kwgpu_option* d_opts,
                          const double* d_premium, size_t n, const kwgpu_surrogate_train_params* tp,
                          float* loss_out) {
    if (!s || !d_opts || !d_premium || !tp || n == 0 || tp->batch < 64 || tp->epochs < 1 ||
        n > 0xffffffffu)
        return kwgpu::ctx_fail(ctx, KWGPU_ERR_INVALID, "kwgpu_surrogate_train: bad argument");
    if (s->per_option)
        return kwgpu::ctx_fail(ctx, KWGPU_ERR_INVALID,
                               "kwgpu_surrogate_train: the surrogate emulates per-option grids; "
                               "use kwgpu_surrogate_fit");
    HIPCHK(hipSetDevice(s->device));
    return train_rows(ctx, s, d_opts, d_premium, n, grid_arg(s, nullptr), nullptr, tp, loss_out);
}

int kwgpu_surrogate_labels(kwgpu_ctx* ctx, const kwgpu_option* d_opts, size_t n,
                           const kwgpu_fd1d_params* fd, kwgpu_option* d_nodes, double* d_premium,
                           uint8_t* d_keep) {
    if (!d_opts || !fd || !d_nodes || !d_premium || !d_keep || n == 0 || n > 0x3fffffffu ||
        (!fd->x_dims) != (!fd->t_dims))
        return kwgpu::ctx_fail(ctx, KWGPU_ERR_INVALID, "kwgpu_surrogate_labels: bad argument");
    if (fd->precision != KWGPU_FP64)
        return kwgpu::ctx_fail(ctx, KWGPU_ERR_INVALID, "kwgpu_surrogate_labels: labels are FP64 FD prices");
    HIPCHK(hipSetDevice(kwgpu::ctx_device(ctx)));
    hipStream_t st = kwgpu::ctx_stream(ctx);
    const size_t n2 = 2 * n;
    int32_t *nx = nullptr, *nt = nullptr, *fst = nullptr;
    double* fdp = nullptr;
    uint8_t* ok = nullptr;
    auto cleanup = [&]() {
        (void)hipStreamSynchronize(st);
        for (void* q : {(void*)nx, (void*)nt, (void*)fst, (void*)fdp, (void*)ok}) (void)hipFree(q);
    };
    const bool per_option = fd->x_dims != nullptr;
    if (hipMalloc(&fst, n2 * sizeof(int32_t)) || hipMalloc(&fdp, n2 * sizeof(double)) ||
        hipMalloc(&ok, n) ||
        (per_option &&
         (hipMalloc(&nx, n2 * sizeof(int32_t)) || hipMalloc(&nt, n2 * sizeof(int32_t))))) {
        cleanup();
        return kwgpu::ctx_fail(ctx, KWGPU_ERR_NOMEM, "kwgpu_surrogate_labels: allocation");
    }
    GridArg g{};
    g.x_dims = fd->x_dims;
    g.t_dims = fd->t_dims;
    g.x_dim = fd->x_dim;
    g.t_dim = fd->t_dim;
    g.density = fd->density;
    g.scale = fd->scale;
    // each option's two bracketing nodes, priced by the FD pricer there
    hipLaunchKernelGGL(node_options_kernel, dim3(nblk(n, 256)), dim3(256), 0, st, d_opts, n, g,
                       d_nodes, nx, nt, ok);
    kwgpu_fd1d_params p = *fd;
    p.x_dims = nx;
    p.t_dims = nt;
    int rc = kwgpu::fd_price_device_prec(ctx, d_nodes, n2, &p, fdp, fst, st, nullptr);
    if (rc == KWGPU_OK) {
        hipLaunchKernelGGL(node_premium_kernel, dim3(nblk(n, 256)), dim3(256), 0, st, d_nodes, n,
                           fdp, fst, ok, d_premium, d_keep);
        if (hipGetLastError() || hipStreamSynchronize(st))
            rc = kwgpu::ctx_fail(ctx, KWGPU_ERR_HIP, "kwgpu_surrogate_labels: premium");
    }
    cleanup();
    return rc;
}

int kwgpu_surrogate_fit(kwgpu_ctx* ctx, kwgpu_surrogate* s, const kwgpu_option* d_opts, size_t n,
                        const kwgpu_fd1d_params* fd, const kwgpu_surrogate_train_params* tp,
                        float* loss_out, size_t* n_pairs) {
    if (n_pairs) *n_pairs = 0;
    if (!s || !d_opts || !fd || !tp || n == 0 || tp->batch < 64 || tp->epochs < 1 ||
        n > 0x3fffffffu || (!fd->x_dims) != (!fd->t_dims))
        return kwgpu::ctx_fail(ctx, KWGPU_ERR_INVALID, "kwgpu_surrogate_fit: bad argument");
    if (fd->precision != KWGPU_FP64)
        return kwgpu::ctx_fail(ctx, KWGPU_ERR_INVALID, "kwgpu_surrogate_fit: labels are FP64 FD prices");
    HIPCHK(hipSetDevice(s->device));
    hipStream_t st = kwgpu::ctx_stream(ctx);
    // the configuration the network will emulate
    s->density = fd->density;
    s->scale = fd->scale;
    s->theta = fd->theta;
    s->x_dim = fd->x_dim;
    s->t_dim = fd->t_dim;
    s->per_option = fd->x_dims != nullptr;
    const size_t n2 = 2 * n;
    kwgpu_option* nodes = nullptr;
    int32_t *nx = nullptr, *nt = nullptr;
    double* prem = nullptr;
    uint8_t* keep = nullptr;
    auto cleanup = [&]() {
        (void)hipStreamSynchronize(st);
        for (void* q : {(void*)nodes, (void*)nx, (void*)nt, (void*)prem, (void*)keep})
            (void)hipFree(q);
    };
    if (hipMalloc(&nodes, n2 * sizeof(kwgpu_option)) || hipMalloc(&prem, n2 * sizeof(double)) ||
        hipMalloc(&keep, n) ||
        (s->per_option &&
         (hipMalloc(&nx, n2 * sizeof(int32_t)) || hipMalloc(&nt, n2 * sizeof(int32_t))))) {
        cleanup();
        return kwgpu::ctx_fail(ctx, KWGPU_ERR_NOMEM, "kwgpu_surrogate_fit: allocation");
    }
    int rc = kwgpu_surrogate_labels(ctx, d_opts, n, fd, nodes, prem, keep);
    std::vector<uint8_t> hkeep(n);
    if (rc == KWGPU_OK && s->per_option) {
        // the node rows' grids: both nodes of option i on option i's grid
        if (hipMemcpyAsync(nx, fd->x_dims, n * sizeof(int32_t), hipMemcpyDeviceToDevice, st) ||
            hipMemcpyAsync(nx + n, fd->x_dims, n * sizeof(int32_t), hipMemcpyDeviceToDevice, st) ||
            hipMemcpyAsync(nt, fd->t_dims, n * sizeof(int32_t), hipMemcpyDeviceToDevice, st) ||
            hipMemcpyAsync(nt + n, fd->t_dims, n * sizeof(int32_t), hipMemcpyDeviceToDevice, st))
            rc = kwgpu::ctx_fail(ctx, KWGPU_ERR_HIP, "kwgpu_surrogate_fit: grids");
    }
    if (rc == KWGPU_OK &&
        (hipMemcpyAsync(hkeep.data(), keep, n, hipMemcpyDeviceToHost, st) || hipStreamSynchronize(st)))
        rc = kwgpu::ctx_fail(ctx, KWGPU_ERR_HIP, "kwgpu_surrogate_fit: labels");
    if (rc == KWGPU_OK) {
        std::vector<uint32_t> rows;
        rows.reserve(n2);
        for (size_t i = 0; i < n; ++i)
            if (hkeep[i]) rows.push_back((uint32_t)i);
        const size_t pairs = rows.size();
        for (size_t k = 0; k < pairs; ++k) rows.push_back(rows[k] + (uint32_t)n);
        if (n_pairs) *n_pairs = pairs;
        GridArg g = grid_arg(s, nullptr);
        g.x_dims = nx;
        g.t_dims = nt;
        rc = train_rows(ctx, s, nodes, prem, n2, g, &rows, tp, loss_out);
    }
    cleanup();
    return rc;
}

int kwgpu_surrogate_predict(kwgpu_ctx* ctx, kwgpu_surrogate* s, const kwgpu_option* d_opts,
                            size_t n, const kwgpu_fd1d_params* fd, double* d_premium,
                            double* d_spread, uint8_t* d_in_box, void* stream) {
    if (!s || !d_opts || !d_premium || !d_spread)
        return kwgpu::ctx_fail(ctx, KWGPU_ERR_INVALID, "kwgpu_surrogate_predict: bad argument");
    if (fd && !grid_matches(s, fd))
        return kwgpu::ctx_fail(ctx, KWGPU_ERR_INVALID,
                               "kwgpu_surrogate_predict: fd is not the configuration the "
                               "surrogate emulates");
    if (!fd && s->per_option)
        return kwgpu::ctx_fail(ctx, KWGPU_ERR_INVALID,
                               "kwgpu_surrogate_predict: per-option grids needed");
    if (n == 0) return KWGPU_OK;
    hipStream_t st = stream ? (hipStream_t)stream : kwgpu::ctx_stream(ctx);
    if (s->done_stream && s->done_stream != st) HIPCHK(hipStreamWaitEvent(st, s->ev_done, 0));
    // two rows per option (the bracketing nodes), in chunks of up to CH options: the
    // features and the members' outputs of one chunk (64 + 32 bytes per option for 4
    // members). A chunk is large so that a batch is seldom split: a small last chunk
    // would pay the full latency of its boundary solves and network launch.
    constexpr size_t CH = size_t(1) << 20;
    const size_t rows = 2 * std::min(n, CH);
    if (s->xp_cap < rows) {
        if (s->xp) HIPCHK(hipFree(s->xp));
        s->xp = nullptr;
        s->xp_cap = 0;
        HIPCHK(hipMalloc(&s->xp, rows * F * sizeof(float)));
        s->xp_cap = rows;
    }
    const size_t need = (size_t)s->members * rows;
    if (s->yall_cap < need) {
        if (s->yall) HIPCHK(hipFree(s->yall));
        s->yall = nullptr;
        s->yall_cap = 0;
        HIPCHK(hipMalloc(&s->yall, need * sizeof(float)));
        s->yall_cap = need;
    }
    const GridArg g0 = grid_arg(s, fd);
    for (size_t c0 = 0; c0 < n; c0 += CH) {
        const size_t cn = std::min(CH, n - c0);
        GridArg g = g0;
        if (g.x_dims) g.x_dims += c0;
        if (g.t_dims) g.t_dims += c0;
        hipLaunchKernelGGL(node_features_kernel, dim3(nblk(cn, 256)), dim3(256), 0, st,
                           d_opts + c0, cn, g, s->xp, d_in_box ? d_in_box + c0 : nullptr);
        // the whole ensemble at once on the shared features, one fused kernel per
        // 128-row tile and member
        hipLaunchKernelGGL(mlp_predict, dim3(nblk(2 * cn, kPredRows), s->members), dim3(256), 0,
                           st, s->xp, s->p[0], s->wb[0], (int)(2 * cn), s->yall, (long)(2 * cn));
        // kGroup lanes per candidate, which together solve its exercise boundary
        hipLaunchKernelGGL(node_combine_kernel, dim3(nblk(kGroup * cn, 64)), dim3(64), 0, st,
                           d_opts + c0, cn, g, s->yall, s->members, 2 * cn, d_premium + c0,
                           d_spread + c0);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipEventRecord(s->ev_done, st));
    s->done_stream = st;
    return KWGPU_OK;
}

}  // extern "C"

namespace {

bool gate_valid(const kwgpu_gate& g) {
    if (g.n_t < 1 || g.n_t > KWGPU_GATE_MAX_T || g.n_m < 1 || g.n_m > KWGPU_GATE_MAX_M)
        return false;
    for (int a = 0; a + 1 < g.n_t; ++a)
        if (!std::isfinite(g.t_edges[a]) || (a > 0 && !(g.t_edges[a] > g.t_edges[a - 1])))
            return false;
    for (int b = 0; b + 1 < g.n_m; ++b)
        if (!std::isfinite(g.m_edges[b]) || (b > 0 && !(g.m_edges[b] > g.m_edges[b - 1])))
            return false;
    return true;
}

int price_gated(kwgpu_ctx* ctx, kwgpu_surrogate* s, const kwgpu_option* d_opts, size_t n,
                const kwgpu_gate& gate, const kwgpu_fd1d_params* fd, double* d_prices,
                int32_t* d_route, size_t* n_fd);

}  // namespace

extern "C" {

int kwgpu_surrogate_price(kwgpu_ctx* ctx, kwgpu_surrogate* s, const kwgpu_option* d_opts,
                          size_t n, double spread_max, double scale_max,
                          const kwgpu_fd1d_params* fd, double* d_prices, int32_t* d_route,
                          size_t* n_fd) {
    kwgpu_gate g{};  // one band each: the two classes are puts and calls
    g.n_t = g.n_m = 1;
    g.spread_max[0] = g.spread_max[1] = spread_max;
    g.scale_max[0] = g.scale_max[1] = scale_max;
    return price_gated(ctx, s, d_opts, n, g, fd, d_prices, d_route, n_fd);
}

int kwgpu_surrogate_price_gated(kwgpu_ctx* ctx, kwgpu_surrogate* s, const kwgpu_option* d_opts,
                                size_t n, const kwgpu_gate* gate, const kwgpu_fd1d_params* fd,
                                double* d_prices, int32_t* d_route, size_t* n_fd) {
    if (!gate || !gate_valid(*gate))
        return kwgpu::ctx_fail(ctx, KWGPU_ERR_INVALID, "kwgpu_surrogate_price_gated: bad gate");
    return price_gated(ctx, s, d_opts, n, *gate, fd, d_prices, d_route, n_fd);
}

}  // extern "C"

namespace {

int price_gated(kwgpu_ctx* ctx, kwgpu_surrogate* s, const kwgpu_option* d_opts, size_t n,
                const kwgpu_gate& gate, const kwgpu_fd1d_params* fd, double* d_prices,
                int32_t* d_route, size_t* n_fd) {
    if (!s || !d_opts || !d_prices || !d_route || !fd)
        return kwgpu::ctx_fail(ctx, KWGPU_ERR_INVALID, "kwgpu_surrogate_price: bad argument");
    if (n_fd) *n_fd = 0;
    if (n == 0) return KWGPU_OK;
    if (n > 0x7fffffffu)
        return kwgpu::ctx_fail(ctx, KWGPU_ERR_INVALID, "kwgpu_surrogate_price: too many options");
    hipStream_t st = kwgpu::ctx_stream(ctx);
    // per-option buffers, carved from one allocation kept with the surrogate
    auto carve = [n](char* base, size_t& at, size_t bytes_per) {
        char* q = base + at;
        at += (bytes_per * n + 255) / 256 * 256;
        return q;
    };
    size_t need = 0;
    for (size_t b : {sizeof(double), sizeof(double), (size_t)1, sizeof(uint32_t), sizeof(double),
                     sizeof(int32_t), sizeof(kwgpu_option), sizeof(int32_t), sizeof(int32_t),
                     sizeof(uint32_t), sizeof(double), sizeof(double), (size_t)1})
        need += (b * n + 255) / 256 * 256;
    need += 256;  // the two counters
    if (s->serve_cap < n) {
        (void)hipStreamSynchronize(st);
        if (s->serve) (void)hipFree(s->serve);
        s->serve = nullptr;
        s->serve_cap = 0;
        if (hipMalloc(&s->serve, need))
            return kwgpu::ctx_fail(ctx, KWGPU_ERR_NOMEM, "kwgpu_surrogate_price: allocation");
        s->serve_cap = n;
    }
    char* base = (char*)s->serve;
    size_t at = 0;
    double* prem = (double*)carve(base, at, sizeof(double));
    double* spr = (double*)carve(base, at, sizeof(double));
    uint8_t* inb = (uint8_t*)carve(base, at, 1);
    uint32_t* idx = (uint32_t*)carve(base, at, sizeof(uint32_t));  // routes 1, then 2
    uint32_t* blk = (uint32_t*)carve(base, at, sizeof(uint32_t));  // 2 per 256 options
    double* sub_p = (double*)carve(base, at, sizeof(double));
    int32_t* sub_s = (int32_t*)carve(base, at, sizeof(int32_t));
    kwgpu_option* sub_o = (kwgpu_option*)carve(base, at, sizeof(kwgpu_option));
    int32_t* sub_x = (int32_t*)carve(base, at, sizeof(int32_t));
    int32_t* sub_t = (int32_t*)carve(base, at, sizeof(int32_t));
    double* prem_c = (double*)carve(base, at, sizeof(double));  // the candidates' predictions
    double* spr_c = (double*)carve(base, at, sizeof(double));
    uint8_t* inb_c = (uint8_t*)carve(base, at, 1);
    int* cnt = (int*)(base + at);
    int rc = KWGPU_OK;
    auto cleanup = [&]() { (void)hipStreamSynchronize(st); };
    // a negative (or NaN) bound passes no option of its class, and an FD configuration
    // other than the one the surrogate emulates passes none at all: with every class
    // closed the ensemble is not evaluated
    bool open = false;
    for (int k = 0; k < 2 * gate.n_t * gate.n_m; ++k) open = open || gate.spread_max[k] >= 0.0;
    const bool serve_none = !open || !grid_matches(s, fd);
    const unsigned nb = nblk(n, kGateBlock);
    if (serve_none && fd->precision == KWGPU_FP64) {
        // nothing to route: the batch goes to the FD pricer as it is, with no compaction
        rc = kwgpu::fd_price_device_prec(ctx, d_opts, n, fd, d_prices, sub_s, st, nullptr);
        if (rc == KWGPU_OK) {
            hipLaunchKernelGGL(all_to_fd, dim3(nblk(n, 256)), dim3(256), 0, st, n, sub_s,
                               d_prices, d_route);
            if (hipStreamSynchronize(st)) rc = kwgpu::ctx_fail(ctx, KWGPU_ERR_HIP, "FD fallback");
        }
        if (n_fd && rc == KWGPU_OK) *n_fd = n;
        cleanup();
        return rc;
    }
    if (!serve_none) {
        // the ensemble runs on the candidates only: listed in index order, gathered with
        // their grid sizes, predicted, and their predictions scattered back
        const GridArg g = grid_arg(s, fd);
        hipLaunchKernelGGL(candidate_kernel, dim3(nblk(n, kCandBlock)), dim3(kCandBlock), 0, st,
                           d_opts, n, g, gate, d_route);
        hipLaunchKernelGGL(count_candidates, dim3(nb), dim3(kGateBlock), 0, st, d_route, n, blk);
        hipLaunchKernelGGL(scan_block_counts, dim3(1), dim3(1024), 0, st, blk, nb, cnt);
        hipLaunchKernelGGL(compact_routes, dim3(nb), dim3(kGateBlock), 0, st, d_route, n, blk,
                           (const int*)cnt, idx);
        hipLaunchKernelGGL(clear_prediction, dim3(nblk(n, 256)), dim3(256), 0, st, n, prem, spr, inb);
        int nc = 0;
        if (hipMemcpyAsync(&nc, cnt, sizeof(int), hipMemcpyDeviceToHost, st) ||
            hipStreamSynchronize(st)) {
            cleanup();
            return kwgpu::ctx_fail(ctx, KWGPU_ERR_HIP, "kwgpu_surrogate_price: candidates");
        }
        if (nc > 0) {
            hipLaunchKernelGGL(gather_opts, dim3(nblk(nc, 256)), dim3(256), 0, st, d_opts, idx, nc,
                               sub_o);
            kwgpu_fd1d_params pc = *fd;
            if (fd->x_dims) {
                hipLaunchKernelGGL(gather_dims, dim3(nblk(nc, 256)), dim3(256), 0, st, fd->x_dims,
                                   idx, nc, sub_x);
                hipLaunchKernelGGL(gather_dims, dim3(nblk(nc, 256)), dim3(256), 0, st, fd->t_dims,
                                   idx, nc, sub_t);
                pc.x_dims = sub_x;
                pc.t_dims = sub_t;
            }
            rc = kwgpu_surrogate_predict(ctx, s, sub_o, (size_t)nc, &pc, prem_c, spr_c, inb_c, st);
            if (rc) {
                cleanup();
                return rc;
            }
            hipLaunchKernelGGL(scatter_prediction, dim3(nblk(nc, 256)), dim3(256), 0, st, idx, nc,
                               prem_c, spr_c, inb_c, prem, spr, inb);
        }
    }
    // an FP32-mixed fallback sends the options beyond its limits to FP64
    const double mixed_scale = fd->precision == KWGPU_FP64 ? 0.0 : fd->scale;
    hipLaunchKernelGGL(gate_kernel, dim3(nb), dim3(kGateBlock), 0, st, d_opts, n,
                       serve_none ? nullptr : prem, spr, inb, gate, mixed_scale, d_prices,
                       d_route, blk);
    hipLaunchKernelGGL(scan_block_counts, dim3(1), dim3(1024), 0, st, blk, nb, cnt);
    hipLaunchKernelGGL(compact_routes, dim3(nb), dim3(kGateBlock), 0, st, d_route, n, blk,
                       (const int*)cnt, idx);
    int hc[2] = {0, 0};
    if (hipMemcpyAsync(hc, cnt, sizeof(hc), hipMemcpyDeviceToHost, st) ||
        hipStreamSynchronize(st)) {
        cleanup();
        return kwgpu::ctx_fail(ctx, KWGPU_ERR_HIP, "kwgpu_surrogate_price: gate");
    }
    // routes 1 and 2 in one FD solve (one uneven-batch plan): the compaction put route
    // 2's options right after route 1's
    const int m = hc[0] + hc[1];
    if (m > 0) {
        hipLaunchKernelGGL(gather_opts, dim3(nblk(m, 256)), dim3(256), 0, st, d_opts, idx, m, sub_o);
        kwgpu_fd1d_params p = *fd;
        p.precision = hc[0] ? fd->precision : KWGPU_FP64;
        // per-option grid sizes follow their options into the compacted subset
        if (fd->x_dims) {
            hipLaunchKernelGGL(gather_dims, dim3(nblk(m, 256)), dim3(256), 0, st, fd->x_dims, idx,
                               m, sub_x);
            p.x_dims = sub_x;
        }
        if (fd->t_dims) {
            hipLaunchKernelGGL(gather_dims, dim3(nblk(m, 256)), dim3(256), 0, st, fd->t_dims, idx,
                               m, sub_t);
            p.t_dims = sub_t;
        }
        std::vector<uint8_t> prec;
        if (hc[0] && hc[1] && fd->precision != KWGPU_FP64) {
            prec.assign((size_t)m, (uint8_t)fd->precision);
            std::fill(prec.begin() + hc[0], prec.end(), (uint8_t)KWGPU_FP64);
        }
        rc = kwgpu::fd_price_device_prec(ctx, sub_o, (size_t)m, &p, sub_p, sub_s, st,
                                         prec.empty() ? nullptr : prec.data());
        if (rc == KWGPU_OK) {
            hipLaunchKernelGGL(scatter_prices, dim3(nblk(m, 256)), dim3(256), 0, st, sub_p, sub_s,
                               idx, m, d_prices);
            if (hipStreamSynchronize(st)) rc = kwgpu::ctx_fail(ctx, KWGPU_ERR_HIP, "FD fallback");
        }
    }
    if (n_fd) *n_fd = (size_t)hc[0] + (size_t)hc[1];
    cleanup();
    return rc;
}

}  // namespace

extern "C" {

int kwgpu_surrogate_save(kwgpu_ctx* ctx, const kwgpu_surrogate* s, const char* path) {
    if (!s || !path) return kwgpu::ctx_fail(ctx, KWGPU_ERR_INVALID, "kwgpu_surrogate_save: bad argument");
    FILE* f = std::fopen(path, "wb");
    if (!f) return kwgpu::ctx_fail(ctx, KWGPU_ERR_INVALID, std::string("cannot write ") + path);
    const uint32_t hdr[4] = {0x4b575355u /* "KWSU" */, kFileVersion, (uint32_t)s->members,
                             (uint32_t)kP};
    std::fwrite(hdr, sizeof(hdr), 1, f);
    // the FD configuration it emulates: density, scale, theta, x and t points, per option
    const double gd[3] = {s->density, s->scale, s->theta};
    const int32_t gi[3] = {s->x_dim, s->t_dim, s->per_option ? 1 : 0};
    std::fwrite(gd, sizeof(gd), 1, f);
    std::fwrite(gi, sizeof(gi), 1, f);
    std::vector<float> w(kP);
    for (int m = 0; m < s->members; ++m) {
        if (hipMemcpy(w.data(), s->p[m], kP * sizeof(float), hipMemcpyDeviceToHost)) {
            std::fclose(f);
            return kwgpu::ctx_fail(ctx, KWGPU_ERR_HIP, "kwgpu_surrogate_save: download");
        }
        std::fwrite(w.data(), sizeof(float), kP, f);
    }
    std::fclose(f);
    return KWGPU_OK;
}

int kwgpu_surrogate_load(kwgpu_ctx* ctx, const char* path, kwgpu_surrogate** out) {
    if (!path || !out) return kwgpu::ctx_fail(ctx, KWGPU_ERR_INVALID, "kwgpu_surrogate_load: bad argument");
    FILE* f = std::fopen(path, "rb");
    if (!f) return kwgpu::ctx_fail(ctx, KWGPU_ERR_INVALID, std::string("cannot read ") + path);
    uint32_t hdr[4] = {};
    if (std::fread(hdr, sizeof(hdr), 1, f) != 1 || hdr[0] != 0x4b575355u || hdr[1] != kFileVersion ||
        hdr[3] != (uint32_t)kP || hdr[2] < 1 || hdr[2] > 16) {
        std::fclose(f);
        return kwgpu::ctx_fail(ctx, KWGPU_ERR_INVALID, "kwgpu_surrogate_load: not a surrogate file");
    }
    double gd[3] = {};
    int32_t gi[3] = {};
    if (std::fread(gd, sizeof(gd), 1, f) != 1 || std::fread(gi, sizeof(gi), 1, f) != 1 ||
        gi[0] < 3 || gi[1] < 2) {
        std::fclose(f);
        return kwgpu::ctx_fail(ctx, KWGPU_ERR_INVALID, "kwgpu_surrogate_load: truncated file");
    }
    int rc = kwgpu_surrogate_create(ctx, (int32_t)hdr[2], 0, out);
    if (rc == KWGPU_OK) {
        (*out)->density = gd[0];
        (*out)->scale = gd[1];
        (*out)->theta = gd[2];
        (*out)->x_dim = gi[0];
        (*out)->t_dim = gi[1];
        (*out)->per_option = gi[2] != 0;
    }
    std::vector<float> w(kP);
    for (uint32_t m = 0; rc == KWGPU_OK && m < hdr[2]; ++m) {
        if (std::fread(w.data(), sizeof(float), kP, f) != (size_t)kP)
            rc = kwgpu::ctx_fail(ctx, KWGPU_ERR_INVALID, "kwgpu_surrogate_load: truncated file");
        else
            rc = kwgpu_surrogate_set_params(ctx, *out, (int32_t)m, w.data());
    }
    std::fclose(f);
    if (rc != KWGPU_OK && *out) {
        kwgpu_surrogate_destroy(*out);
        *out = nullptr;
    }
    return rc;
}

}  // extern "C"
