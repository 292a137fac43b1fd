// kw_oracle.cpp -- CPU oracle (test infrastructure, never part of the product path).
//
// FP64 restatement of the reference FD pricer, written to reproduce its arithmetic
// operation by operation so that it agrees with the compiled reference to the last
// bit on the same compiler flags. Each block cites the reference lines it follows
// (paths relative to the gituliar/kwinto-cuda tree).
//
// Threading follows the reference's model -- one job per deduplicated PDE on a pool
// of hardware_concurrency() workers (src/Math/kwFd1d.cpp:45-55,
// src/kwThreadPool.cpp:9-21) -- but joins with an atomic work counter, so the join
// cannot return before the last PDE is solved.

#include "kw_oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

namespace {

using f64 = double;

constexpr f64 kTheta = 0.5;  // src/Math/kwFd1d.h:31

// KWO_PLANT=n builds a deliberately wrong oracle (tests/test_oracle.py shows that the
// golden-vector check rejects each one). The default build plants nothing.
#ifndef KWO_PLANT
#define KWO_PLANT 0
#endif

struct Pde {  // Fd1dPde, src/Math/kwFd1d.h:15-23
    f64 t, a0, ax, axx;
    bool earlyExercise;
};

// kw::solveTridiagonal, src/Math/kwMath.cpp:16-49 (Numerical-Recipes Thomas with
// gam/bet; the `if (x[j] < 0) continue;` at :39-40 is a no-op and is omitted).
int solve_tridiagonal(int n, const f64* al, const f64* a, const f64* au, const f64* y, f64* x,
                      std::vector<f64>& gam) {
    if (a[0] == 0) return 1;
    if (n <= 2) return 3;
    gam.resize(n);
    f64 bet;
    x[0] = y[0] / (bet = a[0]);
    for (int j = 1; j < n; j++) {
        gam[j] = au[j - 1] / bet;
        bet = a[j] - al[j] * gam[j];
        if (bet == 0) return 2;
        x[j] = (y[j] - al[j] * x[j - 1]) / bet;
    }
    for (int j = n - 2; j >= 0; --j) x[j] -= gam[j + 1] * x[j + 1];
    return 0;
}

// Grid and payoff for one PDE: src/Pricer/kwFd1d.cpp:95-139.
void fill_grids(const kwo_option& a, f64 density, f64 scale, int64_t tDim, int64_t xDim, f64* t,
                f64* x, f64* v) {
    // t-grid, :95-103
    {
        f64 tMin = 0.0, tMax = a.t;
        f64 dt = (tMax - tMin) / (uint64_t)(KWO_PLANT == 3 ? tDim : tDim - 1);
        for (int64_t j = 0; j < tDim; ++j) t[j] = tMin + j * dt;
    }
    // x-grid, :105-125 (sinh mesh centred at log(S/K) = 0)
    {
        const f64 xMid = 0.;
        const f64 xMin = xMid - scale * a.z * std::sqrt(a.t);
        const f64 xMax = xMid + scale * a.z * std::sqrt(a.t);
        const f64 yMin = std::asinh((xMin - xMid) / density);
        const f64 yMax = std::asinh((xMax - xMid) / density);
        const f64 dy = 1. / (uint64_t)(xDim - 1);
        for (int64_t j = 0; j < xDim; ++j) {
            const f64 yj = j * dy;
            x[j] = xMid + density * std::sinh(yMin * (1.0 - yj) + yMax * yj);
        }
    }
    // payoff in units of K, :127-139
    for (int64_t j = 0; j < xDim; ++j) {
        if (a.w < 0 && KWO_PLANT != 5)
            v[j] = std::max<f64>(0, 1. - std::exp(x[j]));
        else
            v[j] = std::max<f64>(0, std::exp(x[j]) - 1.);
    }
}

struct Scratch {
    std::vector<f64> bl, b, bu, w, gam;
};

// Fd1d::solveOne, src/Math/kwFd1d.cpp:61-136. The reference stores the constant
// coefficients a0/ax/axx once per time row (:29-37); the values are identical for
// every row, so they are read from the PDE here. V starts as the payoff (:40-41).
void solve_one(const Pde& pde, int64_t tDim, int64_t xDim, const f64* t, const f64* x,
               const f64* payoff, f64* V, Scratch& s) {
    s.bl.resize(xDim);
    s.b.resize(xDim);
    s.bu.resize(xDim);
    s.w.resize(xDim);
    const f64 th = KWO_PLANT == 1 ? 0.55 : kTheta;
    const f64 a0 = pde.a0, ax = pde.ax, axx = pde.axx;
    for (int64_t j = 0; j < xDim; ++j) V[j] = payoff[j];

    for (int ti = (int)tDim - 2; ti >= 0; --ti) {
        f64 dt = t[ti + 1] - t[ti];  // :72
        {                            // row 0, one-sided, :73-84
            const int64_t xi = 0;
            const f64 inv_dx = 1. / (x[xi + 1] - x[xi]);
            s.bl[xi] = 0;
            s.b[xi] = 1 - th * dt * (a0 - inv_dx * ax);
            s.bu[xi] = -th * dt * (inv_dx * ax);
            s.w[xi] = (1 + (1 - th) * dt * a0) * V[xi] +
                      (1 - th) * dt * (ax * inv_dx) * (V[xi + 1] - V[xi]);
        }
        for (int64_t xi = 1; xi < xDim - 1; ++xi) {  // interior, :86-102
            const f64 inv_dxu = 1. / (x[xi + 1] - x[xi]);
            const f64 inv_dxm = 1. / (x[xi + 1] - x[xi - 1]);
            const f64 inv_dxd = 1. / (x[xi] - x[xi - 1]);
            const f64 inv_dx2u = 2. * (KWO_PLANT == 2 ? inv_dxd : inv_dxu) * inv_dxm;
            const f64 inv_dx2m = 2. * inv_dxd * inv_dxu;
            const f64 inv_dx2l = 2. * (KWO_PLANT == 2 ? inv_dxu : inv_dxd) * inv_dxm;
            s.bl[xi] = -th * dt * (-inv_dxm * ax + inv_dx2l * axx);
            s.b[xi] = 1 - th * dt * (a0 - inv_dx2m * axx);
            s.bu[xi] = -th * dt * (inv_dxm * ax + inv_dx2u * axx);
            s.w[xi] = (1 + (1 - th) * dt * a0) * V[xi] +
                      (1 - th) * dt * (ax * inv_dxm) * (V[xi + 1] - V[xi - 1]) +
                      (1 - th) * dt * (axx) * (inv_dx2u * V[xi + 1] - inv_dx2m * V[xi] + inv_dx2l * V[xi - 1]);
        }
        {  // row N-1, one-sided, :104-114
            const int64_t xi = xDim - 1;
            const f64 inv_dx = 1. / (x[xi] - x[xi - 1]);
            s.bl[xDim - 1] = -th * dt * (-inv_dx * ax);
            s.b[xDim - 1] = 1 - th * dt * (a0 + inv_dx * ax);
            s.bu[xDim - 1] = 0;
            s.w[xDim - 1] = (1 + (1 - th) * dt * a0) * V[xi] +
                            (1 - th) * dt * (ax * inv_dx) * (V[xi] - V[xi - 1]);
        }
        // :118-125 -- the solver's error code is ignored by the reference; on an
        // early return V keeps whatever the solver had written.
        (void)solve_tridiagonal((int)xDim, s.bl.data(), s.b.data(), s.bu.data(), s.w.data(), V,
                                s.gam);
        // post-solve projection on nodes 0..N-2, :129-132
        if (pde.earlyExercise)
            for (int64_t xi = 0; xi < xDim - 1; ++xi) V[xi] = std::max(V[xi], payoff[xi]);
    }
}

// Fd1d::value, src/Math/kwFd1d.cpp:139-158: first node with x_j >= x*, linear interpolation.
bool value_at(const f64* x, const f64* V, int64_t xDim, f64 x_, f64& v, std::string* err) {
    int64_t xi = 0;
    while ((xi < xDim) && (x[xi] < x_)) ++xi;
    if ((xi == 0) || (xi == xDim)) {
        if (err)
            *err = "Fd1d::value: x=" + std::to_string(x_) + " not in range (" +
                   std::to_string(x[0]) + ", " + std::to_string(x[xDim - 1]) + ")";
        return false;
    }
    const int64_t hi = KWO_PLANT == 4 ? std::min<int64_t>(xi + 1, xDim - 1) : xi;
    v = ((x[xi] - x_) * V[xi - 1] + (x_ - x[xi - 1]) * V[hi]) / (x[xi] - x[xi - 1]);
    return true;
}

// PDE compression: src/Pricer/kwFd1d.cpp:28-65 (sort on (t,r,q,z,e,w)).
void dedup(const kwo_option* a, size_t n, std::vector<uint32_t>& asset2pde,
           std::vector<uint32_t>& pde2asset) {
    auto less = [](const kwo_option& l, const kwo_option& r) {
        return std::tie(l.t, l.r, l.q, l.z, l.e, l.w) < std::tie(r.t, r.r, r.q, r.z, r.e, r.w);
    };
    std::vector<uint32_t> sorted(n);
    std::iota(sorted.begin(), sorted.end(), 0u);
    std::sort(sorted.begin(), sorted.end(),
              [&](uint32_t l, uint32_t r) { return less(a[l], a[r]); });
    asset2pde.assign(n, 0);
    pde2asset.clear();
    pde2asset.push_back(sorted[0]);
    for (size_t i = 1, k = 0; i < n; i++) {
        auto l = sorted[i - 1], r = sorted[i];
        if (less(a[l], a[r]) != less(a[r], a[l])) {
            k += 1;
            pde2asset.push_back(r);
        }
        asset2pde[r] = (uint32_t)k;
    }
}

// Fork-join over [0, m) on nthreads workers; returns after every job has finished.
template <class F>
void parallel_for(size_t m, int nthreads, F&& job) {
    if (nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
    nthreads = (int)std::min<size_t>((size_t)nthreads, std::max<size_t>(m, 1));
    std::atomic<size_t> next{0};
    auto worker = [&]() {
        for (;;) {
            size_t i = next.fetch_add(1);
            if (i >= m) break;
            job(i);
        }
    };
    if (nthreads == 1) {
        worker();
        return;
    }
    std::vector<std::thread> pool;
    pool.reserve(nthreads);
    for (int i = 0; i < nthreads; ++i) pool.emplace_back(worker);
    for (auto& th : pool) th.join();
}

// Whole-portfolio FD price at one grid size. mode 0: stop at the first value()
// error with the reference's message; mode 1: per-option status.
int fd1d_price_impl(const kwo_option* opts, size_t n, f64 density, f64 scale, int64_t tDim,
                    int64_t xDim, int nthreads, f64* prices, int32_t* status, std::string* err) {
    if (n == 0) return 0;  // :25-26
    std::vector<uint32_t> asset2pde, pde2asset;
    dedup(opts, n, asset2pde, pde2asset);
    const size_t m = pde2asset.size();

    std::vector<Pde> pdes(m);
    for (size_t i = 0; i < m; ++i) {  // :71-86
        const auto& a = opts[pde2asset[i]];
        pdes[i].t = a.t;
        pdes[i].a0 = -a.r;
        pdes[i].ax = a.r - a.q - a.z * a.z / 2;
        pdes[i].axx = a.z * a.z / 2;
        pdes[i].earlyExercise = a.e;
    }

    std::vector<f64> X((size_t)m * xDim), V((size_t)m * xDim);
    parallel_for(m, nthreads, [&](size_t i) {
        thread_local Scratch s;
        thread_local std::vector<f64> t, payoff;
        t.resize(tDim);
        payoff.resize(xDim);
        f64* x = &X[i * xDim];
        fill_grids(opts[pde2asset[i]], density, scale, tDim, xDim, t.data(), x, payoff.data());
        solve_one(pdes[i], tDim, xDim, t.data(), x, payoff.data(), &V[i * xDim], s);
    });

    for (size_t i = 0; i < n; i++) {  // :146-157
        const auto& a = opts[i];
        f64 v = 0;
        auto x_ = std::log(a.s / a.k);
        std::string e;
        const size_t ni = asset2pde[i];
        if (!value_at(&X[ni * xDim], &V[ni * xDim], xDim, x_, v, status ? nullptr : &e)) {
            if (status) {
                status[i] = 1;
                prices[i] = NAN;
                continue;
            }
            if (err) *err = "Fd1d_Pricer::price " + e;
            return 1;
        }
        if (status) status[i] = 0;
        prices[i] = a.k * v;
    }
    return 0;
}

void copy_err(const std::string& e, char* err, size_t errlen) {
    if (!err || errlen == 0) return;
    std::snprintf(err, errlen, "%s", e.c_str());
}

f64 cdf_normal(f64 x) { return 0.5 * (1 + std::erf(x / std::sqrt(2.0))); }  // src/Math/kwMath.cpp:9-13

// BlackScholes_Pricer::priceOne, src/Pricer/kwBlackScholes.cpp:27-50.
f64 bs_one(const kwo_option& a) {
    if (a.e) return NAN;
    const f64 k = a.k, q = a.q, r = a.r, s = a.s, t = a.t, z = a.z;
    const int8_t w = a.w;
    f64 zt = z * std::sqrt(t);
    f64 d1 = 1 / zt * (std::log(s / k) + (r - q + 0.5 * z * z) * t);
    f64 d2 = d1 - zt;
    return w * (s * cdf_normal(w * d1) * std::exp(-q * t) - k * cdf_normal(w * d2) * std::exp(-r * t));
}

}  // namespace

extern "C" {

int kwo_fd1d_price(const kwo_option* opts, size_t n, double density, double scale, int64_t t_dim,
                   int64_t x_dim, int nthreads, double* prices, char* err, size_t errlen) {
    std::string e;
    int rc = fd1d_price_impl(opts, n, density, scale, t_dim, x_dim, nthreads, prices, nullptr, &e);
    if (rc) copy_err(e, err, errlen);
    return rc;
}

int kwo_fd1d_price_status(const kwo_option* opts, size_t n, double density, double scale,
                          int64_t t_dim, int64_t x_dim, int nthreads, double* prices,
                          int32_t* status) {
    return fd1d_price_impl(opts, n, density, scale, t_dim, x_dim, nthreads, prices, status,
                           nullptr);
}

int kwo_fd1d_price_uneven(const kwo_option* opts, size_t n, double density, double scale,
                          const int32_t* t_dims, const int32_t* x_dims, int nthreads,
                          double* prices, int32_t* status) {
    // One reference call per distinct (x, t) grid; results scattered back.
    std::vector<size_t> idx(n);
    std::iota(idx.begin(), idx.end(), size_t(0));
    std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) {
        return std::tie(x_dims[a], t_dims[a]) < std::tie(x_dims[b], t_dims[b]);
    });
    size_t i = 0;
    std::vector<kwo_option> sub;
    std::vector<f64> p;
    std::vector<int32_t> st;
    while (i < n) {
        size_t j = i;
        while (j < n && x_dims[idx[j]] == x_dims[idx[i]] && t_dims[idx[j]] == t_dims[idx[i]]) ++j;
        sub.clear();
        for (size_t k = i; k < j; ++k) sub.push_back(opts[idx[k]]);
        p.assign(sub.size(), 0);
        st.assign(sub.size(), 0);
        fd1d_price_impl(sub.data(), sub.size(), density, scale, t_dims[idx[i]], x_dims[idx[i]],
                        nthreads, p.data(), st.data(), nullptr);
        for (size_t k = i; k < j; ++k) {
            prices[idx[k]] = p[k - i];
            status[idx[k]] = st[k - i];
        }
        i = j;
    }
    return 0;
}

int kwo_bs_price(const kwo_option* opts, size_t n, double* prices) {
    for (size_t i = 0; i < n; ++i) prices[i] = bs_one(opts[i]);
    return 0;
}

int kwo_fd1d_bs_price(const kwo_option* opts, size_t n, double density, double scale,
                      int64_t t_dim, int64_t x_dim, int nthreads, double* prices, char* err,
                      size_t errlen) {
    // src/Pricer/kwFd1d_BlackScholes.cpp:15-43: FD(amer) + (BS - FD(euro)).
    std::string e;
    if (fd1d_price_impl(opts, n, density, scale, t_dim, x_dim, nthreads, prices, nullptr, &e)) {
        copy_err("Fd1d_BlackScholes_Pricer::price : " + e, err, errlen);
        return 1;
    }
    std::vector<kwo_option> euro(opts, opts + n);
    for (auto& a : euro) a.e = 0;
    std::vector<f64> fd(n), bs(n);
    if (fd1d_price_impl(euro.data(), n, density, scale, t_dim, x_dim, nthreads, fd.data(), nullptr,
                        &e)) {
        copy_err("Fd1d_BlackScholes_Pricer::price : " + e, err, errlen);
        return 1;
    }
    for (size_t i = 0; i < n; ++i) bs[i] = bs_one(euro[i]);
    for (size_t i = 0; i < n; ++i) prices[i] += bs[i] - fd[i];
    return 0;
}

int kwo_fd1d_solve_one(const kwo_option* opt, double density, double scale, int64_t t_dim,
                       int64_t x_dim, double* x_out, double* v_out) {
    Pde pde{opt->t, -opt->r, opt->r - opt->q - opt->z * opt->z / 2, opt->z * opt->z / 2,
            (bool)opt->e};
    std::vector<f64> t(t_dim), payoff(x_dim);
    fill_grids(*opt, density, scale, t_dim, x_dim, t.data(), x_out, payoff.data());
    Scratch s;
    solve_one(pde, t_dim, x_dim, t.data(), x_out, payoff.data(), v_out, s);
    return 0;
}

int kwo_solve_tridiagonal(int n, const double* al, const double* a, const double* au,
                          const double* y, double* x) {
    std::vector<f64> gam;
    return solve_tridiagonal(n, al, a, au, y, x, gam);
}

}  // extern "C"
