/*
 * kw_oracle.h -- CPU oracle for the batched American-option FD pricer.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C++ FP64 restatement of the
 * reference pricer's algorithm, used as the checker by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg. Nothing in the
 * product path (kwinto-cuda_amd/) includes, links or calls it.
 *
 * Every function follows the reference file:line given in its comment
 * (paths relative to the gituliar/kwinto-cuda source tree).
 */
#pragma once

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Byte-identical to kw::Option (src/Core/kwAsset.h:12-21): 56 bytes. */
typedef struct kwo_option {
    double t; /* time to maturity (years)          */
    double k; /* strike                            */
    double z; /* volatility                        */
    double r; /* interest rate                     */
    double q; /* dividend rate                     */
    double s; /* spot                              */
    uint8_t e; /* early exercise: 1 American        */
    int8_t w;  /* parity: -1 put, +1 call           */
} kwo_option;

/*
 * Fd1d_Pricer::price (src/Pricer/kwFd1d.cpp:21-160) with FD1D.DENSITY,
 * FD1D.SCALE, FD1D.T_GRID_SIZE, FD1D.X_GRID_SIZE given explicitly and
 * theta = 0.5 (src/Math/kwFd1d.h:31).
 *
 * Returns 0 on success. On failure returns non-zero and writes the
 * reference's error text into err (if errlen > 0); prices[] then holds the
 * values written before the failure (the reference leaves partial output).
 * nthreads <= 0 uses std::thread::hardware_concurrency().
 */
int kwo_fd1d_price(const kwo_option* opts, size_t n, double density, double scale,
                   int64_t t_dim, int64_t x_dim, int nthreads, double* prices,
                   char* err, size_t errlen);

/*
 * Per-option status variant: never stops at the first failing option.
 * status[i] = 0 ok, 1 spot outside the x-grid (Fd1d::value error,
 * src/Math/kwFd1d.cpp:151-153). prices[i] is NaN when status[i] != 0.
 */
int kwo_fd1d_price_status(const kwo_option* opts, size_t n, double density, double scale,
                          int64_t t_dim, int64_t x_dim, int nthreads, double* prices,
                          int32_t* status);

/*
 * Uneven grids: option i is priced on an x_dims[i] x t_dims[i] grid. Equivalent to
 * calling the reference pricer once per distinct (x, t) pair.
 */
int kwo_fd1d_price_uneven(const kwo_option* opts, size_t n, double density, double scale,
                          const int32_t* t_dims, const int32_t* x_dims, int nthreads,
                          double* prices, int32_t* status);

/* BlackScholes_Pricer::price (src/Pricer/kwBlackScholes.cpp:13-50). */
int kwo_bs_price(const kwo_option* opts, size_t n, double* prices);

/* Fd1d_BlackScholes_Pricer::price (src/Pricer/kwFd1d_BlackScholes.cpp:15-43). */
int kwo_fd1d_bs_price(const kwo_option* opts, size_t n, double density, double scale,
                      int64_t t_dim, int64_t x_dim, int nthreads, double* prices,
                      char* err, size_t errlen);

/* One PDE solve exposed for grid/solution inspection (tests only):
 * writes the x-grid and the solution V (units of K) for one option. */
int kwo_fd1d_solve_one(const kwo_option* opt, double density, double scale, int64_t t_dim,
                       int64_t x_dim, double* x_out, double* v_out);

/* kw::solveTridiagonal (src/Math/kwMath.cpp:16-49); returns 0 or the error number. */
int kwo_solve_tridiagonal(int n, const double* al, const double* a, const double* au,
                          const double* y, double* x);

#ifdef __cplusplus
}
#endif
