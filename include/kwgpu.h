/*
 * kwgpu.h -- C-ABI of the MI355X (gfx950) batched American-option FD pricer.
 *
 * Plain pointers and sizes only. This is the boundary the C++ API in
 * include/kw/ (kw::Pricer, kw::PricerFactory, kw::Portfolio) calls, and the
 * one any FFI (ctypes, cgo, JNI) binds.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the gituliar/kwinto-cuda source tree):
 *   kw::Pricer::init / kw::Pricer::price      src/Pricer/kwPricer.h:12-22
 *   kw::Fd1d_Pricer::init / price            src/Pricer/kwFd1d.cpp:9-160
 *   kw::Fd1d::solve / solveOne / value       src/Math/kwFd1d.cpp:11-158
 *   kw::solveTridiagonal                     src/Math/kwMath.cpp:16-49
 *   kw::BlackScholes_Pricer::price           src/Pricer/kwBlackScholes.cpp:13-50
 *   kw::Fd1d_BlackScholes_Pricer::price      src/Pricer/kwFd1d_BlackScholes.cpp:15-43
 *
 * Threading: one context per host thread. Host-pointer entry points are
 * synchronous (like the reference's blocking price()); *_device entry points
 * are asynchronous on the stream passed in (NULL = the context's stream). Calls
 * on one context may use different streams: an FD solve first waits for the
 * context's previous solve when that one ran on another stream (both reuse the
 * context's device buffers), and a surrogate prediction likewise for the
 * surrogate's previous one.
 */
#ifndef KWGPU_H
#define KWGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Byte-identical to kw::Option (src/Core/kwAsset.h:12-21), 56 bytes:
 * t@0 k@8 z@16 r@24 q@32 s@40 e@48 w@49, so a std::vector<kw::Option>
 * passes without copying. */
typedef struct kwgpu_option {
    double t;  /* time to maturity, years                 */
    double k;  /* strike                                  */
    double z;  /* volatility                              */
    double r;  /* interest rate                           */
    double q;  /* dividend rate                           */
    double s;  /* spot                                    */
    uint8_t e; /* early exercise: 0 European, 1 American  */
    int8_t w;  /* parity: -1 put, +1 call                 */
} kwgpu_option;

/* Return codes. */
enum {
    KWGPU_OK = 0,
    KWGPU_ERR_INVALID = 1,     /* bad argument (null pointer, grid size, params)        */
    KWGPU_ERR_HIP = 2,         /* HIP runtime error                                     */
    KWGPU_ERR_RANGE = 3,       /* an option's log(S/K) lies outside its x-grid          */
    KWGPU_ERR_NOMEM = 4,       /* device allocation failed                              */
    KWGPU_ERR_UNSUPPORTED = 5  /* configuration not built (e.g. layout x precision)     */
};

/* Per-option status written by the pricing entry points. */
enum {
    KWGPU_STATUS_OK = 0,
    KWGPU_STATUS_OUT_OF_GRID = 1, /* Fd1d::value range error, src/Math/kwFd1d.cpp:151-153 */
    KWGPU_STATUS_BAD_INPUT = 2,   /* T<=0, sigma<=0, S<=0, K<=0 or non-finite field        */
    KWGPU_STATUS_NONFINITE = 3,   /* the solve produced a non-finite price                 */
    KWGPU_STATUS_PRECISION = 4    /* FP32-mixed only: the PDE is beyond the mixed march's
                                     drift limit (see KWGPU_FP32_MIXED); solve it in FP64 */
};

/* Arithmetic of the FD march. */
enum {
    KWGPU_FP64 = 0,      /* everything in FP64: matches the reference to <= 1e-8            */
    KWGPU_FP32_MIXED = 1 /* FP64 grid, coefficients, LU pivots, step residual, V
                            accumulation, projection and interpolation; FP32
                            substitutions, carries and fix-ups: <= 1e-4 vs FP64.
                            Two limits, per PDE. Overflow: with scale * sigma *
                            sqrt(T) > 80 a call's FP32 increments can overflow.
                            Drift: with |r - q - sigma^2/2| T > scale * sigma *
                            sqrt(T) the solution drifts past the mesh's half-width
                            and the FP32 substitutions miss the bar (1 % vol over
                            10-30 years at a 25-30 % yield: errors to 4e-3). The
                            host entry points and the surrogate's FP32-mixed
                            fallback solve such PDEs in FP64.
                            kwgpu_fd1d_price_device (and _noalloc) solve as asked:
                            an overflowing call returns KWGPU_STATUS_NONFINITE, a
                            PDE past the drift limit KWGPU_STATUS_PRECISION (price
                            NaN). Validated for sigma 0.02-1.2, T 0.002-15, r
                            -0.03-0.2, q 0-0.25 within both limits.                  */
};

/* Where one option's system lives during the march. */
enum {
    KWGPU_LAYOUT_AUTO = 0,     /* wave layout when the grid fits, else lane layout          */
    KWGPU_LAYOUT_LANE_SOA = 1, /* one option per lane, state in HBM, option-fastest SoA     */
    KWGPU_LAYOUT_WAVE_REG = 2  /* one option per wavefront segment, rows spread over lanes,
                                  factored operator and state resident in VGPRs            */
};

/* Fd1d_Pricer configuration (src/Pricer/kwFd1d.cpp:9-19; theta src/Math/kwFd1d.h:31). */
typedef struct kwgpu_fd1d_params {
    double density;   /* FD1D.DENSITY, default 0.25                                  */
    double scale;     /* FD1D.SCALE, default 50                                      */
    double theta;     /* Crank-Nicolson theta in (0,1], default 0.5                  */
    int32_t x_dim;    /* FD1D.X_GRID_SIZE, points, default 512                       */
    int32_t t_dim;    /* FD1D.T_GRID_SIZE, points (steps = t_dim - 1), default 512   */
    int32_t precision;/* KWGPU_FP64 or KWGPU_FP32_MIXED                              */
    int32_t layout;   /* KWGPU_LAYOUT_*                                              */
    /* Optional per-option grid sizes (uneven batches). NULL = x_dim / t_dim for
     * every option. Host pointers for host entry points, device pointers for
     * *_device entry points. */
    const int32_t* x_dims;
    const int32_t* t_dims;
} kwgpu_fd1d_params;

typedef struct kwgpu_ctx kwgpu_ctx;

/* Fills the reference defaults: density 0.25, scale 50, theta 0.5, 512 x 512, FP64, AUTO. */
void kwgpu_fd1d_params_default(kwgpu_fd1d_params* p);

/* Creates a context on HIP device `device` (owns a stream and device buffers). */
int kwgpu_create(kwgpu_ctx** ctx, int device);
void kwgpu_destroy(kwgpu_ctx* ctx);

/* Text of the last error on this context ("" if none). Same wording as the
 * reference's kw::Error strings where a reference counterpart exists. */
const char* kwgpu_last_error(const kwgpu_ctx* ctx);

/*
 * FD price of a whole portfolio -- replaces Fd1d_Pricer::price
 * (src/Pricer/kwFd1d.cpp:21-160) including its PDE compression (:28-65):
 * options that share (t, r, q, z, e, w) are solved once and interpolated at
 * each spot. Host pointers; synchronous.
 * prices[i] = K * V(log(S/K)). status may be NULL.
 * Returns KWGPU_ERR_RANGE if any option's log(S/K) is outside its grid; the
 * message names the first such option as the reference does, and every other
 * option is still priced (the reference stops there, leaving later ones unset).
 * n == 0 returns KWGPU_OK and touches nothing (:25-26).
 */
int kwgpu_fd1d_price(kwgpu_ctx* ctx, const kwgpu_option* opts, size_t n,
                     const kwgpu_fd1d_params* params, double* prices, int32_t* status);

/*
 * Same pricing on device-resident buffers, one PDE per option (no compression),
 * asynchronous on `stream` (a hipStream_t; NULL = the context's stream).
 * d_status may be NULL. This is the throughput path.
 */
int kwgpu_fd1d_price_device(kwgpu_ctx* ctx, const kwgpu_option* d_opts, size_t n,
                            const kwgpu_fd1d_params* params, double* d_prices,
                            int32_t* d_status, void* stream);

/*
 * kwgpu_fd1d_price_device without any allocation (no counterpart in the reference: it is
 * for callers that cannot allocate device memory, such as code under HIP graph capture,
 * tight serving loops and harnesses that forbid hipMalloc). Same pricing, per-option
 * grids included, on device-resident options, prices and status, asynchronous on
 * `stream`; here params->x_dims / params->t_dims, when set, are HOST arrays of n sizes.
 * The PDE order, the sizes and the bucket counters of the batch live in a fixed device
 * scratch array of the library (12 MiB per device, held until the process ends): at most
 * 1048576 options (more returns KWGPU_ERR_INVALID), wave layout only (x_dim <= 2048, else
 * KWGPU_ERR_UNSUPPORTED). No graph and no side stream: the grid-size buckets run one
 * after another on `stream`. Calls on one context are ordered; calls from different
 * contexts on one GPU share the scratch and race when they overlap.
 */
int kwgpu_fd1d_price_device_noalloc(kwgpu_ctx* ctx, const kwgpu_option* d_opts, size_t n,
                                    const kwgpu_fd1d_params* params, double* d_prices,
                                    int32_t* d_status, void* stream);

/* Closed-form European price -- replaces BlackScholes_Pricer::price
 * (src/Pricer/kwBlackScholes.cpp:13-50): NaN for American options. Host pointers. */
int kwgpu_bs_price(kwgpu_ctx* ctx, const kwgpu_option* opts, size_t n, double* prices);
int kwgpu_bs_price_device(kwgpu_ctx* ctx, const kwgpu_option* d_opts, size_t n,
                          double* d_prices, void* stream);

/* FD1D-BS control variate -- replaces Fd1d_BlackScholes_Pricer::price
 * (src/Pricer/kwFd1d_BlackScholes.cpp:15-43): FD(as given) + BS(euro) - FD(euro).
 * Host pointers; synchronous. */
int kwgpu_fd1d_bs_price(kwgpu_ctx* ctx, const kwgpu_option* opts, size_t n,
                        const kwgpu_fd1d_params* params, double* prices, int32_t* status);

/*
 * Price and Greeks by central bumps of relative size `bump`, as defined by the
 * reference's test/portfolio.py:35-56 (delta, gamma, rho, theta, vega; bump 0.01 there):
 *   delta = (V(s(1+b)) - V(s(1-b))) / (2 b s)
 *   gamma = (V(s(1+b)) - 2 V + V(s(1-b))) / (b s)^2
 *   vega  = (V(z(1+b)) - V(z(1-b))) / (2 b z) / 100
 *   theta = (V(t(1+b)) - V(t(1-b))) / (2 b t) / 360     (derivative in expiry)
 *   rho   = (V(r(1+b)) - V(r(1-b))) / (2 b r) / 100
 * The spot bumps share each option's PDE (the PDE does not depend on s), so delta and
 * gamma come from the solved grid at no extra solve; vega, theta and rho cost two PDEs
 * each. All nine prices of every option are solved in one batch. A zero s, z, t or r
 * gives NaN for the Greek divided by it (the script would raise). `status` (may be
 * NULL) receives the first non-OK status of an option's nine prices; the return value
 * follows kwgpu_fd1d_price, with the failing bumped option in the message.
 */
typedef struct kwgpu_greeks {
    double price, delta, gamma, vega, theta, rho;
} kwgpu_greeks;

int kwgpu_fd1d_greeks(kwgpu_ctx* ctx, const kwgpu_option* opts, size_t n,
                      const kwgpu_fd1d_params* params, double bump, kwgpu_greeks* out,
                      int32_t* status);

/*
 * Autotuner (no counterpart in the reference). For every grid-size bucket of this
 * batch (x-points rounded up to a power of two, 64 .. 2048), times on a sample of the
 * bucket's options each wave-layout variant that holds it (rows per lane x lanes per
 * option: 4, 2 or 1 options per wave, or two waves per option), with a static and with
 * a persistent grid, and in FP64 also the lane layout; keeps the fastest per (bucket,
 * precision) in the context, which later calls use. `report` (may be NULL) receives a
 * JSON object with every timing. Device pointers as for kwgpu_fd1d_price_device (the
 * sampled options' prices are written); synchronous.
 */
int kwgpu_fd1d_autotune(kwgpu_ctx* ctx, const kwgpu_option* d_opts, size_t n,
                        const kwgpu_fd1d_params* params, double* d_prices, int32_t* d_status,
                        char* report, size_t report_len);

/* Saves / loads the context's autotuned choices (text, one "bucket precision variant
 * persistent" line each). Loading skips lines of an unknown precision, variants that
 * do not hold their bucket and lane-layout choices for FP32-mixed. */
int kwgpu_tune_save(kwgpu_ctx* ctx, const char* path);
int kwgpu_tune_load(kwgpu_ctx* ctx, const char* path);

/* Waits for all work queued on the context's stream. */
int kwgpu_synchronize(kwgpu_ctx* ctx);

/* The context's stream (a hipStream_t). */
void* kwgpu_stream(kwgpu_ctx* ctx);

/* ---------------------------------------------------------------------------------
 * Learned early-exercise premium (no counterpart in the reference; it sits in front
 * of the FD pricer above and uses it as its source of labels and as its fallback).
 *
 * A surrogate emulates one FD pricer configuration: density, scale, theta and either
 * one x_dim x t_dim grid or per-option grids. An ensemble of MLPs 8 -> 256 -> 256 ->
 * 256 -> 256 -> 1 with SiLU predicts, at a mesh node x_j of an option's own grid, the
 * FD value there minus the Black-Scholes price, as y = (premium - base) /
 * (base + 1e-4 K), where base is the premium from the option's exercise boundary,
 * solved on the device by the boundary's integral-equation fixed point (a control
 * variate, whose size sets the price scale of the network's error). Features: log(S/K),
 * T, sigma, r, q, w normalised to the box K/S in [0.4, 2], T in [1/52, 5],
 * sigma in [0.1, 0.6], r in [0, 0.1], q in [0, 0.08], and log2 of the grid's x and t
 * points.
 * A price is formed as the FD pricer forms its own (Fd1d::value,
 * src/Math/kwFd1d.cpp:139-158): the network at the two nodes bracketing log(S/K), BS
 * added at each, linear interpolation with the reference's weights. The labels are the
 * FD pricer's values at those nodes.
 * The three 256x256 layers run on BF16 MFMA with FP32 accumulation (forward and
 * backward); the 8-wide input layer and the 1-wide output layer run in FP32; Adam
 * keeps FP32 master weights.
 * --------------------------------------------------------------------------------- */
typedef struct kwgpu_surrogate kwgpu_surrogate;

/* Parameters per member, in this order: W1[256][8], b1[256], W2[256][256], b2[256],
 * W3[256][256], b3[256], W4[256][256], b4[256], W5[256], b5[1]. */
#define KWGPU_SURROGATE_PARAMS 199937

typedef struct kwgpu_surrogate_train_params {
    int32_t epochs;     /* passes over the training set                         */
    int32_t batch;      /* samples per Adam step                                 */
    float lr;           /* initial learning rate (cosine decay to lr_final)      */
    float lr_final;
    float beta1, beta2, eps;
    uint64_t seed;      /* shuffling                                            */
} kwgpu_surrogate_train_params;

void kwgpu_surrogate_train_params_default(kwgpu_surrogate_train_params* p);

/* Creates `members` networks with seeded (He-uniform) initial weights. The surrogate
 * emulates the reference defaults on a 1024 x 1024 grid until kwgpu_surrogate_fit
 * sets the configuration. */
int kwgpu_surrogate_create(kwgpu_ctx* ctx, int32_t members, uint64_t seed,
                           kwgpu_surrogate** out);
void kwgpu_surrogate_destroy(kwgpu_surrogate* s);

/* Labels for fitting a surrogate to the FD pricer configured by `fd` (precision FP64;
 * fd->x_dims / fd->t_dims, when set, device arrays of n per-option grid sizes): option i
 * gives two node options, d_nodes[i] and d_nodes[n + i], equal to it but with spot
 * K e^{x_{j-1}} and K e^{x_j} at the mesh nodes that bracket its log(S/K) on its own grid
 * (the first node x_j >= log(S/K), as Fd1d::value brackets, src/Math/kwFd1d.cpp:139-158);
 * d_premium[r] is the FD price of node option r (the FD value at that node) minus its
 * Black-Scholes European price. d_keep[i] is 0 where option i has no such bracket (spot
 * outside the grid or in its first cell) or a node failed to price. Device pointers
 * (d_nodes 2n, d_premium 2n, d_keep n); synchronous. */
int kwgpu_surrogate_labels(kwgpu_ctx* ctx, const kwgpu_option* d_opts, size_t n,
                           const kwgpu_fd1d_params* fd, kwgpu_option* d_nodes, double* d_premium,
                           uint8_t* d_keep);

/* Fits the surrogate to the FD pricer configured by `fd` (precision FP64; fd->x_dims /
 * fd->t_dims, when set, device arrays of n per-option grid sizes): each option's two
 * node options (kwgpu_surrogate_labels) become two labelled rows; options with
 * d_keep 0 are dropped. Then trains every member on the 2 * (*n_pairs) rows;
 * member m shuffles with seed + m. loss_out (host, may be NULL) receives epochs x
 * members mean training losses; *n_pairs (may be NULL) the options kept. The surrogate
 * records `fd` as the configuration it emulates. Synchronous. */
int kwgpu_surrogate_fit(kwgpu_ctx* ctx, kwgpu_surrogate* s, const kwgpu_option* d_opts, size_t n,
                        const kwgpu_fd1d_params* fd, const kwgpu_surrogate_train_params* p,
                        float* loss_out, size_t* n_pairs);

/* Trains every member on (option, premium) rows held on the device, each row's features
 * taken on the surrogate's single grid (not for a per-option surrogate); loss_out as for
 * kwgpu_surrogate_fit. Synchronous. */
int kwgpu_surrogate_train(kwgpu_ctx* ctx, kwgpu_surrogate* s, const kwgpu_option* d_opts,
                          const double* d_premium, size_t n,
                          const kwgpu_surrogate_train_params* p, float* loss_out);

/* Ensemble inference at each option's spot, formed from the two bracketing nodes as
 * above: premium (price units: the mean surrogate price minus BS at the spot), spread
 * (max - min of the members' prices) and whether the option lies in the training box
 * with a labelled bracket. `fd` (may be NULL for a single-grid surrogate) must be the
 * configuration the surrogate emulates and gives per-option grids. Device pointers,
 * asynchronous on `stream` (NULL = the context's stream); a prediction on another
 * stream than the surrogate's previous one waits for that one. */
int kwgpu_surrogate_predict(kwgpu_ctx* ctx, kwgpu_surrogate* s, const kwgpu_option* d_opts,
                            size_t n, const kwgpu_fd1d_params* fd, double* d_premium,
                            double* d_spread, uint8_t* d_in_box, void* stream);

/* Surrogate-first pricing: options in the box whose ensemble spread is <= spread_max
 * are priced by the surrogate (BS + premium); the rest go to the FD pricer with `fd`.
 * The ensemble runs only on candidates: in-box options with a labelled bracket whose
 * baseline scale (kwgpu_surrogate_baseline's premium + 1e-4 K, the price scale of the
 * network's error) is at most scale_max (INFINITY: every in-box option); the others go to FD directly
 * (fd->precision FP64, or FP32_MIXED with the options beyond the mixed march's limits,
 * see KWGPU_FP32_MIXED, sent to FP64). fd->x_dims / fd->t_dims, when set, are device arrays of
 * n per-option grid sizes, and each routed option is solved on its own. The surrogate
 * serves only when `fd` is the configuration it emulates (same density, scale, theta
 * and grid, or per-option grids for a per-option surrogate); otherwise, or with a
 * negative spread_max, every option goes to FD and the ensemble is skipped.
 * d_route[i] is 0 for the surrogate, 1 for FD at fd->precision, 2 for FD in FP64 (with
 * an FP64 fallback, every option sent to FD). Options sent to FD are solved as one
 * batch: route 1's in index order, then route 2's in index order.
 * *n_fd receives the number routed to FD. Device pointers;
 * synchronous (the FD subset is compacted on the device). Per-option work buffers are
 * kept with `s` and grown on demand. */
int kwgpu_surrogate_price(kwgpu_ctx* ctx, kwgpu_surrogate* s, const kwgpu_option* d_opts,
                          size_t n, double spread_max, double scale_max,
                          const kwgpu_fd1d_params* fd, double* d_prices, int32_t* d_route,
                          size_t* n_fd);

/* The gate by option class. Classes: put/call x n_t maturity bands x n_m log-moneyness
 * bands, split at the inner edges t_edges (T, years) and m_edges (log(S/K)), each
 * increasing; a value equal to an edge falls in the band above it. Class (c, a, b) --
 * c = 0 for puts (w < 0), 1 for calls, band a of T, band b of log(S/K) -- has index
 * (c * n_t + a) * n_m + b. Its in-box options are served under spread_max[index] and
 * scale_max[index] exactly as kwgpu_surrogate_price serves every option under its two
 * bounds; a negative spread_max closes the class. Bounds calibrated on each class apart
 * keep, class by class, what a calibration guarantees (a new option under the bounds
 * exceeds the k held-out options' largest error with probability at most 1/(k+1)),
 * while classes whose errors are small serve more. */
#define KWGPU_GATE_MAX_T 8
#define KWGPU_GATE_MAX_M 4
#define KWGPU_GATE_MAX_CLASSES (2 * KWGPU_GATE_MAX_T * KWGPU_GATE_MAX_M)
typedef struct kwgpu_gate {
    int32_t n_t;                                /* maturity bands, 1..KWGPU_GATE_MAX_T   */
    int32_t n_m;                                /* log-moneyness bands, 1..MAX_M         */
    double t_edges[KWGPU_GATE_MAX_T - 1];       /* n_t - 1 used                          */
    double m_edges[KWGPU_GATE_MAX_M - 1];       /* n_m - 1 used                          */
    double spread_max[KWGPU_GATE_MAX_CLASSES];  /* 2 n_t n_m used                        */
    double scale_max[KWGPU_GATE_MAX_CLASSES];
} kwgpu_gate;

/* kwgpu_surrogate_price with the bounds of `gate`'s classes (no reference counterpart).
 * KWGPU_ERR_INVALID for band counts outside their ranges or edges that are not finite
 * and increasing. With one band each (n_t = n_m = 1: two classes, puts and calls) and
 * the same bounds for both classes it is kwgpu_surrogate_price. */
int kwgpu_surrogate_price_gated(kwgpu_ctx* ctx, kwgpu_surrogate* s, const kwgpu_option* d_opts,
                                size_t n, const kwgpu_gate* gate, const kwgpu_fd1d_params* fd,
                                double* d_prices, int32_t* d_route, size_t* n_fd);

/* Weights of one member as FP32 (KWGPU_SURROGATE_PARAMS values, host memory). */
int kwgpu_surrogate_get_params(kwgpu_ctx* ctx, const kwgpu_surrogate* s, int32_t member,
                               float* host_params);
int kwgpu_surrogate_set_params(kwgpu_ctx* ctx, kwgpu_surrogate* s, int32_t member,
                               const float* host_params);

/* One forward/backward pass of `member` on normalised features (n x 8, FP32) and
 * scaled targets (n, FP32), both host memory, n <= 131072: returns the mean squared
 * error and its gradient with respect to every parameter (host, KWGPU_SURROGATE_PARAMS
 * values). */
int kwgpu_surrogate_loss_grad(kwgpu_ctx* ctx, kwgpu_surrogate* s, int32_t member,
                              const float* host_features, const float* host_targets, size_t n,
                              float* host_loss, float* host_grad);

/* The analytic baseline premium of each option, host memory, price units: Ju & Zhong's
 * correction of Barone-Adesi & Whaley, continuous dividend yield. The gate's scale is
 * this premium + 1e-4 K. */
int kwgpu_surrogate_baseline(kwgpu_ctx* ctx, const kwgpu_option* host_opts, size_t n,
                             double* host_out);

/* Normalised features (n x 8) of options priced on an x_dim x t_dim grid, host memory,
 * as the network sees them. */
int kwgpu_surrogate_features(kwgpu_ctx* ctx, const kwgpu_option* host_opts, size_t n,
                             int32_t x_dim, int32_t t_dim, float* host_features);

/* Saves / loads an ensemble (a small binary file: magic, version, members, the FD
 * configuration it emulates, FP32 weights). */
int kwgpu_surrogate_save(kwgpu_ctx* ctx, const kwgpu_surrogate* s, const char* path);
int kwgpu_surrogate_load(kwgpu_ctx* ctx, const char* path, kwgpu_surrogate** out);

#ifdef __cplusplus
}
#endif

#endif /* KWGPU_H */
