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
 * Threading: one context per host thread; every call on a context is
 * serialised on that context's HIP stream. Host-pointer entry points are
 * synchronous (like the reference's blocking price()); *_device entry points
 * are asynchronous on the stream passed in (NULL = the context's stream).
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
    KWGPU_STATUS_NONFINITE = 3    /* the solve produced a non-finite price                 */
};

/* Arithmetic of the FD march. */
enum {
    KWGPU_FP64 = 0,      /* everything in FP64: matches the reference to <= 1e-8            */
    KWGPU_FP32_MIXED = 1 /* FP64 grid, coefficients, LU pivots, V accumulation and
                            interpolation; FP32 RHS, substitutions and correction        */
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

/* Waits for all work queued on the context's stream. */
int kwgpu_synchronize(kwgpu_ctx* ctx);

/* The context's stream (a hipStream_t). */
void* kwgpu_stream(kwgpu_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* KWGPU_H */
