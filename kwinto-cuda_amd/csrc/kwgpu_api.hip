// kwgpu_api.hip -- host implementation of the C-ABI in include/kwgpu.h.
//
// Owns the per-context HIP stream and device buffers, reproduces the reference
// pricer's host-side logic around the solve (config defaults, PDE compression,
// error text) and dispatches the FD kernels by layout and grid size.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <tuple>
#include <vector>

#include "kwgpu_internal.h"

namespace {

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
};

}  // namespace

struct kwgpu_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    DevBuf opts, prices, status, rep, off, list, order, xd, td, ws;
    DevBuf opts2, prices2, status2;  // second leg of FD1D-BS
};

namespace {

using kwgpu::FdBatch;

int fail(kwgpu_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

int hip_fail(kwgpu_ctx* c, hipError_t e, const char* where) {
    return fail(c, KWGPU_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

hipError_t ensure(DevBuf& b, size_t bytes) {
    if (bytes == 0) bytes = 1;
    if (b.cap >= bytes) return hipSuccess;
    if (b.p) {
        hipError_t e = hipFree(b.p);
        b.p = nullptr;
        b.cap = 0;
        if (e != hipSuccess) return e;
    }
    size_t want = bytes + bytes / 4;
    hipError_t e = hipMalloc(&b.p, want);
    if (e != hipSuccess) {
        b.p = nullptr;
        return e;
    }
    b.cap = want;
    return hipSuccess;
}

void free_buf(DevBuf& b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
}

int check_params(kwgpu_ctx* c, const kwgpu_fd1d_params* p) {
    if (!p) return fail(c, KWGPU_ERR_INVALID, "kwgpu: null params");
    if (!(p->density > 0) || !std::isfinite(p->density))
        return fail(c, KWGPU_ERR_INVALID, "kwgpu: FD1D.DENSITY must be > 0");
    if (!(p->scale > 0) || !std::isfinite(p->scale))
        return fail(c, KWGPU_ERR_INVALID, "kwgpu: FD1D.SCALE must be > 0");
    if (!(p->theta > 0) || !(p->theta <= 1))
        return fail(c, KWGPU_ERR_INVALID, "kwgpu: theta must lie in (0, 1]");
    if (!p->x_dims && p->x_dim < 3)
        return fail(c, KWGPU_ERR_INVALID, "kwgpu: FD1D.X_GRID_SIZE must be >= 3");
    if (!p->t_dims && p->t_dim < 2)
        return fail(c, KWGPU_ERR_INVALID, "kwgpu: FD1D.T_GRID_SIZE must be >= 2");
    if (p->precision != KWGPU_FP64 && p->precision != KWGPU_FP32_MIXED)
        return fail(c, KWGPU_ERR_INVALID, "kwgpu: unknown precision");
    if (p->layout < KWGPU_LAYOUT_AUTO || p->layout > KWGPU_LAYOUT_WAVE_REG)
        return fail(c, KWGPU_ERR_INVALID, "kwgpu: unknown layout");
    return KWGPU_OK;
}

int check_dims(kwgpu_ctx* c, const int32_t* xd, const int32_t* td, size_t n) {
    for (size_t i = 0; i < n; ++i) {
        if (xd && xd[i] < 3) return fail(c, KWGPU_ERR_INVALID, "kwgpu: x_dims[i] must be >= 3");
        if (td && td[i] < 2) return fail(c, KWGPU_ERR_INVALID, "kwgpu: t_dims[i] must be >= 2");
    }
    return KWGPU_OK;
}

// Wave-layout instantiation a grid size maps to (0 = does not fit).
int wave_bucket(int x) {
    if (x <= 64) return 64;
    if (x <= 128) return 128;
    if (x <= 256) return 256;
    if (x <= 512) return 512;
    if (x <= 1024) return 1024;
    if (x <= 2048) return 2048;
    return 0;
}

// Runs the FD solve of `b` (device buffers already set) on stream s. pde_x / pde_t are
// host copies of each PDE's grid sizes, used to group PDEs into launches; empty means
// every PDE uses b.x_dim / b.t_dim.
int run_fd(kwgpu_ctx* c, FdBatch b, const std::vector<int32_t>& pde_x,
           const std::vector<int32_t>& pde_t, int precision, int layout, hipStream_t s) {
    const bool uniform = pde_x.empty();
    if (precision == KWGPU_FP32_MIXED && layout == KWGPU_LAYOUT_LANE_SOA)
        return fail(c, KWGPU_ERR_UNSUPPORTED, "kwgpu: FP32-mixed runs in the wave layout only");
    auto use_wave = [&](int x) {
        if (layout == KWGPU_LAYOUT_LANE_SOA) return false;
        return kwgpu::wave_layout_fits(x, precision);
    };
    if (layout == KWGPU_LAYOUT_WAVE_REG) {
        int mx = b.x_dim;
        if (!uniform) mx = *std::max_element(pde_x.begin(), pde_x.end());
        if (!kwgpu::wave_layout_fits(mx, precision))
            return fail(c, KWGPU_ERR_UNSUPPORTED, "kwgpu: wave layout supports x_dim <= 2048");
    }

    auto launch_lane = [&](FdBatch bb, int max_x) -> int {
        if (precision != KWGPU_FP64)
            return fail(c, KWGPU_ERR_UNSUPPORTED,
                        "kwgpu: FP32-mixed supports x_dim <= 2048 (wave layout)");
        const size_t need = kwgpu::lane_workspace_doubles(bb.m, max_x) * sizeof(double);
        if (hipError_t e = ensure(c->ws, need); e != hipSuccess)
            return fail(c, KWGPU_ERR_NOMEM, "kwgpu: lane workspace allocation failed");
        bb.ws = (double*)c->ws.p;
        bb.ws_stride = bb.m;
        bb.ws_rows = max_x;
        if (hipError_t e = kwgpu::launch_fd1d_lane(bb, s); e != hipSuccess)
            return hip_fail(c, e, "kwgpu: lane kernel launch");
        return KWGPU_OK;
    };

    if (uniform) {
        if (use_wave(b.x_dim)) {
            if (hipError_t e = kwgpu::launch_fd1d_wave(b, b.x_dim, precision, s); e != hipSuccess)
                return hip_fail(c, e, "kwgpu: wave kernel launch");
            return KWGPU_OK;
        }
        return launch_lane(b, b.x_dim);
    }

    // Uneven grids: one launch per wave instantiation (PDEs sorted by t-points so the
    // segments packed into one wave march for about the same number of steps), and one
    // lane-layout launch for grids the wave layout does not cover.
    const int m = b.m;
    std::vector<uint32_t> ord(m);
    std::iota(ord.begin(), ord.end(), 0u);
    auto key = [&](uint32_t p) {
        const int wb = use_wave(pde_x[p]) ? wave_bucket(pde_x[p]) : 1 << 30;
        return std::make_tuple(wb, -pde_t[p], p);
    };
    std::sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t z) { return key(a) < key(z); });
    if (hipError_t e = ensure(c->order, sizeof(uint32_t) * m); e != hipSuccess)
        return fail(c, KWGPU_ERR_NOMEM, "kwgpu: order buffer allocation failed");
    if (hipError_t e = hipMemcpyAsync(c->order.p, ord.data(), sizeof(uint32_t) * m,
                                      hipMemcpyHostToDevice, s);
        e != hipSuccess)
        return hip_fail(c, e, "kwgpu: order upload");
    size_t i = 0;
    while (i < (size_t)m) {
        const int wb = std::get<0>(key(ord[i]));
        size_t j = i;
        int mx = 0;
        while (j < (size_t)m && std::get<0>(key(ord[j])) == wb) mx = std::max(mx, pde_x[ord[j++]]);
        FdBatch bb = b;
        bb.order = (const uint32_t*)c->order.p + i;
        bb.m = (int)(j - i);
        int rc;
        if (wb != 1 << 30) {
            hipError_t e = kwgpu::launch_fd1d_wave(bb, mx, precision, s);
            rc = e == hipSuccess ? KWGPU_OK : hip_fail(c, e, "kwgpu: wave kernel launch");
        } else {
            rc = launch_lane(bb, mx);
        }
        if (rc) return rc;
        i = j;
    }
    return KWGPU_OK;
}

// Host-pointer FD price with the reference's PDE compression; results are left in
// the device buffers `d_prices` / `d_status` (sized n). Returns KWGPU_OK or an error.
int fd_price_host(kwgpu_ctx* c, const kwgpu_option* opts, size_t n, const kwgpu_fd1d_params* p,
                  DevBuf& d_opts, DevBuf& d_prices, DevBuf& d_status) {
    hipStream_t s = c->stream;
    // PDE compression, src/Pricer/kwFd1d.cpp:28-65: options equal in (t, r, q, z, e, w)
    // -- and, for uneven batches, in grid size -- share one solve.
    auto gx = [&](size_t i) { return p->x_dims ? p->x_dims[i] : p->x_dim; };
    auto gt = [&](size_t i) { return p->t_dims ? p->t_dims[i] : p->t_dim; };
    auto key = [&](size_t i) {
        const auto& a = opts[i];
        return std::make_tuple(a.t, a.r, a.q, a.z, a.e, a.w, gx(i), gt(i));
    };
    std::vector<uint32_t> sorted(n);
    std::iota(sorted.begin(), sorted.end(), 0u);
    std::stable_sort(sorted.begin(), sorted.end(),
                     [&](uint32_t l, uint32_t r) { return key(l) < key(r); });
    std::vector<uint32_t> rep, off, list(sorted);
    for (size_t i = 0; i < n; ++i) {
        if (i == 0 || key(sorted[i - 1]) < key(sorted[i])) {
            rep.push_back(sorted[i]);
            off.push_back((uint32_t)i);
        }
    }
    off.push_back((uint32_t)n);
    const size_t m = rep.size();

    hipError_t e;
    if ((e = ensure(d_opts, n * sizeof(kwgpu_option))) != hipSuccess ||
        (e = ensure(d_prices, n * sizeof(double))) != hipSuccess ||
        (e = ensure(d_status, n * sizeof(int32_t))) != hipSuccess ||
        (e = ensure(c->rep, m * sizeof(uint32_t))) != hipSuccess ||
        (e = ensure(c->off, (m + 1) * sizeof(uint32_t))) != hipSuccess ||
        (e = ensure(c->list, n * sizeof(uint32_t))) != hipSuccess)
        return fail(c, KWGPU_ERR_NOMEM, "kwgpu: device allocation failed");
    if ((e = hipMemcpyAsync(d_opts.p, opts, n * sizeof(kwgpu_option), hipMemcpyHostToDevice, s)) ||
        (e = hipMemcpyAsync(c->rep.p, rep.data(), m * sizeof(uint32_t), hipMemcpyHostToDevice, s)) ||
        (e = hipMemcpyAsync(c->off.p, off.data(), (m + 1) * sizeof(uint32_t), hipMemcpyHostToDevice,
                            s)) ||
        (e = hipMemcpyAsync(c->list.p, list.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, s)))
        return hip_fail(c, e, "kwgpu: upload");

    FdBatch b{};
    b.opts = (const kwgpu_option*)d_opts.p;
    b.rep = (const uint32_t*)c->rep.p;
    b.off = (const uint32_t*)c->off.p;
    b.list = (const uint32_t*)c->list.p;
    b.m = (int)m;
    b.x_dim = p->x_dim;
    b.t_dim = p->t_dim;
    b.density = p->density;
    b.scale = p->scale;
    b.theta = p->theta;
    b.prices = (double*)d_prices.p;
    b.status = (int32_t*)d_status.p;
    std::vector<int32_t> pde_x, pde_t;
    if (p->x_dims || p->t_dims) {
        // per-option sizes travel with the options (indexed by option)
        if ((e = ensure(c->xd, n * sizeof(int32_t))) != hipSuccess ||
            (e = ensure(c->td, n * sizeof(int32_t))) != hipSuccess)
            return fail(c, KWGPU_ERR_NOMEM, "kwgpu: device allocation failed");
        std::vector<int32_t> hx(n), ht(n);
        for (size_t i = 0; i < n; ++i) {
            hx[i] = gx(i);
            ht[i] = gt(i);
        }
        if ((e = hipMemcpyAsync(c->xd.p, hx.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, s)) ||
            (e = hipMemcpyAsync(c->td.p, ht.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, s)))
            return hip_fail(c, e, "kwgpu: upload");
        b.x_dims = (const int32_t*)c->xd.p;
        b.t_dims = (const int32_t*)c->td.p;
        pde_x.resize(m);
        pde_t.resize(m);
        for (size_t k = 0; k < m; ++k) {
            pde_x[k] = hx[rep[k]];
            pde_t[k] = ht[rep[k]];
        }
        // the host vectors must outlive the async copies
        if ((e = hipStreamSynchronize(s)) != hipSuccess) return hip_fail(c, e, "kwgpu: sync");
    }
    int rc = run_fd(c, b, pde_x, pde_t, p->precision, p->layout, s);
    // the host vectors above must outlive every copy queued from them
    e = hipStreamSynchronize(s);
    if (rc) return rc;
    if (e != hipSuccess) return hip_fail(c, e, "kwgpu: FD solve");
    return KWGPU_OK;
}

// The reference's Fd1d::value error text for option i (src/Math/kwFd1d.cpp:151-153),
// prefixed as Fd1d_Pricer::price does (src/Pricer/kwFd1d.cpp:154-155).
std::string range_error(const kwgpu_option& a, double density, double scale, int x_dim) {
    const double xMin = 0.0 - scale * a.z * std::sqrt(a.t);
    const double xMax = 0.0 + scale * a.z * std::sqrt(a.t);
    const double yMin = std::asinh(xMin / density), yMax = std::asinh(xMax / density);
    const double dy = 1. / (double)(x_dim - 1);
    auto xj = [&](int j) {
        const double yj = j * dy;
        return density * std::sinh(yMin * (1.0 - yj) + yMax * yj);
    };
    return "Fd1d_Pricer::price Fd1d::value: x=" + std::to_string(std::log(a.s / a.k)) +
           " not in range (" + std::to_string(xj(0)) + ", " + std::to_string(xj(x_dim - 1)) + ")";
}

int collect(kwgpu_ctx* c, const kwgpu_option* opts, size_t n, const kwgpu_fd1d_params* p,
            const std::vector<int32_t>& st, const char* prefix) {
    for (size_t i = 0; i < n; ++i) {
        if (st[i] == KWGPU_STATUS_OK) continue;
        const int xd = p->x_dims ? p->x_dims[i] : p->x_dim;
        std::string msg;
        int code = KWGPU_ERR_RANGE;
        if (st[i] == KWGPU_STATUS_OUT_OF_GRID) {
            msg = range_error(opts[i], p->density, p->scale, xd);
        } else if (st[i] == KWGPU_STATUS_BAD_INPUT) {
            msg = "Fd1d_Pricer::price: invalid option #" + std::to_string(i);
            code = KWGPU_ERR_INVALID;
        } else {
            msg = "Fd1d_Pricer::price: non-finite price for option #" + std::to_string(i);
            code = KWGPU_ERR_RANGE;
        }
        return fail(c, code, std::string(prefix) + msg);
    }
    return KWGPU_OK;
}

}  // namespace

extern "C" {

void kwgpu_fd1d_params_default(kwgpu_fd1d_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->density = 0.25;   // src/Pricer/kwFd1d.cpp:12
    p->scale = 50.;      // :13
    p->theta = 0.5;      // src/Math/kwFd1d.h:31
    p->t_dim = 512;      // :15
    p->x_dim = 512;      // :16
    p->precision = KWGPU_FP64;
    p->layout = KWGPU_LAYOUT_AUTO;
}

int kwgpu_create(kwgpu_ctx** out, int device) {
    if (!out) return KWGPU_ERR_INVALID;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
        return KWGPU_ERR_HIP;
    if (hipSetDevice(device) != hipSuccess) return KWGPU_ERR_HIP;
    auto* c = new kwgpu_ctx();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return KWGPU_ERR_HIP;
    }
    *out = c;
    return KWGPU_OK;
}

void kwgpu_destroy(kwgpu_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (DevBuf* b : {&c->opts, &c->prices, &c->status, &c->rep, &c->off, &c->list, &c->order,
                      &c->xd, &c->td, &c->ws, &c->opts2, &c->prices2, &c->status2})
        free_buf(*b);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* kwgpu_last_error(const kwgpu_ctx* c) { return c ? c->err.c_str() : "kwgpu: null context"; }

void* kwgpu_stream(kwgpu_ctx* c) { return c ? (void*)c->stream : nullptr; }

int kwgpu_synchronize(kwgpu_ctx* c) {
    if (!c) return KWGPU_ERR_INVALID;
    if (hipError_t e = hipStreamSynchronize(c->stream); e != hipSuccess)
        return hip_fail(c, e, "kwgpu_synchronize");
    return KWGPU_OK;
}

int kwgpu_fd1d_price(kwgpu_ctx* c, const kwgpu_option* opts, size_t n,
                     const kwgpu_fd1d_params* p, double* prices, int32_t* status) {
    if (!c) return KWGPU_ERR_INVALID;
    c->err.clear();
    if (int rc = check_params(c, p)) return rc;
    if (n == 0) return KWGPU_OK;
    if (!opts || !prices) return fail(c, KWGPU_ERR_INVALID, "kwgpu: null pointer");
    if (n > 0x7fffffffu) return fail(c, KWGPU_ERR_INVALID, "kwgpu: too many options");
    if (int rc = check_dims(c, p->x_dims, p->t_dims, n)) return rc;
    if (hipError_t e = hipSetDevice(c->device); e != hipSuccess) return hip_fail(c, e, "hipSetDevice");
    if (int rc = fd_price_host(c, opts, n, p, c->opts, c->prices, c->status)) return rc;
    std::vector<int32_t> st(n);
    hipError_t e;
    if ((e = hipMemcpy(prices, c->prices.p, n * sizeof(double), hipMemcpyDeviceToHost)) ||
        (e = hipMemcpy(st.data(), c->status.p, n * sizeof(int32_t), hipMemcpyDeviceToHost)))
        return hip_fail(c, e, "kwgpu: download");
    if (status) std::memcpy(status, st.data(), n * sizeof(int32_t));
    return collect(c, opts, n, p, st, "");
}

int kwgpu_fd1d_price_device(kwgpu_ctx* c, const kwgpu_option* d_opts, size_t n,
                            const kwgpu_fd1d_params* p, double* d_prices, int32_t* d_status,
                            void* stream) {
    if (!c) return KWGPU_ERR_INVALID;
    c->err.clear();
    if (int rc = check_params(c, p)) return rc;
    if (n == 0) return KWGPU_OK;
    if (!d_opts || !d_prices) return fail(c, KWGPU_ERR_INVALID, "kwgpu: null pointer");
    if (n > 0x7fffffffu) return fail(c, KWGPU_ERR_INVALID, "kwgpu: too many options");
    if (hipError_t e = hipSetDevice(c->device); e != hipSuccess) return hip_fail(c, e, "hipSetDevice");
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    FdBatch b{};
    b.opts = d_opts;
    b.m = (int)n;
    b.x_dim = p->x_dim;
    b.t_dim = p->t_dim;
    b.x_dims = p->x_dims;
    b.t_dims = p->t_dims;
    b.density = p->density;
    b.scale = p->scale;
    b.theta = p->theta;
    b.prices = d_prices;
    b.status = d_status;
    std::vector<int32_t> hx, ht;
    if (p->x_dims || p->t_dims) {
        // grouping needs the sizes on the host
        hx.assign(n, p->x_dim);
        ht.assign(n, p->t_dim);
        hipError_t e;
        if (p->x_dims &&
            (e = hipMemcpyAsync(hx.data(), p->x_dims, n * sizeof(int32_t), hipMemcpyDeviceToHost, s)))
            return hip_fail(c, e, "kwgpu: x_dims download");
        if (p->t_dims &&
            (e = hipMemcpyAsync(ht.data(), p->t_dims, n * sizeof(int32_t), hipMemcpyDeviceToHost, s)))
            return hip_fail(c, e, "kwgpu: t_dims download");
        if ((e = hipStreamSynchronize(s)) != hipSuccess) return hip_fail(c, e, "kwgpu: sync");
        if (int rc = check_dims(c, hx.data(), ht.data(), n)) return rc;
    }
    return run_fd(c, b, hx, ht, p->precision, p->layout, s);
}

int kwgpu_bs_price(kwgpu_ctx* c, const kwgpu_option* opts, size_t n, double* prices) {
    if (!c) return KWGPU_ERR_INVALID;
    c->err.clear();
    if (n == 0) return KWGPU_OK;
    if (!opts || !prices) return fail(c, KWGPU_ERR_INVALID, "kwgpu: null pointer");
    if (hipError_t e = hipSetDevice(c->device); e != hipSuccess) return hip_fail(c, e, "hipSetDevice");
    hipError_t e;
    if ((e = ensure(c->opts, n * sizeof(kwgpu_option))) || (e = ensure(c->prices, n * sizeof(double))))
        return fail(c, KWGPU_ERR_NOMEM, "kwgpu: device allocation failed");
    if ((e = hipMemcpyAsync(c->opts.p, opts, n * sizeof(kwgpu_option), hipMemcpyHostToDevice,
                            c->stream)))
        return hip_fail(c, e, "kwgpu: upload");
    if ((e = kwgpu::launch_bs((const kwgpu_option*)c->opts.p, n, (double*)c->prices.p, c->stream)))
        return hip_fail(c, e, "kwgpu: BS kernel launch");
    if ((e = hipMemcpyAsync(prices, c->prices.p, n * sizeof(double), hipMemcpyDeviceToHost,
                            c->stream)) ||
        (e = hipStreamSynchronize(c->stream)))
        return hip_fail(c, e, "kwgpu: download");
    return KWGPU_OK;
}

int kwgpu_bs_price_device(kwgpu_ctx* c, const kwgpu_option* d_opts, size_t n, double* d_prices,
                          void* stream) {
    if (!c) return KWGPU_ERR_INVALID;
    c->err.clear();
    if (n == 0) return KWGPU_OK;
    if (!d_opts || !d_prices) return fail(c, KWGPU_ERR_INVALID, "kwgpu: null pointer");
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    if (hipError_t e = kwgpu::launch_bs(d_opts, n, d_prices, s); e != hipSuccess)
        return hip_fail(c, e, "kwgpu: BS kernel launch");
    return KWGPU_OK;
}

int kwgpu_fd1d_bs_price(kwgpu_ctx* c, const kwgpu_option* opts, size_t n,
                        const kwgpu_fd1d_params* p, double* prices, int32_t* status) {
    if (!c) return KWGPU_ERR_INVALID;
    c->err.clear();
    if (int rc = check_params(c, p)) return rc;
    if (n == 0) return KWGPU_OK;
    if (!opts || !prices) return fail(c, KWGPU_ERR_INVALID, "kwgpu: null pointer");
    if (n > 0x7fffffffu) return fail(c, KWGPU_ERR_INVALID, "kwgpu: too many options");
    if (int rc = check_dims(c, p->x_dims, p->t_dims, n)) return rc;
    if (hipError_t e = hipSetDevice(c->device); e != hipSuccess) return hip_fail(c, e, "hipSetDevice");
    const char* prefix = "Fd1d_BlackScholes_Pricer::price : ";
    // leg 1: FD on the options as given (src/Pricer/kwFd1d_BlackScholes.cpp:18)
    if (int rc = fd_price_host(c, opts, n, p, c->opts, c->prices, c->status)) return rc;
    std::vector<int32_t> st(n), st2(n);
    hipError_t e;
    if ((e = hipMemcpy(st.data(), c->status.p, n * sizeof(int32_t), hipMemcpyDeviceToHost)))
        return hip_fail(c, e, "kwgpu: download");
    // leg 2: FD on European copies (:21-28)
    std::vector<kwgpu_option> euro(opts, opts + n);
    for (auto& a : euro) a.e = 0;
    if (int rc = fd_price_host(c, euro.data(), n, p, c->opts2, c->prices2, c->status2)) return rc;
    if ((e = hipMemcpy(st2.data(), c->status2.p, n * sizeof(int32_t), hipMemcpyDeviceToHost)))
        return hip_fail(c, e, "kwgpu: download");
    // leg 3 and combine: price += BS(euro) - FD(euro) (:30-40)
    if ((e = kwgpu::launch_fd1d_bs_combine((const kwgpu_option*)c->opts.p, n, (double*)c->prices.p,
                                           (const double*)c->prices2.p, c->stream)))
        return hip_fail(c, e, "kwgpu: combine kernel launch");
    if ((e = hipMemcpyAsync(prices, c->prices.p, n * sizeof(double), hipMemcpyDeviceToHost,
                            c->stream)) ||
        (e = hipStreamSynchronize(c->stream)))
        return hip_fail(c, e, "kwgpu: download");
    for (size_t i = 0; i < n; ++i)
        if (st[i] == KWGPU_STATUS_OK && st2[i] != KWGPU_STATUS_OK) st[i] = st2[i];
    if (status) std::memcpy(status, st.data(), n * sizeof(int32_t));
    return collect(c, opts, n, p, st, prefix);
}

}  // extern "C"
