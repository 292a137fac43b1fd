// kw/kwPricer.h -- the pricer interface (reference: src/Pricer/kwPricer.h:12-22) and its
// MI355X implementations, each a thin wrapper over the kwgpu C-ABI (include/kwgpu.h).
//
//   BlackScholes_Pricer        PRICER=BS       src/Pricer/kwBlackScholes.cpp:13-50
//   Fd1d_Pricer                PRICER=FD1D     src/Pricer/kwFd1d.cpp:9-160
//   Fd1d_BlackScholes_Pricer   PRICER=FD1D-BS  src/Pricer/kwFd1d_BlackScholes.cpp:15-43
//
// Config keys read by init(): FD1D.DENSITY (0.25), FD1D.SCALE (50.), FD1D.T_GRID_SIZE
// (512, integer), FD1D.X_GRID_SIZE (512, integer) as in the reference, plus
// FD1D.PRECISION ("FP64" | "FP32-MIXED"), FD1D.LAYOUT ("AUTO" | "LANE" | "WAVE") and
// GPU.DEVICE (0). price() is synchronous and, as in the reference, not re-entrant per
// instance: each instance owns one kwgpu context (a HIP stream and device buffers).
#pragma once

#include <vector>

#include "kw/kwAsset.h"
#include "kw/kwConfig.h"
#include "kwgpu.h"

namespace kw {

class Pricer {
public:
    virtual ~Pricer() = default;
    virtual Error init(const Config& config) = 0;
    virtual Error price(const std::vector<Option>& assets, std::vector<double>& prices) = 0;
};

namespace detail {

/// Owns one kwgpu context.
class GpuContext {
    kwgpu_ctx* m_ctx = nullptr;

public:
    GpuContext() = default;
    GpuContext(const GpuContext&) = delete;
    GpuContext& operator=(const GpuContext&) = delete;
    ~GpuContext() { kwgpu_destroy(m_ctx); }

    Error open(int device) {
        if (m_ctx) return "";
        if (int rc = kwgpu_create(&m_ctx, device); rc != KWGPU_OK)
            return "kwgpu_create(device=" + std::to_string(device) + ") failed, code " +
                   std::to_string(rc);
        return "";
    }
    kwgpu_ctx* get() const { return m_ctx; }
    Error error() const { return kwgpu_last_error(m_ctx); }
};

inline Error read_fd1d_config(const Config& config, kwgpu_fd1d_params& p, int& device) {
    kwgpu_fd1d_params_default(&p);
    p.density = config.get("FD1D.DENSITY", 0.25);
    p.scale = config.get("FD1D.SCALE", 50.);
    p.t_dim = (int32_t)config.get("FD1D.T_GRID_SIZE", 512);
    p.x_dim = (int32_t)config.get("FD1D.X_GRID_SIZE", 512);
    const std::string prec = config.get("FD1D.PRECISION", "FP64");
    if (prec == "FP64")
        p.precision = KWGPU_FP64;
    else if (prec == "FP32-MIXED")
        p.precision = KWGPU_FP32_MIXED;
    else
        return "Unknown FD1D.PRECISION = " + prec;
    const std::string layout = config.get("FD1D.LAYOUT", "AUTO");
    if (layout == "AUTO")
        p.layout = KWGPU_LAYOUT_AUTO;
    else if (layout == "LANE")
        p.layout = KWGPU_LAYOUT_LANE_SOA;
    else if (layout == "WAVE")
        p.layout = KWGPU_LAYOUT_WAVE_REG;
    else
        return "Unknown FD1D.LAYOUT = " + layout;
    device = (int)config.get("GPU.DEVICE", 0);
    return "";
}

}  // namespace detail

/// Closed-form European price; NaN for American options.
class BlackScholes_Pricer : public Pricer {
    detail::GpuContext m_gpu;

public:
    Error init(const Config& config) final {
        return m_gpu.open((int)config.get("GPU.DEVICE", 0));
    }
    Error price(const std::vector<Option>& assets, std::vector<f64>& prices) final {
        prices.resize(assets.size());
        if (assets.empty()) return "";
        if (!m_gpu.get()) return "BlackScholes_Pricer::price : not initialised";
        if (kwgpu_bs_price(m_gpu.get(), reinterpret_cast<const kwgpu_option*>(assets.data()),
                           assets.size(), prices.data()) != KWGPU_OK)
            return "BlackScholes_Pricer::price : " + m_gpu.error();
        return "";
    }
};

/// Crank-Nicolson FD pricer on a sinh-stretched log-moneyness grid.
class Fd1d_Pricer : public Pricer {
    detail::GpuContext m_gpu;
    kwgpu_fd1d_params m_params{};
    std::vector<int32_t> m_status;

public:
    Error init(const Config& config) final {
        int device = 0;
        if (auto err = detail::read_fd1d_config(config, m_params, device); !err.empty())
            return "Fd1d_Pricer::init: " + err;
        return m_gpu.open(device);
    }
    Error price(const std::vector<Option>& assets, std::vector<f64>& prices) final {
        const auto n = assets.size();
        if (n == 0) return "";
        if (!m_gpu.get()) return "Fd1d_Pricer::price: not initialised";
        prices.resize(n);
        m_status.resize(n);
        if (kwgpu_fd1d_price(m_gpu.get(), reinterpret_cast<const kwgpu_option*>(assets.data()), n,
                             &m_params, prices.data(), m_status.data()) != KWGPU_OK)
            return m_gpu.error();
        return "";
    }
    /// Price and bump Greeks (test/portfolio.py:35-56) of every option, all nine bumped
    /// prices per option solved in one batch; spot bumps share the option's PDE.
    Error greeks(const std::vector<Option>& assets, std::vector<kwgpu_greeks>& out,
                 f64 bump = 0.01) {
        const auto n = assets.size();
        out.resize(n);
        if (n == 0) return "";
        if (!m_gpu.get()) return "Fd1d_Pricer::greeks: not initialised";
        m_status.resize(n);
        if (kwgpu_fd1d_greeks(m_gpu.get(), reinterpret_cast<const kwgpu_option*>(assets.data()), n,
                              &m_params, bump, out.data(), m_status.data()) != KWGPU_OK)
            return m_gpu.error();
        return "";
    }
    /// Per-option status of the last price() or greeks() call (KWGPU_STATUS_*).
    const std::vector<int32_t>& status() const { return m_status; }
};

/// Control variate: FD(as given) + (BS - FD(European)) on one grid.
class Fd1d_BlackScholes_Pricer : public Pricer {
    detail::GpuContext m_gpu;
    kwgpu_fd1d_params m_params{};
    std::vector<int32_t> m_status;

public:
    Error init(const Config& config) final {
        int device = 0;
        if (auto err = detail::read_fd1d_config(config, m_params, device); !err.empty())
            return "Fd1d_BlackScholes_Pricer::init : " + err;
        return m_gpu.open(device);
    }
    Error price(const std::vector<Option>& assets, std::vector<f64>& prices) final {
        const auto n = assets.size();
        if (n == 0) return "";
        if (!m_gpu.get()) return "Fd1d_BlackScholes_Pricer::price : not initialised";
        prices.resize(n);
        m_status.resize(n);
        if (kwgpu_fd1d_bs_price(m_gpu.get(), reinterpret_cast<const kwgpu_option*>(assets.data()),
                                n, &m_params, prices.data(), m_status.data()) != KWGPU_OK)
            return m_gpu.error();
        return "";
    }
};

}  // namespace kw
