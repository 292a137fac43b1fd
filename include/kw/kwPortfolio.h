// kw/kwPortfolio.h -- a portfolio of options priced in one call (reference:
// src/Utils/kwPortfolio.{h,cpp}). load() reads the reference's header-mapped CSV
// (expiry, spot, strike, volatility, interest_rate, dividend_rate, parity c|p,
// exercise a|e, price); price() goes through PricerFactory; printPricesStats() prints
// the reference's statistics over reference prices >= 0.5.
#pragma once

#include <charconv>
#include <cmath>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "kw/kwAsset.h"
#include "kw/kwConfig.h"
#include "kw/kwPricerFactory.h"

namespace kw {

class Portfolio {
private:
    std::vector<Option> m_assets;
    std::vector<f64> m_prices;

    static std::vector<std::string> split(const std::string& src, char delim) {
        std::vector<std::string> out;
        std::stringstream buf(src);
        for (std::string item; std::getline(buf, item, delim);) out.push_back(item);
        return out;
    }
    static Error parse(const std::string& src, f64& value) {
        const char* b = src.data();
        const char* e = b + src.size();
        if (std::from_chars(b, e, value).ec != std::errc{})
            return "kw::fromString : Fail to parse '" + src + "'";
        return "";
    }

public:
    const auto& assets() const { return m_assets; }
    const auto& prices() const { return m_prices; }

    Error load(const std::string& srcPath_) {
        const auto srcPath = std::filesystem::absolute(srcPath_);
        std::ifstream src(srcPath);
        if (!src.is_open()) return "Portfolio::load : Failed to open " + srcPath.string();
        int e = -1, k = -1, q = -1, r = -1, s = -1, t = -1, v = -1, w = -1, z = -1;
        {
            std::string header;
            std::getline(src, header);
            if (!header.empty() && header.back() == '\r') header.pop_back();
            int i = 0;
            for (const auto& c : split(header, ',')) {
                if (c == "exercise") e = i;
                else if (c == "strike") k = i;
                else if (c == "dividend_rate") q = i;
                else if (c == "interest_rate") r = i;
                else if (c == "spot") s = i;
                else if (c == "expiry") t = i;
                else if (c == "price") v = i;
                else if (c == "parity") w = i;
                else if (c == "volatility") z = i;
                ++i;
            }
            if (e == -1 || k == -1 || q == -1 || r == -1 || s == -1 || t == -1 || v == -1 ||
                w == -1 || z == -1) {
                std::stringstream error;
                error << "Portfolio::load : Some option data is missing: e=" << e << ", k=" << k
                      << ", q=" << q << ", r=" << r << ", s=" << s << ", t=" << t << ", v=" << v
                      << ", w=" << w << ", z=" << z;
                return error.str();
            }
        }
        const int need = std::max({e, k, q, r, s, t, v, w, z});
        for (std::string line; std::getline(src, line);) {
            if (!line.empty() && line.back() == '\r') line.pop_back();
            if (line.empty()) continue;
            auto vals = split(line, ',');
            if ((int)vals.size() <= need) return "Portfolio::load : short row '" + line + "'";
            Option a{};
            f64 price = 0;
            for (auto [col, dst] : {std::pair{k, &a.k}, {q, &a.q}, {r, &a.r}, {t, &a.t},
                                    {s, &a.s}, {z, &a.z}, {v, &price}})
                if (auto err = parse(vals[col], *dst); !err.empty()) return "Portfolio::load : " + err;
            a.e = vals[e] == "a" ? 1 : 0;
            a.w = vals[w] == "c" ? +1 : -1;
            m_assets.push_back(a);
            m_prices.push_back(price);
        }
        return "";
    }

    Error price(const Config& config, std::vector<f64>& prices) {
        sPtr<Pricer> pricer;
        if (auto err = PricerFactory::create(config, pricer); !err.empty())
            return "Portfolio::price : " + err;
        if (auto err = pricer->price(m_assets, prices); !err.empty())
            return "Portfolio::price : " + err;
        return "";
    }

    /// RMSE here is, as in the reference, the standard deviation of |diff|.
    Error printPricesStats(const std::vector<f64>& prices, f64 tolerance = 0.5) const {
        f64 a1 = 0, a2 = 0, r1 = 0, r2 = 0, mae = 0, mre = 0;
        u64 na = 0;
        Option maeAsset{}, mreAsset{};
        for (size_t j = 0; j < m_prices.size() && j < prices.size(); j++) {
            const f64 want = m_prices[j], got = prices[j];
            if (want < tolerance) continue;
            const f64 ad = std::abs(want - got), rd = ad / want;
            if (ad > mae) { mae = ad; maeAsset = m_assets[j]; }
            if (rd > mre) { mre = rd; mreAsset = m_assets[j]; }
            a1 += ad; a2 += ad * ad; r1 += rd; r2 += rd * rd;
            na++;
        }
        const f64 am = a1 / na, rm = r1 / na;
        std::cout << "Price Statistics\n" << std::scientific;
        std::cout << "       RMSE : " << std::sqrt(a2 / na - am * am) << "\n";
        std::cout << "      RRMSE : " << std::sqrt(r2 / na - rm * rm) << "\n";
        std::cout << "        MAE : " << mae << "\n";
        std::cout << "        MRE : " << mre << "\n";
        std::cout << "  MAE Asset : " << maeAsset.asString() << "\n";
        std::cout << "  MRE Asset : " << mreAsset.asString() << "\n" << std::fixed;
        std::cout << "      total : " << na << " options\n" << std::endl;
        return "";
    }
};

}  // namespace kw
