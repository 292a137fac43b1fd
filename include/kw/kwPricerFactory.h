// kw/kwPricerFactory.h -- string-keyed pricer factory (reference:
// src/Pricer/kwPricerFactory.h:15-41), same keys and error text.
#pragma once

#include "kw/kwConfig.h"
#include "kw/kwPricer.h"

namespace kw {

class PricerFactory {
public:
    static Error create(const Config& config, sPtr<Pricer>& pricer) {
        std::string mode = config.get("PRICER", "");
        if (mode.empty()) return "PricerFactory: Missing PRICER key";

        if (mode == "BS")
            pricer = make_sPtr<BlackScholes_Pricer>();
        else if (mode == "FD1D")
            pricer = make_sPtr<Fd1d_Pricer>();
        else if (mode == "FD1D-BS")
            pricer = make_sPtr<Fd1d_BlackScholes_Pricer>();
        else
            return "PricerFactory: Unknown PRICER = " + mode;

        if (auto err = pricer->init(config); !err.empty()) return "PricerFactory: " + err;
        return "";
    }
};

}  // namespace kw
