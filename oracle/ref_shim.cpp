// ref_shim.cpp -- C-ABI shim over the reference pricer, compiled together with the
// reference's own library sources (in place, from the reference tree) into
// oracle/_ref/libkwref.so. Test infrastructure only: it is the checker used to pin
// the oracle restatement (kw_oracle.cpp) and to write the golden vectors; nothing in
// the product path loads it.
//
// Calls kw::PricerFactory::create (src/Pricer/kwPricerFactory.h:15-41) and
// Pricer::price (src/Pricer/kwPricer.h:12-22) with the Config keys read by
// Fd1d_Pricer::init (src/Pricer/kwFd1d.cpp:9-19).
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "Pricer/kwPricerFactory.h"

extern "C" int kwref_price(const char* pricer, const void* opts, size_t n, double density,
                           double scale, long long t_dim, long long x_dim, double* prices,
                           char* err, size_t errlen) {
    static_assert(sizeof(kw::Option) == 56, "kw::Option layout");
    kw::Config config;
    config.set("PRICER", std::string(pricer));
    config.set("FD1D.DENSITY", density);
    config.set("FD1D.SCALE", scale);
    config.set("FD1D.T_GRID_SIZE", (int64_t)t_dim);
    config.set("FD1D.X_GRID_SIZE", (int64_t)x_dim);

    const kw::Option* a = static_cast<const kw::Option*>(opts);
    std::vector<kw::Option> assets(a, a + n);
    kw::sPtr<kw::Pricer> p;
    std::string e = kw::PricerFactory::create(config, p);
    std::vector<double> out;
    if (e.empty()) e = p->price(assets, out);
    for (size_t i = 0; i < n; ++i) prices[i] = i < out.size() ? out[i] : NAN;
    if (!e.empty()) {
        if (err && errlen) std::snprintf(err, errlen, "%s", e.c_str());
        return 1;
    }
    return 0;
}

// Worker threads of the reference's ThreadPool (src/kwThreadPool.cpp:9-16): one per
// hardware thread.
extern "C" int kwref_threads() { return (int)std::thread::hardware_concurrency(); }
