// kw_api_test.cpp -- the reference's own test cases, written against the kw:: API of
// this build (include/kw/) and run on the GPU. Mirrors test/kwPricer_test.cpp:10-108
// and test/kwPortfolio_test.cpp:29-58 without GoogleTest: each check prints a line and
// the exit status is the number of failed checks.
//
//   kw_api_test <portfolio_qdfp.csv>
#include <cmath>
#include <cstdio>
#include <string>
#include <utility>
#include <vector>

#include "kw/kwPortfolio.h"
#include "kw/kwPricerFactory.h"

using namespace kw;

static int g_failed = 0;
#define EXPECT(cond, ...)                                   \
    do {                                                    \
        if (!(cond)) {                                      \
            ++g_failed;                                     \
            std::printf("FAIL %s:%d: ", __FILE__, __LINE__); \
            std::printf(__VA_ARGS__);                       \
            std::printf("\n");                              \
        }                                                   \
    } while (0)

// test/kwPricer_test.cpp:16-28
static std::vector<std::pair<Option, f64>> testData() {
    return {
        {{1.0, 100., 0.2, 0.06, 0.02, 90., 0, -1}, 10.627},
        {{1.0, 100., 0.2, 0.06, 0.02, 100., 0, -1}, 5.885},
        {{1.0, 100., 0.2, 0.06, 0.02, 110., 0, -1}, 2.987},
        {{1.0, 100., 0.2, 0.06, 0.02, 90., 0, +1}, 4.668},
        {{1.0, 100., 0.2, 0.06, 0.02, 100., 0, +1}, 9.729},
        {{1.0, 100., 0.2, 0.06, 0.02, 110., 0, +1}, 16.633},
        {{1.0, 100., 0.2, 0.06, 0.08, 90., 1, -1}, 13.988121682},
        {{1.0, 100., 0.2, 0.06, 0.08, 100., 1, -1}, 8.409190396},
        {{1.0, 100., 0.2, 0.06, 0.08, 110., 1, -1}, 4.6592955111},
        {{1.0, 100., 0.2, 0.06, 0.08, 90., 1, +1}, 2.9472036256},
        {{1.0, 100., 0.2, 0.06, 0.08, 100., 1, +1}, 6.8422540642},
        {{1.0, 100., 0.2, 0.06, 0.08, 110., 1, +1}, 12.7940107108},
    };
}

static void pricerTest(const char* name, const char* key, bool skipNan) {
    std::vector<Option> assets;
    std::vector<f64> want;
    for (const auto& t : testData()) {
        assets.push_back(t.first);
        want.push_back(t.second);
    }
    Config config;
    config.set("PRICER", key);
    sPtr<Pricer> pricer;
    auto err = PricerFactory::create(config, pricer);
    EXPECT(err.empty(), "%s create: %s", name, err.c_str());
    if (!err.empty()) return;
    std::vector<f64> prices;
    err = pricer->price(assets, prices);
    EXPECT(err.empty(), "%s price: %s", name, err.c_str());
    EXPECT(prices.size() == assets.size(), "%s size", name);
    for (size_t i = 0; i < want.size() && i < prices.size(); ++i) {
        if (skipNan && std::isnan(prices[i])) continue;
        EXPECT(std::abs(want[i] - prices[i]) <= 1.3e-3, "%s #%zu want %.6f got %.6f", name, i,
               want[i], prices[i]);
    }
    std::printf("kwPricerTest.%s done\n", name);
}

static void portfolioTest(const char* csv) {
    Portfolio portfolio;
    auto err = portfolio.load(csv);
    EXPECT(err.empty(), "load: %s", err.c_str());
    EXPECT(portfolio.assets().size() == 6000, "size %zu", portfolio.assets().size());
    std::vector<Option> assets;
    for (const auto& a : portfolio.assets())
        if (a.e) assets.push_back(a);
    Config config;
    config.set("PRICER", "FD1D");
    config.set("FD1D.T_GRID_SIZE", 512);
    config.set("FD1D.X_GRID_SIZE", 512);
    sPtr<Pricer> engine;
    err = PricerFactory::create(config, engine);
    EXPECT(err.empty(), "create: %s", err.c_str());
    std::vector<double> prices;
    err = engine->price(assets, prices);
    EXPECT(err.empty(), "price: %s", err.c_str());
    f64 mae = 0;
    for (size_t i = 0; i < assets.size(); i++) {
        const f64 d = std::abs(portfolio.prices()[i] - prices[i]);
        mae = std::max(mae, d);
        EXPECT(d <= 0.005, "asset %s want %.6f got %.6f", assets[i].asString().c_str(),
               portfolio.prices()[i], prices[i]);
    }
    std::printf("kwPortfolioTest.Fd1d done, max abs error vs QuantLib %.3e\n", mae);
    // Portfolio::price through the factory, FD1D-BS (the CLI default pricer)
    config.set("PRICER", "FD1D-BS");
    std::vector<double> p2;
    err = portfolio.price(config, p2);
    EXPECT(err.empty(), "Portfolio::price: %s", err.c_str());
    portfolio.printPricesStats(p2);
}

static void errorTest() {
    Config config;
    sPtr<Pricer> p;
    EXPECT(PricerFactory::create(config, p) == "PricerFactory: Missing PRICER key", "missing key");
    config.set("PRICER", "XYZ");
    EXPECT(PricerFactory::create(config, p) == "PricerFactory: Unknown PRICER = XYZ", "unknown");
    // log(S/K) outside the x-grid fails the whole call with the reference's message
    config.set("PRICER", "FD1D");
    config.set("FD1D.SCALE", 0.05);
    config.set("FD1D.X_GRID_SIZE", 64);
    config.set("FD1D.T_GRID_SIZE", 32);
    EXPECT(PricerFactory::create(config, p).empty(), "create");
    std::vector<Option> a = {{1.0, 100., 0.2, 0.05, 0.0, 150., 1, -1}};
    std::vector<f64> prices;
    const auto err = p->price(a, prices);
    EXPECT(err.rfind("Fd1d_Pricer::price Fd1d::value: x=0.405465 not in range (", 0) == 0,
           "range error text: %s", err.c_str());
    // empty portfolio: no error, output untouched (src/Pricer/kwFd1d.cpp:25-26)
    std::vector<f64> keep = {1.0};
    EXPECT(p->price({}, keep).empty() && keep.size() == 1, "empty");
    std::printf("errors done\n");
}

// Fd1d_Pricer::greeks against the bump formulas (test/portfolio.py:35-56) applied to
// separate price() calls on the bumped options.
static void greeksTest() {
    std::vector<Option> assets;
    for (const auto& t : testData()) assets.push_back(t.first);
    Config config;
    auto pricer = make_sPtr<Fd1d_Pricer>();
    auto err = pricer->init(config);
    EXPECT(err.empty(), "greeks init: %s", err.c_str());
    if (!err.empty()) return;
    std::vector<kwgpu_greeks> g;
    const f64 b = 0.01;
    err = pricer->greeks(assets, g, b);
    EXPECT(err.empty() && g.size() == assets.size(), "greeks: %s", err.c_str());
    if (!err.empty()) return;
    std::vector<f64> v[9];
    for (int k = 0; k < 9; ++k) {
        auto a = assets;
        for (auto& o : a) {
            const f64 f = k == 0 ? 1.0 : (k - 1) % 2 == 0 ? 1.0 + b : 1.0 - b;
            f64* x[4] = {&o.s, &o.z, &o.t, &o.r};
            if (k > 0) *x[(k - 1) / 2] = f * *x[(k - 1) / 2];
        }
        err = pricer->price(a, v[k]);
        EXPECT(err.empty(), "bumped price %d: %s", k, err.c_str());
        if (!err.empty()) return;
    }
    for (size_t i = 0; i < assets.size(); ++i) {
        const auto& o = assets[i];
        auto d = [&](int k, f64 x) { return (v[k][i] - v[k + 1][i]) / (2 * b * x); };
        const f64 want[6] = {v[0][i], d(1, o.s), (v[1][i] - 2 * v[0][i] + v[2][i]) / ((b * o.s) * (b * o.s)),
                             d(3, o.z) / 100, d(5, o.t) / 360, d(7, o.r) / 100};
        const f64 got[6] = {g[i].price, g[i].delta, g[i].gamma, g[i].vega, g[i].theta, g[i].rho};
        for (int j = 0; j < 6; ++j)
            EXPECT(std::abs(want[j] - got[j]) <= 1e-9 * std::max(1.0, std::abs(want[j])),
                   "greek %d of #%zu want %.12g got %.12g", j, i, want[j], got[j]);
        // European puts and calls: Black-Scholes delta within the grid's accuracy
        if (!o.e) {
            const f64 sq = o.z * std::sqrt(o.t);
            const f64 d1 = (std::log(o.s / o.k) + (o.r - o.q + 0.5 * o.z * o.z) * o.t) / sq;
            const f64 delta = o.w * std::exp(-o.q * o.t) * 0.5 * std::erfc(-o.w * d1 / std::sqrt(2.));
            EXPECT(std::abs(delta - g[i].delta) <= 2e-3, "BS delta #%zu want %.6f got %.6f", i,
                   delta, g[i].delta);
        }
    }
    std::printf("greeks done\n");
}

int main(int argc, char** argv) {
    pricerTest("BlackScholes", "BS", true);
    pricerTest("Fd1d", "FD1D", false);
    pricerTest("Fd1dBlackScholes", "FD1D-BS", false);
    errorTest();
    greeksTest();
    if (argc > 1) portfolioTest(argv[1]);
    std::printf("%d failed\n", g_failed);
    return g_failed;
}
