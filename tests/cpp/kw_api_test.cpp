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

int main(int argc, char** argv) {
    pricerTest("BlackScholes", "BS", true);
    pricerTest("Fd1d", "FD1D", false);
    pricerTest("Fd1dBlackScholes", "FD1D-BS", false);
    errorTest();
    if (argc > 1) portfolioTest(argv[1]);
    std::printf("%d failed\n", g_failed);
    return g_failed;
}
