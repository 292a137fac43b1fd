// kwinto -- the command-line front end of the reference (src/kwinto.cpp:15-85) over
// the kw:: API of this build: `kwinto price [options] <path>` loads a portfolio CSV,
// prices it on the GPU through PricerFactory and prints the price statistics against
// the CSV's reference prices. Same options, defaults and error texts; the argument
// parser is written here (no docopt).
#include <charconv>
#include <cstdint>
#include <iostream>
#include <map>
#include <string>

#include "kw/kwPortfolio.h"
#include "kw/kwPricerFactory.h"

namespace {

constexpr char kUsage[] = R"(
kwinto - Options Pricing Analytics

Usage:
    kwinto price [options] <path>

Arguments:
    <path>          CSV file with options data

Options:
    --density <num> Density of the x-grid distribution [default: 0.25]
    --scale <num>   Scale of the x-grid distribution [default: 50.0]
    -e <num>        Reject option prices less than <num> from RRMS error stats [default: 0.5]
    -p <name>       Pricer name [default: FD1D-BS]
    -t <num>        Use <num> points for t-grid [default: 512]
    -x <num>        Use <num> points for x-grid [default: 512]
    -v              Show extra details, be verbose

    -h --help       Print this screen
    --version       Print Kwinto version
)";

constexpr char kVersion[] = "kwinto-cuda_amd (gfx950, HIP)";

bool parse_f64(const std::string& s, double& v) {
    const char* b = s.data();
    const char* e = b + s.size();
    auto r = std::from_chars(b, e, v);
    return !s.empty() && r.ec == std::errc{} && r.ptr == e;
}

bool parse_i64(const std::string& s, int64_t& v) {
    const char* b = s.data();
    const char* e = b + s.size();
    auto r = std::from_chars(b, e, v);
    return !s.empty() && r.ec == std::errc{} && r.ptr == e;
}

// src/kwinto.cpp:40-83
kw::Error cmdPrice(const std::map<std::string, std::string>& args) {
    kw::Config config;
    config.set("PRICER", args.at("-p"));

    double density;
    if (!parse_f64(args.at("--density"), density))
        return "cmdBench: Fail to parse '--density <num>': kw::fromString : Fail to parse '" +
               args.at("--density") + "'";
    config.set("FD1D.DENSITY", density);

    double scale;
    if (!parse_f64(args.at("--scale"), scale))
        return "cmdBench: Fail to parse '--scale <num>': kw::fromString : Fail to parse '" +
               args.at("--scale") + "'";
    config.set("FD1D.SCALE", scale);

    int64_t t_dim, x_dim;
    if (!parse_i64(args.at("-t"), t_dim)) return "cmdPrice: Fail to parse '-t <num>'";
    if (!parse_i64(args.at("-x"), x_dim)) return "cmdPrice: Fail to parse '-x <num>'";
    config.set("FD1D.T_GRID_SIZE", t_dim);
    config.set("FD1D.X_GRID_SIZE", x_dim);

    double tolerance;
    if (!parse_f64(args.at("-e"), tolerance))
        return "cmdBench: Fail to parse '-e <num>': kw::fromString : Fail to parse '" +
               args.at("-e") + "'";

    kw::Portfolio portfolio;
    if (auto err = portfolio.load(args.at("<path>")); !err.empty()) return "cmdPrice: " + err;

    std::cout << "Portfolio" << std::endl;
    std::cout << "    Assets : " << portfolio.assets().size() << std::endl;
    std::cout << std::endl;

    std::vector<kw::f64> prices;
    if (auto err = portfolio.price(config, prices); !err.empty()) return "cmdPrice: " + err;

    // the reference parses -e but always rejects below 0.5; here -e is applied
    if (auto err = portfolio.printPricesStats(prices, tolerance); !err.empty())
        return "cmdPrice: " + err;
    return "";
}

}  // namespace

int main(int argc, char** argv) {
    std::map<std::string, std::string> args = {
        {"--density", "0.25"}, {"--scale", "50.0"}, {"-e", "0.5"}, {"-p", "FD1D-BS"},
        {"-t", "512"},         {"-x", "512"},       {"-v", "false"}};
    bool price = false;
    std::string path;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "-h" || a == "--help") {
            std::cout << kUsage << std::endl;
            return 0;
        }
        if (a == "--version") {
            std::cout << kVersion << std::endl;
            return 0;
        }
        if (a == "-v") {
            args["-v"] = "true";
        } else if (args.count(a)) {
            if (i + 1 >= argc) {
                std::cerr << kUsage << std::endl;
                return 1;
            }
            args[a] = argv[++i];
        } else if (a.rfind("--density=", 0) == 0 || a.rfind("--scale=", 0) == 0) {
            const auto eq = a.find('=');
            args[a.substr(0, eq)] = a.substr(eq + 1);
        } else if (a == "price" && !price) {
            price = true;
        } else if (!a.empty() && a[0] != '-' && path.empty() && price) {
            path = a;
        } else {
            std::cerr << kUsage << std::endl;
            return 1;
        }
    }
    if (!price || path.empty()) {
        std::cerr << kUsage << std::endl;
        return 1;
    }
    args["<path>"] = path;

    std::cout << kVersion << '\n' << std::endl;
    if (args["-v"] == "true") {
        std::cout << "Command-Line Arguments" << std::endl;
        for (const auto& [k, v] : args) std::cout << "    " << k << ": " << v << std::endl;
        std::cout << std::endl;
    }
    if (auto error = cmdPrice(args); !error.empty()) {
        std::cerr << error << std::endl;
        return 1;
    }
    return 0;
}
