// Drives every host-side path of the C-ABI in the order the GPU tests use them, on one
// context, against the stand-in runtime of hip_mock.cpp (no kernel runs): uneven-batch
// planning, replays and re-plans, PDE grouping, FD1D-BS and Greeks batches, the lane
// workspace, the autotuner and its file, error returns, and the surrogate's training,
// prediction and gated serving. Exits non-zero on an unexpected return code or on
// allocations, streams, events or graphs left alive at the end. Test infrastructure only.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

#include "kwgpu.h"

extern "C" int mock_live();
// the stand-in runtime's streams (hipStream_t is a pointer; hipError_t an int enum)
extern "C" int hipStreamCreateWithFlags(void** s, unsigned flags);
extern "C" int hipStreamDestroy(void* s);

namespace {

int failures = 0;

void expect(int rc, int want, kwgpu_ctx* c, const char* what) {
    if (rc != want) {
        std::fprintf(stderr, "FAIL %s: rc %d (want %d) %s\n", what, rc, want,
                     c ? kwgpu_last_error(c) : "");
        ++failures;
    }
}

std::vector<kwgpu_option> options(size_t n, unsigned seed) {
    std::mt19937_64 g(seed);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    std::vector<kwgpu_option> o(n);
    for (auto& x : o) {
        x.t = 1.0 / 12 + u(g) * (3.0 - 1.0 / 12);
        x.k = 50 + 100 * u(g);
        x.z = 0.1 + 0.5 * u(g);
        x.r = 0.1 * u(g);
        x.q = 0.08 * u(g);
        x.s = 100;
        x.e = 1;
        x.w = u(g) < 0.5 ? -1 : 1;
    }
    return o;
}

struct Sizes {
    std::vector<int32_t> x, t;
};

Sizes sizes(size_t n, std::initializer_list<int> xs, int t0, int tspan, unsigned salt) {
    Sizes s;
    std::vector<int> xv(xs);
    for (size_t i = 0; i < n; ++i) {
        s.x.push_back(xv[(i * 7 + salt) % xv.size()]);
        s.t.push_back(t0 + (int)((i * 13 + salt) % (size_t)tspan));
    }
    return s;
}

int host_price(kwgpu_ctx* c, const std::vector<kwgpu_option>& o, kwgpu_fd1d_params p) {
    std::vector<double> pr(o.size());
    std::vector<int32_t> st(o.size());
    return kwgpu_fd1d_price(c, o.data(), o.size(), &p, pr.data(), st.data());
}

int device_price(kwgpu_ctx* c, const std::vector<kwgpu_option>& o, const kwgpu_fd1d_params& p,
                 void* stream = nullptr) {
    std::vector<double> pr(o.size());
    std::vector<int32_t> st(o.size());
    return kwgpu_fd1d_price_device(c, o.data(), o.size(), &p, pr.data(), st.data(), stream);
}

}  // namespace

int main(int argc, char** argv) {
    const std::string dir = argc > 1 ? argv[1] : ".";
    kwgpu_ctx* c = nullptr;
    expect(kwgpu_create(&c, 0), KWGPU_OK, c, "create");
    kwgpu_fd1d_params p;

    {  // FD1D-BS with per-option grids: one solve of 2n PDEs
        auto o = options(90, 23);
        Sizes s = sizes(90, {256, 300, 512}, 40, 25, 0);
        kwgpu_fd1d_params_default(&p);
        p.x_dims = s.x.data();
        p.t_dims = s.t.data();
        std::vector<double> pr(90);
        std::vector<int32_t> st(90);
        expect(kwgpu_fd1d_bs_price(c, o.data(), 90, &p, pr.data(), st.data()), KWGPU_OK, c,
               "fd1d_bs uneven");
    }
    {  // Greeks: 9 bumped copies per option, uniform and per-option grids
        auto o = options(96, 17);
        std::vector<kwgpu_greeks> g(96);
        std::vector<int32_t> st(96);
        kwgpu_fd1d_params_default(&p);
        expect(kwgpu_fd1d_greeks(c, o.data(), 96, &p, 0.01, g.data(), st.data()), KWGPU_OK, c,
               "greeks");
        Sizes s = sizes(96, {256, 512}, 128, 73, 1);
        p.x_dims = s.x.data();
        p.t_dims = s.t.data();
        expect(kwgpu_fd1d_greeks(c, o.data(), 96, &p, 0.01, g.data(), st.data()), KWGPU_OK, c,
               "greeks uneven");
    }
    {  // uneven host batches, grids above 2048 points in the lane layout (FP64 only)
        auto o = options(40, 2);
        Sizes s = sizes(40, {256, 512, 1024, 2048, 3000}, 32, 300, 2);
        kwgpu_fd1d_params_default(&p);
        p.x_dims = s.x.data();
        p.t_dims = s.t.data();
        expect(host_price(c, o, p), KWGPU_OK, c, "uneven with lane bucket");
        p.precision = KWGPU_FP32_MIXED;
        expect(host_price(c, o, p), KWGPU_ERR_UNSUPPORTED, c, "mixed above 2048 points");
        Sizes s2 = sizes(40, {256, 512, 1024, 2048}, 32, 300, 3);
        p.x_dims = s2.x.data();
        p.t_dims = s2.t.data();
        expect(host_price(c, o, p), KWGPU_OK, c, "mixed uneven");
        // FP32-mixed with PDEs that go to FP64 (scale sigma sqrt T > 80): two solves
        for (size_t i = 0; i < o.size(); i += 3) o[i].z = 1.5, o[i].t = 10.0;
        expect(host_price(c, o, p), KWGPU_OK, c, "mixed uneven with FP64 PDEs");
        kwgpu_fd1d_params_default(&p);
        p.precision = KWGPU_FP32_MIXED;
        expect(host_price(c, o, p), KWGPU_OK, c, "mixed uniform with FP64 PDEs");
        for (auto& x : o) x.z = 2.0, x.t = 20.0;
        expect(host_price(c, o, p), KWGPU_OK, c, "mixed, every PDE in FP64");
    }
    for (int x : {3, 17, 64, 100, 129, 256, 300, 511, 1000, 1024, 2048, 4096})  // uniform grids
        for (int lay : {KWGPU_LAYOUT_AUTO, KWGPU_LAYOUT_WAVE_REG, KWGPU_LAYOUT_LANE_SOA}) {
            kwgpu_fd1d_params_default(&p);
            p.x_dim = x;
            p.t_dim = 20;
            p.layout = lay;
            const int want = lay == KWGPU_LAYOUT_WAVE_REG && x > 2048 ? KWGPU_ERR_UNSUPPORTED : KWGPU_OK;
            expect(host_price(c, options(37, 1000 + x), p), want, c, "uniform sizes");
        }
    {  // device entry point: a replayed batch, then more batches than any cache would hold
        auto o = options(3000, 77);
        Sizes s = sizes(3000, {256, 512, 1024, 2048}, 32, 33, 4);
        kwgpu_fd1d_params_default(&p);
        p.x_dims = s.x.data();
        p.t_dims = s.t.data();
        for (int k = 0; k < 3; ++k) expect(device_price(c, o, p), KWGPU_OK, c, "device uneven");
        std::vector<std::vector<kwgpu_option>> os;
        std::vector<Sizes> ss;
        for (int b = 0; b < 6; ++b) {
            os.push_back(options(700 + 50 * b, 80 + b));
            ss.push_back(sizes(700 + 50 * b, {256, 512, 1024, 2048}, 20, 21, 5 + b));
        }
        for (int b : {0, 1, 0, 1, 2, 3, 0, 1, 2, 3, 4, 5, 0, 5, 4, 1}) {
            p.x_dims = ss[b].x.data();
            p.t_dims = ss[b].t.data();
            expect(device_price(c, os[b], p), KWGPU_OK, c, "device batches in turn");
        }
        // autotuner on the first batch, its file, and pricing with the loaded choices
        p.x_dims = s.x.data();
        p.t_dims = s.t.data();
        std::vector<double> pr(o.size());
        std::vector<int32_t> st(o.size());
        std::vector<char> report(1 << 16);
        expect(kwgpu_fd1d_autotune(c, o.data(), o.size(), &p, pr.data(), st.data(), report.data(),
                                   report.size()),
               KWGPU_OK, c, "autotune");
        const std::string path = dir + "/tune.txt";
        expect(kwgpu_tune_save(c, path.c_str()), KWGPU_OK, c, "tune save");
        expect(device_price(c, o, p), KWGPU_OK, c, "tuned price");
        kwgpu_ctx* c2 = nullptr;
        expect(kwgpu_create(&c2, 0), KWGPU_OK, c2, "second context");
        expect(kwgpu_tune_load(c2, path.c_str()), KWGPU_OK, c2, "tune load");
        expect(device_price(c2, o, p), KWGPU_OK, c2, "loaded choices");
        // a replay on the context's stream, then the same batch on two other streams: each
        // solve waits for the previous one (same graph, same PDE order and counters)
        void* sa = nullptr;
        void* sb = nullptr;
        hipStreamCreateWithFlags(&sa, 1);
        hipStreamCreateWithFlags(&sb, 1);
        std::vector<double> rp(o.size());
        std::vector<int32_t> rs(o.size());
        for (void* st : {(void*)nullptr, sa, sb, sa})
            expect(kwgpu_fd1d_price_device(c2, o.data(), o.size(), &p, rp.data(), rs.data(), st),
                   KWGPU_OK, c2, "replay on a stream");
        // loading choices drops the graph whose launch may still be queued
        expect(kwgpu_tune_load(c2, path.c_str()), KWGPU_OK, c2, "tune load after a launch");
        // a file with a foreign precision and a lane-layout choice for FP32-mixed: both
        // lines are skipped, and FP32-mixed batches still price
        const std::string bad = dir + "/tune_bad.txt";
        if (FILE* f = std::fopen(bad.c_str(), "w")) {
            std::fprintf(f, "256 7 2 0\n512 %d -1 0\n1024 %d -1 1\n-1 9 1 0\n",
                         KWGPU_FP32_MIXED, KWGPU_FP32_MIXED);
            std::fclose(f);
        }
        expect(kwgpu_tune_load(c2, bad.c_str()), KWGPU_OK, c2, "tune load, bad lines");
        kwgpu_fd1d_params pm = p;
        pm.precision = KWGPU_FP32_MIXED;
        expect(device_price(c2, o, pm, sb), KWGPU_OK, c2, "mixed after a bad tune file");
        kwgpu_destroy(c2);
        hipStreamDestroy(sa);
        hipStreamDestroy(sb);
    }
    {  // the allocation-free entry point: host sizes, two calls, the second smaller;
       // uniform grids; the size limit and the lane layout refused
        auto o = options(2500, 91);
        Sizes s = sizes(2500, {64, 256, 512, 1024, 2048}, 32, 300, 9);
        kwgpu_fd1d_params_default(&p);
        p.x_dims = s.x.data();
        p.t_dims = s.t.data();
        std::vector<double> pr(o.size());
        std::vector<int32_t> st(o.size());
        expect(kwgpu_fd1d_price_device_noalloc(c, o.data(), o.size(), &p, pr.data(), st.data(),
                                               nullptr),
               KWGPU_OK, c, "noalloc uneven");
        expect(kwgpu_fd1d_price_device_noalloc(c, o.data(), 700, &p, pr.data(), st.data(), nullptr),
               KWGPU_OK, c, "noalloc uneven, smaller");
        kwgpu_fd1d_params_default(&p);
        p.x_dim = 1000;
        p.t_dim = 50;
        expect(kwgpu_fd1d_price_device_noalloc(c, o.data(), o.size(), &p, pr.data(), st.data(),
                                               nullptr),
               KWGPU_OK, c, "noalloc uniform");
        p.x_dim = 3000;
        expect(kwgpu_fd1d_price_device_noalloc(c, o.data(), 10, &p, pr.data(), st.data(), nullptr),
               KWGPU_ERR_UNSUPPORTED, c, "noalloc above 2048 points");
        p.x_dim = 512;
        expect(kwgpu_fd1d_price_device_noalloc(c, o.data(), (size_t(1) << 20) + 1, &p, pr.data(),
                                               st.data(), nullptr),
               KWGPU_ERR_INVALID, c, "noalloc above its maximum");
    }
    {  // PDE grouping: ten (spot, strike) pairs per PDE
        auto base = options(60, 7);
        std::vector<kwgpu_option> book;
        for (const auto& b : base)
            for (int j = 0; j < 10; ++j) {
                kwgpu_option x = b;
                x.s = 80 + 5 * j;
                x.k = 115 - 2.5 * j;
                book.push_back(x);
            }
        kwgpu_fd1d_params_default(&p);
        expect(host_price(c, book, p), KWGPU_OK, c, "grouped book");
    }
    {  // invalid inputs
        auto o = options(8, 9);
        kwgpu_fd1d_params_default(&p);
        p.x_dim = 2;
        expect(host_price(c, o, p), KWGPU_ERR_INVALID, c, "x_dim 2");
        kwgpu_fd1d_params_default(&p);
        Sizes s = sizes(8, {256}, 32, 1, 0);
        s.x[3] = 1;
        p.x_dims = s.x.data();
        p.t_dims = s.t.data();
        expect(host_price(c, o, p), KWGPU_ERR_INVALID, c, "x_dims[3] = 1");
        expect(device_price(c, o, p), KWGPU_ERR_INVALID, c, "device x_dims[3] = 1");
    }
    {  // surrogate: training, prediction, gated serving with per-option grids
        kwgpu_surrogate* s = nullptr;
        expect(kwgpu_surrogate_create(c, 2, 5, &s), KWGPU_OK, c, "surrogate create");
        const size_t n = 600;
        auto o = options(n, 31);
        std::vector<double> prem(n, 0.01);
        kwgpu_surrogate_train_params tp;
        kwgpu_surrogate_train_params_default(&tp);
        tp.epochs = 2;
        tp.batch = 128;
        std::vector<float> loss(2 * 2);
        expect(kwgpu_surrogate_train(c, s, o.data(), prem.data(), n, &tp, loss.data()), KWGPU_OK,
               c, "surrogate train");
        std::vector<double> pm(n), sp(n);
        std::vector<uint8_t> ib(n);
        expect(kwgpu_surrogate_predict(c, s, o.data(), n, nullptr, pm.data(), sp.data(), ib.data(),
                                       nullptr),
               KWGPU_OK, c, "surrogate predict");
        // fitted to the FD pricer: node labels on one grid, then on per-option grids. No
        // kernel runs here, so no node is labelled (the node grids read back as zeros) and
        // the fit stops before training with KWGPU_ERR_INVALID; what is checked is the
        // path up to there, its buffers and the configuration it records.
        Sizes g = sizes(n, {64, 128, 256, 512}, 20, 40, 6);
        size_t pairs = 0;
        kwgpu_fd1d_params_default(&p);
        p.x_dim = 128;
        p.t_dim = 32;
        expect(kwgpu_surrogate_fit(c, s, o.data(), n, &p, &tp, loss.data(), &pairs),
               KWGPU_ERR_INVALID, c, "surrogate fit");
        expect(kwgpu_surrogate_predict(c, s, o.data(), n, &p, pm.data(), sp.data(), ib.data(),
                                       nullptr),
               KWGPU_OK, c, "predict on the fitted grid");
        kwgpu_fd1d_params other = p;
        other.x_dim = 256;
        expect(kwgpu_surrogate_predict(c, s, o.data(), n, &other, pm.data(), sp.data(), ib.data(),
                                       nullptr),
               KWGPU_ERR_INVALID, c, "predict on another grid");
        kwgpu_surrogate* sg = nullptr;
        expect(kwgpu_surrogate_create(c, 2, 6, &sg), KWGPU_OK, c, "second surrogate");
        p.x_dims = g.x.data();
        p.t_dims = g.t.data();
        expect(kwgpu_surrogate_fit(c, sg, o.data(), n, &p, &tp, loss.data(), &pairs),
               KWGPU_ERR_INVALID, c, "surrogate fit, per-option grids");
        expect(kwgpu_surrogate_predict(c, sg, o.data(), n, nullptr, pm.data(), sp.data(), ib.data(),
                                       nullptr),
               KWGPU_ERR_INVALID, c, "per-option surrogate without grids");
        expect(kwgpu_surrogate_predict(c, sg, o.data(), n, &p, pm.data(), sp.data(), ib.data(),
                                       nullptr),
               KWGPU_OK, c, "per-option surrogate predict");
        std::vector<double> pr(n);
        std::vector<int32_t> route(n);
        size_t nfd = 0;
        for (kwgpu_surrogate* sx : {s, sg})
            for (int prec : {KWGPU_FP64, KWGPU_FP32_MIXED})
                for (double bound : {-1.0, 0.0, 1e300})
                    for (bool per_option : {false, true}) {
                        kwgpu_fd1d_params_default(&p);
                        p.x_dim = 128;
                        p.t_dim = 32;
                        p.precision = prec;
                        if (per_option) {
                            p.x_dims = g.x.data();
                            p.t_dims = g.t.data();
                        }
                        // a smaller batch after a larger one reuses the serving buffers
                        for (size_t m : {n, n / 3})
                            expect(kwgpu_surrogate_price(c, sx, o.data(), m, bound, 0.05, &p,
                                                         pr.data(), route.data(), &nfd),
                                   KWGPU_OK, c, "surrogate price");
                        // the same by class: 2 x 3 x 2 classes, every other one closed
                        kwgpu_gate gt{};
                        gt.n_t = 3;
                        gt.n_m = 2;
                        gt.t_edges[0] = 0.5;
                        gt.t_edges[1] = 2.0;
                        gt.m_edges[0] = 0.0;
                        for (int k = 0; k < 12; ++k) {
                            gt.spread_max[k] = k % 2 ? bound : -1.0;
                            gt.scale_max[k] = 0.05;
                        }
                        expect(kwgpu_surrogate_price_gated(c, sx, o.data(), n, &gt, &p,
                                                           pr.data(), route.data(), &nfd),
                               KWGPU_OK, c, "surrogate price by class");
                    }
        {
            kwgpu_gate bad{};
            bad.n_t = 2;
            bad.n_m = 1;
            bad.t_edges[0] = NAN;
            expect(kwgpu_surrogate_price_gated(c, s, o.data(), n, &bad, &p, pr.data(), route.data(),
                                               &nfd),
                   KWGPU_ERR_INVALID, c, "gate with a NaN edge");
            bad.t_edges[0] = 1.0;
            bad.n_m = KWGPU_GATE_MAX_M + 1;
            expect(kwgpu_surrogate_price_gated(c, s, o.data(), n, &bad, &p, pr.data(), route.data(),
                                               &nfd),
                   KWGPU_ERR_INVALID, c, "gate with too many bands");
            expect(kwgpu_surrogate_price_gated(c, s, o.data(), n, nullptr, &p, pr.data(),
                                               route.data(), &nfd),
                   KWGPU_ERR_INVALID, c, "no gate");
        }
        kwgpu_surrogate_destroy(sg);
        const std::string path = dir + "/sur.bin";
        expect(kwgpu_surrogate_save(c, s, path.c_str()), KWGPU_OK, c, "surrogate save");
        kwgpu_surrogate* s2 = nullptr;
        expect(kwgpu_surrogate_load(c, path.c_str(), &s2), KWGPU_OK, c, "surrogate load");
        kwgpu_surrogate_destroy(s2);
        kwgpu_surrogate_destroy(s);
    }
    kwgpu_destroy(c);
    const int live = mock_live();
    if (live) {
        std::fprintf(stderr, "FAIL %d handles or allocations left alive\n", live);
        ++failures;
    }
    std::printf("host logic: %s\n", failures ? "FAILED" : "ok");
    return failures ? 1 : 0;
}
