// kwgpu_api.hip -- host implementation of the C-ABI in include/kwgpu.h.
//
// Owns the per-context HIP stream and device buffers, reproduces the reference
// pricer's host-side logic around the solve (config defaults, PDE compression,
// error text) and dispatches the FD kernels by layout and grid size.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <map>
#include <mutex>
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

// Kernel choice for one grid-size bucket: a wave-layout variant (or -1 for the lane
// layout) and whether its grid is persistent.
struct Choice {
    int variant = -1;
    bool persist = true;
};

// One launch of an uneven batch: PDEs order[begin, begin + count) of one bucket.
struct Bucket {
    int cap;         // grid-size bucket: power of two >= x_dim (64 .. 2048), or kLaneCap
    Choice ch;
    size_t begin, count;
    int max_x;
    int prec = KWGPU_FP64;  // the march's precision for this bucket's PDEs
};

constexpr int kLaneCap = 1 << 30;
constexpr int kSide = 4;  // side streams an uneven batch forks onto

// Device scratch of kwgpu_fd1d_price_device_noalloc: one uneven batch's PDE order, its
// per-option x and t points, and its bucket counters, for at most kNoAllocMaxOptions
// options (12 MiB + 256 B, held by every device the library runs on until the process
// ends). Host code never uses the name: it takes the device address from
// hipGetSymbolAddress, once per device, when a context is created there.
constexpr size_t kNoAllocMaxOptions = size_t(1) << 20;
constexpr int kNoAllocCounters = 64;
__device__ __align__(256) uint32_t scratch[3 * kNoAllocMaxOptions + kNoAllocCounters];
constexpr int kMaxDevices = 64;
uint32_t* g_scratch[kMaxDevices] = {};  // device address of `scratch`, per device
std::mutex g_scratch_mu;

// With `device` current: looks up and keeps its address of `scratch` (first call per
// device only).
hipError_t init_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[device]) return hipSuccess;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[device] = static_cast<uint32_t*>(ptr);
    return hipSuccess;
}

struct kwgpu_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    DevBuf opts, prices, status, rep, off, list, order, xd, td, ws, counters;
    // autotuned choices per (bucket, precision), and per precision whether an uneven
    // batch's buckets run one after another on the caller's stream (true) or
    // concurrently on the side streams (false, the default)
    std::map<std::pair<int, int>, Choice> tuned;
    std::map<int, bool> serial;
    // uneven batches: side streams, fork/join events, the cached plan and its graph
    hipStream_t side[kSide] = {};
    hipEvent_t ev_fork = nullptr, ev_join[kSide] = {};
    bool plan_valid = false;
    uint64_t plan_key = 0;              // hash of everything but the sizes
    std::vector<int32_t> plan_x, plan_t;  // the sizes the plan was made for (compared exactly)
    std::vector<uint8_t> plan_p;          // ... and the per-PDE precisions (empty: one for all)
    int plan_prec = -1, plan_layout = -1;  // ... and the precision and layout it bucketed for
    int32_t* hsizes = nullptr;          // pinned staging for device-resident sizes
    size_t hsizes_cap = 0;
    std::vector<Bucket> plan;
    hipGraphExec_t graph = nullptr;
    bool use_graphs = true;
    // the last launch of `graph` (drop_graph waits for it before destroying the graph)
    hipEvent_t ev_graph = nullptr;
    bool graph_in_flight = false;
    // end of the last FD solve on this context and its stream: a solve on another stream
    // waits for it before it reuses the context's PDE order, counters and workspace
    hipEvent_t ev_done = nullptr;
    hipStream_t done_stream = nullptr;
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

// Grid-size bucket of an x_dim: the power of two >= x_dim (at least 64), or kLaneCap
// for grids the wave layout does not hold.
int bucket_of(int x, int precision, int layout) {
    if (layout == KWGPU_LAYOUT_LANE_SOA || !kwgpu::wave_layout_fits(x, precision)) return kLaneCap;
    int c = 64;
    while (c < x) c <<= 1;
    return c;
}

Choice choice_for(const kwgpu_ctx* c, int cap, int precision, int layout) {
    Choice ch;
    if (cap == kLaneCap) return ch;  // lane layout
    auto it = c->tuned.find({cap, precision});
    if (it != c->tuned.end() && !(layout == KWGPU_LAYOUT_WAVE_REG && it->second.variant < 0))
        return it->second;
    ch.variant = kwgpu::default_wave_variant(cap);
    ch.persist = true;
    return ch;
}

uint64_t fnv(uint64_t h, const void* p, size_t n) {
    const unsigned char* b = (const unsigned char*)p;
    for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
    return h;
}

// Hashes of the fields themselves: the structs hold padding whose bytes are unspecified
// after aggregate initialisation or a copy, so hashing their storage would change the
// key from call to call.
template <class... T>
uint64_t fnv_fields(uint64_t h, const T&... v) {
    ((h = fnv(h, &v, sizeof(v))), ...);
    return h;
}

uint64_t fnv_batch(uint64_t h, const FdBatch& b) {
    return fnv_fields(h, b.opts, b.rep, b.off, b.list, b.order, b.x_dims, b.t_dims, b.m, b.x_dim,
                      b.t_dim, b.density, b.scale, b.theta, b.prices, b.status, b.ws, b.ws_stride,
                      b.ws_rows);
}

uint64_t fnv_bucket(uint64_t h, const Bucket& k) {
    const int persist = k.ch.persist ? 1 : 0;
    return fnv_fields(h, k.cap, k.ch.variant, persist, k.begin, k.count, k.max_x, k.prec);
}

int lane_launch(kwgpu_ctx* c, FdBatch bb, int max_x, int precision, hipStream_t s, double* ws) {
    if (precision != KWGPU_FP64)
        return fail(c, KWGPU_ERR_UNSUPPORTED, "kwgpu: FP32-mixed runs in the wave layout only");
    bb.ws = ws;
    bb.ws_stride = bb.m;
    bb.ws_rows = max_x;
    if (hipError_t e = kwgpu::launch_fd1d_lane(bb, s); e != hipSuccess)
        return hip_fail(c, e, "kwgpu: lane kernel launch");
    return KWGPU_OK;
}

// Launches the buckets of an uneven batch: in order on the caller's stream when
// `serial`, else bucket i on side stream i % kSide (forked
// from and joined back into s), persistent buckets taking work from counters[i].
int launch_buckets(kwgpu_ctx* c, const FdBatch& b, const std::vector<Bucket>& plan,
                   hipStream_t s, bool serial) {
    hipError_t e;
    if ((e = kwgpu::launch_zero_counters((int*)c->counters.p, (int)plan.size(), s)))
        return hip_fail(c, e, "kwgpu: counter reset");
    if (!serial && (e = hipEventRecord(c->ev_fork, s))) return hip_fail(c, e, "kwgpu: fork");
    size_t ws_off = 0;
    bool used[kSide] = {};
    for (size_t i = 0; i < plan.size(); ++i) {
        const Bucket& k = plan[i];
        hipStream_t ss = serial ? s : c->side[i % kSide];
        if (!serial && !used[i % kSide] && (e = hipStreamWaitEvent(ss, c->ev_fork, 0)))
            return hip_fail(c, e, "kwgpu: fork wait");
        if (!serial) used[i % kSide] = true;
        FdBatch bb = b;
        bb.order = (const uint32_t*)c->order.p + k.begin;
        bb.m = (int)k.count;
        if (k.ch.variant < 0) {
            double* ws = (double*)c->ws.p + ws_off;
            ws_off += kwgpu::lane_workspace_doubles(bb.m, k.max_x);
            if (int rc = lane_launch(c, bb, k.max_x, k.prec, ss, ws)) return rc;
        } else if ((e = kwgpu::launch_fd1d_wave_variant(bb, k.ch.variant, k.prec, ss,
                                                        k.ch.persist ? (int*)c->counters.p + i
                                                                     : nullptr))) {
            return hip_fail(c, e, "kwgpu: wave kernel launch");
        }
    }
    for (int k = 0; k < kSide; ++k) {
        if (!used[k]) continue;
        if ((e = hipEventRecord(c->ev_join[k], c->side[k])) || (e = hipStreamWaitEvent(s, c->ev_join[k], 0)))
            return hip_fail(c, e, "kwgpu: join");
    }
    return KWGPU_OK;
}

int ensure_side_streams(kwgpu_ctx* c) {
    if (c->ev_fork) return KWGPU_OK;
    hipError_t e;
    for (int k = 0; k < kSide; ++k)
        if ((e = hipStreamCreateWithFlags(&c->side[k], hipStreamNonBlocking)) ||
            (e = hipEventCreateWithFlags(&c->ev_join[k], hipEventDisableTiming)))
            return hip_fail(c, e, "kwgpu: side streams");
    if ((e = hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming)))
        return hip_fail(c, e, "kwgpu: side streams");
    return KWGPU_OK;
}

// Destroys the cached graph once its last launch, which may still be queued on the
// stream it was launched on, has finished.
void drop_graph(kwgpu_ctx* c) {
    if (c->graph) {
        if (c->graph_in_flight) (void)hipEventSynchronize(c->ev_graph);
        (void)hipGraphExecDestroy(c->graph);
    }
    c->graph = nullptr;
    c->graph_in_flight = false;
    c->plan_valid = false;
}

// Runs the FD solve of `b` (device buffers already set) on stream s. pde_x / pde_t are
// host copies of each PDE's grid sizes (b.m each); null means every PDE uses b.x_dim /
// b.t_dim. pde_prec (host, b.m, optional) gives each PDE its own precision; the buckets
// then split by precision too, and the batch is still one plan and one graph.
//
// Uneven batches: PDEs are grouped by grid-size bucket, each bucket sorted by t-points
// (descending) so that co-packed segments march for about the same number of steps and
// a persistent grid takes the longest first; each bucket's kernel (autotuned variant)
// runs on its own side stream. The whole launch sequence is captured once into a
// hipGraph and replayed while the batch (sizes, buffers, parameters) is unchanged.
int run_fd_on(kwgpu_ctx* c, FdBatch b, const int32_t* pde_x, const int32_t* pde_t, int precision,
              int layout, hipStream_t s, const uint8_t* pde_prec) {
    // per-PDE precisions go through the planned path, with the uniform sizes spelled out
    std::vector<int32_t> ux, ut;
    if (pde_prec && !pde_x) {
        ux.assign((size_t)b.m, b.x_dim);
        ut.assign((size_t)b.m, b.t_dim);
        pde_x = ux.data();
        pde_t = ut.data();
    }
    const bool uniform = pde_x == nullptr;
    auto prec_of = [&](int p) { return pde_prec ? (int)pde_prec[p] : precision; };
    // a batch with exactly the cached plan's sizes, precisions and layout has the plan's
    // buckets: checked first, so a replay does no per-PDE pass but these compares
    const size_t bytes = sizeof(int32_t) * (size_t)b.m;
    const bool same_sizes =
        !uniform && c->plan_valid && c->plan_prec == precision && c->plan_layout == layout &&
        c->plan_x.size() == (size_t)b.m && !std::memcmp(c->plan_x.data(), pde_x, bytes) &&
        !std::memcmp(c->plan_t.data(), pde_t, bytes) &&
        (pde_prec ? c->plan_p.size() == (size_t)b.m && !std::memcmp(c->plan_p.data(), pde_prec, b.m)
                  : c->plan_p.empty());
    if (!same_sizes) {  // (a replayed plan passed these checks when it was made)
        int mx = b.x_dim, mx_mixed = 0;
        if (!uniform) {
            mx = 0;
            for (int p = 0; p < b.m; ++p) {
                mx = std::max(mx, pde_x[p]);
                if (prec_of(p) == KWGPU_FP32_MIXED) mx_mixed = std::max(mx_mixed, pde_x[p]);
            }
        } else if (precision == KWGPU_FP32_MIXED) {
            mx_mixed = b.x_dim;
        }
        if (mx_mixed && layout == KWGPU_LAYOUT_LANE_SOA)
            return fail(c, KWGPU_ERR_UNSUPPORTED, "kwgpu: FP32-mixed runs in the wave layout only");
        if (layout == KWGPU_LAYOUT_WAVE_REG && !kwgpu::wave_layout_fits(mx, precision))
            return fail(c, KWGPU_ERR_UNSUPPORTED, "kwgpu: wave layout supports x_dim <= 2048");
        if (mx_mixed && !kwgpu::wave_layout_fits(mx_mixed, KWGPU_FP32_MIXED))
            return fail(c, KWGPU_ERR_UNSUPPORTED, "kwgpu: FP32-mixed supports x_dim <= 2048");
    }

    if (uniform) {
        const int cap = bucket_of(b.x_dim, precision, layout);
        Choice ch = choice_for(c, cap, precision, layout);
        if (ch.variant < 0) {
            const size_t need = kwgpu::lane_workspace_doubles(b.m, b.x_dim) * sizeof(double);
            if (c->ws.cap < need) drop_graph(c);  // a captured graph may hold the old buffer
            if (ensure(c->ws, need) != hipSuccess)
                return fail(c, KWGPU_ERR_NOMEM, "kwgpu: lane workspace allocation failed");
            return lane_launch(c, b, b.x_dim, precision, s, (double*)c->ws.p);
        }
        // uniform batches launch directly: one kernel, nothing to fork or reorder
        if (hipError_t e = kwgpu::launch_fd1d_wave_variant(b, ch.variant, precision, s); e != hipSuccess)
            return hip_fail(c, e, "kwgpu: wave kernel launch");
        return KWGPU_OK;
    }

    const int m = b.m;
    // buckets in ascending capacity (64 .. 2048, then the lane layout), FP64 ones first,
    // then FP32-mixed ones: counts and widest grid, without sorting the PDEs
    constexpr int kSlots = 7;
    auto slot_of = [&](int p) {
        const int cap = bucket_of(pde_x[p], precision, layout);
        const int k = cap == kLaneCap ? kSlots - 1 : 31 - __builtin_clz((unsigned)cap) - 6;
        return prec_of(p) == KWGPU_FP32_MIXED ? kSlots + k : k;
    };
    std::vector<Bucket> plan;
    if (same_sizes) {
        plan = c->plan;
        for (Bucket& k : plan) k.ch = choice_for(c, k.cap, k.prec, layout);
    } else {
        size_t count[2 * kSlots] = {};
        int widest[2 * kSlots] = {};
        for (int p = 0; p < m; ++p) {
            const int k = slot_of(p);
            ++count[k];
            widest[k] = std::max(widest[k], pde_x[p]);
        }
        size_t begin = 0;
        for (int k = 0; k < 2 * kSlots; ++k) {
            if (!count[k]) continue;
            const int kc = k % kSlots;
            const int cap = kc == kSlots - 1 ? kLaneCap : 64 << kc;
            const int pr = k < kSlots ? KWGPU_FP64 : KWGPU_FP32_MIXED;
            plan.push_back(Bucket{cap, choice_for(c, cap, pr, layout), begin, count[k], widest[k], pr});
            begin += count[k];
        }
    }
    // the batch's identity: buffers, parameters and kernel choices (hashed) and the sizes
    // (compared exactly with the plan's own copy)
    uint64_t key = 1469598103934665603ull;
    key = fnv_batch(key, b);
    key = fnv_fields(key, precision);
    for (const Bucket& k : plan) key = fnv_bucket(key, k);
    const auto sit = c->serial.find(precision);
    const int serial = sit != c->serial.end() && sit->second ? 1 : 0;
    key = fnv_fields(key, serial);

    // ... and the context buffers a captured graph would reference
    auto with_buffers = [&](uint64_t k) {
        const void* bufs[3] = {c->order.p, c->counters.p, c->ws.p};
        return fnv(k, bufs, sizeof(bufs));
    };
    if (int rc = ensure_side_streams(c)) return rc;
    const bool rebuild = !same_sizes || c->plan_key != with_buffers(key);
    if (rebuild) {
        drop_graph(c);
        size_t ws = 0;
        for (const Bucket& k : plan)
            if (k.ch.variant < 0) ws += kwgpu::lane_workspace_doubles((int)k.count, k.max_x);
        if (ensure(c->order, sizeof(uint32_t) * m) != hipSuccess ||
            ensure(c->counters, sizeof(int) * plan.size()) != hipSuccess ||
            (ws && ensure(c->ws, ws * sizeof(double)) != hipSuccess))
            return fail(c, KWGPU_ERR_NOMEM, "kwgpu: uneven batch allocation failed");
        key = with_buffers(key);
        // PDEs by bucket, longest march first within a bucket (ties by index)
        std::vector<int> slot(m);
        for (int p = 0; p < m; ++p) slot[p] = slot_of(p);
        std::vector<uint32_t> ord(m);
        std::iota(ord.begin(), ord.end(), 0u);
        std::sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t z) {
            return std::make_tuple(slot[a], -pde_t[a], a) < std::make_tuple(slot[z], -pde_t[z], z);
        });
        hipError_t e;
        if ((e = hipMemcpyAsync(c->order.p, ord.data(), sizeof(uint32_t) * m, hipMemcpyHostToDevice, s)) ||
            (e = hipStreamSynchronize(s)))
            return hip_fail(c, e, "kwgpu: order upload");
        c->plan = plan;
        c->plan_key = key;
        c->plan_x.assign(pde_x, pde_x + m);
        c->plan_t.assign(pde_t, pde_t + m);
        if (pde_prec)
            c->plan_p.assign(pde_prec, pde_prec + m);
        else
            c->plan_p.clear();
        c->plan_prec = precision;
        c->plan_layout = layout;
        c->plan_valid = true;
    }
    if (!c->use_graphs) return launch_buckets(c, b, c->plan, s, serial);
    if (!c->graph) {
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            const int rc = launch_buckets(c, b, c->plan, s, serial);
            hipError_t e2 = hipStreamEndCapture(s, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            e = e2;
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
        if (g) (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            // streams that cannot be captured (e.g. the legacy null stream): launch directly
            (void)hipGetLastError();
            c->graph = nullptr;
            return launch_buckets(c, b, c->plan, s, serial);
        }
    }
    if (hipError_t e = hipGraphLaunch(c->graph, s); e != hipSuccess)
        return hip_fail(c, e, "kwgpu: graph launch");
    if (hipError_t e = hipEventRecord(c->ev_graph, s); e != hipSuccess)
        return hip_fail(c, e, "kwgpu: graph event");
    c->graph_in_flight = true;
    return KWGPU_OK;
}

// run_fd_on, ordered after the context's previous solve: a solve on a stream other than
// the previous one's first waits for that solve, whose kernels may still read the
// context's PDE order, counters and lane workspace that this one rewrites; its own end
// is recorded for the next.
int run_fd(kwgpu_ctx* c, FdBatch b, const int32_t* pde_x, const int32_t* pde_t, int precision,
           int layout, hipStream_t s, const uint8_t* pde_prec = nullptr) {
    if (c->done_stream && c->done_stream != s)
        if (hipError_t e = hipStreamWaitEvent(s, c->ev_done, 0); e != hipSuccess)
            return hip_fail(c, e, "kwgpu: order after the previous solve");
    const int rc = run_fd_on(c, b, pde_x, pde_t, precision, layout, s, pde_prec);
    if (hipError_t e = hipEventRecord(c->ev_done, s); e != hipSuccess)
        return rc ? rc : hip_fail(c, e, "kwgpu: solve event");
    c->done_stream = s;
    return rc;
}

// PDE compression (src/Pricer/kwFd1d.cpp:28-65): options equal in (t, r, q, z, e, w) and
// grid size share one solve. The reference sorts the options by that key; only the
// grouping matters here (each PDE is solved from its first option's fields), so the key
// goes into an open-addressing hash table, O(n), and a counting pass lays out each PDE's
// options (rep = first option of each PDE, off = CSR offsets into list). Fields compare
// with ==, so a NaN field never matches and -0.0 matches +0.0, as under the reference's
// `<`; options whose key holds a NaN each get a PDE of their own (they fail as bad
// input either way).
void group_pdes(const kwgpu_option* opts, size_t n, const int32_t* xd, const int32_t* td, int x_dim,
                int t_dim, std::vector<uint32_t>& rep, std::vector<uint32_t>& off,
                std::vector<uint32_t>& list) {
    auto gx = [&](size_t i) { return xd ? xd[i] : x_dim; };
    auto gt = [&](size_t i) { return td ? td[i] : t_dim; };
    auto same = [&](size_t i, size_t j) {
        const kwgpu_option &a = opts[i], &b = opts[j];
        return a.t == b.t && a.r == b.r && a.q == b.q && a.z == b.z && a.e == b.e && a.w == b.w &&
               gx(i) == gx(j) && gt(i) == gt(j);
    };
    auto mix = [](uint64_t h, double v) {
        uint64_t u;
        v = v == 0.0 ? 0.0 : v;  // -0.0 hashes as +0.0
        std::memcpy(&u, &v, sizeof(u));
        h ^= u + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
        return h;
    };
    size_t cap = 16;
    while (cap < 2 * n) cap <<= 1;
    std::vector<int32_t> table(cap, -1);  // PDE id per slot
    std::vector<uint32_t> pde(n), count;
    rep.clear();
    for (size_t i = 0; i < n; ++i) {
        const kwgpu_option& a = opts[i];
        uint64_t h = 0;
        h = mix(h, a.t);
        h = mix(h, a.r);
        h = mix(h, a.q);
        h = mix(h, a.z);
        h = mix(h, (double)(a.e * 4 + a.w));
        h = mix(h, (double)gx(i) * 65536.0 + (double)gt(i));
        h ^= h >> 31;
        h *= 0xbf58476d1ce4e5b9ull;
        h ^= h >> 29;
        size_t slot = h & (cap - 1);
        for (;;) {
            const int32_t id = table[slot];
            if (id < 0) {
                table[slot] = (int32_t)rep.size();
                pde[i] = (uint32_t)rep.size();
                rep.push_back((uint32_t)i);
                count.push_back(1);
                break;
            }
            if (same(rep[id], i)) {
                pde[i] = (uint32_t)id;
                ++count[id];
                break;
            }
            slot = (slot + 1) & (cap - 1);
        }
    }
    const size_t m = rep.size();
    off.assign(m + 1, 0);
    for (size_t k = 0; k < m; ++k) off[k + 1] = off[k] + count[k];
    list.resize(n);
    std::vector<uint32_t> fill(off.begin(), off.end() - 1);
    for (size_t i = 0; i < n; ++i) list[fill[pde[i]]++] = (uint32_t)i;  // ascending per PDE
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
    std::vector<uint32_t> rep, off, list;
    group_pdes(opts, n, p->x_dims, p->t_dims, p->x_dim, p->t_dim, rep, off, list);
    const size_t m = rep.size();
    // FP32-mixed: a PDE beyond the mixed march's limits (FP32 increments that could
    // overflow, or a drift past the mesh) is solved in FP64, as the serving gate does,
    // in the same solve
    std::vector<uint8_t> pprec;
    if (p->precision == KWGPU_FP32_MIXED) {
        bool any = false;
        pprec.resize(m);
        for (size_t k = 0; k < m; ++k) {
            const kwgpu_option& a = opts[rep[k]];
            const bool f64 =
                kwgpu::mixed_overflow_risk(a, p->scale) || kwgpu::mixed_drift_risk(a, p->scale);
            pprec[k] = f64 ? KWGPU_FP64 : KWGPU_FP32_MIXED;
            any = any || f64;
        }
        if (!any) pprec.clear();
    }

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
    const int rc = run_fd(c, b, pde_x.empty() ? nullptr : pde_x.data(),
                          pde_t.empty() ? nullptr : pde_t.data(), p->precision, p->layout, s,
                          pprec.empty() ? nullptr : pprec.data());
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

namespace kwgpu {
int ctx_fail(kwgpu_ctx* c, int code, const std::string& msg) { return fail(c, code, msg); }
hipStream_t ctx_stream(kwgpu_ctx* c) { return c ? c->stream : nullptr; }
int ctx_device(kwgpu_ctx* c) { return c ? c->device : 0; }

// kwgpu_fd1d_price_device with an optional precision per option (host array of n,
// KWGPU_FP64 or KWGPU_FP32_MIXED; null: params->precision for all): one plan, one graph.
int fd_price_device_prec(kwgpu_ctx* c, const kwgpu_option* d_opts, size_t n,
                         const kwgpu_fd1d_params* p, double* d_prices, int32_t* d_status,
                         void* stream, const uint8_t* host_prec) {
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
    if (!p->x_dims && !p->t_dims)
        return run_fd(c, b, nullptr, nullptr, p->precision, p->layout, s, host_prec);
    // grouping needs the sizes on the host: pinned staging owned by the context
    if (c->hsizes_cap < 2 * n) {
        if (c->hsizes) (void)hipHostFree(c->hsizes);
        c->hsizes = nullptr;
        c->hsizes_cap = 0;
        if (hipHostMalloc((void**)&c->hsizes, 2 * n * sizeof(int32_t), hipHostMallocDefault) !=
            hipSuccess)
            return fail(c, KWGPU_ERR_NOMEM, "kwgpu: pinned staging allocation failed");
        c->hsizes_cap = 2 * n;
    }
    int32_t* hx = c->hsizes;
    int32_t* ht = c->hsizes + n;
    hipError_t e;
    if (!p->x_dims) std::fill(hx, hx + n, p->x_dim);
    if (!p->t_dims) std::fill(ht, ht + n, p->t_dim);
    if (p->x_dims &&
        (e = hipMemcpyAsync(hx, p->x_dims, n * sizeof(int32_t), hipMemcpyDeviceToHost, s)))
        return hip_fail(c, e, "kwgpu: x_dims download");
    if (p->t_dims &&
        (e = hipMemcpyAsync(ht, p->t_dims, n * sizeof(int32_t), hipMemcpyDeviceToHost, s)))
        return hip_fail(c, e, "kwgpu: t_dims download");
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return hip_fail(c, e, "kwgpu: sync");
    if (int rc = check_dims(c, hx, ht, n)) return rc;
    return run_fd(c, b, hx, ht, p->precision, p->layout, s, host_prec);
}
}  // namespace kwgpu

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
    // the allocation-free entry point's scratch on this device (a failure only disables
    // that entry point, which then reports it)
    (void)init_scratch(device);
    auto* c = new kwgpu_ctx();
    c->device = device;
    if (const char* g = std::getenv("KWGPU_NO_GRAPH")) c->use_graphs = g[0] == '0';
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return KWGPU_ERR_HIP;
    }
    if (hipEventCreateWithFlags(&c->ev_graph, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_done, hipEventDisableTiming) != hipSuccess) {
        if (c->ev_graph) (void)hipEventDestroy(c->ev_graph);
        (void)hipStreamDestroy(c->stream);
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
    if (c->done_stream) (void)hipEventSynchronize(c->ev_done);  // a solve on a caller's stream
    drop_graph(c);
    for (int k = 0; k < kSide; ++k) {
        if (c->side[k]) (void)hipStreamDestroy(c->side[k]);
        if (c->ev_join[k]) (void)hipEventDestroy(c->ev_join[k]);
    }
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    (void)hipEventDestroy(c->ev_graph);
    (void)hipEventDestroy(c->ev_done);
    if (c->hsizes) (void)hipHostFree(c->hsizes);
    for (DevBuf* b : {&c->opts, &c->prices, &c->status, &c->rep, &c->off, &c->list, &c->order,
                      &c->xd, &c->td, &c->ws, &c->counters})
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
    return kwgpu::fd_price_device_prec(c, d_opts, n, p, d_prices, d_status, stream, nullptr);
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
    if (n > 0x3fffffffu)  // both legs are solved as one batch of 2n
        return fail(c, KWGPU_ERR_INVALID, "kwgpu: too many options");
    if (int rc = check_dims(c, p->x_dims, p->t_dims, n)) return rc;
    if (hipError_t e = hipSetDevice(c->device); e != hipSuccess) return hip_fail(c, e, "hipSetDevice");
    const char* prefix = "Fd1d_BlackScholes_Pricer::price : ";
    // legs 1 and 2 in one solve: the options as given (src/Pricer/kwFd1d_BlackScholes.cpp:18)
    // followed by their European copies (:21-28), 2n options on the same grids
    std::vector<kwgpu_option> both(2 * n);
    std::copy(opts, opts + n, both.begin());
    std::copy(opts, opts + n, both.begin() + n);
    for (size_t i = n; i < 2 * n; ++i) both[i].e = 0;
    kwgpu_fd1d_params p2 = *p;
    std::vector<int32_t> xd2, td2;
    if (p->x_dims) {
        xd2.assign(p->x_dims, p->x_dims + n);
        xd2.insert(xd2.end(), p->x_dims, p->x_dims + n);
        p2.x_dims = xd2.data();
    }
    if (p->t_dims) {
        td2.assign(p->t_dims, p->t_dims + n);
        td2.insert(td2.end(), p->t_dims, p->t_dims + n);
        p2.t_dims = td2.data();
    }
    if (int rc = fd_price_host(c, both.data(), 2 * n, &p2, c->opts, c->prices, c->status)) return rc;
    hipError_t e;
    // leg 3 and combine: price += BS(euro) - FD(euro) (:30-40)
    if ((e = kwgpu::launch_fd1d_bs_combine((const kwgpu_option*)c->opts.p, n, (double*)c->prices.p,
                                           (const double*)c->prices.p + n, c->stream)))
        return hip_fail(c, e, "kwgpu: combine kernel launch");
    std::vector<int32_t> st(2 * n);
    if ((e = hipMemcpyAsync(prices, c->prices.p, n * sizeof(double), hipMemcpyDeviceToHost,
                            c->stream)) ||
        (e = hipMemcpyAsync(st.data(), c->status.p, 2 * n * sizeof(int32_t), hipMemcpyDeviceToHost,
                            c->stream)) ||
        (e = hipStreamSynchronize(c->stream)))
        return hip_fail(c, e, "kwgpu: download");
    for (size_t i = 0; i < n; ++i)
        if (st[i] == KWGPU_STATUS_OK && st[n + i] != KWGPU_STATUS_OK) st[i] = st[n + i];
    st.resize(n);
    if (status) std::memcpy(status, st.data(), n * sizeof(int32_t));
    return collect(c, opts, n, p, st, prefix);
}

int kwgpu_fd1d_greeks(kwgpu_ctx* c, const kwgpu_option* opts, size_t n,
                      const kwgpu_fd1d_params* p, double bump, kwgpu_greeks* out,
                      int32_t* status) {
#pragma clang fp contract(off)
    constexpr size_t kSets = 9;  // as given, then s, z, t, r times (1 + bump), (1 - bump)
    if (!c) return KWGPU_ERR_INVALID;
    c->err.clear();
    if (int rc = check_params(c, p)) return rc;
    if (n == 0) return KWGPU_OK;
    if (!opts || !out) return fail(c, KWGPU_ERR_INVALID, "kwgpu: null pointer");
    if (!(bump > 0.0 && bump < 1.0)) return fail(c, KWGPU_ERR_INVALID, "kwgpu: bump not in (0, 1)");
    if (n > 0x7fffffffu / kSets) return fail(c, KWGPU_ERR_INVALID, "kwgpu: too many options");
    if (int rc = check_dims(c, p->x_dims, p->t_dims, n)) return rc;
    if (hipError_t e = hipSetDevice(c->device); e != hipSuccess) return hip_fail(c, e, "hipSetDevice");
    // option i's nine prices sit at kSets * i + k
    std::vector<kwgpu_option> all(n * kSets);
    std::vector<int32_t> xd, td;
    kwgpu_fd1d_params p9 = *p;
    if (p->x_dims) xd.resize(n * kSets);
    if (p->t_dims) td.resize(n * kSets);
    for (size_t i = 0; i < n; ++i) {
        for (size_t k = 0; k < kSets; ++k) {
            kwgpu_option a = opts[i];
            if (k > 0) {
                const double f = (k - 1) % 2 == 0 ? 1.0 + bump : 1.0 - bump;
                switch ((k - 1) / 2) {
                    case 0: a.s = f * a.s; break;
                    case 1: a.z = f * a.z; break;
                    case 2: a.t = f * a.t; break;
                    default: a.r = f * a.r; break;
                }
            }
            all[kSets * i + k] = a;
            if (p->x_dims) xd[kSets * i + k] = p->x_dims[i];
            if (p->t_dims) td[kSets * i + k] = p->t_dims[i];
        }
    }
    if (p->x_dims) p9.x_dims = xd.data();
    if (p->t_dims) p9.t_dims = td.data();
    const size_t n9 = n * kSets;
    if (int rc = fd_price_host(c, all.data(), n9, &p9, c->opts, c->prices, c->status)) return rc;
    std::vector<double> v(n9);
    std::vector<int32_t> st(n9);
    hipError_t e;
    if ((e = hipMemcpy(v.data(), c->prices.p, n9 * sizeof(double), hipMemcpyDeviceToHost)) ||
        (e = hipMemcpy(st.data(), c->status.p, n9 * sizeof(int32_t), hipMemcpyDeviceToHost)))
        return hip_fail(c, e, "kwgpu: download");
    for (size_t i = 0; i < n; ++i) {
        const double* w = &v[kSets * i];
        const kwgpu_option& a = opts[i];
        auto first = [&](int k, double x) { return (w[k] - w[k + 1]) / (2.0 * bump * x) + 0.0; };
        const double bs = bump * a.s;
        kwgpu_greeks g;
        g.price = w[0];
        g.delta = first(1, a.s);
        g.gamma = (w[1] - 2.0 * w[0] + w[2]) / (bs * bs) + 0.0;
        g.vega = first(3, a.z) / 100.0;
        g.theta = first(5, a.t) / 360.0;
        g.rho = first(7, a.r) / 100.0;
        out[i] = g;
        if (status) {
            status[i] = KWGPU_STATUS_OK;
            for (size_t k = 0; k < kSets; ++k)
                if (st[kSets * i + k] != KWGPU_STATUS_OK) {
                    status[i] = st[kSets * i + k];
                    break;
                }
        }
    }
    return collect(c, all.data(), n9, &p9, st, "kwgpu_fd1d_greeks: ");
}

int kwgpu_fd1d_price_device_noalloc(kwgpu_ctx* c, const kwgpu_option* d_opts, size_t n,
                                    const kwgpu_fd1d_params* p, double* d_prices,
                                    int32_t* d_status, void* stream) {
    if (!c) return KWGPU_ERR_INVALID;
    c->err.clear();
    if (int rc = check_params(c, p)) return rc;
    if (n == 0) return KWGPU_OK;
    if (!d_opts || !d_prices) return fail(c, KWGPU_ERR_INVALID, "kwgpu: null pointer");
    if (n > kNoAllocMaxOptions)
        return fail(c, KWGPU_ERR_INVALID,
                    "kwgpu_fd1d_price_device_noalloc: " + std::to_string(n) +
                        " options exceed its maximum of " + std::to_string(kNoAllocMaxOptions));
    if (int rc = check_dims(c, p->x_dims, p->t_dims, n)) return rc;
    const int32_t* hx = p->x_dims;  // host arrays here
    const int32_t* ht = p->t_dims;
    int mx = p->x_dim;
    if (hx) mx = *std::max_element(hx, hx + n);
    if (p->layout == KWGPU_LAYOUT_LANE_SOA || !kwgpu::wave_layout_fits(mx, p->precision))
        return fail(c, KWGPU_ERR_UNSUPPORTED,
                    "kwgpu_fd1d_price_device_noalloc: the wave layout only (x_dim <= 2048)");
    if (hipError_t e = hipSetDevice(c->device); e != hipSuccess) return hip_fail(c, e, "hipSetDevice");
    uint32_t* base = c->device < kMaxDevices ? g_scratch[c->device] : nullptr;
    if (!base)
        return fail(c, KWGPU_ERR_HIP, "kwgpu_fd1d_price_device_noalloc: no scratch on this device");
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    // the previous solve on this context may still read the scratch (calls from other
    // contexts on this device are not ordered: they race)
    if (c->done_stream && c->done_stream != s)
        if (hipError_t e = hipStreamWaitEvent(s, c->ev_done, 0); e != hipSuccess)
            return hip_fail(c, e, "kwgpu: order after the previous solve");
    uint32_t* d_order = base;
    int32_t* d_x = reinterpret_cast<int32_t*>(base + kNoAllocMaxOptions);
    int32_t* d_t = reinterpret_cast<int32_t*>(base + 2 * kNoAllocMaxOptions);
    int* d_counters = reinterpret_cast<int*>(base + 3 * kNoAllocMaxOptions);
    FdBatch b{};
    b.opts = d_opts;
    b.m = (int)n;
    b.x_dim = p->x_dim;
    b.t_dim = p->t_dim;
    b.density = p->density;
    b.scale = p->scale;
    b.theta = p->theta;
    b.prices = d_prices;
    b.status = d_status;
    // buckets: (x-points rounded up to a power of two) in ascending order, PDEs by
    // t-points descending within each (ties by index), as for uneven batches elsewhere
    std::vector<Bucket> plan;
    hipError_t e = hipSuccess;
    if (!hx && !ht) {
        const int cap = bucket_of(p->x_dim, p->precision, p->layout);
        plan.push_back(Bucket{cap, choice_for(c, cap, p->precision, p->layout), 0, n, p->x_dim,
                              p->precision});
    } else {
        auto gx = [&](size_t i) { return hx ? hx[i] : p->x_dim; };
        auto gt = [&](size_t i) { return ht ? ht[i] : p->t_dim; };
        std::vector<uint32_t> ord(n);
        std::iota(ord.begin(), ord.end(), 0u);
        std::vector<int> cap(n);
        for (size_t i = 0; i < n; ++i) cap[i] = bucket_of(gx(i), p->precision, p->layout);
        std::sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t z) {
            return std::make_tuple(cap[a], -gt(a), a) < std::make_tuple(cap[z], -gt(z), z);
        });
        for (size_t k = 0; k < n;) {
            size_t k1 = k;
            int widest = 0;
            while (k1 < n && cap[ord[k1]] == cap[ord[k]]) widest = std::max(widest, gx(ord[k1++]));
            const int cp = cap[ord[k]];
            plan.push_back(Bucket{cp, choice_for(c, cp, p->precision, p->layout), k, k1 - k,
                                  widest, p->precision});
            k = k1;
        }
        std::vector<int32_t> fx, ft;
        if (!hx) fx.assign(n, p->x_dim), hx = fx.data();
        if (!ht) ft.assign(n, p->t_dim), ht = ft.data();
        if ((e = hipMemcpyAsync(d_order, ord.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, s)) ||
            (e = hipMemcpyAsync(d_x, hx, n * sizeof(int32_t), hipMemcpyHostToDevice, s)) ||
            (e = hipMemcpyAsync(d_t, ht, n * sizeof(int32_t), hipMemcpyHostToDevice, s)) ||
            (e = hipStreamSynchronize(s)))  // the host arrays must outlive the copies
            return hip_fail(c, e, "kwgpu_fd1d_price_device_noalloc: upload");
        b.order = d_order;
        b.x_dims = d_x;
        b.t_dims = d_t;
    }
    if (plan.size() > (size_t)kNoAllocCounters)
        return fail(c, KWGPU_ERR_INVALID, "kwgpu_fd1d_price_device_noalloc: too many buckets");
    // a fill kernel, not a memset, clears the counters the persistent grids take work from
    if ((e = kwgpu::launch_zero_counters(d_counters, (int)plan.size(), s)))
        return hip_fail(c, e, "kwgpu_fd1d_price_device_noalloc: counter reset");
    for (size_t i = 0; i < plan.size(); ++i) {
        const Bucket& k = plan[i];
        FdBatch bb = b;
        if (bb.order) bb.order += k.begin;
        bb.m = (int)k.count;
        if ((e = kwgpu::launch_fd1d_wave_variant(bb, k.ch.variant, k.prec, s,
                                                 k.ch.persist ? d_counters + i : nullptr)))
            return hip_fail(c, e, "kwgpu_fd1d_price_device_noalloc: launch");
    }
    if ((e = hipEventRecord(c->ev_done, s))) return hip_fail(c, e, "kwgpu: solve event");
    c->done_stream = s;
    return KWGPU_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------------
// Autotuner
// ---------------------------------------------------------------------------------
extern "C" {

int kwgpu_fd1d_autotune(kwgpu_ctx* c, const kwgpu_option* d_opts, size_t n,
                        const kwgpu_fd1d_params* p, double* d_prices, int32_t* d_status,
                        char* report, size_t report_len) {
    if (!c) return KWGPU_ERR_INVALID;
    c->err.clear();
    if (int rc = check_params(c, p)) return rc;
    if (!d_opts || !d_prices || n == 0 || n > 0x7fffffffu)
        return fail(c, KWGPU_ERR_INVALID, "kwgpu_fd1d_autotune: bad argument");
    if (hipError_t e = hipSetDevice(c->device); e != hipSuccess) return hip_fail(c, e, "hipSetDevice");
    hipStream_t s = c->stream;
    std::vector<int32_t> hx(n, p->x_dim), ht(n, p->t_dim);
    hipError_t e;
    if ((p->x_dims && (e = hipMemcpy(hx.data(), p->x_dims, n * sizeof(int32_t), hipMemcpyDeviceToHost))) ||
        (p->t_dims && (e = hipMemcpy(ht.data(), p->t_dims, n * sizeof(int32_t), hipMemcpyDeviceToHost))))
        return hip_fail(c, e, "kwgpu_fd1d_autotune: sizes");
    if (int rc = check_dims(c, hx.data(), ht.data(), n)) return rc;
    const int prec = p->precision;
    std::map<int, std::vector<uint32_t>> buckets;
    for (size_t i = 0; i < n; ++i)
        buckets[bucket_of(hx[i], prec, KWGPU_LAYOUT_AUTO)].push_back((uint32_t)i);

    FdBatch b{};
    b.opts = d_opts;
    b.x_dim = p->x_dim;
    b.t_dim = p->t_dim;
    b.x_dims = p->x_dims;
    b.t_dims = p->t_dims;
    b.density = p->density;
    b.scale = p->scale;
    b.theta = p->theta;
    b.prices = d_prices;
    b.status = d_status;
    DevBuf sample, cnt, ws;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    std::string js = "{\"precision\": \"" + std::string(prec == KWGPU_FP64 ? "fp64" : "fp32-mixed") +
                     "\", \"buckets\": [";
    int rc = KWGPU_OK;
    if ((e = hipEventCreate(&e0)) || (e = hipEventCreate(&e1))) rc = hip_fail(c, e, "events");
    bool first_bucket = true;
    for (auto& [cap, ids] : buckets) {
        if (rc) break;
        if (cap == kLaneCap) continue;  // nothing to choose: only the lane layout holds it
        // a strided sample in longest-first order
        const size_t S = std::min<size_t>(ids.size(), 8192);
        std::vector<uint32_t> smp;
        for (size_t k = 0; k < S; ++k) smp.push_back(ids[k * ids.size() / S]);
        std::sort(smp.begin(), smp.end(), [&](uint32_t a, uint32_t z) { return ht[a] > ht[z]; });
        int bx = 0;
        for (uint32_t i : smp) bx = std::max(bx, hx[i]);
        if (ensure(sample, S * sizeof(uint32_t)) || ensure(cnt, sizeof(int)) ||
            (e = hipMemcpy(sample.p, smp.data(), S * sizeof(uint32_t), hipMemcpyHostToDevice))) {
            rc = fail(c, KWGPU_ERR_NOMEM, "kwgpu_fd1d_autotune: sample upload");
            break;
        }
        FdBatch bb = b;
        bb.order = (const uint32_t*)sample.p;
        bb.m = (int)S;
        struct Cand {
            Choice ch;
            float ms;
            bool pruned = false;  // ms is a lower bound; the candidate was not run
        };
        std::vector<Cand> cands;
        for (int v = 0; v < kwgpu::wave_variant_count(); ++v)
            if (kwgpu::wave_variant(v).capacity() == cap)
                for (bool per : {false, true}) cands.push_back({Choice{v, per}, 0.f});
        if (prec == KWGPU_FP64) cands.push_back({Choice{-1, false}, 0.f});
        // the lane layout streams its state through HBM and runs ~100x slower than the
        // wave variants: it is timed on the first kLaneSample sampled options and scaled
        constexpr size_t kLaneSample = 256;
        FdBatch bl = bb;
        bl.m = (int)std::min(S, kLaneSample);
        // the lane layout's lower bound: one option runs as one thread's serial chain, so
        // the bucket takes at least (the longest option's point-steps) x (the time per
        // point-step of a single option run alone, on the cheapest option of the sample)
        size_t kmin = 0;
        double pts_min = 0.0, pts_max = 0.0;
        for (size_t k = 0; k < S; ++k) {
            const double pts = (double)hx[smp[k]] * (ht[smp[k]] - 1);
            if (k == 0 || pts < pts_min) pts_min = pts, kmin = k;
            pts_max = std::max(pts_max, pts);
        }
        for (auto& cd : cands) {
            if (cd.ch.variant < 0 && pts_min > 0.0) {
                float best_wave = 0.f;
                for (const auto& w : cands)
                    if (w.ch.variant >= 0 && (best_wave == 0.f || w.ms < best_wave)) best_wave = w.ms;
                FdBatch bp = bb;
                bp.m = 1;
                bp.order = (const uint32_t*)sample.p + kmin;
                if (ensure(ws, kwgpu::lane_workspace_doubles(1, bx) * sizeof(double))) {
                    rc = fail(c, KWGPU_ERR_NOMEM, "kwgpu_fd1d_autotune: workspace");
                    break;
                }
                float pilot = 0.f;
                for (int r = 0; r < 2 && !rc; ++r) {  // the first run loads the code object
                    if ((e = hipEventRecord(e0, s))) { rc = hip_fail(c, e, "event"); break; }
                    if ((rc = lane_launch(c, bp, bx, prec, s, (double*)ws.p))) break;
                    if ((e = hipEventRecord(e1, s)) || (e = hipEventSynchronize(e1)) ||
                        (e = hipEventElapsedTime(&pilot, e0, e1)))
                        rc = hip_fail(c, e, "event");
                }
                if (rc) break;
                const float bound = (float)(pilot / pts_min * pts_max);
                if (best_wave > 0.f && bound >= best_wave) {
                    cd.ms = bound;
                    cd.pruned = true;
                    continue;
                }
            }
            auto run = [&]() -> int {
                if (cd.ch.variant < 0) {
                    if (ensure(ws, kwgpu::lane_workspace_doubles(bl.m, bx) * sizeof(double)))
                        return fail(c, KWGPU_ERR_NOMEM, "kwgpu_fd1d_autotune: workspace");
                    return lane_launch(c, bl, bx, prec, s, (double*)ws.p);
                }
                int* q = nullptr;
                if (cd.ch.persist) {
                    q = (int*)cnt.p;
                    if (hipError_t e2 = kwgpu::launch_zero_counters(q, 1, s)) return hip_fail(c, e2, "reset");
                }
                if (hipError_t e2 = kwgpu::launch_fd1d_wave_variant(bb, cd.ch.variant, prec, s, q))
                    return hip_fail(c, e2, "kwgpu_fd1d_autotune: launch");
                return KWGPU_OK;
            };
            if ((rc = run())) break;  // warm-up
            // the median of three timed runs (one for the lane layout, which is slow)
            const int reps = cd.ch.variant < 0 ? 1 : 3;
            float t[3] = {0.f, 0.f, 0.f};
            for (int r = 0; r < reps && !rc; ++r) {
                if ((e = hipEventRecord(e0, s))) { rc = hip_fail(c, e, "event"); break; }
                if ((rc = run())) break;
                if ((e = hipEventRecord(e1, s)) || (e = hipEventSynchronize(e1)) ||
                    (e = hipEventElapsedTime(&t[r], e0, e1)))
                    rc = hip_fail(c, e, "event");
            }
            if (rc) break;
            if (reps == 3) std::sort(t, t + 3);
            cd.ms = reps == 3 ? t[1] : t[0];
            if (cd.ch.variant < 0) cd.ms *= (float)S / (float)bl.m;
        }
        if (rc) break;
        // a later candidate must be faster by 2 % to displace an earlier one: within
        // that margin sample timings are noise, and the order puts the 16-row,
        // static-grid variant first
        const Cand* best = &cands[0];
        for (const auto& cd : cands)
            if (cd.ms < 0.98f * best->ms) best = &cd;
        c->tuned[{cap, prec}] = best->ch;
        char buf[256];
        js += first_bucket ? "" : ", ";
        first_bucket = false;
        std::snprintf(buf, sizeof(buf), "{\"bucket\": %d, \"options\": %zu, \"sampled\": %zu, \"candidates\": [",
                      cap, ids.size(), S);
        js += buf;
        for (size_t k = 0; k < cands.size(); ++k) {
            const auto& cd = cands[k];
            if (cd.ch.variant < 0) {
                std::snprintf(buf, sizeof(buf), "%s{\"layout\": \"lane\", \"ms\": %.4f, \"pruned\": %s}",
                              k ? ", " : "", cd.ms, cd.pruned ? "true" : "false");
            } else {
                const auto wv = kwgpu::wave_variant(cd.ch.variant);
                std::snprintf(buf, sizeof(buf),
                              "%s{\"layout\": \"wave\", \"rows_per_lane\": %d, \"lanes\": %d, \"waves\": %d, "
                              "\"persistent\": %s, \"ms\": %.4f}",
                              k ? ", " : "", wv.M, wv.L * wv.W, wv.W, cd.ch.persist ? "true" : "false", cd.ms);
            }
            js += buf;
        }
        std::snprintf(buf, sizeof(buf), "], \"chosen\": %d}", (int)(best - &cands[0]));
        js += buf;
    }
    js += "]";
    // the bucket schedule: the whole batch with the chosen kernels, buckets concurrent on
    // the side streams and one after another on the caller's stream
    if (rc == KWGPU_OK && buckets.size() > 1) {
        b.m = (int)n;
        const bool graphs = c->use_graphs;
        c->use_graphs = false;
        float ms[2] = {0.f, 0.f};
        for (int serial = 0; serial < 2 && rc == KWGPU_OK; ++serial) {
            c->serial[prec] = serial != 0;
            if ((rc = run_fd(c, b, hx.data(), ht.data(), prec, KWGPU_LAYOUT_AUTO, s))) break;
            if ((e = hipEventRecord(e0, s))) { rc = hip_fail(c, e, "event"); break; }
            for (int r = 0; r < 2 && rc == KWGPU_OK; ++r)
                rc = run_fd(c, b, hx.data(), ht.data(), prec, KWGPU_LAYOUT_AUTO, s);
            if (rc) break;
            if ((e = hipEventRecord(e1, s)) || (e = hipEventSynchronize(e1)) ||
                (e = hipEventElapsedTime(&ms[serial], e0, e1))) {
                rc = hip_fail(c, e, "event");
                break;
            }
            ms[serial] /= 2;
        }
        c->use_graphs = graphs;
        if (rc == KWGPU_OK) {
            c->serial[prec] = ms[1] < ms[0];
            char buf[160];
            std::snprintf(buf, sizeof(buf),
                          ", \"schedule\": {\"concurrent_ms\": %.4f, \"serial_ms\": %.4f, "
                          "\"chosen\": \"%s\"}",
                          ms[0], ms[1], ms[1] < ms[0] ? "serial" : "concurrent");
            js += buf;
        }
    }
    js += "}";
    (void)hipStreamSynchronize(s);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    free_buf(sample);
    free_buf(cnt);
    free_buf(ws);
    drop_graph(c);  // choices changed
    if (report && report_len) std::snprintf(report, report_len, "%s", js.c_str());
    return rc;
}

int kwgpu_tune_save(kwgpu_ctx* c, const char* path) {
    if (!c || !path) return KWGPU_ERR_INVALID;
    FILE* f = std::fopen(path, "w");
    if (!f) return fail(c, KWGPU_ERR_INVALID, std::string("cannot write ") + path);
    for (const auto& [k, ch] : c->tuned)
        std::fprintf(f, "%d %d %d %d\n", k.first, k.second, ch.variant, ch.persist ? 1 : 0);
    for (const auto& [prec, serial] : c->serial)  // bucket -1: the bucket schedule
        std::fprintf(f, "%d %d %d %d\n", -1, prec, serial ? 1 : 0, 0);
    std::fclose(f);
    return KWGPU_OK;
}

int kwgpu_tune_load(kwgpu_ctx* c, const char* path) {
    if (!c || !path) return KWGPU_ERR_INVALID;
    FILE* f = std::fopen(path, "r");
    if (!f) return fail(c, KWGPU_ERR_INVALID, std::string("cannot read ") + path);
    int cap, prec, v, per;
    while (std::fscanf(f, "%d %d %d %d", &cap, &prec, &v, &per) == 4) {
        // lines of another precision, and lane-layout choices for FP32-mixed (which runs
        // in the wave layout only), are skipped: they could only make later calls fail
        if (prec != KWGPU_FP64 && prec != KWGPU_FP32_MIXED) continue;
        if (cap == -1) {
            c->serial[prec] = v != 0;
            continue;
        }
        if (v >= kwgpu::wave_variant_count() || (v >= 0 && kwgpu::wave_variant(v).capacity() != cap)) continue;
        if (v < 0 && prec != KWGPU_FP64) continue;
        c->tuned[{cap, prec}] = Choice{v, per != 0};
    }
    std::fclose(f);
    drop_graph(c);
    return KWGPU_OK;
}

}  // extern "C"
