// A stand-in for the HIP runtime, for checking the library's host code on a machine
// without a GPU: device and pinned allocations come from calloc, copies are memmove,
// kernels are not run (their launch shapes are checked), and streams, events, graphs and
// graph executables are heap tokens that abort on use after destroy, double destroy or
// an unknown handle. Built with AddressSanitizer, this turns host-side memory errors of
// the library (its uneven-batch plans, grouping, staging, serving buffers) into test
// failures.
//
// Work queued on a stream is modelled as still running until something the host waits
// on covers it: hipStreamSynchronize of that stream (or of a stream that waited on an
// event recorded after it), hipEventSynchronize of an event recorded after it, or
// hipDeviceSynchronize. Destroying a graph executable whose launch is still running
// aborts, as does a hipStreamWaitEvent on an event never recorded, and so does a launch
// of a graph executable on one stream while its launch on another may still run and the
// first stream's work was not waited for. Test infrastructure only.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <set>

static std::set<void*> g_dev, g_host, g_graph, g_exec, g_stream, g_event;
static bool g_capturing = false;
static long g_launches = 0, g_graph_launches = 0;
// graph executables whose launches may still run: per stream, and as captured by each
// recorded event
static std::map<void*, std::set<void*>> g_pending;
static std::map<void*, std::set<void*>> g_event_work;
static std::set<void*> g_recorded;

static void complete(const std::set<void*>& done) {
    for (auto& [s, work] : g_pending)
        for (void* x : done) work.erase(x);
    for (auto& [e, work] : g_event_work)
        for (void* x : done) work.erase(x);
}

static void* token(std::set<void*>& s) {
    void* p = std::malloc(16);
    s.insert(p);
    return p;
}
static void untoken(std::set<void*>& s, void* p, const char* what) {
    if (!p) return;
    if (!s.count(p)) {
        std::fprintf(stderr, "MOCK: %s of unknown handle %p\n", what, p);
        std::abort();
    }
    s.erase(p);
    std::free(p);
}

extern "C" {
void** __hipRegisterFatBinary(const void*) {
    static void* m = nullptr;
    return &m;
}
void __hipRegisterFunction(void**, const void*, char*, const char*, unsigned int, uint3*, uint3*,
                           dim3*, dim3*, int*) {}
void __hipUnregisterFatBinary(void**) {}
void __hipRegisterVar(void**, void*, char*, const char*, int, size_t, int, int) {}
// a __device__ variable's host shadow (full size in a host-only build) stands in for it
hipError_t hipGetSymbolAddress(void** p, const void* sym) {
    *p = const_cast<void*>(sym);
    return hipSuccess;
}
static dim3 g_grid, g_block;
static size_t g_shm;
static hipStream_t g_cs;
hipError_t __hipPushCallConfiguration(dim3 g, dim3 b, size_t s, hipStream_t st) {
    g_grid = g, g_block = b, g_shm = s, g_cs = st;
    return hipSuccess;
}
hipError_t __hipPopCallConfiguration(dim3* g, dim3* b, size_t* s, hipStream_t* st) {
    *g = g_grid, *b = g_block, *s = g_shm, *st = g_cs;
    return hipSuccess;
}
hipError_t hipLaunchKernel(const void*, dim3 g, dim3 b, void**, size_t shm, hipStream_t) {
    const unsigned long long threads = (unsigned long long)b.x * b.y * b.z;
    if (g.x == 0 || g.y == 0 || g.z == 0 || threads == 0 || threads > 1024 || g.y > 65535 ||
        g.z > 65535 || shm > 160 * 1024) {
        std::fprintf(stderr, "MOCK: bad launch shape grid %u %u %u block %u %u %u shm %zu\n", g.x,
                     g.y, g.z, b.x, b.y, b.z, shm);
        std::abort();
    }
    ++g_launches;
    return hipSuccess;
}
hipError_t hipDeviceSynchronize() {
    g_pending.clear();
    g_event_work.clear();
    return hipSuccess;
}
hipError_t hipGetLastError() { return hipSuccess; }
const char* hipGetErrorString(hipError_t) { return "mock error"; }
hipError_t hipGetDeviceCount(int* n) {
    *n = 1;
    return hipSuccess;
}
hipError_t hipGetDevice(int* d) {
    *d = 0;
    return hipSuccess;
}
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t a, int) {
    *v = a == hipDeviceAttributeMultiprocessorCount ? 256 : 64;
    return hipSuccess;
}
hipError_t hipMalloc(void** p, size_t n) {
    *p = std::calloc(1, n ? n : 1);
    g_dev.insert(*p);
    return hipSuccess;
}
hipError_t hipFree(void* p) {
    if (!p) return hipSuccess;
    if (!g_dev.count(p)) {
        std::fprintf(stderr, "MOCK: hipFree of unknown %p\n", p);
        std::abort();
    }
    g_dev.erase(p);
    std::free(p);
    return hipSuccess;
}
hipError_t hipHostMalloc(void** p, size_t n, unsigned int) {
    *p = std::calloc(1, n ? n : 1);
    g_host.insert(*p);
    return hipSuccess;
}
hipError_t hipHostFree(void* p) {
    if (!p) return hipSuccess;
    if (!g_host.count(p)) {
        std::fprintf(stderr, "MOCK: hipHostFree of unknown %p\n", p);
        std::abort();
    }
    g_host.erase(p);
    std::free(p);
    return hipSuccess;
}
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) {
    std::memmove(d, s, n);
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t) {
    if (g_capturing) {
        std::fprintf(stderr, "MOCK: memcpy during capture\n");
        std::abort();
    }
    std::memmove(d, s, n);
    return hipSuccess;
}
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned int) {
    *s = (hipStream_t)token(g_stream);
    return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) {
    untoken(g_stream, s, "hipStreamDestroy");
    g_pending.erase(s);
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t s) {
    if (g_capturing) {
        std::fprintf(stderr, "MOCK: sync during capture\n");
        std::abort();
    }
    const std::set<void*> done = g_pending[s];
    complete(done);
    return hipSuccess;
}
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned int) {
    if (e && !g_event.count(e)) {
        std::fprintf(stderr, "MOCK: wait on unknown event\n");
        std::abort();
    }
    if (e && !g_recorded.count(e)) {
        std::fprintf(stderr, "MOCK: wait on an event never recorded\n");
        std::abort();
    }
    // s's later work runs after the event's, so waiting for s covers it too
    if (e) g_pending[s].insert(g_event_work[e].begin(), g_event_work[e].end());
    return hipSuccess;
}
hipError_t hipEventCreate(hipEvent_t* e) {
    *e = (hipEvent_t)token(g_event);
    return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e) {
    untoken(g_event, e, "hipEventDestroy");
    g_event_work.erase(e);
    g_recorded.erase(e);
    return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s) {
    if (!g_event.count(e)) {
        std::fprintf(stderr, "MOCK: record of unknown event\n");
        std::abort();
    }
    g_recorded.insert(e);
    g_event_work[e] = g_pending[s];
    return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t e) {
    const std::set<void*> done = g_event_work[e];
    complete(done);
    return hipSuccess;
}
hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) {
    *ms = 1.0f;
    return hipSuccess;
}
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) {
    if (g_capturing) {
        std::fprintf(stderr, "MOCK: nested capture\n");
        std::abort();
    }
    g_capturing = true;
    return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g) {
    g_capturing = false;
    *g = (hipGraph_t)token(g_graph);
    return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) {
    untoken(g_graph, g, "hipGraphDestroy");
    return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* x, hipGraph_t g, hipGraphNode_t*, char*, size_t) {
    if (!g_graph.count(g)) {
        std::fprintf(stderr, "MOCK: instantiate of unknown graph\n");
        std::abort();
    }
    *x = (hipGraphExec_t)token(g_exec);
    return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t x) {
    for (const auto& [s, work] : g_pending)
        if (work.count(x)) {
            std::fprintf(stderr, "MOCK: hipGraphExecDestroy of %p while a launch of it may "
                                 "still run on stream %p\n", (void*)x, s);
            std::abort();
        }
    untoken(g_exec, x, "hipGraphExecDestroy");
    return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t x, hipStream_t s) {
    if (!g_exec.count(x)) {
        std::fprintf(stderr, "MOCK: launch of unknown graph exec %p\n", x);
        std::abort();
    }
    for (const auto& [other, work] : g_pending)
        if (other != (void*)s && work.count(x) && !g_pending[s].count(x)) {
            std::fprintf(stderr, "MOCK: graph exec %p launched on stream %p while its launch on "
                                 "stream %p may still run, with no wait between them\n",
                         (void*)x, (void*)s, other);
            std::abort();
        }
    ++g_graph_launches;
    g_pending[s].insert(x);
    return hipSuccess;
}
}

// Number of allocations, streams, events, graphs and executables still alive.
extern "C" int mock_live() {
    const size_t live = g_dev.size() + g_host.size() + g_graph.size() + g_exec.size() +
                        g_stream.size() + g_event.size();
    std::fprintf(stderr, "mock runtime: %ld kernel launches, %ld graph launches, %zu left alive\n",
                 g_launches, g_graph_launches, live);
    return (int)live;
}
