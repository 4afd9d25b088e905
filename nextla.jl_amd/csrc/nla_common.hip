// Host-side infrastructure: thread-local error messages, per-device scratch pool and
// the small utility entry points of the C ABI.
#include "nla_common.hpp"

#include <map>
#include <utility>
#include <vector>

namespace nla {

static thread_local std::string g_error;

void set_error(const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_error = buf;
}

int arg_error(int pos, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_error = buf;
    return -pos;
}

int wait_timeout_error(const char *what) {
    set_error("%s: a dependency wait exceeded its time limit (results are invalid)", what);
    return STATUS_WAIT_TIMEOUT;
}

int hip_error(hipError_t e, const char *what) {
    set_error("HIP error %d (%s) in %s", int(e), hipGetErrorString(e), what);
    return int(e) > 0 ? int(e) : 1;
}

namespace {
constexpr int MAX_DEV = 64;
struct Pool {
    void *ptr[NUM_SLOTS] = {};
    size_t size[NUM_SLOTS] = {};
};
std::mutex g_pool_mu;
// One pool per (device, stream): calls enqueued on different streams never share scratch, so
// they may run concurrently; calls on one stream are ordered and reuse the stream's buffers.
std::map<std::pair<int, hipStream_t>, Pool> g_pool;
}  // namespace

void *scratch(ScratchSlot slot, size_t bytes, hipStream_t stream, int *status) {
    *status = 0;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= MAX_DEV) {
        *status = hip_error(e == hipSuccess ? hipErrorInvalidDevice : e, "hipGetDevice");
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_pool_mu);
    Pool &p = g_pool[std::make_pair(dev, stream)];
    if (bytes <= p.size[slot] && p.ptr[slot]) return p.ptr[slot];
    size_t want = bytes < (1u << 20) ? (1u << 20) : bytes;
    want = (want + 255) & ~size_t(255);
    if (p.ptr[slot]) {
        // only this stream's work can use the old buffer: drain it, then replace the buffer
        e = hipStreamSynchronize(stream);
        if (e != hipSuccess) { *status = hip_error(e, "scratch sync"); return nullptr; }
        (void)hipFree(p.ptr[slot]);
        p.ptr[slot] = nullptr;
        p.size[slot] = 0;
    }
    e = hipMalloc(&p.ptr[slot], want);
    if (e != hipSuccess) {
        p.ptr[slot] = nullptr;
        *status = hip_error(e, "hipMalloc(scratch)");
        return nullptr;
    }
    p.size[slot] = want;
    return p.ptr[slot];
}

// ---------------------------------------------------------------------------------
namespace {
std::mutex g_prof_mu;
bool g_prof_enabled = false;
struct ProfRec { int kind; hipEvent_t a, b; };
std::vector<ProfRec> g_prof;
std::vector<hipEvent_t> g_prof_free;
const char *g_prof_names[PK_COUNT] = {"gemm", "gemm128", "trsm_base", "trmm_base", "ge_apply", "ts_apply",
                                      "panel_qr", "ts_panel", "tile_qr", "other"};
}  // namespace

// Calls that must not allocate (the no-allocation entry point) switch the profiler off on
// their thread: creating its events could allocate.
static thread_local int g_prof_suppress = 0;
void prof_suppress(bool on) { g_prof_suppress += on ? 1 : -1; }
bool prof_on() { return g_prof_enabled && g_prof_suppress == 0; }

hipEvent_t prof_event() {
    std::lock_guard<std::mutex> lock(g_prof_mu);
    if (!g_prof_free.empty()) { hipEvent_t e = g_prof_free.back(); g_prof_free.pop_back(); return e; }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

void prof_push(int kind, hipEvent_t a, hipEvent_t b) {
    std::lock_guard<std::mutex> lock(g_prof_mu);
    g_prof.push_back({kind, a, b});
}

}  // namespace nla

NLA_API int nla_prof_enable(int on) {
    nla::g_prof_enabled = on != 0;
    return 0;
}

NLA_API void nla_prof_reset(void) {
    std::lock_guard<std::mutex> lock(nla::g_prof_mu);
    for (auto &r : nla::g_prof) {
        (void)hipEventSynchronize(r.b);
        nla::g_prof_free.push_back(r.a);
        nla::g_prof_free.push_back(r.b);
    }
    nla::g_prof.clear();
}

NLA_API int nla_prof_read(const char *kernel, double *total_ms, int64_t *launches) {
    std::lock_guard<std::mutex> lock(nla::g_prof_mu);
    double ms = 0;
    int64_t n = 0;
    for (auto &r : nla::g_prof) {
        if (strcmp(nla::g_prof_names[r.kind], kernel) != 0) continue;
        hipError_t e = hipEventSynchronize(r.b);
        if (e != hipSuccess) return nla::hip_error(e, "hipEventSynchronize(prof)");
        float t = 0;
        e = hipEventElapsedTime(&t, r.a, r.b);
        if (e != hipSuccess) return nla::hip_error(e, "hipEventElapsedTime(prof)");
        ms += t;
        ++n;
    }
    *total_ms = ms;
    *launches = n;
    return 0;
}

NLA_API const char *nla_last_error(void) { return nla::g_error.c_str(); }

NLA_API int nla_version(void) { return 100; }

NLA_API int nla_synchronize(void *stream) {
    hipError_t e = hipStreamSynchronize(nla::as_stream(stream));
    if (e != hipSuccess) return nla::hip_error(e, "hipStreamSynchronize");
    return 0;
}

NLA_API int64_t nla_scratch_bytes(void) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= nla::MAX_DEV) return 0;
    std::lock_guard<std::mutex> lock(nla::g_pool_mu);
    int64_t s = 0;
    for (auto &kv : nla::g_pool)
        if (kv.first.first == dev)
            for (int i = 0; i < nla::NUM_SLOTS; ++i) s += int64_t(kv.second.size[i]);
    return s;
}
