// Host-side infrastructure: thread-local error messages, per-device scratch pool and
// the small utility entry points of the C ABI.
#include "nla_common.hpp"

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
Pool g_pool[MAX_DEV];
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
    Pool &p = g_pool[dev];
    if (bytes <= p.size[slot] && p.ptr[slot]) return p.ptr[slot];
    size_t want = bytes < (1u << 20) ? (1u << 20) : bytes;
    want = (want + 255) & ~size_t(255);
    if (p.ptr[slot]) {
        e = hipStreamSynchronize(stream);
        if (e == hipSuccess) e = hipDeviceSynchronize();
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

}  // namespace nla

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
    for (int i = 0; i < nla::NUM_SLOTS; ++i) s += int64_t(nla::g_pool[dev].size[i]);
    return s;
}
