// Shared host/device infrastructure of libnextla_hip: status handling, the per-device
// scratch pool, wavefront reductions and the CDNA4 MFMA wrappers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#define NLA_API extern "C" __attribute__((visibility("default")))

namespace nla {

using i64 = int64_t;

// ---------------------------------------------------------------------------------
// Status handling: 0 ok, -i illegal argument i (message kept per thread), >0 HIP error,
// STATUS_WAIT_TIMEOUT a device-side dependency wait of a persistent launch ran out of time.
constexpr int STATUS_WAIT_TIMEOUT = 10001;
void set_error(const char *fmt, ...);
int wait_timeout_error(const char *what);
int arg_error(int pos, const char *fmt, ...);
int hip_error(hipError_t e, const char *what);

#define NLA_HIP(expr)                                                   \
    do {                                                                \
        hipError_t nla_e_ = (expr);                                     \
        if (nla_e_ != hipSuccess) return ::nla::hip_error(nla_e_, #expr); \
    } while (0)

#define NLA_LAUNCH_CHECK() NLA_HIP(hipGetLastError())

// Propagate a library status (0 ok, -i argument error, > 0 HIP or library error) unchanged.
#define NLA_ST(expr)                                \
    do {                                            \
        int nla_st_ = (expr);                       \
        if (nla_st_ != 0) return nla_st_;           \
    } while (0)

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// Scratch pool per (device, stream), slots keyed by purpose.  Grows (never shrinks) under a
// mutex; growth synchronises the stream first so its in-flight kernels never see the old
// buffer freed.  Kernels of one call share the slot they asked for; calls on different
// streams use different buffers and may overlap.
enum ScratchSlot { SLOT_W = 0, SLOT_AUX = 1, SLOT_SCHED = 2, SLOT_MIXED = 3, SLOT_TRSM = 4, SLOT_RHS = 5, SLOT_LS = 6, SLOT_LPIMG = 7, SLOT_SHADOW = 8, NUM_SLOTS = 9 };
void *scratch(ScratchSlot slot, size_t bytes, hipStream_t stream, int *status);

// Launch-time profiler: when enabled, every kernel family brackets its launches with a
// pair of hipEvents on the launch stream; nla_prof_read() sums the elapsed times.
bool prof_on();
void prof_suppress(bool on);
void prof_push(int kind, hipEvent_t a, hipEvent_t b);
hipEvent_t prof_event();
enum ProfKind { PK_GEMM = 0, PK_GEMM128, PK_TRSM_BASE, PK_TRMM_BASE, PK_GE_APPLY, PK_TS_APPLY, PK_PANEL, PK_TS_PANEL,
                PK_TILE_QR, PK_OTHER, PK_COUNT };
struct ProfScope {
    int kind;
    hipStream_t s;
    hipEvent_t a = nullptr;
    ProfScope(int k, hipStream_t st) : kind(k), s(st) {
        if (prof_on()) { a = prof_event(); if (a) (void)hipEventRecord(a, s); }
    }
    ~ProfScope() {
        if (a) {
            hipEvent_t b = prof_event();
            if (b) { (void)hipEventRecord(b, s); prof_push(kind, a, b); }
        }
    }
};

inline bool is_trans(char t) { return t == 'T' || t == 'C' || t == 't' || t == 'c'; }
inline char up(char c) { return (c >= 'a' && c <= 'z') ? char(c - 32) : c; }

template <typename T> struct Tname;
template <> struct Tname<double> { static constexpr const char *s = "Float64"; };
template <> struct Tname<float> { static constexpr const char *s = "Float32"; };

// ---------------------------------------------------------------------------------
// Device helpers.
constexpr int WAVE = 64;

template <typename T> __device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}

template <typename T> __device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        T o = __shfl_xor(v, off, WAVE);
        v = v > o ? v : o;
    }
    return v;
}

// Block-wide sum for blockDim.x a multiple of 64 (<= 1024).  red must hold 16 entries.
template <typename T> __device__ __forceinline__ T block_sum(T v, T *red) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    v = wave_sum(v);
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    T s = 0;
    for (int i = 0; i < nw; ++i) s += red[i];
    return s;
}

// MFMA 16x16x4 (FP64 / FP32).  Fragment layouts (one 64-lane wave), measured on gfx950:
//   A (16 x 4):  lane l holds A[l % 16][l / 16]
//   B (4 x 16):  lane l holds B[l / 16][l % 16]
//   D (16 x 16): lane l, element r holds D[acc_row(l / 16, r)][l % 16] with
//                acc_row = l/16 + 4 r      for v_mfma_f64_16x16x4f64
//                acc_row = 4 (l/16) + r    for v_mfma_f32_16x16x4f32
typedef double dx4 __attribute__((ext_vector_type(4)));
typedef float fx4 __attribute__((ext_vector_type(4)));
template <typename T> struct Acc;
template <> struct Acc<double> { typedef dx4 t; };
template <> struct Acc<float> { typedef fx4 t; };

__device__ __forceinline__ void mfma16(dx4 &c, double a, double b) {
    c = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ void mfma16(fx4 &c, float a, float b) {
    c = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

template <typename T> __device__ __forceinline__ int acc_row(int lk, int r);
template <> __device__ __forceinline__ int acc_row<double>(int lk, int r) { return lk + 4 * r; }
template <> __device__ __forceinline__ int acc_row<float>(int lk, int r) { return 4 * lk + r; }

template <typename T> __device__ __forceinline__ typename Acc<T>::t acc_zero() {
    typename Acc<T>::t z;
    z[0] = 0; z[1] = 0; z[2] = 0; z[3] = 0;
    return z;
}

}  // namespace nla
