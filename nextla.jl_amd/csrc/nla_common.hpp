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
enum ScratchSlot { SLOT_W = 0, SLOT_AUX = 1, SLOT_SCHED = 2, SLOT_MIXED = 3, SLOT_TRSM = 4, SLOT_RHS = 5, SLOT_LS = 6, SLOT_LPIMG = 7, SLOT_SHADOW = 8, SLOT_LU = 9, SLOT_HOST = 10, NUM_SLOTS = 11 };
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

// ---------------------------------------------------------------------------------
// Complex element type of the QR-family routines (ComplexF64 / ComplexF32: two reals, real
// part first, the layout Julia, numpy and torch use).  The real types use the same helper
// names, so one kernel body serves both: conj_ is the identity on reals.
template <typename R> struct alignas(2 * sizeof(R)) cplx {
    R re, im;
    cplx() = default;
    __host__ __device__ constexpr cplx(R r) : re(r), im(0) {}
    __host__ __device__ constexpr cplx(R r, R i) : re(r), im(i) {}
    __host__ __device__ cplx &operator+=(cplx b) { re += b.re; im += b.im; return *this; }
    __host__ __device__ cplx &operator-=(cplx b) { re -= b.re; im -= b.im; return *this; }
    __host__ __device__ cplx &operator*=(cplx b) {
        const R r = re * b.re - im * b.im;
        im = re * b.im + im * b.re;
        re = r;
        return *this;
    }
    __host__ __device__ cplx &operator/=(cplx b);
};
template <typename R> __host__ __device__ inline cplx<R> operator+(cplx<R> a, cplx<R> b) { return {a.re + b.re, a.im + b.im}; }
template <typename R> __host__ __device__ inline cplx<R> operator-(cplx<R> a, cplx<R> b) { return {a.re - b.re, a.im - b.im}; }
template <typename R> __host__ __device__ inline cplx<R> operator-(cplx<R> a) { return {-a.re, -a.im}; }
template <typename R> __host__ __device__ inline cplx<R> operator*(cplx<R> a, cplx<R> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename R> __host__ __device__ inline cplx<R> operator*(R a, cplx<R> b) { return {a * b.re, a * b.im}; }
template <typename R> __host__ __device__ inline cplx<R> operator*(cplx<R> a, R b) { return {a.re * b, a.im * b}; }
template <typename R> __host__ __device__ inline cplx<R> operator/(cplx<R> a, cplx<R> b) {
    const R d = b.re * b.re + b.im * b.im;
    return {(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}
template <typename R> __host__ __device__ inline cplx<R> operator/(cplx<R> a, R b) { return {a.re / b, a.im / b}; }
template <typename R> __host__ __device__ inline cplx<R> &cplx<R>::operator/=(cplx<R> b) { return *this = *this / b; }
template <typename R> __host__ __device__ inline bool operator==(cplx<R> a, cplx<R> b) { return a.re == b.re && a.im == b.im; }
template <typename R> __host__ __device__ inline bool operator!=(cplx<R> a, cplx<R> b) { return !(a == b); }

template <typename T> struct RealOf { typedef T t; };
template <typename R> struct RealOf<cplx<R>> { typedef R t; };
template <typename T> struct IsCplx { static constexpr bool v = false; };
template <typename R> struct IsCplx<cplx<R>> { static constexpr bool v = true; };

__host__ __device__ inline double conj_(double x) { return x; }
__host__ __device__ inline float conj_(float x) { return x; }
template <typename R> __host__ __device__ inline cplx<R> conj_(cplx<R> x) { return {x.re, -x.im}; }
__host__ __device__ inline double re_(double x) { return x; }
__host__ __device__ inline float re_(float x) { return x; }
template <typename R> __host__ __device__ inline R re_(cplx<R> x) { return x.re; }
__host__ __device__ inline double im_(double) { return 0; }
__host__ __device__ inline float im_(float) { return 0; }
template <typename R> __host__ __device__ inline R im_(cplx<R> x) { return x.im; }
__host__ __device__ inline double abs2_(double x) { return x * x; }
__host__ __device__ inline float abs2_(float x) { return x * x; }
template <typename R> __host__ __device__ inline R abs2_(cplx<R> x) { return x.re * x.re + x.im * x.im; }
// largest |component|: a scale factor for overflow-safe norms
__host__ __device__ inline double absmax_(double x) { return x < 0 ? -x : x; }
__host__ __device__ inline float absmax_(float x) { return x < 0 ? -x : x; }
template <typename R> __host__ __device__ inline R absmax_(cplx<R> x) {
    const R a = x.re < 0 ? -x.re : x.re, b = x.im < 0 ? -x.im : x.im;
    return a > b ? a : b;
}
// element built from real and imaginary parts (the imaginary part is dropped for reals)
template <typename T> __host__ __device__ inline T mkc_(typename RealOf<T>::t re, typename RealOf<T>::t im);
template <> __host__ __device__ inline double mkc_<double>(double re, double) { return re; }
template <> __host__ __device__ inline float mkc_<float>(float re, float) { return re; }
template <> __host__ __device__ inline cplx<double> mkc_<cplx<double>>(double re, double im) { return {re, im}; }
template <> __host__ __device__ inline cplx<float> mkc_<cplx<float>>(float re, float im) { return {re, im}; }

template <typename T> struct Tname;
template <> struct Tname<double> { static constexpr const char *s = "Float64"; };
template <> struct Tname<float> { static constexpr const char *s = "Float32"; };
template <> struct Tname<cplx<double>> { static constexpr const char *s = "ComplexF64"; };
template <> struct Tname<cplx<float>> { static constexpr const char *s = "ComplexF32"; };

// ---------------------------------------------------------------------------------
// Device helpers.
constexpr int WAVE = 64;

template <typename T> __device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    return v;
}
template <typename R> __device__ __forceinline__ cplx<R> wave_sum(cplx<R> v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        v.re += __shfl_xor(v.re, off, WAVE);
        v.im += __shfl_xor(v.im, off, WAVE);
    }
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
    T s = T(0);
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
