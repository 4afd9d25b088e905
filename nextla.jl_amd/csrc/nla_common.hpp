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

// Global-memory view of a generic pointer.  Pointers read from memory (work lists, problem
// descriptors) are generic, and generic loads are flat loads, which also count against the LDS
// wait counter: every wait for LDS data would then wait for all outstanding memory loads too.
template <typename T> using GPtr = __attribute__((address_space(1))) T *;
template <typename T> __device__ __forceinline__ GPtr<T> to_global(T *p) { return (GPtr<T>)p; }

// Workgroup barrier ordering LDS accesses only: global stores issued before it may still be in
// flight (a __syncthreads() waits for them), so use it only where no thread reads memory that
// another thread wrote to global memory since the last full barrier.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Cross-lane exchanges of gfx950 without the LDS crossbar.  Lane layouts (checked on the
// hardware): v_permlane32_swap(x, y) leaves x = [x(0:31), y(0:31)], y = [x(32:63), y(32:63)];
// v_permlane16_swap does the same per 32-lane half with 16-lane rows; DPP row_ror:8 gives lane
// i the value of lane i ^ 8 of its row, row_shl:4 / row_shr:4 that of lane i + 4 / i - 4,
// quad_perm [2,3,0,1] / [1,0,3,2] that of lane i ^ 2 / i ^ 1.
__device__ __forceinline__ void pl32_swap(float &x, float &y) {
    auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(y), false, false);
    x = __uint_as_float(r[0]);
    y = __uint_as_float(r[1]);
}
__device__ __forceinline__ void pl16_swap(float &x, float &y) {
    auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(y), false, false);
    x = __uint_as_float(r[0]);
    y = __uint_as_float(r[1]);
}
__device__ __forceinline__ void pl32_swap(double &x, double &y) {
    const unsigned long long xb = (unsigned long long)__double_as_longlong(x), yb = (unsigned long long)__double_as_longlong(y);
    auto lo = __builtin_amdgcn_permlane32_swap((unsigned)xb, (unsigned)yb, false, false);
    auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(xb >> 32), (unsigned)(yb >> 32), false, false);
    x = __longlong_as_double((long long)(((unsigned long long)hi[0] << 32) | lo[0]));
    y = __longlong_as_double((long long)(((unsigned long long)hi[1] << 32) | lo[1]));
}
__device__ __forceinline__ void pl16_swap(double &x, double &y) {
    const unsigned long long xb = (unsigned long long)__double_as_longlong(x), yb = (unsigned long long)__double_as_longlong(y);
    auto lo = __builtin_amdgcn_permlane16_swap((unsigned)xb, (unsigned)yb, false, false);
    auto hi = __builtin_amdgcn_permlane16_swap((unsigned)(xb >> 32), (unsigned)(yb >> 32), false, false);
    x = __longlong_as_double((long long)(((unsigned long long)hi[0] << 32) | lo[0]));
    y = __longlong_as_double((long long)(((unsigned long long)hi[1] << 32) | lo[1]));
}
template <int CTRL> __device__ __forceinline__ float dpp_mov(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL> __device__ __forceinline__ double dpp_mov(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
constexpr int DPP_ROW_SHL4 = 0x104, DPP_ROW_SHR4 = 0x114, DPP_ROW_ROR4 = 0x124, DPP_ROW_ROR8 = 0x128;
constexpr int DPP_QUAD_X2 = 0x4E, DPP_QUAD_X1 = 0xB1;

// (x, y) -> x summed over lanes {l, l ^ 32} in lanes < 32, y summed likewise in lanes >= 32
template <typename T> __device__ __forceinline__ T fold32(T x, T y) {
    pl32_swap(x, y);
    return x + y;
}
// (x, y) -> x summed over lanes {l, l ^ 16} where bit 4 of l is clear, y where it is set
template <typename T> __device__ __forceinline__ T fold16(T x, T y) {
    pl16_swap(x, y);
    return x + y;
}
// sum over the wave; lane 0 holds it (the other lanes hold sums of the same terms)
template <typename T> __device__ __forceinline__ T wave_sum_fast(T s) {
    s = fold32(s, s);
    s = fold16(s, s);
    s += dpp_mov<DPP_ROW_ROR8>(s);
    s += dpp_mov<DPP_ROW_ROR4>(s);
    s += dpp_mov<DPP_QUAD_X2>(s);
    s += dpp_mov<DPP_QUAD_X1>(s);
    return s;
}

}  // namespace nla
