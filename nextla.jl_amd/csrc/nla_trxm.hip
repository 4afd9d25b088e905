// Recursive TRSM/TRMM (unified_rectrxm!, src/rectrxm.jl:43-198) with its base-case
// kernels (src/trsm.jl:5-150, src/trmm.jl:43-389) and MFMA GEMM updates.
//
// Base kernels are written once, in "left" form: op'(A) X = X (solve) or X := op'(A) X
// (multiply) on an n x m right-hand side X.  A right-side call X op(A) = B is the left
// form on X^T with op(A)^T (B read through a transposed view), and an upper triangle is a
// lower one in reversed index space, so one kernel body serves all eight variants.
//
// The solve kernel keeps a chunk of CW right-hand sides in LDS, divided by the diagonal
// on load as the reference does (x = b/d - sum (a/d) x, src/trsm.jl:17-25).  It walks the
// triangle in 64-row diagonal blocks: one wavefront per right-hand side solves the block
// by register shuffles (no workgroup barrier per row, unlike the reference's n barriers),
// then the whole workgroup applies the block to the rows below.
#include "nla_common.hpp"
#include "nla_gemm.hpp"

namespace nla {

template <typename T, bool UP, bool TRA> struct TriView {
    const T *a;
    i64 lda;
    int n;
    __device__ __forceinline__ T operator()(int i, int j) const {
        const int oi = UP ? n - 1 - i : i, oj = UP ? n - 1 - j : j;
        return TRA ? a[oj + (i64)oi * lda] : a[oi + (i64)oj * lda];
    }
};

template <typename T, bool UP, bool TRB> struct RhsView {
    T *b;
    i64 ldb;
    int n;
    __device__ __forceinline__ T &operator()(int r, int c) const {
        const int orow = UP ? n - 1 - r : r;
        return TRB ? b[c + (i64)orow * ldb] : b[orow + (i64)c * ldb];
    }
};

constexpr int TRSM_BS = 64;

template <typename T, bool UP, bool TRA, bool TRB>
__global__ __launch_bounds__(256) void trsm_base_kernel(int n, int m, int cwmax, const T *__restrict__ A, i64 lda,
                                                        T *__restrict__ B, i64 ldb) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *X = reinterpret_cast<T *>(smem_raw);             // [cw][n]
    T *D = X + (size_t)cwmax * n;                        // [64][65] scaled diagonal block
    const TriView<T, UP, TRA> a{A, lda, n};
    const int c0 = blockIdx.x * cwmax;
    const int cw = min(cwmax, m - c0);
    const RhsView<T, UP, TRB> bv{B, ldb, n};
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;

    for (int idx = t; idx < n * cw; idx += 256) {
        const int r = idx % n, c = idx / n;
        X[c * n + r] = bv(r, c0 + c) / a(r, r);
    }
    for (int b0 = 0; b0 < n; b0 += TRSM_BS) {
        const int bs = min(TRSM_BS, n - b0);
        __syncthreads();
        for (int idx = t; idx < bs * bs; idx += 256) {
            const int i = idx % bs, j = idx / bs;
            if (j < i) D[j * 65 + i] = a(b0 + i, b0 + j) / a(b0 + i, b0 + i);
        }
        __syncthreads();
        for (int c = w; c < cw; c += 4) {
            T x = lane < bs ? X[c * n + b0 + lane] : T(0);
            for (int i = 0; i < bs - 1; ++i) {
                const T xi = __shfl(x, i, 64);
                if (lane > i && lane < bs) x -= D[i * 65 + lane] * xi;
            }
            if (lane < bs) X[c * n + b0 + lane] = x;
        }
        __syncthreads();
        // rows below the block: X[r] -= (sum_j a(r, b0+j) X[b0+j]) / a(r, r)
        for (int r = b0 + bs + t; r < n; r += 256) {
            T acc[16];
#pragma unroll
            for (int c = 0; c < 16; ++c) acc[c] = 0;
            for (int j = 0; j < bs; ++j) {
                const T av = a(r, b0 + j);
#pragma unroll
                for (int c = 0; c < 16; ++c)
                    if (c < cw) acc[c] += av * X[c * n + b0 + j];
            }
            const T inv = T(1) / a(r, r);
#pragma unroll
            for (int c = 0; c < 16; ++c)
                if (c < cw) X[c * n + r] -= acc[c] * inv;
        }
    }
    __syncthreads();
    for (int idx = t; idx < n * cw; idx += 256) {
        const int r = idx % n, c = idx / n;
        bv(r, c0 + c) = X[c * n + r];
    }
}

// X := op'(A) X for a lower (reversed: upper) triangle.
template <typename T, bool UP, bool TRA, bool TRB>
__global__ __launch_bounds__(256) void trmm_base_kernel(int n, int m, int cwmax, const T *__restrict__ A, i64 lda,
                                                        T *__restrict__ B, i64 ldb) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *X = reinterpret_cast<T *>(smem_raw);  // [cw][n]
    const TriView<T, UP, TRA> a{A, lda, n};
    const int c0 = blockIdx.x * cwmax;
    const int cw = min(cwmax, m - c0);
    const RhsView<T, UP, TRB> bv{B, ldb, n};
    const int t = threadIdx.x;
    for (int idx = t; idx < n * cw; idx += 256) {
        const int r = idx % n, c = idx / n;
        X[c * n + r] = bv(r, c0 + c);
    }
    __syncthreads();
    for (int idx = t; idx < n * cw; idx += 256) {
        const int r = idx % n, c = idx / n;
        T acc = 0;
        for (int q = 0; q <= r; ++q) acc += a(r, q) * X[c * n + q];
        bv(r, c0 + c) = acc;
    }
}

template <typename T>
__global__ void scale_kernel(int rows, int cols, T alpha, T *__restrict__ B, i64 ldb) {
    const i64 total = (i64)rows * cols;
    for (i64 idx = blockIdx.x * (i64)blockDim.x + threadIdx.x; idx < total; idx += (i64)gridDim.x * blockDim.x) {
        const int r = idx % rows, c = idx / rows;
        B[r + (i64)c * ldb] *= alpha;
    }
}

// Right-hand sides per workgroup: keeps cw*n + 64*65 elements of LDS under 64 KiB.
template <typename T> static int base_cw(int n) {
    const int budget = (65536 / (int)sizeof(T)) - 64 * 65;
    int cw = 16;
    while (cw > 1 && cw * n > budget) cw >>= 1;
    return cw;
}

// Left-form base call: uplo/trA describe op'(A); trB = B seen transposed.
template <typename T>
static hipError_t base_launch(char func, bool up, bool trA, bool trB, int n, int m, const T *A, i64 lda, T *B, i64 ldb,
                              hipStream_t s) {
    if (n <= 0 || m <= 0) return hipSuccess;
    const int cw = base_cw<T>(n);
    const dim3 grid((m + cw - 1) / cw), block(256);
    size_t sh = (size_t)cw * n * sizeof(T);
    ProfScope prof(func == 'S' ? PK_TRSM_BASE : PK_TRMM_BASE, s);
    if (func == 'S') {
        sh += 64 * 65 * sizeof(T);
#define NLA_TRSM_CASE(U, TA, TB)                                                                      \
    if (up == U && trA == TA && trB == TB)                                                            \
        trsm_base_kernel<T, U, TA, TB><<<grid, block, sh, s>>>(n, m, cw, A, lda, B, ldb);
        NLA_TRSM_CASE(false, false, false) NLA_TRSM_CASE(false, false, true) NLA_TRSM_CASE(false, true, false)
        NLA_TRSM_CASE(false, true, true) NLA_TRSM_CASE(true, false, false) NLA_TRSM_CASE(true, false, true)
        NLA_TRSM_CASE(true, true, false) NLA_TRSM_CASE(true, true, true)
#undef NLA_TRSM_CASE
    } else {
#define NLA_TRMM_CASE(U, TA, TB)                                                                      \
    if (up == U && trA == TA && trB == TB)                                                            \
        trmm_base_kernel<T, U, TA, TB><<<grid, block, sh, s>>>(n, m, cw, A, lda, B, ldb);
        NLA_TRMM_CASE(false, false, false) NLA_TRMM_CASE(false, false, true) NLA_TRMM_CASE(false, true, false)
        NLA_TRMM_CASE(false, true, true) NLA_TRMM_CASE(true, false, false) NLA_TRMM_CASE(true, false, true)
        NLA_TRMM_CASE(true, true, false) NLA_TRMM_CASE(true, true, true)
#undef NLA_TRMM_CASE
    }
    return hipGetLastError();
}

// Base case of unified_rec for (side, uplo) of op(A), where op(A)(i,j) = tr ? A(j,i) : A(i,j).
template <typename T>
static hipError_t base_case(char func, char side, char uplo, bool tr, int n, int m, const T *A, i64 lda, T *B, i64 ldb,
                            hipStream_t s) {
    if (side == 'L') return base_launch<T>(func, uplo == 'U', tr, false, n, m, A, lda, B, ldb, s);
    // X op(A) = B  <=>  op(A)^T X^T = B^T ; op(A)^T(i,j) = op(A)(j,i)
    return base_launch<T>(func, uplo != 'U', !tr, true, n, m, A, lda, B, ldb, s);
}

// Pointer to block (i0, j0) of op(A).
template <typename T> static inline const T *opA_block(const T *A, i64 lda, bool tr, i64 i0, i64 j0) {
    return tr ? A + j0 + i0 * lda : A + i0 + j0 * lda;
}

// unified_rec: src/rectrxm.jl:101-198 (same split rule, same forward/backward orders).
template <typename T>
static hipError_t rec(char func, char side, char uplo, bool tr, i64 n, i64 m, const T *A, i64 lda, T *B, i64 ldb,
                      i64 threshold, hipStream_t s) {
    if (n <= threshold) return base_case<T>(func, side, uplo, tr, (int)n, (int)m, A, lda, B, ldb, s);
    i64 mid;
    if ((n & (n - 1)) == 0) mid = n / 2;
    else { mid = 1; while (mid * 2 <= n) mid *= 2; }
    const i64 rem = n - mid;
    const T *A11 = A;
    const T *A22 = opA_block(A, lda, tr, mid, mid);
    const T *A21 = opA_block(A, lda, tr, mid, 0);
    const T *A12 = opA_block(A, lda, tr, 0, mid);
    const char ta = tr ? 'T' : 'N';
    T *B1 = B;
    T *B2 = (side == 'L') ? B + mid : B + mid * ldb;
    const bool forward = (side == 'L' && uplo == 'L' && func == 'S') || (side == 'R' && uplo == 'U' && func == 'S') ||
                         (side == 'L' && uplo == 'U' && func == 'M') || (side == 'R' && uplo == 'L' && func == 'M');
    hipError_t e;
    if (forward) {
        if ((e = rec<T>(func, side, uplo, tr, mid, m, A11, lda, B1, ldb, threshold, s)) != hipSuccess) return e;
        if (side == 'L') {
            if (func == 'S') e = gemm<T>(ta, 'N', rem, m, mid, T(-1), A21, lda, B1, ldb, T(1), B2, ldb, s);  // B2 -= A21 B1
            else e = gemm<T>(ta, 'N', mid, m, rem, T(1), A12, lda, B2, ldb, T(1), B1, ldb, s);              // B1 += A12 B2
        } else {
            if (func == 'S') e = gemm<T>('N', ta, m, rem, mid, T(-1), B1, ldb, A12, lda, T(1), B2, ldb, s);  // B2 -= B1 A12
            else e = gemm<T>('N', ta, m, mid, rem, T(1), B2, ldb, A21, lda, T(1), B1, ldb, s);              // B1 += B2 A21
        }
        if (e != hipSuccess) return e;
        return rec<T>(func, side, uplo, tr, rem, m, A22, lda, B2, ldb, threshold, s);
    }
    if ((e = rec<T>(func, side, uplo, tr, rem, m, A22, lda, B2, ldb, threshold, s)) != hipSuccess) return e;
    if (side == 'L') {
        if (func == 'S') e = gemm<T>(ta, 'N', mid, m, rem, T(-1), A12, lda, B2, ldb, T(1), B1, ldb, s);  // B1 -= A12 B2
        else e = gemm<T>(ta, 'N', rem, m, mid, T(1), A21, lda, B1, ldb, T(1), B2, ldb, s);              // B2 += A21 B1
    } else {
        if (func == 'S') e = gemm<T>('N', ta, m, mid, rem, T(-1), B2, ldb, A21, lda, T(1), B1, ldb, s);  // B1 -= B2 A21
        else e = gemm<T>('N', ta, m, rem, mid, T(1), B1, ldb, A12, lda, T(1), B2, ldb, s);              // B2 += B1 A12
    }
    if (e != hipSuccess) return e;
    return rec<T>(func, side, uplo, tr, mid, m, A11, lda, B1, ldb, threshold, s);
}

template <typename T>
static hipError_t scale(i64 rows, i64 cols, T alpha, T *B, i64 ldb, hipStream_t s) {
    if (alpha == T(1) || rows <= 0 || cols <= 0) return hipSuccess;
    i64 total = rows * cols;
    int blocks = (int)((total + 255) / 256);
    if (blocks > 4096) blocks = 4096;
    scale_kernel<T><<<blocks, 256, 0, s>>>((int)rows, (int)cols, alpha, B, ldb);
    return hipGetLastError();
}

template <typename T>
static int rectrxm_entry(char side, char uplo, char trans, T alpha, char func, i64 n, i64 m, const T *A, i64 lda, T *B,
                         i64 ldb, i64 cutoff, void *stream) {
    side = up(side); uplo = up(uplo); trans = up(trans); func = up(func);
    if (side != 'L' && side != 'R') return arg_error(1, "side must be 'L' or 'R', got '%c'", side);
    if (uplo != 'L' && uplo != 'U') return arg_error(2, "uplo must be 'L' or 'U', got '%c'", uplo);
    if (trans != 'N' && trans != 'T' && trans != 'C') return arg_error(3, "transpose must be 'N', 'T' or 'C', got '%c'", trans);
    if (func != 'S' && func != 'M') return arg_error(5, "func must be 'S' or 'M', got '%c'", func);
    if (n < 0) return arg_error(6, "n must be non-negative, got %lld", (long long)n);
    if (m < 0) return arg_error(7, "m must be non-negative, got %lld", (long long)m);
    if (n == 0 || m == 0) return 0;
    if (lda < n) return arg_error(9, "lda must be >= n");
    const i64 brows = side == 'L' ? n : m;
    if (ldb < brows) return arg_error(11, "ldb must be >= %lld", (long long)brows);
    if (cutoff < 0) return arg_error(12, "cutoff must be non-negative");
    if (n > INT32_MAX || m > INT32_MAX) return arg_error(6, "dimension exceeds 2^31-1");
    i64 thr = cutoff == 0 ? (func == 'S' ? 256 : 16) : cutoff;
    if (thr > 1024) return arg_error(12, "cutoff must be <= 1024 (base kernels hold n <= 1024 rows, src/trsm.jl:10-12)");
    bool tr = false;
    if (trans == 'T' || trans == 'C') { tr = true; uplo = (uplo == 'L') ? 'U' : 'L'; }  // src/rectrxm.jl:56-59
    hipStream_t s = as_stream(stream);
    const i64 rows = side == 'L' ? n : m, cols = side == 'L' ? m : n;
    if (func == 'S') NLA_HIP(scale<T>(rows, cols, alpha, B, ldb, s));
    NLA_HIP(rec<T>(func, side, uplo, tr, n, m, A, lda, B, ldb, thr, s));
    if (func == 'M') NLA_HIP(scale<T>(rows, cols, alpha, B, ldb, s));
    return 0;
}

template <typename T>
static int base_entry(char func, char side, char uplo, i64 n, i64 m, const T *A, i64 lda, T *B, i64 ldb, void *stream) {
    side = up(side); uplo = up(uplo);
    if (side != 'L' && side != 'R') return arg_error(1, "side must be 'L' or 'R', got '%c'", side);
    if (uplo != 'L' && uplo != 'U') return arg_error(2, "uplo must be 'L' or 'U', got '%c'", uplo);
    if (n < 0) return arg_error(3, "n must be non-negative");
    if (m < 0) return arg_error(4, "m must be non-negative");
    if (n > 1024) return arg_error(3, "base kernels hold at most 1024 rows (src/trsm.jl:10-12), got %lld", (long long)n);
    if (n == 0 || m == 0) return 0;
    if (lda < n) return arg_error(6, "lda must be >= n");
    if (ldb < (side == 'L' ? n : m)) return arg_error(8, "ldb too small");
    NLA_HIP(base_case<T>(func, side, uplo, false, (int)n, (int)m, A, lda, B, ldb, as_stream(stream)));
    return 0;
}

// Exposed for the mixed-precision and differentiation code.
template <typename T>
hipError_t rectrxm_async(char side, char uplo, char trans, T alpha, char func, i64 n, i64 m, const T *A, i64 lda, T *B,
                         i64 ldb, i64 cutoff, hipStream_t s) {
    int st = rectrxm_entry<T>(side, uplo, trans, alpha, func, n, m, A, lda, B, ldb, cutoff, (void *)s);
    return st == 0 ? hipSuccess : hipErrorInvalidValue;
}
template hipError_t rectrxm_async<double>(char, char, char, double, char, i64, i64, const double *, i64, double *, i64,
                                          i64, hipStream_t);
template hipError_t rectrxm_async<float>(char, char, char, float, char, i64, i64, const float *, i64, float *, i64, i64,
                                         hipStream_t);

}  // namespace nla

NLA_API int nla_drectrxm(char side, char uplo, char trans, double alpha, char func, int64_t n, int64_t m,
                         const double *A, int64_t lda, double *B, int64_t ldb, int64_t cutoff, void *stream) {
    return nla::rectrxm_entry<double>(side, uplo, trans, alpha, func, n, m, A, lda, B, ldb, cutoff, stream);
}
NLA_API int nla_srectrxm(char side, char uplo, char trans, float alpha, char func, int64_t n, int64_t m,
                         const float *A, int64_t lda, float *B, int64_t ldb, int64_t cutoff, void *stream) {
    return nla::rectrxm_entry<float>(side, uplo, trans, alpha, func, n, m, A, lda, B, ldb, cutoff, stream);
}
NLA_API int nla_dtrsm_base(char side, char uplo, int64_t n, int64_t m, const double *A, int64_t lda, double *B,
                           int64_t ldb, void *stream) {
    return nla::base_entry<double>('S', side, uplo, n, m, A, lda, B, ldb, stream);
}
NLA_API int nla_strsm_base(char side, char uplo, int64_t n, int64_t m, const float *A, int64_t lda, float *B,
                           int64_t ldb, void *stream) {
    return nla::base_entry<float>('S', side, uplo, n, m, A, lda, B, ldb, stream);
}
NLA_API int nla_dtrmm_base(char side, char uplo, int64_t n, int64_t m, const double *A, int64_t lda, double *B,
                           int64_t ldb, void *stream) {
    return nla::base_entry<double>('M', side, uplo, n, m, A, lda, B, ldb, stream);
}
NLA_API int nla_strmm_base(char side, char uplo, int64_t n, int64_t m, const float *A, int64_t lda, float *B,
                           int64_t ldb, void *stream) {
    return nla::base_entry<float>('M', side, uplo, n, m, A, lda, B, ldb, stream);
}
