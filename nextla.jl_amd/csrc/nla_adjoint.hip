// Reverse-mode rules (adjoints) of the recursive triangular solve / multiply and of the thin
// tile QR.  The reference has no differentiation code; these are the kernels a
// ChainRulesCore.rrule for unified_rectrxm! (src/rectrxm.jl:43-76) and for a thin QR built
// from geqrt!/tsqrt! would call.  Everything is composed from the library's own recursive
// TRSM/TRMM, MFMA GEMM and tile-QR apply kernels plus three small element-wise kernels.
//
// Triangular solve, left side, op(A) X = alpha B (A n x n, X and B n x m):
//   Y = op(A)^-T Xbar,   Bbar = alpha Y,   op(A)bar = -Y X^T,
//   Abar = tri(-Y X^T) for op = N, tri(-X Y^T) for op = T   (tri = the stored triangle).
// Right side, X op(A) = alpha B: Y = Xbar op(A)^-T, Bbar = alpha Y, op(A)bar = -X^T Y.
// Triangular multiply, left, Yf = alpha op(A) B: Bbar = alpha op(A)^T Ybar,
//   op(A)bar = alpha Ybar B^T; right, Yf = alpha B op(A): Bbar = alpha Ybar op(A)^T,
//   op(A)bar = alpha B^T Ybar.
// Thin QR, A = Q R (A m x n, Q m x n, R n x n upper):
//   M = R Rbar^T - Qbar^T Q,   Abar = (Qbar + Q copyltu(M)) R^-T,
// with copyltu(M) = tril(M) + tril(M, -1)^T; Q is applied from the stored compact-WY
// factors (V, T) of the tile QR and never formed.
#include "nla_common.hpp"
#include "nla_gemm.hpp"

NLA_API int nla_dtile_ormqr(char trans, int64_t m, int64_t n, int64_t nrhs, int64_t nb, int64_t ib, const double *A,
                            int64_t lda, const double *T, int64_t ldt, double *C, int64_t ldc, void *stream);
NLA_API int nla_stile_ormqr(char trans, int64_t m, int64_t n, int64_t nrhs, int64_t nb, int64_t ib, const float *A,
                            int64_t lda, const float *T, int64_t ldt, float *C, int64_t ldc, void *stream);

namespace nla {

template <typename T>
int rectrxm_async(char side, char uplo, char trans, T alpha, char func, i64 n, i64 m, const T *A, i64 lda, T *B,
                         i64 ldb, i64 cutoff, hipStream_t s);

// C (n x n) := tri(C) (keep the uplo triangle with its diagonal, zero the other one).
template <typename T> __global__ void tri_keep_kernel(int n, bool upper, T *C, i64 ldc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= n || j >= n) return;
    if (upper ? i > j : i < j) C[i + (i64)j * ldc] = T(0);
}
// Dst (m x n) := [triu(Src[0:n, 0:n]); 0] (upper = true) or [Src[0:n, 0:n]; 0] (upper = false).
template <typename T>
__global__ void embed_kernel(int m, int n, bool upper, const T *Src, i64 lds, T *Dst, i64 ldd) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= m || j >= n) return;
    T v = T(0);
    if (i < n && (!upper || i <= j)) v = Src[i + (i64)j * lds];
    Dst[i + (i64)j * ldd] = v;
}
// Y (m x n) += X (m x n)
template <typename T> __global__ void add_kernel(int m, int n, const T *X, i64 ldx, T *Y, i64 ldy) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= m || j >= n) return;
    Y[i + (i64)j * ldy] += X[i + (i64)j * ldx];
}
// Y (m x n) -= X (m x n)
template <typename T> __global__ void sub_kernel(int m, int n, const T *X, i64 ldx, T *Y, i64 ldy) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= m || j >= n) return;
    Y[i + (i64)j * ldy] -= X[i + (i64)j * ldx];
}
// Dst (m x n) := [S; 0] with S[i][j] = Mt[min(i,j)][max(i,j)] (Mt n x n)
template <typename T> __global__ void mirror_embed_kernel(int m, int n, const T *Mt, i64 ldm, T *Dst, i64 ldd) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= m || j >= n) return;
    T v = T(0);
    if (i < n) v = i <= j ? Mt[i + (i64)j * ldm] : Mt[j + (i64)i * ldm];
    Dst[i + (i64)j * ldd] = v;
}

static inline dim3 grid2(i64 m, i64 n) { return dim3((unsigned)((m + 255) / 256), (unsigned)n); }

static inline char flip(char t) { return is_trans(t) ? 'N' : 'T'; }

template <typename T> static int ormqr(char trans, i64 m, i64 n, i64 nrhs, i64 nb, i64 ib, const T *A, i64 lda,
                                       const T *Tm, i64 ldt, T *C, i64 ldc, hipStream_t s);
template <> int ormqr<double>(char trans, i64 m, i64 n, i64 nrhs, i64 nb, i64 ib, const double *A, i64 lda,
                              const double *Tm, i64 ldt, double *C, i64 ldc, hipStream_t s) {
    return nla_dtile_ormqr(trans, m, n, nrhs, nb, ib, A, lda, Tm, ldt, C, ldc, s);
}
template <> int ormqr<float>(char trans, i64 m, i64 n, i64 nrhs, i64 nb, i64 ib, const float *A, i64 lda,
                             const float *Tm, i64 ldt, float *C, i64 ldc, hipStream_t s) {
    return nla_stile_ormqr(trans, m, n, nrhs, nb, ib, A, lda, Tm, ldt, C, ldc, s);
}

static int check_trx(char side, char uplo, char trans, i64 n, i64 m) {
    if (side != 'L' && side != 'R') return arg_error(1, "side must be 'L' or 'R'");
    if (uplo != 'L' && uplo != 'U') return arg_error(2, "uplo must be 'L' or 'U'");
    if (trans != 'N' && trans != 'T' && trans != 'C') return arg_error(3, "transpose must be 'N', 'T' or 'C'");
    if (n < 0) return arg_error(5, "n must be non-negative");
    if (m < 0) return arg_error(6, "m must be non-negative");
    if (n > INT32_MAX || m > INT32_MAX) return arg_error(5, "dimension exceeds 2^31-1");
    return 0;
}

// Abar (n x n) := tri(alpha * P Q^T) (left; P, Q n x m) or tri(alpha * P^T Q) (right; m x n).
template <typename T>
static hipError_t tri_product(bool left, bool upper, i64 n, i64 m, T alpha, const T *P, i64 ldp, const T *Q, i64 ldq,
                              T *Abar, i64 ldab, hipStream_t s) {
    if (left) {
        hipError_t e = gemm<T>('N', 'T', n, n, m, alpha, P, ldp, Q, ldq, T(0), Abar, ldab, s);
        if (e != hipSuccess) return e;
    } else {
        hipError_t e = gemm<T>('T', 'N', n, n, m, alpha, P, ldp, Q, ldq, T(0), Abar, ldab, s);
        if (e != hipSuccess) return e;
    }
    tri_keep_kernel<T><<<grid2(n, n), 256, 0, s>>>((int)n, upper, Abar, ldab);
    return hipGetLastError();
}

template <typename T>
static int trsm_rrule_entry(char side, char uplo, char trans, T alpha, i64 n, i64 m, const T *A, i64 lda, const T *X,
                            i64 ldx, T *Xbar, i64 ldxb, T *Abar, i64 ldab, i64 cutoff, void *stream) {
    side = up(side), uplo = up(uplo), trans = up(trans);
    if (int r = check_trx(side, uplo, trans, n, m)) return r;
    const i64 rows = side == 'L' ? n : m;
    if (lda < (n > 1 ? n : 1)) return arg_error(8, "lda must be >= max(1,n)");
    if (ldx < (rows > 1 ? rows : 1)) return arg_error(10, "ldx must be >= the rows of X");
    if (ldxb < (rows > 1 ? rows : 1)) return arg_error(12, "ldxbar must be >= the rows of X");
    if (ldab < (n > 1 ? n : 1)) return arg_error(14, "ldabar must be >= max(1,n)");
    if (n == 0 || m == 0) return 0;
    hipStream_t s = as_stream(stream);
    const bool left = side == 'L', tr = is_trans(trans), upper = uplo == 'U';
    // Y = op(A)^-T Xbar (left) / Xbar op(A)^-T (right), in place of Xbar
    NLA_ST(rectrxm_async<T>(side, uplo, flip(trans), T(1), 'S', n, m, A, lda, Xbar, ldxb, cutoff, s));
    // op(A)bar = -Y X^T (left) / -X^T Y (right); Abar is its (transposed for op = T) triangle
    if (left) {
        if (!tr) NLA_HIP(tri_product<T>(true, upper, n, m, T(-1), Xbar, ldxb, X, ldx, Abar, ldab, s));
        else NLA_HIP(tri_product<T>(true, upper, n, m, T(-1), X, ldx, Xbar, ldxb, Abar, ldab, s));
    } else {
        if (!tr) NLA_HIP(tri_product<T>(false, upper, n, m, T(-1), X, ldx, Xbar, ldxb, Abar, ldab, s));
        else NLA_HIP(tri_product<T>(false, upper, n, m, T(-1), Xbar, ldxb, X, ldx, Abar, ldab, s));
    }
    // Bbar = alpha Y
    if (alpha != T(1)) NLA_HIP(gemm<T>('N', 'N', rows, left ? m : n, 0, T(0), nullptr, 1, nullptr, 1, alpha, Xbar, ldxb, s));
    return 0;
}

template <typename T>
static int trmm_rrule_entry(char side, char uplo, char trans, T alpha, i64 n, i64 m, const T *A, i64 lda, const T *B,
                            i64 ldb, T *Ybar, i64 ldyb, T *Abar, i64 ldab, i64 cutoff, void *stream) {
    side = up(side), uplo = up(uplo), trans = up(trans);
    if (int r = check_trx(side, uplo, trans, n, m)) return r;
    const i64 rows = side == 'L' ? n : m;
    if (lda < (n > 1 ? n : 1)) return arg_error(8, "lda must be >= max(1,n)");
    if (ldb < (rows > 1 ? rows : 1)) return arg_error(10, "ldb must be >= the rows of B");
    if (ldyb < (rows > 1 ? rows : 1)) return arg_error(12, "ldybar must be >= the rows of B");
    if (ldab < (n > 1 ? n : 1)) return arg_error(14, "ldabar must be >= max(1,n)");
    if (n == 0 || m == 0) return 0;
    hipStream_t s = as_stream(stream);
    const bool left = side == 'L', tr = is_trans(trans), upper = uplo == 'U';
    // op(A)bar = alpha Ybar B^T (left) / alpha B^T Ybar (right), before Ybar is overwritten
    if (left) {
        if (!tr) NLA_HIP(tri_product<T>(true, upper, n, m, alpha, Ybar, ldyb, B, ldb, Abar, ldab, s));
        else NLA_HIP(tri_product<T>(true, upper, n, m, alpha, B, ldb, Ybar, ldyb, Abar, ldab, s));
    } else {
        if (!tr) NLA_HIP(tri_product<T>(false, upper, n, m, alpha, B, ldb, Ybar, ldyb, Abar, ldab, s));
        else NLA_HIP(tri_product<T>(false, upper, n, m, alpha, Ybar, ldyb, B, ldb, Abar, ldab, s));
    }
    // Bbar = alpha op(A)^T Ybar (left) / alpha Ybar op(A)^T (right), in place of Ybar
    NLA_ST(rectrxm_async<T>(side, uplo, flip(trans), alpha, 'M', n, m, A, lda, Ybar, ldyb, cutoff, s));
    return 0;
}

template <typename T>
static int qr_rrule_entry(i64 m, i64 n, i64 nb, i64 ib, const T *AF, i64 lda, const T *Tm, i64 ldt, T *Qbar, i64 ldq,
                          const T *Rbar, i64 ldr, void *stream) {
    if (m < 0) return arg_error(1, "m must be non-negative");
    if (n < 0) return arg_error(2, "n must be non-negative");
    if (nb <= 0) return arg_error(3, "nb must be positive");
    if (ib <= 0 || ib > nb) return arg_error(4, "ib must be in 1..nb");
    if (m % nb || n % nb) return arg_error(1, "m and n must be multiples of nb");
    if (n > m) return arg_error(2, "the thin QR needs n <= m");
    if (m == 0 || n == 0) return 0;
    if (lda < m) return arg_error(6, "lda must be >= m");
    if (ldt < (m / nb) * ib) return arg_error(8, "ldt must be >= (m/nb)*ib");
    if (ldq < m) return arg_error(10, "ldqbar must be >= m");
    if (ldr < n) return arg_error(12, "ldrbar must be >= n");
    if (m > INT32_MAX / 2) return arg_error(1, "dimension too large");
    hipStream_t s = as_stream(stream);
    int st = 0;
    // workspace: W (m x n), Rm (n x n), Mt (n x n)
    T *W = (T *)scratch(SLOT_MIXED, sizeof(T) * ((size_t)m * n + 2 * (size_t)n * n), s, &st);
    if (!W) return st;
    T *Rm = W + (size_t)m * n, *Mt = Rm + (size_t)n * n;
    // W = Q_full^T Qbar, whose top n x n block is Q^T Qbar = (Qbar^T Q)^T
    NLA_HIP(hipMemcpy2DAsync(W, sizeof(T) * m, Qbar, sizeof(T) * ldq, sizeof(T) * m, n, hipMemcpyDeviceToDevice, s));
    if (int r = ormqr<T>('T', m, n, n, nb, ib, AF, lda, Tm, ldt, W, m, s)) return r;
    // R = triu(AF[0:n, 0:n])
    embed_kernel<T><<<grid2(n, n), 256, 0, s>>>((int)n, (int)n, true, AF, lda, Rm, n);
    NLA_LAUNCH_CHECK();
    // M^T = Rbar R^T - Q^T Qbar
    NLA_HIP(gemm<T>('N', 'T', n, n, n, T(1), Rbar, ldr, Rm, n, T(0), Mt, n, s));
    sub_kernel<T><<<grid2(n, n), 256, 0, s>>>((int)n, (int)n, W, m, Mt, n);
    NLA_LAUNCH_CHECK();
    // W = [copyltu(M); 0]: copyltu(M)[i][j] = M[max(i,j)][min(i,j)] = M^T[min(i,j)][max(i,j)]
    mirror_embed_kernel<T><<<grid2(m, n), 256, 0, s>>>((int)m, (int)n, Mt, n, W, m);
    NLA_LAUNCH_CHECK();
    // W = Q_full [copyltu(M); 0] = Q copyltu(M);  Qbar += W
    if (int r = ormqr<T>('N', m, n, n, nb, ib, AF, lda, Tm, ldt, W, m, s)) return r;
    add_kernel<T><<<grid2(m, n), 256, 0, s>>>((int)m, (int)n, W, m, Qbar, ldq);
    NLA_LAUNCH_CHECK();
    // Abar = (Qbar + Q copyltu(M)) R^-T, in place of Qbar
    NLA_ST(rectrxm_async<T>('R', 'U', 'T', T(1), 'S', n, m, AF, lda, Qbar, ldq, 0, s));
    return 0;
}

}  // namespace nla

NLA_API int nla_dtrsm_rrule(char side, char uplo, char trans, double alpha, int64_t n, int64_t m, const double *A,
                            int64_t lda, const double *X, int64_t ldx, double *Xbar, int64_t ldxbar, double *Abar,
                            int64_t ldabar, int64_t cutoff, void *stream) {
    return nla::trsm_rrule_entry<double>(side, uplo, trans, alpha, n, m, A, lda, X, ldx, Xbar, ldxbar, Abar, ldabar,
                                         cutoff, stream);
}
NLA_API int nla_strsm_rrule(char side, char uplo, char trans, float alpha, int64_t n, int64_t m, const float *A,
                            int64_t lda, const float *X, int64_t ldx, float *Xbar, int64_t ldxbar, float *Abar,
                            int64_t ldabar, int64_t cutoff, void *stream) {
    return nla::trsm_rrule_entry<float>(side, uplo, trans, alpha, n, m, A, lda, X, ldx, Xbar, ldxbar, Abar, ldabar,
                                        cutoff, stream);
}
NLA_API int nla_dtrmm_rrule(char side, char uplo, char trans, double alpha, int64_t n, int64_t m, const double *A,
                            int64_t lda, const double *B, int64_t ldb, double *Ybar, int64_t ldybar, double *Abar,
                            int64_t ldabar, int64_t cutoff, void *stream) {
    return nla::trmm_rrule_entry<double>(side, uplo, trans, alpha, n, m, A, lda, B, ldb, Ybar, ldybar, Abar, ldabar,
                                         cutoff, stream);
}
NLA_API int nla_strmm_rrule(char side, char uplo, char trans, float alpha, int64_t n, int64_t m, const float *A,
                            int64_t lda, const float *B, int64_t ldb, float *Ybar, int64_t ldybar, float *Abar,
                            int64_t ldabar, int64_t cutoff, void *stream) {
    return nla::trmm_rrule_entry<float>(side, uplo, trans, alpha, n, m, A, lda, B, ldb, Ybar, ldybar, Abar, ldabar,
                                        cutoff, stream);
}
NLA_API int nla_dqr_rrule(int64_t m, int64_t n, int64_t nb, int64_t ib, const double *AF, int64_t lda, const double *T,
                          int64_t ldt, double *Qbar, int64_t ldqbar, const double *Rbar, int64_t ldrbar, void *stream) {
    return nla::qr_rrule_entry<double>(m, n, nb, ib, AF, lda, T, ldt, Qbar, ldqbar, Rbar, ldrbar, stream);
}
NLA_API int nla_sqr_rrule(int64_t m, int64_t n, int64_t nb, int64_t ib, const float *AF, int64_t lda, const float *T,
                          int64_t ldt, float *Qbar, int64_t ldqbar, const float *Rbar, int64_t ldrbar, void *stream) {
    return nla::qr_rrule_entry<float>(m, n, nb, ib, AF, lda, T, ldt, Qbar, ldqbar, Rbar, ldrbar, stream);
}
