// Mixed-precision recursive triangular solve with iterative refinement (no reference
// function: the reference's unified_rectrxm!, src/rectrxm.jl:43-76, runs in one precision).
//   X0 = op(A)^-1 (alpha B) solved in Float32 storage: the diagonal-block inverses and base
//        solves in Float32, the recursion's GEMM updates (the GEMM_SUB! call sites,
//        src/rectrxm.jl:164,176,183,192) with Float32 ('S', v_mfma_f32_16x16x4f32), FP16 ('H')
//        or BF16 ('B', v_mfma_f32_16x16x32_{f16,bf16}) operands and Float32 accumulation;
//   R_k = alpha B - op(A) X_k in Float64 (a Float64 recursive TRMM, cut-off 256);
//   X_{k+1} = X_k + op(A)^-1 R_k solved in Float32.
// Each step multiplies the error by about kappa(A) u_32.  The Float64 residual costs as
// many flops (n^2 m) as a Float64 solve, so on hardware whose Float64 and Float32 rates
// differ by 2x this path cannot beat the plain Float64 solve at Float64 accuracy; it is
// the measured comparison point for that claim.
#include "nla_common.hpp"
#include "nla_gemm.hpp"

#include <cmath>
#include <vector>

namespace nla {

template <typename T>
int rectrxm_async(char side, char uplo, char trans, T alpha, char func, i64 n, i64 m, const T *A, i64 lda, T *B,
                         i64 ldb, i64 cutoff, hipStream_t s);

// Dst (rows x cols, To) = alpha * Src (From); only the uplo triangle when tri != 0 ('L'/'U')
template <typename To, typename From>
__global__ void convert_kernel(int rows, int cols, char tri, From alpha, const From *Src, i64 lds, To *Dst, i64 ldd) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= rows || j >= cols) return;
    From v = Src[i + (i64)j * lds] * alpha;
    if ((tri == 'L' && i < j) || (tri == 'U' && i > j)) v = From(0);
    Dst[i + (i64)j * ldd] = (To)v;
}
// R = B0 - Y (Float64), R32 = (float)R; accumulates nothing
__global__ void residual_kernel(int rows, int cols, const double *B0, i64 ldb0, const double *Y, i64 ldy, float *R32,
                                i64 ldr) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i >= rows || j >= cols) return;
    R32[i + (i64)j * ldr] = (float)(B0[i + (i64)j * ldb0] - Y[i + (i64)j * ldy]);
}
// X += (double) D32;  sums[0] += ||D||^2, sums[1] += ||X||^2 over this block
__global__ void update_kernel(int rows, int cols, const float *D32, i64 ldd, double *X, i64 ldx, double *sums) {
    __shared__ double red[2][4];
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    double dd = 0, xx = 0;
    if (i < rows && j < cols) {
        const double d = D32[i + (i64)j * ldd];
        const double x = X[i + (i64)j * ldx] + d;
        X[i + (i64)j * ldx] = x;
        dd = d * d;
        xx = x * x;
    }
    dd = wave_sum(dd);
    xx = wave_sum(xx);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) red[0][w] = dd, red[1][w] = xx;
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0, b = 0;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) a += red[0][k], b += red[1][k];
        atomicAdd(&sums[0], a);
        atomicAdd(&sums[1], b);
    }
}

static int rectrsm_mixed_entry(char side, char uplo, char trans, double alpha, i64 n, i64 m, const double *A, i64 lda,
                               double *B, i64 ldb, i64 cutoff, i64 maxit, double tol, char prec, i64 *iters,
                               double *last, void *stream) {
    side = up(side), uplo = up(uplo), trans = up(trans), prec = up(prec);
    if (prec != 'S' && prec != 'H' && prec != 'B') return arg_error(14, "prec must be 'S', 'H' or 'B', got '%c'", prec);
    if (side != 'L' && side != 'R') return arg_error(1, "side must be 'L' or 'R', got '%c'", side);
    if (uplo != 'L' && uplo != 'U') return arg_error(2, "uplo must be 'L' or 'U', got '%c'", uplo);
    if (trans != 'N' && trans != 'T' && trans != 'C') return arg_error(3, "transpose must be 'N', 'T' or 'C'");
    if (n < 0) return arg_error(5, "n must be non-negative");
    if (m < 0) return arg_error(6, "m must be non-negative");
    if (maxit < 0) return arg_error(12, "maxit must be non-negative");
    if (iters) *iters = 0;
    if (last) *last = 0;
    if (n == 0 || m == 0) return 0;
    if (lda < n) return arg_error(8, "lda must be >= n");
    const i64 rows = side == 'L' ? n : m, cols = side == 'L' ? m : n;
    if (ldb < rows) return arg_error(10, "ldb must be >= %lld", (long long)rows);
    if (n > INT32_MAX || m > INT32_MAX) return arg_error(5, "dimension exceeds 2^31-1");
    hipStream_t s = as_stream(stream);
    int st = 0;
    // scratch: A32 (n x n) | R32 (rows x cols) | B0 (rows x cols, double) | Y (double) | sums
    const size_t bytes = sizeof(float) * ((size_t)n * n + (size_t)rows * cols) + sizeof(double) * (2 * (size_t)rows * cols + 2);
    char *base = (char *)scratch(SLOT_MIXED, bytes, s, &st);
    if (!base) return st;
    float *A32 = (float *)base, *R32 = A32 + (size_t)n * n;
    double *B0 = (double *)(R32 + (size_t)rows * cols), *Y = B0 + (size_t)rows * cols, *sums = Y + (size_t)rows * cols;
    const dim3 ga((unsigned)((n + 255) / 256), (unsigned)n), gb((unsigned)((rows + 255) / 256), (unsigned)cols);
    convert_kernel<float, double><<<ga, 256, 0, s>>>((int)n, (int)n, uplo, 1.0, A, lda, A32, n);
    NLA_LAUNCH_CHECK();
    // B0 = alpha B (Float64, the right-hand side of every residual); R32 = (float) B0
    convert_kernel<double, double><<<gb, 256, 0, s>>>((int)rows, (int)cols, 0, alpha, B, ldb, B0, rows);
    NLA_LAUNCH_CHECK();
    convert_kernel<float, double><<<gb, 256, 0, s>>>((int)rows, (int)cols, 0, 1.0, B0, rows, R32, rows);
    NLA_LAUNCH_CHECK();
    {
        GemmLowp lowp(prec);
        NLA_ST(rectrxm_async<float>(side, uplo, trans, 1.0f, 'S', n, m, A32, n, R32, rows, cutoff, s));
    }
    convert_kernel<double, float><<<gb, 256, 0, s>>>((int)rows, (int)cols, 0, 1.0f, R32, rows, B, ldb);
    NLA_LAUNCH_CHECK();
    double prev = 0;
    for (i64 it = 0; it < maxit; ++it) {
        // R = B0 - op(A) X in Float64
        NLA_HIP(hipMemcpy2DAsync(Y, sizeof(double) * rows, B, sizeof(double) * ldb, sizeof(double) * rows, cols,
                                 hipMemcpyDeviceToDevice, s));
        NLA_ST(rectrxm_async<double>(side, uplo, trans, 1.0, 'M', n, m, A, lda, Y, rows, 256, s));
        residual_kernel<<<gb, 256, 0, s>>>((int)rows, (int)cols, B0, rows, Y, rows, R32, rows);
        NLA_LAUNCH_CHECK();
        // D = op(A)^-1 R in Float32 storage with the chosen update precision; X += D
        {
            GemmLowp lowp(prec);
            NLA_ST(rectrxm_async<float>(side, uplo, trans, 1.0f, 'S', n, m, A32, n, R32, rows, cutoff, s));
        }
        NLA_HIP(hipMemsetAsync(sums, 0, 2 * sizeof(double), s));
        update_kernel<<<gb, 256, 0, s>>>((int)rows, (int)cols, R32, rows, B, ldb, sums);
        NLA_LAUNCH_CHECK();
        double h[2] = {0, 0};
        NLA_HIP(hipMemcpyAsync(h, sums, sizeof h, hipMemcpyDeviceToHost, s));
        NLA_HIP(hipStreamSynchronize(s));
        const double ratio = h[1] > 0 ? std::sqrt(h[0] / h[1]) : std::sqrt(h[0]);
        if (iters) *iters = it + 1;
        if (last) *last = ratio;
        if (ratio <= tol) break;
        if (it > 0 && ratio > 0.5 * prev) break;  // stagnation: the Float64 residual's floor
        prev = ratio;
    }
    return 0;
}

}  // namespace nla

NLA_API int nla_drectrsm_mixed(char side, char uplo, char trans, double alpha, int64_t n, int64_t m, const double *A,
                               int64_t lda, double *B, int64_t ldb, int64_t cutoff, int64_t maxit, double tol, char prec,
                               int64_t *iters, double *last_correction, void *stream) {
    return nla::rectrsm_mixed_entry(side, uplo, trans, alpha, n, m, A, lda, B, ldb, cutoff, maxit, tol, prec, iters,
                                    last_correction, stream);
}
