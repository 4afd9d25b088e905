// Tile-QR driver: flat-tree PLASMA-order composition of geqrt!/unmqr!/tsqrt!/tsmqr!
// (no reference function: the reference leaves this loop to its caller).
//   for k: geqrt(A_kk); unmqr('L','C', A_kk, A_k,j>k); for i>k: tsqrt(A_kk, A_ik);
//          tsmqr('L','C', A_k,j>k, A_i,j>k, V=A_ik, T_ik)
// Each unmqr/tsmqr step is one strip launch over the whole trailing tile row.
#include "nla_common.hpp"
#include "nla_gemm.hpp"
#include "nla_qr.hpp"

#include <cmath>
#include <vector>

namespace nla {

template <typename T>
int tile_geqrf_dag(i64 m, i64 n, i64 nb, i64 ib, T *A, i64 lda, T *Tm, i64 ldt, hipStream_t s, bool *handled);
int tile_geqrf_dag_mixed(i64 m, i64 n, i64 nb, i64 ib, double *A, i64 lda, double *Tm, i64 ldt, hipStream_t s,
                         bool *handled);
template <typename T>
int tile_apply_qt_dag(i64 m, i64 n, i64 nb, i64 ib, const T *A, i64 lda, const T *Tm, i64 ldt, T *C, i64 ldc,
                      i64 ncols, hipStream_t s, bool *handled, bool qt);
template <typename T>
hipError_t rectrxm_async(char side, char uplo, char trans, T alpha, char func, i64 n, i64 m, const T *A, i64 lda, T *B,
                         i64 ldb, i64 cutoff, hipStream_t s);

template <typename T>
static int tile_geqrf_entry(i64 m, i64 n, i64 nb, i64 ib, T *A, i64 lda, T *Tm, i64 ldt, void *stream) {
    if (m < 0) return arg_error(1, "m must be non-negative");
    if (n < 0) return arg_error(2, "n must be non-negative");
    if (nb <= 0) return arg_error(3, "nb must be positive");
    if (ib <= 0 || ib > nb) return arg_error(4, "ib must be in 1..nb");
    if (m % nb || n % nb) return arg_error(1, "m and n must be multiples of nb");
    if (m == 0 || n == 0) return 0;
    if (lda < m) return arg_error(6, "lda must be >= m");
    const i64 mt = m / nb, nt = n / nb, kt = mt < nt ? mt : nt;
    if (ldt < mt * ib) return arg_error(8, "ldt must be >= (m/nb)*ib");
    if (ib > 256) return arg_error(4, "ib must be <= 256");
    hipStream_t s = as_stream(stream);
    {
        bool handled = false;
        const int r = tile_geqrf_dag<T>(m, n, nb, ib, A, lda, Tm, ldt, s, &handled);
        if (handled) return r < 0 ? 1 : r;
    }
    // other (nb, ib): host loop over per-step launches
    int st = 0;
    T *tau = (T *)scratch(SLOT_AUX, sizeof(T) * (size_t)nb, s, &st);
    if (!tau) return st;
    for (i64 k = 0; k < kt; ++k) {
        T *Akk = A + k * nb + k * nb * lda;
        T *Tkk = Tm + k * ib + k * nb * ldt;
        NLA_HIP(geqrt_async<T>(nb, nb, ib, Akk, lda, Tkk, ldt, tau, s));
        const i64 ntr = n - (k + 1) * nb;
        T *Akj = A + k * nb + (k + 1) * nb * lda;
        if (ntr > 0) NLA_HIP(launch_ge_apply<T>(false, nb, ntr, nb, ib, Akk, lda, Tkk, ldt, false, true, Akj, lda, s));
        for (i64 i = k + 1; i < mt; ++i) {
            T *Aik = A + i * nb + k * nb * lda;
            T *Tik = Tm + i * ib + k * nb * ldt;
            NLA_HIP(tsqrt_async<T>(nb, nb, ib, nb, Akk, lda, Aik, lda, Tik, ldt, tau, s));
            if (ntr > 0)
                NLA_HIP(launch_ts_apply<T>(false, nb, ntr, nb, ib, Aik, lda, Tik, ldt, false, true, Akj, lda,
                                           A + i * nb + (k + 1) * nb * lda, lda, s));
        }
    }
    return 0;
}

template <typename T>
static int tile_ormqr_entry(char trans, i64 m, i64 n, i64 nrhs, i64 nb, i64 ib, const T *A, i64 lda, const T *Tm,
                            i64 ldt, T *C, i64 ldc, void *stream) {
    trans = up(trans);
    if (trans != 'N' && trans != 'T' && trans != 'C') return arg_error(1, "trans must be 'N', 'C', or 'T'");
    if (m < 0 || n < 0 || nrhs < 0) return arg_error(2, "dimensions must be non-negative");
    if (nb <= 0) return arg_error(5, "nb must be positive");
    if (ib <= 0 || ib > nb) return arg_error(6, "ib must be in 1..nb");
    if (m % nb || n % nb) return arg_error(2, "m and n must be multiples of nb");
    if (m == 0 || n == 0 || nrhs == 0) return 0;
    if (lda < m) return arg_error(8, "lda must be >= m");
    if (ldc < m) return arg_error(12, "ldc must be >= m");
    const i64 mt = m / nb, nt = n / nb, kt = mt < nt ? mt : nt;
    if (ldt < mt * ib) return arg_error(10, "ldt must be >= (m/nb)*ib");
    const bool notrans = trans == 'N';
    hipStream_t s = as_stream(stream);
    {
        // Q^T C / Q C by the dataflow kernel's walkers alone (C padded to a multiple of 64 columns)
        constexpr i64 TWC = 64;
        const i64 ncols = (nrhs + TWC - 1) / TWC * TWC;
        bool handled = false;
        if (ncols == nrhs) {
            const int r = tile_apply_qt_dag<T>(m, n, nb, ib, A, lda, Tm, ldt, C, ldc, ncols, s, &handled, !notrans);
            if (handled) return r < 0 ? 1 : r;
        } else {
            int st = 0;
            T *P = (T *)scratch(SLOT_RHS, sizeof(T) * (size_t)m * ncols, s, &st);
            if (!P) return st;
            NLA_HIP(hipMemsetAsync(P, 0, sizeof(T) * (size_t)m * ncols, s));
            NLA_HIP(hipMemcpy2DAsync(P, sizeof(T) * m, C, sizeof(T) * ldc, sizeof(T) * m, nrhs, hipMemcpyDeviceToDevice, s));
            const int r = tile_apply_qt_dag<T>(m, n, nb, ib, A, lda, Tm, ldt, P, m, ncols, s, &handled, !notrans);
            if (handled) {
                if (r) return r < 0 ? 1 : r;
                NLA_HIP(hipMemcpy2DAsync(C, sizeof(T) * ldc, P, sizeof(T) * m, sizeof(T) * m, nrhs, hipMemcpyDeviceToDevice, s));
                return 0;
            }
        }
    }
    for (i64 kk = 0; kk < kt; ++kk) {
        const i64 k = notrans ? kt - 1 - kk : kk;
        const T *Akk = A + k * nb + k * nb * lda;
        const T *Tkk = Tm + k * ib + k * nb * ldt;
        T *Ck = C + k * nb;
        if (!notrans) NLA_HIP(launch_ge_apply<T>(false, nb, nrhs, nb, ib, Akk, lda, Tkk, ldt, false, true, Ck, ldc, s));
        for (i64 ii = 0; ii < mt - k - 1; ++ii) {
            const i64 i = notrans ? mt - 1 - ii : k + 1 + ii;
            NLA_HIP(launch_ts_apply<T>(false, nb, nrhs, nb, ib, A + i * nb + k * nb * lda, lda, Tm + i * ib + k * nb * ldt,
                                       ldt, notrans, !notrans, Ck, ldc, C + i * nb, ldc, s));
        }
        if (notrans) NLA_HIP(launch_ge_apply<T>(false, nb, nrhs, nb, ib, Akk, lda, Tkk, ldt, true, false, Ck, ldc, s));
    }
    return 0;
}

// Float64 tile QR with Float32 trailing updates (prec 'S') or plain ('D').
static int tile_geqrf_mixed_entry(i64 m, i64 n, i64 nb, i64 ib, double *A, i64 lda, double *Tm, i64 ldt, char prec,
                                  void *stream) {
    prec = up(prec);
    if (prec != 'D' && prec != 'S') return arg_error(9, "prec must be 'D' or 'S'");
    if (prec == 'D') return tile_geqrf_entry<double>(m, n, nb, ib, A, lda, Tm, ldt, stream);
    if (m < 0) return arg_error(1, "m must be non-negative");
    if (n < 0) return arg_error(2, "n must be non-negative");
    if (nb <= 0) return arg_error(3, "nb must be positive");
    if (ib <= 0 || ib > nb) return arg_error(4, "ib must be in 1..nb");
    if (m % nb || n % nb) return arg_error(1, "m and n must be multiples of nb");
    if (m == 0 || n == 0) return 0;
    if (lda < m) return arg_error(6, "lda must be >= m");
    if (ldt < (m / nb) * ib) return arg_error(8, "ldt must be >= (m/nb)*ib");
    bool handled = false;
    const int r = tile_geqrf_dag_mixed(m, n, nb, ib, A, lda, Tm, ldt, as_stream(stream), &handled);
    if (!handled) return arg_error(3, "mixed precision needs nb in {64, 128}, ib = 32 and 16-byte aligned columns");
    return r < 0 ? 1 : r;
}

// y := y + x (n)
template <typename T> __global__ void axpy1_kernel(i64 n, const T *x, T *y) {
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] += x[i];
}

// f := b - r (m)
template <typename T> __global__ void bmr_kernel(i64 m, const T *b, const T *r, T *f) {
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) f[i] = b[i] - r[i];
}
// u := d - h (n)
template <typename T> __global__ void dmh_kernel(i64 n, const T *d, const T *h, T *u) {
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) u[i] = d[i] - h[i];
}

// Least squares min ||A x - b|| (A m x n, m >= n, full rank) through the tile QR A = Q R,
// refined on the augmented system [I A; A^T 0] [r; x] = [b; 0] (Bjorck): with the residuals
//   f = b - r - A x,  g = -A^T r   (working precision, unmodified A),
// the correction is  h = R^-T g,  d = Q^T f,  dx = R^-1 (d[0:n] - h),  dr = Q [h; d[n:m]].
// Classical refinement (x += R^-1 (Q^T (b - A x))[0:n]) would stall at the Q error times the
// least-squares residual; the augmented form converges to the least-squares solution while
// kappa(A) times the factorization's precision stays below one.  prec 'S' (Float64 data
// only) factors with Float32 trailing updates.  Stops after maxit steps or when
// ||dx|| <= tol ||x||.
template <typename T>
static int tile_gels_entry(i64 m, i64 n, i64 nb, i64 ib, const T *A, i64 lda, const T *b, T *x, char prec,
                           i64 maxit, double tol, i64 *iters, double *last, void *stream) {
    prec = up(prec);
    if (m < 0) return arg_error(1, "m must be non-negative");
    if (n < 0 || n > m) return arg_error(2, "n must be in 0..m");
    if (nb <= 0) return arg_error(3, "nb must be positive");
    if (ib <= 0 || ib > nb) return arg_error(4, "ib must be in 1..nb");
    if (m % nb || n % nb) return arg_error(1, "m and n must be multiples of nb");
    if (lda < m) return arg_error(6, "lda must be >= m");
    if (prec != 'D' && prec != 'S') return arg_error(9, "prec must be 'D' or 'S'");
    if (prec == 'S' && sizeof(T) != 8) return arg_error(9, "prec 'S' needs Float64 data");
    if (maxit < 0) return arg_error(10, "maxit must be non-negative");
    if (iters) *iters = 0;
    if (last) *last = 0;
    if (m == 0 || n == 0) return 0;
    hipStream_t s = as_stream(stream);
    const i64 mt = m / nb, ldt = mt * ib;
    constexpr i64 W = 64;  // right-hand side block width of the Q / Q^T walkers
    int st = 0;
    const size_t ne = (size_t)m * n + (size_t)ldt * n + (size_t)m * W + (size_t)m + 2 * (size_t)n;
    T *AF = (T *)scratch(SLOT_LS, sizeof(T) * ne, s, &st);
    if (!AF) return st;
    T *Tm = AF + (size_t)m * n, *Cb = Tm + (size_t)ldt * n, *r = Cb + (size_t)m * W;
    T *h = r + m, *u = h + n;
    const unsigned gm = (unsigned)((m + 255) / 256), gn = (unsigned)((n + 255) / 256);
    NLA_HIP(hipMemcpy2DAsync(AF, sizeof(T) * m, A, sizeof(T) * lda, sizeof(T) * m, n, hipMemcpyDeviceToDevice, s));
    NLA_HIP(hipMemsetAsync(Tm, 0, sizeof(T) * (size_t)ldt * n, s));
    int rc;
    if (prec == 'S') rc = tile_geqrf_mixed_entry(m, n, nb, ib, (double *)AF, m, (double *)Tm, ldt, 'S', s);
    else rc = tile_geqrf_entry<T>(m, n, nb, ib, AF, m, Tm, ldt, s);
    if (rc) return rc;
    NLA_HIP(hipMemsetAsync(Cb, 0, sizeof(T) * (size_t)m * W, s));
    // x0 = R^-1 (Q^T b)[0:n];  r0 = b - A x0
    NLA_HIP(hipMemcpyAsync(Cb, b, sizeof(T) * m, hipMemcpyDeviceToDevice, s));
    if ((rc = tile_ormqr_entry<T>('T', m, n, W, nb, ib, AF, m, Tm, ldt, Cb, m, s))) return rc;
    NLA_HIP(rectrxm_async<T>('L', 'U', 'N', T(1), 'S', n, 1, AF, m, Cb, m, 0, s));
    NLA_HIP(hipMemcpyAsync(x, Cb, sizeof(T) * n, hipMemcpyDeviceToDevice, s));
    NLA_HIP(hipMemcpyAsync(r, b, sizeof(T) * m, hipMemcpyDeviceToDevice, s));
    NLA_HIP(gemm<T>('N', 'N', m, 1, n, T(-1), A, lda, x, n, T(1), r, m, s));
    std::vector<T> hx(n), hd(n);
    double prev = 0;
    for (i64 it = 0; it < maxit; ++it) {
        // f = b - r - A x (in Cb's first column, the other columns zero); g = -A^T r (in h)
        NLA_HIP(hipMemsetAsync(Cb, 0, sizeof(T) * (size_t)m * W, s));
        bmr_kernel<T><<<gm, 256, 0, s>>>(m, b, r, Cb);
        NLA_LAUNCH_CHECK();
        NLA_HIP(gemm<T>('N', 'N', m, 1, n, T(-1), A, lda, x, n, T(1), Cb, m, s));
        NLA_HIP(gemm<T>('T', 'N', n, 1, m, T(-1), A, lda, r, m, T(0), h, n, s));
        // h = R^-T g;  d = Q^T f (in Cb);  dx = R^-1 (d[0:n] - h) (in u)
        NLA_HIP(rectrxm_async<T>('L', 'U', 'T', T(1), 'S', n, 1, AF, m, h, n, 0, s));
        if ((rc = tile_ormqr_entry<T>('T', m, n, W, nb, ib, AF, m, Tm, ldt, Cb, m, s))) return rc;
        dmh_kernel<T><<<gn, 256, 0, s>>>(n, Cb, h, u);
        NLA_LAUNCH_CHECK();
        NLA_HIP(rectrxm_async<T>('L', 'U', 'N', T(1), 'S', n, 1, AF, m, u, n, 0, s));
        // dr = Q [h; d[n:m]] (in Cb);  x += dx;  r += dr
        NLA_HIP(hipMemcpyAsync(Cb, h, sizeof(T) * n, hipMemcpyDeviceToDevice, s));
        if ((rc = tile_ormqr_entry<T>('N', m, n, W, nb, ib, AF, m, Tm, ldt, Cb, m, s))) return rc;
        axpy1_kernel<T><<<gn, 256, 0, s>>>(n, u, x);
        NLA_LAUNCH_CHECK();
        axpy1_kernel<T><<<gm, 256, 0, s>>>(m, Cb, r);
        NLA_LAUNCH_CHECK();
        NLA_HIP(hipMemcpyAsync(hd.data(), u, sizeof(T) * n, hipMemcpyDeviceToHost, s));
        NLA_HIP(hipMemcpyAsync(hx.data(), x, sizeof(T) * n, hipMemcpyDeviceToHost, s));
        NLA_HIP(hipStreamSynchronize(s));
        double nd = 0, nx = 0;
        for (i64 i = 0; i < n; ++i) nd += (double)hd[i] * hd[i], nx += (double)hx[i] * hx[i];
        const double ratio = nx > 0 ? std::sqrt(nd / nx) : std::sqrt(nd);
        if (iters) *iters = it + 1;
        if (last) *last = ratio;
        if (ratio <= tol) break;
        if (it > 0 && ratio > 0.5 * prev) break;  // stagnation: the working precision's floor
        prev = ratio;
    }
    NLA_HIP(hipStreamSynchronize(s));
    return 0;
}

}  // namespace nla

NLA_API int nla_dtile_geqrf_mixed(int64_t m, int64_t n, int64_t nb, int64_t ib, double *A, int64_t lda, double *T,
                                  int64_t ldt, char prec, void *stream) {
    return nla::tile_geqrf_mixed_entry(m, n, nb, ib, A, lda, T, ldt, prec, stream);
}
NLA_API int nla_dtile_gels(int64_t m, int64_t n, int64_t nb, int64_t ib, const double *A, int64_t lda, const double *b,
                           double *x, char prec, int64_t maxit, double tol, int64_t *iters, double *last_correction,
                           void *stream) {
    return nla::tile_gels_entry<double>(m, n, nb, ib, A, lda, b, x, prec, maxit, tol, iters, last_correction, stream);
}
NLA_API int nla_stile_gels(int64_t m, int64_t n, int64_t nb, int64_t ib, const float *A, int64_t lda, const float *b,
                           float *x, char prec, int64_t maxit, double tol, int64_t *iters, double *last_correction,
                           void *stream) {
    return nla::tile_gels_entry<float>(m, n, nb, ib, A, lda, b, x, prec, maxit, tol, iters, last_correction, stream);
}

NLA_API int nla_dtile_geqrf(int64_t m, int64_t n, int64_t nb, int64_t ib, double *A, int64_t lda, double *T,
                            int64_t ldt, void *stream) {
    return nla::tile_geqrf_entry<double>(m, n, nb, ib, A, lda, T, ldt, stream);
}
NLA_API int nla_stile_geqrf(int64_t m, int64_t n, int64_t nb, int64_t ib, float *A, int64_t lda, float *T, int64_t ldt,
                            void *stream) {
    return nla::tile_geqrf_entry<float>(m, n, nb, ib, A, lda, T, ldt, stream);
}
NLA_API int nla_dtile_ormqr(char trans, int64_t m, int64_t n, int64_t nrhs, int64_t nb, int64_t ib, const double *A,
                            int64_t lda, const double *T, int64_t ldt, double *C, int64_t ldc, void *stream) {
    return nla::tile_ormqr_entry<double>(trans, m, n, nrhs, nb, ib, A, lda, T, ldt, C, ldc, stream);
}
NLA_API int nla_stile_ormqr(char trans, int64_t m, int64_t n, int64_t nrhs, int64_t nb, int64_t ib, const float *A,
                            int64_t lda, const float *T, int64_t ldt, float *C, int64_t ldc, void *stream) {
    return nla::tile_ormqr_entry<float>(trans, m, n, nrhs, nb, ib, A, lda, T, ldt, C, ldc, stream);
}
