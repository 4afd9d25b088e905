// Tile-QR driver: flat-tree PLASMA-order composition of geqrt!/unmqr!/tsqrt!/tsmqr!
// (no reference function: the reference leaves this loop to its caller).
//   for k: geqrt(A_kk); unmqr('L','C', A_kk, A_k,j>k); for i>k: tsqrt(A_kk, A_ik);
//          tsmqr('L','C', A_k,j>k, A_i,j>k, V=A_ik, T_ik)
// Each unmqr/tsmqr step is one strip launch over the whole trailing tile row.
#include "nla_common.hpp"
#include "nla_qr.hpp"

namespace nla {

template <typename T>
int tile_geqrf_dag(i64 m, i64 n, i64 nb, i64 ib, T *A, i64 lda, T *Tm, i64 ldt, hipStream_t s, bool *handled);

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

}  // namespace nla

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
