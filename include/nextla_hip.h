/*
 * nextla_hip.h — C ABI of libnextla_hip.so, the MI355X (gfx950) implementation of
 * NextLA.jl's tile-QR kernels and recursive TRSM/TRMM.
 *
 * Conventions (all entry points):
 *  - Matrices are column-major DEVICE pointers with an explicit leading dimension,
 *    exactly what a Julia view with stride(A,1) == 1 hands over (pointer(A), stride(A,2)).
 *  - Sizes are int64_t.  Char flags take the reference's values ('L'/'R', 'N'/'C'/'T', ...).
 *    For real types 'T' means the same as 'C' (src/larfb.jl:79-85, src/parfb.jl:96-100).
 *  - `stream` is a hipStream_t passed as void*; NULL means the default stream.  Every
 *    call is stream-ordered.  Calls return without synchronising unless their comment says
 *    "Synchronises the stream"; a caller that needs the reference's synchronous semantics
 *    calls nla_synchronize(stream).
 *  - Return status: 0 = success; -i = the i-th argument is illegal (the reference would
 *    throw ArgumentError; nla_last_error() returns the reference's message text);
 *    NLA_STATUS_WAIT_TIMEOUT = a device-side dependency wait of a persistent launch ran
 *    out of time (results invalid); any other value > 0 = a HIP runtime error code.
 *  - The caller owns every buffer.  Workspace arguments accept the reference's sizes;
 *    the library never needs more than the reference asks for, and it keeps any extra
 *    device scratch in a pool of its own per (device, stream), so calls on different
 *    streams may run concurrently.
 *  - Quick returns and k clamping are those of the reference routine cited on each line.
 * 'd' = Float64, 's' = Float32.
 */
#ifndef NEXTLA_HIP_H
#define NEXTLA_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NLA_STATUS_WAIT_TIMEOUT 10001

/* Message of the last non-zero status on this thread (the reference's ArgumentError text). */
const char *nla_last_error(void);
/* Library ABI version (major*100 + minor). */
int nla_version(void);
/* hipStreamSynchronize(stream); returns a HIP error code or 0. */
int nla_synchronize(void *stream);
/* Bytes of device scratch the library holds on the current device. */
int64_t nla_scratch_bytes(void);

/* Launch profiler (no reference counterpart; measurement hook for bench.py): while enabled,
 * each kernel family brackets its launches with hipEvents on the launch stream.
 * nla_prof_read sums elapsed time and launch count for one family: "gemm" (every GEMM
 * launch), "gemm128" (the 128 x 128-tile GEMM launches alone), "trsm_base",
 * "trmm_base", "ge_apply", "ts_apply", "panel_qr", "ts_panel", "tile_qr". */
int nla_prof_enable(int on);
void nla_prof_reset(void);
int nla_prof_read(const char *kernel, double *total_ms, int64_t *launches);

/* ---- Element types of the QR-family routines ------------------------------------------
 * Each routine below has four entry points: 'd' Float64, 's' Float32, 'z' ComplexF64,
 * 'c' ComplexF32.  Complex elements are two reals, real part first (Julia's Complex{T},
 * numpy's complex128/complex64).  For complex types 'C' means the conjugate transpose and
 * 'T' is rejected; for real types 'T' means the same as 'C'. */
typedef struct nla_zcomplex { double re, im; } nla_zcomplex;
typedef struct nla_ccomplex { float re, im; } nla_ccomplex;

/* ---- Householder generation ------------------------------------------------------ */
/* larfg!(n, alpha, x, incx, tau): src/larfg.jl:43-106.  alpha and tau are DEVICE scalars;
 * on return *alpha = beta, x = v[2:n], *tau = tau.  (Julia shim returns (alpha, tau).) */
int nla_dlarfg(int64_t n, double *alpha, double *x, int64_t incx, double *tau, void *stream);
int nla_slarfg(int64_t n, float *alpha, float *x, int64_t incx, float *tau, void *stream);
int nla_zlarfg(int64_t n, nla_zcomplex *alpha, nla_zcomplex *x, int64_t incx, nla_zcomplex *tau, void *stream);
int nla_clarfg(int64_t n, nla_ccomplex *alpha, nla_ccomplex *x, int64_t incx, nla_ccomplex *tau, void *stream);

/* ---- Panel QR and T factor --------------------------------------------------------- */
/* geqr2!(m, n, A, tau, work): src/geqr2.jl:107-121.  work (length n) is accepted, unused. */
int nla_dgeqr2(int64_t m, int64_t n, double *A, int64_t lda, double *tau, double *work, void *stream);
int nla_sgeqr2(int64_t m, int64_t n, float *A, int64_t lda, float *tau, float *work, void *stream);
int nla_zgeqr2(int64_t m, int64_t n, nla_zcomplex *A, int64_t lda, nla_zcomplex *tau, nla_zcomplex *work, void *stream);
int nla_cgeqr2(int64_t m, int64_t n, nla_ccomplex *A, int64_t lda, nla_ccomplex *tau, nla_ccomplex *work, void *stream);

/* larft!(direct, storev, n, k, V, tau, T): src/larft.jl:205-213.  direct='F' with storev 'C'
 * (V n x k) or 'R' (V k x n); direct='B' returns -1. */
int nla_dlarft(char direct, char storev, int64_t n, int64_t k, const double *V, int64_t ldv,
               const double *tau, double *T, int64_t ldt, void *stream);
int nla_slarft(char direct, char storev, int64_t n, int64_t k, const float *V, int64_t ldv,
               const float *tau, float *T, int64_t ldt, void *stream);
int nla_zlarft(char direct, char storev, int64_t n, int64_t k, const nla_zcomplex *V, int64_t ldv,
               const nla_zcomplex *tau, nla_zcomplex *T, int64_t ldt, void *stream);
int nla_clarft(char direct, char storev, int64_t n, int64_t k, const nla_ccomplex *V, int64_t ldv,
               const nla_ccomplex *tau, nla_ccomplex *T, int64_t ldt, void *stream);

/* geqrt!(m, n, ib, A, T, tau, work): src/geqrt.jl:22-68.  T is ib x min(m,n) (ldt >= ib). */
int nla_dgeqrt(int64_t m, int64_t n, int64_t ib, double *A, int64_t lda, double *T, int64_t ldt,
               double *tau, double *work, void *stream);
int nla_sgeqrt(int64_t m, int64_t n, int64_t ib, float *A, int64_t lda, float *T, int64_t ldt,
               float *tau, float *work, void *stream);
int nla_zgeqrt(int64_t m, int64_t n, int64_t ib, nla_zcomplex *A, int64_t lda, nla_zcomplex *T, int64_t ldt,
               nla_zcomplex *tau, nla_zcomplex *work, void *stream);
int nla_cgeqrt(int64_t m, int64_t n, int64_t ib, nla_ccomplex *A, int64_t lda, nla_ccomplex *T, int64_t ldt,
               nla_ccomplex *tau, nla_ccomplex *work, void *stream);

/* ---- Block-reflector application ----------------------------------------------------- */
/* larfb!(side, trans, direct, storev, m, n, k, V, ldv, T, C, work): src/larfb.jl:37-142.
 * direct='F', storev='C' only, as in the reference (src/larfb.jl:42-44).  Blocks wider than
 * 64 reflectors are applied as 64-wide sub-blocks, which assumes T is larft!'s factor of V. */
int nla_dlarfb(char side, char trans, char direct, char storev, int64_t m, int64_t n, int64_t k,
               const double *V, int64_t ldv, const double *T, int64_t ldt, double *C, int64_t ldc,
               double *work, int64_t ldwork, void *stream);
int nla_slarfb(char side, char trans, char direct, char storev, int64_t m, int64_t n, int64_t k,
               const float *V, int64_t ldv, const float *T, int64_t ldt, float *C, int64_t ldc,
               float *work, int64_t ldwork, void *stream);
int nla_zlarfb(char side, char trans, char direct, char storev, int64_t m, int64_t n, int64_t k,
               const nla_zcomplex *V, int64_t ldv, const nla_zcomplex *T, int64_t ldt, nla_zcomplex *C, int64_t ldc,
               nla_zcomplex *work, int64_t ldwork, void *stream);
int nla_clarfb(char side, char trans, char direct, char storev, int64_t m, int64_t n, int64_t k,
               const nla_ccomplex *V, int64_t ldv, const nla_ccomplex *T, int64_t ldt, nla_ccomplex *C, int64_t ldc,
               nla_ccomplex *work, int64_t ldwork, void *stream);

/* unmqr!(side, trans, m, n, k, ib, A, lda, T, C, work): src/unmqr.jl:52-149.
 * a_cols = size(A,2), used by the clamp k = min(k, nq, size(A,2)) (src/unmqr.jl:83). */
int nla_dunmqr(char side, char trans, int64_t m, int64_t n, int64_t k, int64_t ib, const double *A,
               int64_t lda, int64_t a_cols, const double *T, int64_t ldt, double *C, int64_t ldc,
               double *work, void *stream);
int nla_sunmqr(char side, char trans, int64_t m, int64_t n, int64_t k, int64_t ib, const float *A,
               int64_t lda, int64_t a_cols, const float *T, int64_t ldt, float *C, int64_t ldc,
               float *work, void *stream);
int nla_zunmqr(char side, char trans, int64_t m, int64_t n, int64_t k, int64_t ib, const nla_zcomplex *A,
               int64_t lda, int64_t a_cols, const nla_zcomplex *T, int64_t ldt, nla_zcomplex *C, int64_t ldc,
               nla_zcomplex *work, void *stream);
int nla_cunmqr(char side, char trans, int64_t m, int64_t n, int64_t k, int64_t ib, const nla_ccomplex *A,
               int64_t lda, int64_t a_cols, const nla_ccomplex *T, int64_t ldt, nla_ccomplex *C, int64_t ldc,
               nla_ccomplex *work, void *stream);

/* ---- Triangle-on-square and triangle-on-triangle kernels ---------------------------- */
/* tsqrt!(m, n, ib, A1, A2, T, tau, work): src/tsqrt.jl:128-156.  m1 = size(A1,1);
 * k = min(m1, n) reflectors (src/tsqrt.jl:139-140). */
int nla_dtsqrt(int64_t m, int64_t n, int64_t ib, double *A1, int64_t lda1, int64_t m1, double *A2,
               int64_t lda2, double *T, int64_t ldt, double *tau, double *work, void *stream);
int nla_stsqrt(int64_t m, int64_t n, int64_t ib, float *A1, int64_t lda1, int64_t m1, float *A2,
               int64_t lda2, float *T, int64_t ldt, float *tau, float *work, void *stream);
int nla_ztsqrt(int64_t m, int64_t n, int64_t ib, nla_zcomplex *A1, int64_t lda1, int64_t m1, nla_zcomplex *A2,
               int64_t lda2, nla_zcomplex *T, int64_t ldt, nla_zcomplex *tau, nla_zcomplex *work, void *stream);
int nla_ctsqrt(int64_t m, int64_t n, int64_t ib, nla_ccomplex *A1, int64_t lda1, int64_t m1, nla_ccomplex *A2,
               int64_t lda2, nla_ccomplex *T, int64_t ldt, nla_ccomplex *tau, nla_ccomplex *work, void *stream);

/* tsmqr!(side, trans, m1, n1, m2, n2, k, ib, A1, A2, V, T, work): src/tsmqr.jl:47-146.
 * v_cols = size(V,2), t_cols = size(T,2) for the clamp at src/tsmqr.jl:94-96.  ib <= 256. */
int nla_dtsmqr(char side, char trans, int64_t m1, int64_t n1, int64_t m2, int64_t n2, int64_t k,
               int64_t ib, double *A1, int64_t lda1, double *A2, int64_t lda2, const double *V, int64_t ldv,
               int64_t v_cols, const double *T, int64_t ldt, int64_t t_cols, double *work, void *stream);
int nla_stsmqr(char side, char trans, int64_t m1, int64_t n1, int64_t m2, int64_t n2, int64_t k,
               int64_t ib, float *A1, int64_t lda1, float *A2, int64_t lda2, const float *V, int64_t ldv,
               int64_t v_cols, const float *T, int64_t ldt, int64_t t_cols, float *work, void *stream);
int nla_ztsmqr(char side, char trans, int64_t m1, int64_t n1, int64_t m2, int64_t n2, int64_t k,
               int64_t ib, nla_zcomplex *A1, int64_t lda1, nla_zcomplex *A2, int64_t lda2, const nla_zcomplex *V, int64_t ldv,
               int64_t v_cols, const nla_zcomplex *T, int64_t ldt, int64_t t_cols, nla_zcomplex *work, void *stream);
int nla_ctsmqr(char side, char trans, int64_t m1, int64_t n1, int64_t m2, int64_t n2, int64_t k,
               int64_t ib, nla_ccomplex *A1, int64_t lda1, nla_ccomplex *A2, int64_t lda2, const nla_ccomplex *V, int64_t ldv,
               int64_t v_cols, const nla_ccomplex *T, int64_t ldt, int64_t t_cols, nla_ccomplex *work, void *stream);

/* ttqrt!(m, n, ib, A1, A2, T, tau, work): src/ttqrt.jl:81-118.  QR of [A1; A2] with both
 * m x n upper triangular (trapezoidal); column r's reflector has rows 0..min(r+1, m)-1 of A2. */
int nla_dttqrt(int64_t m, int64_t n, int64_t ib, double *A1, int64_t lda1, double *A2, int64_t lda2,
               double *T, int64_t ldt, double *tau, double *work, void *stream);
int nla_sttqrt(int64_t m, int64_t n, int64_t ib, float *A1, int64_t lda1, float *A2, int64_t lda2,
               float *T, int64_t ldt, float *tau, float *work, void *stream);
int nla_zttqrt(int64_t m, int64_t n, int64_t ib, nla_zcomplex *A1, int64_t lda1, nla_zcomplex *A2, int64_t lda2,
               nla_zcomplex *T, int64_t ldt, nla_zcomplex *tau, nla_zcomplex *work, void *stream);
int nla_cttqrt(int64_t m, int64_t n, int64_t ib, nla_ccomplex *A1, int64_t lda1, nla_ccomplex *A2, int64_t lda2,
               nla_ccomplex *T, int64_t ldt, nla_ccomplex *tau, nla_ccomplex *work, void *stream);

/* ttmqr!(side, trans, m1, n1, m2, n2, k, ib, A1, A2, V, T, work): src/ttmqr.jl:1-95 (block i
 * uses the first min(i+kb, m2) rows -- columns for side 'R' -- of A2 and V, pentagon order
 * min(kb, m2 - i - 1) as src/ttmqr.jl:70-74, 83-87).  ib <= 256. */
int nla_dttmqr(char side, char trans, int64_t m1, int64_t n1, int64_t m2, int64_t n2, int64_t k,
               int64_t ib, double *A1, int64_t lda1, double *A2, int64_t lda2, const double *V, int64_t ldv,
               const double *T, int64_t ldt, double *work, void *stream);
int nla_sttmqr(char side, char trans, int64_t m1, int64_t n1, int64_t m2, int64_t n2, int64_t k,
               int64_t ib, float *A1, int64_t lda1, float *A2, int64_t lda2, const float *V, int64_t ldv,
               const float *T, int64_t ldt, float *work, void *stream);
int nla_zttmqr(char side, char trans, int64_t m1, int64_t n1, int64_t m2, int64_t n2, int64_t k,
               int64_t ib, nla_zcomplex *A1, int64_t lda1, nla_zcomplex *A2, int64_t lda2, const nla_zcomplex *V, int64_t ldv,
               const nla_zcomplex *T, int64_t ldt, nla_zcomplex *work, void *stream);
int nla_cttmqr(char side, char trans, int64_t m1, int64_t n1, int64_t m2, int64_t n2, int64_t k,
               int64_t ib, nla_ccomplex *A1, int64_t lda1, nla_ccomplex *A2, int64_t lda2, const nla_ccomplex *V, int64_t ldv,
               const nla_ccomplex *T, int64_t ldt, nla_ccomplex *work, void *stream);

/* parfb!(side, trans, direct, storev, m1, n1, m2, n2, k, l, A1, A2, V, T, work):
 * src/parfb.jl:44-191.  Every side / trans / direct / storev with pentagon order l; applies
 * I - V T V^H exactly for any T of the stored triangle (upper 'F', lower 'B').  k <= 256. */
int nla_dparfb(char side, char trans, char direct, char storev, int64_t m1, int64_t n1, int64_t m2,
               int64_t n2, int64_t k, int64_t l, double *A1, int64_t lda1, double *A2, int64_t lda2,
               const double *V, int64_t ldv, const double *T, int64_t ldt, double *work, int64_t ldwork,
               void *stream);
int nla_sparfb(char side, char trans, char direct, char storev, int64_t m1, int64_t n1, int64_t m2,
               int64_t n2, int64_t k, int64_t l, float *A1, int64_t lda1, float *A2, int64_t lda2,
               const float *V, int64_t ldv, const float *T, int64_t ldt, float *work, int64_t ldwork,
               void *stream);
int nla_zparfb(char side, char trans, char direct, char storev, int64_t m1, int64_t n1, int64_t m2,
               int64_t n2, int64_t k, int64_t l, nla_zcomplex *A1, int64_t lda1, nla_zcomplex *A2, int64_t lda2,
               const nla_zcomplex *V, int64_t ldv, const nla_zcomplex *T, int64_t ldt, nla_zcomplex *work, int64_t ldwork,
               void *stream);
int nla_cparfb(char side, char trans, char direct, char storev, int64_t m1, int64_t n1, int64_t m2,
               int64_t n2, int64_t k, int64_t l, nla_ccomplex *A1, int64_t lda1, nla_ccomplex *A2, int64_t lda2,
               const nla_ccomplex *V, int64_t ldv, const nla_ccomplex *T, int64_t ldt, nla_ccomplex *work, int64_t ldwork,
               void *stream);

/* pamm!(op, side, storev, direct, m, n, k, l, A1, A2, V, W): src/pamm.jl:49-100.  op 'W':
 * W = A1 + op(V) A2 (side L, W k x n) / A1 + A2 op(V) (side R, W m x k); op 'A':
 * A2 -= op(V) W / W op(V); m x n = size(A2); V pentagonal of order l.  k <= 256. */
int nla_dpamm(char op, char side, char storev, char direct, int64_t m, int64_t n, int64_t k,
              int64_t l, const double *A1, int64_t lda1, double *A2, int64_t lda2, const double *V, int64_t ldv,
              double *W, int64_t ldw, void *stream);
int nla_spamm(char op, char side, char storev, char direct, int64_t m, int64_t n, int64_t k,
              int64_t l, const float *A1, int64_t lda1, float *A2, int64_t lda2, const float *V, int64_t ldv,
              float *W, int64_t ldw, void *stream);
int nla_zpamm(char op, char side, char storev, char direct, int64_t m, int64_t n, int64_t k,
              int64_t l, const nla_zcomplex *A1, int64_t lda1, nla_zcomplex *A2, int64_t lda2, const nla_zcomplex *V, int64_t ldv,
              nla_zcomplex *W, int64_t ldw, void *stream);
int nla_cpamm(char op, char side, char storev, char direct, int64_t m, int64_t n, int64_t k,
              int64_t l, const nla_ccomplex *A1, int64_t lda1, nla_ccomplex *A2, int64_t lda2, const nla_ccomplex *V, int64_t ldv,
              nla_ccomplex *W, int64_t ldw, void *stream);

/* pemv!(trans, storev, m, n, l, alpha, A, X, beta, Y, work): src/pemv.jl:39-181.
 * storev 'C' with trans 'T'/'C': Y (n) = alpha op(A) X + beta Y; storev 'R' with trans 'N':
 * Y (m) = alpha A X + beta Y; A m x n pentagonal of order l (l = 1 counts as 0).  alpha and
 * beta are passed by value; beta = 0 does not read Y; alpha = beta = 0 leaves Y unchanged.
 * work is accepted, unused. */
int nla_dpemv(char trans, char storev, int64_t m, int64_t n, int64_t l, double alpha, const double *A,
              int64_t lda, const double *x, double beta, double *y, double *work, void *stream);
int nla_spemv(char trans, char storev, int64_t m, int64_t n, int64_t l, float alpha, const float *A,
              int64_t lda, const float *x, float beta, float *y, float *work, void *stream);
int nla_zpemv(char trans, char storev, int64_t m, int64_t n, int64_t l, nla_zcomplex alpha, const nla_zcomplex *A,
              int64_t lda, const nla_zcomplex *x, nla_zcomplex beta, nla_zcomplex *y, nla_zcomplex *work, void *stream);
int nla_cpemv(char trans, char storev, int64_t m, int64_t n, int64_t l, nla_ccomplex alpha, const nla_ccomplex *A,
              int64_t lda, const nla_ccomplex *x, nla_ccomplex beta, nla_ccomplex *y, nla_ccomplex *work, void *stream);

/* ---- Recursive TRSM / TRMM ------------------------------------------------------ */
/* unified_rectrxm!(side, uplo, trans, alpha, func, A, B): src/rectrxm.jl:43-76.
 * func 'S' = solve op(A) X = alpha B / X op(A) = alpha B, 'M' = B := alpha op(A) B / alpha B op(A).
 * n = size(A,1); m = the other dimension of B (B is n x m for side 'L', m x n for 'R').
 * cutoff = recursion threshold (src/rectrxm.jl:52,63); 0 selects this library's default 256
 * for both functions (the reference's thresholds are 256 for 'S' and 16 for 'M'; 16 is still
 * accepted).  With a cut-off in [256, 512) the base blocks of 256 rows run as MFMA triangular
 * products.  Only the uplo triangle of A is read. */
int nla_drectrxm(char side, char uplo, char trans, double alpha, char func, int64_t n, int64_t m,
                 const double *A, int64_t lda, double *B, int64_t ldb, int64_t cutoff, void *stream);
int nla_srectrxm(char side, char uplo, char trans, float alpha, char func, int64_t n, int64_t m,
                 const float *A, int64_t lda, float *B, int64_t ldb, int64_t cutoff, void *stream);

/* No reference counterpart: nla_drectrxm for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results as nla_drectrxm; it calls no hipMalloc/hipFree and allocates nothing:
 * its workspace is a library-owned 18 MiB device array per GPU (held until the process ends),
 * enough for n <= 8192 with the default cut-off 256 (n <= 73728 with other cut-offs).  A call
 * that would need more returns -6 with a message naming the sizes.  All calls on one GPU share
 * that array: calls that overlap on one GPU (e.g. on two streams) race. */
int nla_drectrxm_noalloc(char side, char uplo, char trans, double alpha, char func, int64_t n, int64_t m,
                         const double *A, int64_t lda, double *B, int64_t ldb, int64_t cutoff, void *stream);

/* Host-pointer forms (NextLAMatrix wraps a host Matrix, src/NextLAMatrix.jl:4-32; the corrected
 * 4-Char BLAS.trsm!/trmm! forwards of src/NextLAMatrix.jl:27-28 call these): A and B are HOST
 * column-major arrays; they are copied to the device, solved / multiplied as nla_?rectrxm and
 * B is copied back.  Synchronises the stream. */
int nla_hdrectrxm(char side, char uplo, char trans, double alpha, char func, int64_t n, int64_t m,
                  const double *A, int64_t lda, double *B, int64_t ldb, int64_t cutoff, void *stream);
int nla_hsrectrxm(char side, char uplo, char trans, float alpha, char func, int64_t n, int64_t m,
                  const float *A, int64_t lda, float *B, int64_t ldb, int64_t cutoff, void *stream);

/* Base-case kernels LeftLowerTRSM!/LeftUpperTRSM!/RightLowerTRSM!/RightUpperTRSM!
 * (src/trsm.jl:128-150) and the TRMM counterparts (src/trmm.jl:332-389): one diagonal
 * block, B n x m (side 'L') or m x n (side 'R'); n <= 1024 as in src/trsm.jl:10-12. */
int nla_dtrsm_base(char side, char uplo, int64_t n, int64_t m, const double *A, int64_t lda, double *B,
                   int64_t ldb, void *stream);
int nla_strsm_base(char side, char uplo, int64_t n, int64_t m, const float *A, int64_t lda, float *B,
                   int64_t ldb, void *stream);
int nla_dtrmm_base(char side, char uplo, int64_t n, int64_t m, const double *A, int64_t lda, double *B,
                   int64_t ldb, void *stream);
int nla_strmm_base(char side, char uplo, int64_t n, int64_t m, const float *A, int64_t lda, float *B,
                   int64_t ldb, void *stream);

/* GEMM_ADD!(A,B,C): C += A*B and GEMM_SUB!(A,B,C): A -= B*C (src/matmul.jl:69-81) as one
 * primitive: C[N x M] += alpha * X[N x R] * Y[R x M].
 * GEMM_ADD!(A,B,C) = nla_dgemm_acc(N=size(A,1), R=size(B,1), M=size(B,2),  1, A, B, C);
 * GEMM_SUB!(A,B,C) = nla_dgemm_acc(N=size(A,1), R=size(C,1), M=size(A,2), -1, B, C, A). */
int nla_dgemm_acc(int64_t N, int64_t R, int64_t M, double alpha, const double *X, int64_t ldx,
                  const double *Y, int64_t ldy, double *C, int64_t ldc, void *stream);
int nla_sgemm_acc(int64_t N, int64_t R, int64_t M, float alpha, const float *X, int64_t ldx,
                  const float *Y, int64_t ldy, float *C, int64_t ldc, void *stream);

/* ---- Recursive LU (src/lu.jl) --------------------------------------------------------- */
/* getrf2!(A, ipiv, info): src/lu.jl:185-300 (LAPACK xgetrf2's recursion, partial pivoting).
 * A (m x n) is overwritten with L (unit lower, below the diagonal) and U; ipiv (DEVICE,
 * min(m,n) entries) receives the 1-based pivot rows; info (DEVICE, one entry) receives the
 * first i with U(i,i) exactly zero (1-based), 0 if none.  The off-diagonal updates run on
 * nla_?rectrxm's TRSM and the MFMA GEMM.  Leaves of more than 2048 rows (Float32: 4096) run on
 * several workgroups with a bounded wait between them; then the call synchronises the stream
 * to read its timeout flag. */
int nla_dgetrf2(int64_t m, int64_t n, double *A, int64_t lda, int64_t *ipiv, int64_t *info, void *stream);
int nla_sgetrf2(int64_t m, int64_t n, float *A, int64_t lda, int64_t *ipiv, int64_t *info, void *stream);
/* laswp(A, first, last, ipiv, incx): src/lu.jl:470-522.  Rows first..last (1-based) of the n
 * columns of A interchanged with rows ipiv(k) (DEVICE, 1-based), in order for incx > 0 and in
 * reverse for incx < 0; incx = 0 does nothing. */
int nla_dlaswp(int64_t n, double *A, int64_t lda, int64_t k1, int64_t k2, const int64_t *ipiv, int64_t incx,
               void *stream);
int nla_slaswp(int64_t n, float *A, int64_t lda, int64_t k1, int64_t k2, const int64_t *ipiv, int64_t incx,
               void *stream);

/* ---- Tile-QR driver (no reference function: the composition of geqrt!/unmqr!/tsqrt!/
 * tsmqr! that config "tile QR" describes, flat-tree PLASMA order) ------------------------- */
/* Factor A (m x n, both multiples of nb) in nb x nb tiles with inner block ib.
 * T receives one ib x nb block per tile: tile (i,k) at T[i*ib, k*nb], ldt >= (m/nb)*ib.
 * With ib = 32, nb in {64, 128} and 16-byte aligned columns the factorization is one
 * persistent launch that synchronises the stream (it reads the launch's wait-timeout flag);
 * other shapes are enqueued as per-step launches without synchronising. */
int nla_dtile_geqrf(int64_t m, int64_t n, int64_t nb, int64_t ib, double *A, int64_t lda, double *T,
                    int64_t ldt, void *stream);
int nla_stile_geqrf(int64_t m, int64_t n, int64_t nb, int64_t ib, float *A, int64_t lda, float *T,
                    int64_t ldt, void *stream);
/* The task list (claim order) nla_?tile_geqrf's persistent launch uses for an m x n matrix
 * with nb in {64, 128}: code = type << 30 | k << 18 | s, type 0 = PANEL(step k, block stage s),
 * type 1 = WALK(step k, 64-column strip group s).  Writes min(count, cap) codes to the host
 * array `codes` and returns the count, or -1 for a shape the persistent launch does not take.
 * (nla_dtile_geqrf_mixed with nb = 128 uses the same order over 128-column groups.)
 * Host-only; exposed so the order can be checked against the tasks' dependencies. */
int64_t nla_tile_qr_schedule(int64_t m, int64_t n, int64_t nb, int32_t *codes, int64_t cap);
/* Apply Q (trans 'N') or Q^T ('T'/'C') of nla_?tile_geqrf from the left to C (m x nrhs).
 * Synchronises the stream when the factorization's shape takes the persistent launch. */
int nla_dtile_ormqr(char trans, int64_t m, int64_t n, int64_t nrhs, int64_t nb, int64_t ib,
                    const double *A, int64_t lda, const double *T, int64_t ldt, double *C, int64_t ldc,
                    void *stream);
int nla_stile_ormqr(char trans, int64_t m, int64_t n, int64_t nrhs, int64_t nb, int64_t ib,
                    const float *A, int64_t lda, const float *T, int64_t ldt, float *C, int64_t ldc,
                    void *stream);

/* The same factorization (same A and T layout as nla_?tile_geqrf) issued the way the
 * reference's callers issue it: the flat-tree loop of one geqrt!/unmqr!/tsqrt!/tsmqr! call per
 * tile or tile pair (about (m/nb)(n/nb)^2/3 tsmqr! calls), each through the per-call entry
 * point above (nla_?geqrt, nla_?unmqr, nla_?tsqrt, nla_?tsmqr), enqueued without
 * synchronising.  No reference function: it is the caller loop of the reference's tile-QR
 * use, provided to run and time that call sequence natively. */
int nla_dtile_geqrf_calls(int64_t m, int64_t n, int64_t nb, int64_t ib, double *A, int64_t lda, double *T,
                          int64_t ldt, void *stream);
int nla_stile_geqrf_calls(int64_t m, int64_t n, int64_t nb, int64_t ib, float *A, int64_t lda, float *T,
                          int64_t ldt, void *stream);
int nla_ztile_geqrf_calls(int64_t m, int64_t n, int64_t nb, int64_t ib, nla_zcomplex *A, int64_t lda,
                          nla_zcomplex *T, int64_t ldt, void *stream);
int nla_ctile_geqrf_calls(int64_t m, int64_t n, int64_t nb, int64_t ib, nla_ccomplex *A, int64_t lda,
                          nla_ccomplex *T, int64_t ldt, void *stream);

/* ---- Mixed precision and least squares (no reference function: config "mixed-precision
 * tile QR + least squares"; the reference stops at the tile kernels) --------------------- */
/* nla_dtile_geqrf with the trailing updates (unmqr!/tsmqr!) on MFMA with Float32
 * accumulation: prec 'S' Float32 operands, 'H' FP16 operands, 'B' BF16 operands, 'X' split
 * BF16 (each operand as two BF16 parts, hi = bf16(x), lo = bf16(x - hi), three products
 * hi*hi + hi*lo + lo*hi: about 2^-16 per product at three BF16 products' cost) (the strips
 * stay Float32 between products; V, T and each product's operands are rounded); the panels,
 * larfg norms/scaling and the T factors stay Float64.  prec 'D' is nla_dtile_geqrf.  'H'
 * assumes the matrix's entries are within FP16 range.  Needs nb in {64, 128}, ib = 32 and
 * 16-byte aligned columns.  Holds in the device scratch pool a Float32 copy of the trailing
 * matrix (4 lda n bytes) and the panels' per-block V/T images (~21 KB per ib-block of each
 * panel tile; 'X': ~42 KB).  Synchronises the stream. */
int nla_dtile_geqrf_mixed(int64_t m, int64_t n, int64_t nb, int64_t ib, double *A, int64_t lda, double *T,
                          int64_t ldt, char prec, void *stream);
/* min ||A x - b|| for A m x n (m >= n, multiples of nb, full rank), A and b unchanged, x (n)
 * out.  Tile QR of a copy of A (prec 'D' plain; 'S' / 'H' / 'B' / 'X' as nla_dtile_geqrf_mixed,
 * Float64 data only), the direct solution x0 = R^-1 (Q^T b)[0:n] (returned for maxit = 0; with
 * prec 'D' and maxit > 0, x0 = R^-1 R^-T (A^T b), the seminormal equations), then up to maxit
 * iterations of LSQR (Paige & Saunders) on the correction, right-preconditioned by R:
 * min ||(A R^-1) y - (b - A x0)||, x = x0 + R^-1 y, with every product by A in the working
 * precision.  Stops when LSQR's estimate of ||M^T r|| / (||M||_F ||r||) (M = A R^-1) reaches
 * tol or the residual estimate falls to tol ||b - A x0||.  With a 16-bit factor ('H', 'B', 'X')
 * the contraction of that estimate is watched from iteration 12 on: when the iterations
 * still needed at the observed rate exceed what is left of maxit, A is factored again with
 * Float32 updates and LSQR restarts from that factor's direct solution.  iters / last_correction / prec_used
 * (may be NULL) receive the iterations taken over all restarts, the final estimate and the
 * update precision of the factor that produced x.  Synchronises the stream. */
int nla_dtile_gels(int64_t m, int64_t n, int64_t nb, int64_t ib, const double *A, int64_t lda, const double *b,
                   double *x, char prec, int64_t maxit, double tol, int64_t *iters, double *last_correction,
                   char *prec_used, void *stream);
int nla_stile_gels(int64_t m, int64_t n, int64_t nb, int64_t ib, const float *A, int64_t lda, const float *b,
                   float *x, char prec, int64_t maxit, double tol, int64_t *iters, double *last_correction,
                   char *prec_used, void *stream);
/* Triangular solve op(A) X = alpha B / X op(A) = alpha B (arguments as nla_drectrxm with
 * func 'S'; X overwrites B) in Float32 storage with Float64 iterative refinement: X0 from the
 * Float32 recursion (cutoff as there) whose GEMM updates take Float32 (prec 'S'), FP16 ('H')
 * or BF16 ('B') operands with Float32 accumulation (diagonal blocks stay Float32), then up to
 * maxit steps X += op(A)^-1 (alpha B - op(A) X) with the residual in Float64 and the correction
 * solved as X0 was, stopping when ||dX|| <= tol ||X|| or when a step shrinks ||dX|| by less than
 * half.  'H' needs |A| and the corrections below the FP16 range (65504).  iters /
 * last_correction may be NULL.  Synchronises the stream.  No reference function: the
 * reference's unified_rectrxm! (src/rectrxm.jl:43-76) runs in one precision. */
int nla_drectrsm_mixed(char side, char uplo, char trans, double alpha, int64_t n, int64_t m, const double *A,
                       int64_t lda, double *B, int64_t ldb, int64_t cutoff, int64_t maxit, double tol, char prec,
                       int64_t *iters, double *last_correction, void *stream);

/* ---- Batched uneven problems (no reference function: the batched workload) ------------- */
/* One work-list entry.  QR part: geqrt!(m, n, ib, A, T) (src/geqrt.jl:22-68) of the m x n
 * matrix A (n <= m, 1 <= ib <= 64; A -> R and V, T ib x n).  Solve part: X = L^-1 B with L
 * nt x nt lower triangular (non-unit) and B nt x nrhs overwritten with X (nt <= 512;
 * nt = 0 or nrhs = 0: no solve).  All pointers are device pointers. */
typedef struct nla_dbatch_problem {
    int64_t m, n, ib;
    double *A;
    int64_t lda;
    double *T;
    int64_t ldt;
    int64_t nt, nrhs;
    const double *L;
    int64_t ldl;
    double *B;
    int64_t ldb;
} nla_dbatch_problem;
typedef struct nla_sbatch_problem {
    int64_t m, n, ib;
    float *A;
    int64_t lda;
    float *T;
    int64_t ldt;
    int64_t nt, nrhs;
    const float *L;
    int64_t ldl;
    float *B;
    int64_t ldb;
} nla_sbatch_problem;
/* Every QR and every solve of the `count` entries of the DEVICE array `problems` in one
 * persistent launch (a workgroup per task, tasks claimed from the device int `counter`, which
 * the call clears with a kernel).  validate != 0 copies the list to the host and checks it
 * first (synchronising); validate = 0 only enqueues two kernels, allocates nothing and issues
 * no memset, so the call can be captured into a hipGraph and replayed. */
int nla_dbatched_qr_trsm(int64_t count, const void *problems, int *counter, int validate, void *stream);
int nla_sbatched_qr_trsm(int64_t count, const void *problems, int *counter, int validate, void *stream);
/* Autotuner of the batched QR's ib: for each class c (m_values[c] x n_values[c], host
 * arrays) times the batched launch on `samples` random problems for each of the ncand
 * candidates and writes the fastest to ib_out[c] (host) and its time in ms to ms_out[c]
 * (host, may be NULL).  Allocates and frees its own buffers; synchronises. */
int nla_dbatched_autotune(int64_t nclass, const int64_t *m_values, const int64_t *n_values, int64_t samples,
                          const int64_t *candidates, int64_t ncand, int64_t *ib_out, double *ms_out, void *stream);
int nla_sbatched_autotune(int64_t nclass, const int64_t *m_values, const int64_t *n_values, int64_t samples,
                          const int64_t *candidates, int64_t ncand, int64_t *ib_out, double *ms_out, void *stream);

/* ---- Adjoints (no reference function: the reference has no differentiation code; these
 * are what a ChainRulesCore.rrule for unified_rectrxm! and for a thin tile QR calls) ---- */
/* Solve X = op(A)^-1 (alpha B) (side 'L') / (alpha B) op(A)^-1 ('R') given X and Xbar:
 * Xbar is overwritten with Bbar, Abar (n x n) receives the adjoint of A on its uplo
 * triangle (zeros elsewhere).  Shapes as nla_?rectrxm; cutoff as there. */
int nla_dtrsm_rrule(char side, char uplo, char trans, double alpha, int64_t n, int64_t m, const double *A,
                    int64_t lda, const double *X, int64_t ldx, double *Xbar, int64_t ldxbar, double *Abar,
                    int64_t ldabar, int64_t cutoff, void *stream);
int nla_strsm_rrule(char side, char uplo, char trans, float alpha, int64_t n, int64_t m, const float *A,
                    int64_t lda, const float *X, int64_t ldx, float *Xbar, int64_t ldxbar, float *Abar,
                    int64_t ldabar, int64_t cutoff, void *stream);
/* Multiply Y = alpha op(A) B ('L') / alpha B op(A) ('R') given the input B and Ybar:
 * Ybar is overwritten with Bbar, Abar as above. */
int nla_dtrmm_rrule(char side, char uplo, char trans, double alpha, int64_t n, int64_t m, const double *A,
                    int64_t lda, const double *B, int64_t ldb, double *Ybar, int64_t ldybar, double *Abar,
                    int64_t ldabar, int64_t cutoff, void *stream);
int nla_strmm_rrule(char side, char uplo, char trans, float alpha, int64_t n, int64_t m, const float *A,
                    int64_t lda, const float *B, int64_t ldb, float *Ybar, int64_t ldybar, float *Abar,
                    int64_t ldabar, int64_t cutoff, void *stream);
/* Thin QR A = Q R (m x n, n <= m, both multiples of nb) from nla_?tile_geqrf's output (AF, T):
 * given Qbar (m x n, overwritten with Abar) and Rbar (n x n).  Synchronises the stream
 * (its Q applications are persistent tile-QR launches). */
int nla_dqr_rrule(int64_t m, int64_t n, int64_t nb, int64_t ib, const double *AF, int64_t lda, const double *T,
                  int64_t ldt, double *Qbar, int64_t ldqbar, const double *Rbar, int64_t ldrbar, void *stream);
int nla_sqr_rrule(int64_t m, int64_t n, int64_t nb, int64_t ib, const float *AF, int64_t lda, const float *T,
                  int64_t ldt, float *Qbar, int64_t ldqbar, const float *Rbar, int64_t ldrbar, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* NEXTLA_HIP_H */
