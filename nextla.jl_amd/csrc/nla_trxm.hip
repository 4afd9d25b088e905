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

// ------------------------------------------------------------------------------------
// Blocked base solve (cut-off >= 32).  The recursion peels powers of two (>= 32) off the
// front, so every base block starts at a multiple of 32 of op(A) and only the block that
// ends at n can have a ragged tail.  Hence the 32 x 32 diagonal sub-blocks of every base
// call are the 32-aligned diagonal blocks of op(A): one launch inverts them all, then each
// base call is a left-looking (lower) / right-looking (upper) chain of MFMA products:
//   X_s = inv(A_ss) (B_s - sum_{j done} A_sj X_j).
// op(A)(i,j) = TRA ? A[j + i lda] : A[i + j lda]; no index reversal here.
constexpr int DB = 32;   // diagonal sub-block
constexpr int BCW = 16;  // right-hand sides per workgroup

template <typename T, bool TRA> __device__ __forceinline__ T opa(const T *A, i64 lda, int i, int j) {
    return TRA ? A[j + (i64)i * lda] : A[i + (i64)j * lda];
}

// inverse of each DB x DB diagonal block of op(A) (lower or upper), column-major into
// Dinv[b * DB * DB]; entries outside the triangle are zero.  One wave per block, lane = column.
template <typename T, bool UP, bool TRA>
__global__ __launch_bounds__(64) void diag_inv_kernel(int n, const T *__restrict__ A, i64 lda, T *__restrict__ Dinv) {
    __shared__ T L[DB][DB + 1];
    const int b = blockIdx.x, o = b * DB, bs = min(DB, n - o);
    const int j = threadIdx.x;
    for (int idx = j; idx < DB * DB; idx += 64) {
        const int r = idx % DB, c = idx / DB;
        L[r][c] = (r < bs && c < bs && (UP ? r <= c : r >= c)) ? opa<T, TRA>(A, lda, o + r, o + c) : T(0);
    }
    __syncthreads();
    if (j < DB) {
        T x[DB];
#pragma unroll
        for (int i = 0; i < DB; ++i) x[i] = 0;
        if (j < bs) {
            if (!UP) {
#pragma unroll
                for (int i = 0; i < DB; ++i) {
                    if (i < j || i >= bs) continue;
                    T s = (i == j) ? T(1) : T(0);
#pragma unroll
                    for (int k = 0; k < DB; ++k)
                        if (k >= j && k < i) s -= L[i][k] * x[k];
                    x[i] = s / L[i][i];
                }
            } else {
#pragma unroll
                for (int ii = DB - 1; ii >= 0; --ii) {
                    if (ii > j) continue;
                    T s = (ii == j) ? T(1) : T(0);
#pragma unroll
                    for (int k = 0; k < DB; ++k)
                        if (k > ii && k <= j) s -= L[ii][k] * x[k];
                    x[ii] = s / L[ii][ii];
                }
            }
        }
        T *D = Dinv + (size_t)b * DB * DB;
#pragma unroll
        for (int i = 0; i < DB; ++i) D[i + DB * j] = x[i];
    }
}

// One base block (n <= 1024 rows of op(A) starting at the pointer A; B likewise), BCW
// right-hand sides per workgroup.  Dinv points at this block's first inverted sub-block.
// blockIdx.y selects one of several equal base blocks: A, B and Dinv advance by the strides.
template <typename T, bool UP, bool TRA, bool TRB>
__global__ __launch_bounds__(256) void trsm_blk_kernel(int n, int m, const T *__restrict__ A, i64 lda,
                                                       T *__restrict__ B, i64 ldb, const T *__restrict__ Dinv,
                                                       i64 a_step, i64 b_step, i64 d_step) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    A += blockIdx.y * a_step;
    B += blockIdx.y * b_step;
    Dinv += blockIdx.y * d_step;
    const int npad = (n + DB - 1) / DB * DB;
    T *Xs = reinterpret_cast<T *>(smem_raw);   // [npad][BCW + 1]
    T *P = Xs + (size_t)npad * (BCW + 1);       // [4 waves][2 tiles][16][17]
    T *Ys = P + 4 * 2 * 16 * 17;                // [DB][BCW + 1]
    constexpr int XS = BCW + 1;
    const int c0 = blockIdx.x * BCW;
    const int cw = min(BCW, m - c0);
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int lr = lane & 15, lk = lane >> 4;

    for (int idx = t; idx < npad * BCW; idx += 256) {
        int r, c;
        if (TRB) { c = idx % BCW; r = idx / BCW; } else { r = idx % npad; c = idx / npad; }
        T v = 0;
        if (r < n && c < cw) v = TRB ? B[(c0 + c) + (i64)r * ldb] : B[r + (i64)(c0 + c) * ldb];
        Xs[r * XS + c] = v;
    }
    const int nsb = npad / DB;
    // Software pipeline over steps: the A fragments (first chunk of the K range) and the
    // inverted diagonal block of step s+1 are requested while step s reduces and applies
    // its own inverse, so global-memory latency is paid once per call, not once per step.
    constexpr int MAXIT = 16;
    T ab0[MAXIT], ab1[MAXIT], dv[DB / 4];
    auto issue = [&](int step) {
        const int sb = UP ? nsb - 1 - step : step;
        const int r0 = sb * DB;
        const int kb = UP ? r0 + DB : 0, ke = UP ? n : r0;
        const int kc0 = kb + 4 * w;
#pragma unroll
        for (int it = 0; it < MAXIT; ++it) {
            const int kc = kc0 + 16 * it + lk;
            T a0 = 0, a1 = 0;
            if (kc0 + 16 * it < ke && kc < n) {
                if (r0 + lr < n) a0 = opa<T, TRA>(A, lda, r0 + lr, kc);
                if (r0 + 16 + lr < n) a1 = opa<T, TRA>(A, lda, r0 + 16 + lr, kc);
            }
            ab0[it] = a0;
            ab1[it] = a1;
        }
        if (w < 2) {
            const T *D = Dinv + (size_t)sb * DB * DB;
#pragma unroll
            for (int q = 0; q < DB / 4; ++q) dv[q] = D[(16 * w + lr) + DB * (4 * q + lk)];
        }
    };
    issue(0);
    for (int step = 0; step < nsb; ++step) {
        const int sb = UP ? nsb - 1 - step : step;
        const int r0 = sb * DB;
        const int kb = UP ? r0 + DB : 0, ke = UP ? n : r0;
        typename Acc<T>::t acc0 = acc_zero<T>(), acc1 = acc_zero<T>();
        __syncthreads();
        const int kc0 = kb + 4 * w;
#pragma unroll
        for (int it = 0; it < MAXIT; ++it) {
            if (kc0 + 16 * it < ke) {
                const T bf = Xs[(kc0 + 16 * it + lk) * XS + lr];  // rows >= n are zero in Xs
                mfma16(acc0, ab0[it], bf);
                mfma16(acc1, ab1[it], bf);
            }
        }
        // K ranges longer than one chunk (base blocks over 256 rows): remaining chunks load in place
        for (int kc1 = kc0 + 16 * MAXIT; kc1 < ke; kc1 += 16) {
            const int kc = kc1 + lk;
            T a0 = 0, a1 = 0;
            if (kc < n) {
                if (r0 + lr < n) a0 = opa<T, TRA>(A, lda, r0 + lr, kc);
                if (r0 + 16 + lr < n) a1 = opa<T, TRA>(A, lda, r0 + 16 + lr, kc);
            }
            const T bf = Xs[kc * XS + lr];
            mfma16(acc0, a0, bf);
            mfma16(acc1, a1, bf);
        }
        T dcur[DB / 4];
#pragma unroll
        for (int q = 0; q < DB / 4; ++q) dcur[q] = dv[q];
        if (step + 1 < nsb) issue(step + 1);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = acc_row<T>(lk, r);
            P[((w * 2 + 0) * 16 + row) * 17 + lr] = acc0[r];
            P[((w * 2 + 1) * 16 + row) * 17 + lr] = acc1[r];
        }
        __syncthreads();
        for (int idx = t; idx < DB * BCW; idx += 256) {
            const int row = idx / BCW, c = idx % BCW, tile = row >> 4, rr = row & 15;
            T s = Xs[(r0 + row) * XS + c];
#pragma unroll
            for (int ww = 0; ww < 4; ++ww) s -= P[((ww * 2 + tile) * 16 + rr) * 17 + c];
            Ys[row * XS + c] = s;
        }
        __syncthreads();
        if (w < 2) {
            // X_s rows 16w.. = Dinv_s[16w.., :] * Y
            typename Acc<T>::t acc = acc_zero<T>();
#pragma unroll
            for (int q = 0; q < DB / 4; ++q) mfma16(acc, dcur[q], Ys[(4 * q + lk) * XS + lr]);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = r0 + 16 * w + acc_row<T>(lk, r);
                Xs[row * XS + lr] = row < n ? acc[r] : T(0);
            }
        }
    }
    __syncthreads();
    for (int idx = t; idx < n * BCW; idx += 256) {
        int r, c;
        if (TRB) { c = idx % BCW; r = idx / BCW; } else { r = idx % n; c = idx / n; }
        if (c < cw) {
            if (TRB) B[(c0 + c) + (i64)r * ldb] = Xs[r * XS + c];
            else B[r + (i64)(c0 + c) * ldb] = Xs[r * XS + c];
        }
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

template <typename T> static size_t blk_lds_bytes(int n) {
    const int npad = (n + DB - 1) / DB * DB;
    return ((size_t)npad * (BCW + 1) + 4 * 2 * 16 * 17 + DB * (BCW + 1)) * sizeof(T);
}
// The blocked solve holds a base block of n rows in LDS (<= 160 KiB).
template <typename T> static bool blk_ok(int n) { return blk_lds_bytes<T>(n) <= 160 * 1024; }

template <typename T>
static hipError_t diag_inv_launch(bool up, bool trA, int n, const T *A, i64 lda, T *Dinv, hipStream_t s) {
    const int nb = (n + DB - 1) / DB;
    if (!up && !trA) diag_inv_kernel<T, false, false><<<nb, 64, 0, s>>>(n, A, lda, Dinv);
    else if (!up && trA) diag_inv_kernel<T, false, true><<<nb, 64, 0, s>>>(n, A, lda, Dinv);
    else if (up && !trA) diag_inv_kernel<T, true, false><<<nb, 64, 0, s>>>(n, A, lda, Dinv);
    else diag_inv_kernel<T, true, true><<<nb, 64, 0, s>>>(n, A, lda, Dinv);
    return hipGetLastError();
}

template <typename T, bool U, bool TA, bool TB>
static hipError_t blk_launch_one(int n, int m, const T *A, i64 lda, T *B, i64 ldb, const T *Dinv, hipStream_t s,
                                 int batch = 1, i64 a_step = 0, i64 b_step = 0, i64 d_step = 0) {
    const size_t sh = blk_lds_bytes<T>(n);
    if (sh > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&trsm_blk_kernel<T, U, TA, TB>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
        if (e != hipSuccess) return e;
    }
    const dim3 grid((unsigned)((m + BCW - 1) / BCW), (unsigned)batch);
    trsm_blk_kernel<T, U, TA, TB><<<grid, 256, sh, s>>>(n, m, A, lda, B, ldb, Dinv, a_step, b_step, d_step);
    return hipGetLastError();
}

template <typename T>
static hipError_t blk_launch(bool up, bool trA, bool trB, int n, int m, const T *A, i64 lda, T *B, i64 ldb,
                             const T *Dinv, hipStream_t s) {
    if (n <= 0 || m <= 0) return hipSuccess;
    ProfScope prof(PK_TRSM_BASE, s);
    if (!up && !trA && !trB) return blk_launch_one<T, false, false, false>(n, m, A, lda, B, ldb, Dinv, s);
    if (!up && !trA && trB) return blk_launch_one<T, false, false, true>(n, m, A, lda, B, ldb, Dinv, s);
    if (!up && trA && !trB) return blk_launch_one<T, false, true, false>(n, m, A, lda, B, ldb, Dinv, s);
    if (!up && trA && trB) return blk_launch_one<T, false, true, true>(n, m, A, lda, B, ldb, Dinv, s);
    if (up && !trA && !trB) return blk_launch_one<T, true, false, false>(n, m, A, lda, B, ldb, Dinv, s);
    if (up && !trA && trB) return blk_launch_one<T, true, false, true>(n, m, A, lda, B, ldb, Dinv, s);
    if (up && trA && !trB) return blk_launch_one<T, true, true, false>(n, m, A, lda, B, ldb, Dinv, s);
    return blk_launch_one<T, true, true, true>(n, m, A, lda, B, ldb, Dinv, s);
}

// ---------------------------------------------------------------------------------
// Base blocks of IBS rows by an explicit inverse.  The blocked substitution above runs its
// IBS/32 steps one after another inside each workgroup (barriers, a growing K loop, the
// next step's loads behind the current one); with the dense inverse Inv of the block's
// triangle, formed once per call for every aligned IBS-block on the diagonal, the base solve
// is the triangular product X = Inv B with no sequential dependence.  Inverted diagonal
// blocks of this size are the usual GPU TRSM scheme; the error of applying Inv instead of
// substituting is of the order cond(A_bb) u, like the substitution's forward error.
constexpr int IBS = 256;

// X = Inv B for one base block of IBS rows, in place: a workgroup owns BCW right-hand sides
// and every row (it reads them all into LDS before writing any).  Inv is dense column-major
// (ld IBS), zero outside its lower (UP false) or upper triangle.  With NW = 4 waves, wave W
// computes the output row tiles {W, 15-W} and {4+W, 11-W} (34 K steps of 16 per wave); with
// NW = 8 (two waves per SIMD, the launches below) the pair {W, 15-W} (17 steps): equal work
// for every wave either way.
//
// A pair's two tiles need 17 K steps of 16 between them (tile x: first step f(x), count
// c(x); lower f = 0, c = x + 1, upper f = x, c = 16 - x; c(x) + c(15 - x) = 17), so each pair
// is one fixed 17-step sequence: the first c(a) steps accumulate tile a, the rest tile b.
// W is a template parameter, so every tile and step index is a constant and each tile has
// its own accumulator: the steps are straight-line loads and MFMAs (the loads of step v + PF
// issued while step v multiplies), with four waves the two pairs' chains interleaved.
template <typename T, bool UP, bool TRB, int W, int NW = 4>
__device__ __forceinline__ void inv_apply_wave(const T *__restrict__ Inv, T *__restrict__ B, i64 ldb,
                                               const T (*Xs)[BCW + 1], int c0, int cw, T (*Ys)[BCW + 1] = nullptr) {
    constexpr int NT = IBS / 16, PF = 3, NS = NT + 1;
    static_assert(NT == 16, "the pairing assumes 16 row tiles");
    const int lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
    static_assert(NW == 4 || NW == 8, "four or eight waves");
    constexpr int NP = 8 / NW;  // pairs per wave: {W + NW q, 15 - (W + NW q)}
    constexpr int TA[2] = {W, NP > 1 ? NW + W : W}, TB[2] = {NT - 1 - TA[0], NT - 1 - TA[1]};
    constexpr int FA[2] = {UP ? TA[0] : 0, UP ? TA[1] : 0}, FB[2] = {UP ? TB[0] : 0, UP ? TB[1] : 0};
    constexpr int CA[2] = {UP ? NT - TA[0] : TA[0] + 1, UP ? NT - TA[1] : TA[1] + 1};
    auto tile_of = [&](int p, int v) { return v < CA[p] ? TA[p] : TB[p]; };
    auto step_of = [&](int p, int v) { return v < CA[p] ? FA[p] + v : FB[p] + v - CA[p]; };
    typename Acc<T>::t acc[2][2];
#pragma unroll
    for (int p = 0; p < NP; ++p) acc[p][0] = acc[p][1] = acc_zero<T>();
    T a[PF][2][4];
    auto load_step = [&](int v, T (&dst)[2][4]) {
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            const T *Ap = Inv + (16 * tile_of(p, v) + lr) + (i64)(16 * step_of(p, v) + lk) * IBS;
#pragma unroll
            for (int u = 0; u < 4; ++u) dst[p][u] = Ap[(i64)(4 * u) * IBS];
        }
    };
#pragma unroll
    for (int v = 0; v < PF; ++v) load_step(v, a[v]);
#pragma unroll
    for (int v = 0; v < NS; ++v) {
        T cur[2][4];
#pragma unroll
        for (int p = 0; p < NP; ++p)
#pragma unroll
            for (int u = 0; u < 4; ++u) cur[p][u] = a[v % PF][p][u];
        if (v + PF < NS) load_step(v + PF, a[v % PF]);
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int p = 0; p < NP; ++p)
                mfma16(acc[p][v < CA[p] ? 0 : 1], cur[p][u], Xs[16 * step_of(p, v) + lk + 4 * u][lr]);
    }
#pragma unroll
    for (int p = 0; p < NP; ++p)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int it = h == 0 ? TA[p] : TB[p];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = 16 * it + acc_row<T>(lk, r);
                if (Ys) Ys[row][lr] = acc[p][h][r];  // a copy of X for a following product
                if (lr < cw) {
                    if (TRB) B[(c0 + lr) + (i64)row * ldb] = acc[p][h][r];
                    else B[row + (i64)(c0 + lr) * ldb] = acc[p][h][r];
                }
            }
        }
}

// Xo -= A21 Xk on the wave's row tiles {W, W+4, W+8, W+12} (16 rows each) of an IBS x IBS block
// A21 (column-major, lda) times the IBS x BCW panel Xk, both panels in LDS: 16 K-steps of 16
// per tile, four tiles' chains interleaved, the A21 loads of step v + PF issued while step v
// multiplies (the same operand roles as the inverse apply above).
template <typename T, int W, int NW = 4>
__device__ __forceinline__ void sub_prod_wave(const T *__restrict__ A21, i64 lda, const T (*Xk)[BCW + 1],
                                              T (*Xo)[BCW + 1]) {
    constexpr int NKS = IBS / 16, PF = 3, NTW = 16 / NW;  // row tiles W, W + NW, ...;
    const int lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
    typename Acc<T>::t acc[NTW];
#pragma unroll
    for (int p = 0; p < NTW; ++p) acc[p] = acc_zero<T>();
    T a[PF][NTW][4];
    auto load_step = [&](int ks, T (&dst)[NTW][4]) {
#pragma unroll
        for (int p = 0; p < NTW; ++p) {
            const T *Ap = A21 + (16 * (W + NW * p) + lr) + (i64)(16 * ks + lk) * lda;
#pragma unroll
            for (int u = 0; u < 4; ++u) dst[p][u] = Ap[(i64)(4 * u) * lda];
        }
    };
#pragma unroll
    for (int v = 0; v < PF; ++v) load_step(v, a[v]);
#pragma unroll
    for (int ks = 0; ks < NKS; ++ks) {
        T cur[NTW][4];
#pragma unroll
        for (int p = 0; p < NTW; ++p)
#pragma unroll
            for (int u = 0; u < 4; ++u) cur[p][u] = a[ks % PF][p][u];
        if (ks + PF < NKS) load_step(ks + PF, a[ks % PF]);
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int p = 0; p < NTW; ++p) mfma16(acc[p], cur[p][u], Xk[16 * ks + lk + 4 * u][lr]);
    }
#pragma unroll
    for (int p = 0; p < NTW; ++p)
#pragma unroll
        for (int r = 0; r < 4; ++r) Xo[16 * (W + NW * p) + acc_row<T>(lk, r)][lr] -= acc[p][r];
}

// A lower, left, untransposed 2*IBS-row block of the recursion in one launch, instead of base,
// GEMM, base: X1 = Inv1 B1; B2 -= A21 X1; X2 = Inv2 B2 (Inv1, Inv2 the inverses of the two
// diagonal blocks, A21 the block below the first).  A workgroup owns BCW right-hand sides
// and keeps their 2*IBS rows (Xs) and X1 (Ys) in LDS throughout, so B is read and written
// once and the three steps need no launches between them.  Dynamic LDS: 3 IBS (BCW+1) values.
template <typename T>
__global__ __launch_bounds__(512) void trsm_inv2_kernel(int m, const T *__restrict__ Inv1, const T *__restrict__ Inv2,
                                                        const T *__restrict__ A21, i64 lda, T *__restrict__ B, i64 ldb) {
    extern __shared__ __align__(16) unsigned char inv2_smem[];
    T (*Xs)[BCW + 1] = reinterpret_cast<T (*)[BCW + 1]>(inv2_smem);  // 2 IBS rows
    T (*Ys)[BCW + 1] = Xs + 2 * IBS;                                  // IBS rows
    const int c0 = blockIdx.x * BCW, cw = min(BCW, m - c0);
    const int t = threadIdx.x;
    for (int idx = t; idx < 2 * IBS * BCW; idx += 512) {
        const int r = idx % (2 * IBS), c = idx / (2 * IBS);
        Xs[r][c] = c < cw ? B[r + (i64)(c0 + c) * ldb] : T(0);
    }
    __syncthreads();
    // eight waves, two per SIMD: each carries half the tiles of the four-wave kernel
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);
#define NLA_INV2_WAVES(STMT)                                                                                       \
    switch (w) {                                                                                                   \
    case 0: STMT(0); break;                                                                                        \
    case 1: STMT(1); break;                                                                                        \
    case 2: STMT(2); break;                                                                                        \
    case 3: STMT(3); break;                                                                                        \
    case 4: STMT(4); break;                                                                                        \
    case 5: STMT(5); break;                                                                                        \
    case 6: STMT(6); break;                                                                                        \
    default: STMT(7); break;                                                                                       \
    }
#define NLA_INV2_P1(W) inv_apply_wave<T, false, false, W, 8>(Inv1, B, ldb, Xs, c0, cw, Ys)
#define NLA_INV2_P2(W) sub_prod_wave<T, W, 8>(A21, lda, Ys, Xs + IBS)
#define NLA_INV2_P3(W) inv_apply_wave<T, false, false, W, 8>(Inv2, B + IBS, ldb, Xs + IBS, c0, cw)
    NLA_INV2_WAVES(NLA_INV2_P1)
    __syncthreads();
    NLA_INV2_WAVES(NLA_INV2_P2)
    __syncthreads();
    NLA_INV2_WAVES(NLA_INV2_P3)
#undef NLA_INV2_P1
#undef NLA_INV2_P2
#undef NLA_INV2_P3
#undef NLA_INV2_WAVES
}

template <typename T, bool UP, bool TRB>
__global__ __launch_bounds__(512) void trsm_inv_apply_kernel(int m, const T *__restrict__ Inv, T *__restrict__ B,
                                                             i64 ldb) {
    __shared__ T Xs[IBS][BCW + 1];
    const int c0 = blockIdx.x * BCW, cw = min(BCW, m - c0);
    const int t = threadIdx.x;
    for (int idx = t; idx < IBS * BCW; idx += 512) {
        int r, c;
        if (TRB) { c = idx % BCW; r = idx / BCW; } else { r = idx % IBS; c = idx / IBS; }
        T v = 0;
        if (c < cw) v = TRB ? B[(c0 + c) + (i64)r * ldb] : B[r + (i64)(c0 + c) * ldb];
        Xs[r][c] = v;
    }
    __syncthreads();
    // eight waves, two per SIMD, one tile pair each; the wave index is uniform across the
    // wave: one scalar branch picks the wave's body
    switch (__builtin_amdgcn_readfirstlane(t >> 6)) {
    case 0: inv_apply_wave<T, UP, TRB, 0, 8>(Inv, B, ldb, Xs, c0, cw); break;
    case 1: inv_apply_wave<T, UP, TRB, 1, 8>(Inv, B, ldb, Xs, c0, cw); break;
    case 2: inv_apply_wave<T, UP, TRB, 2, 8>(Inv, B, ldb, Xs, c0, cw); break;
    case 3: inv_apply_wave<T, UP, TRB, 3, 8>(Inv, B, ldb, Xs, c0, cw); break;
    case 4: inv_apply_wave<T, UP, TRB, 4, 8>(Inv, B, ldb, Xs, c0, cw); break;
    case 5: inv_apply_wave<T, UP, TRB, 5, 8>(Inv, B, ldb, Xs, c0, cw); break;
    case 6: inv_apply_wave<T, UP, TRB, 6, 8>(Inv, B, ldb, Xs, c0, cw); break;
    default: inv_apply_wave<T, UP, TRB, 7, 8>(Inv, B, ldb, Xs, c0, cw); break;
    }
}

// The aligned IBS x IBS diagonal blocks of the left-form matrix op'(A) themselves (zero outside
// the triangle), dense column-major: the TRMM base B_b := op'(A)_bb B_b is then the same
// triangular MFMA product as the solve's X = Inv B (trsm_inv_apply_kernel).
template <typename T, bool UP, bool TRA>
__global__ void tri_blocks_kernel(int nblocks, const T *__restrict__ A, i64 lda, T *__restrict__ Out) {
    const i64 total = (i64)nblocks * IBS * IBS;
    for (i64 idx = blockIdx.x * (i64)blockDim.x + threadIdx.x; idx < total; idx += (i64)gridDim.x * blockDim.x) {
        const i64 b = idx / ((i64)IBS * IBS), e = idx % ((i64)IBS * IBS);
        const int r = (int)(e % IBS), c = (int)(e / IBS);
        const i64 o = b * IBS;
        Out[idx] = (UP ? r <= c : r >= c) ? opa<T, TRA>(A, lda, (int)(o + r), (int)(o + c)) : T(0);
    }
}

template <typename T>
static hipError_t tri_blocks(bool up, bool trA, int nblocks, const T *A, i64 lda, T *Out, hipStream_t s) {
    if (!up && !trA) tri_blocks_kernel<T, false, false><<<2048, 256, 0, s>>>(nblocks, A, lda, Out);
    else if (!up && trA) tri_blocks_kernel<T, false, true><<<2048, 256, 0, s>>>(nblocks, A, lda, Out);
    else if (up && !trA) tri_blocks_kernel<T, true, false><<<2048, 256, 0, s>>>(nblocks, A, lda, Out);
    else tri_blocks_kernel<T, true, true><<<2048, 256, 0, s>>>(nblocks, A, lda, Out);
    return hipGetLastError();
}

template <typename T> __global__ void identity_blocks_kernel(int nblocks, T *Inv) {
    const i64 total = (i64)nblocks * IBS * IBS;
    for (i64 idx = blockIdx.x * (i64)blockDim.x + threadIdx.x; idx < total; idx += (i64)gridDim.x * blockDim.x) {
        const i64 e = idx % ((i64)IBS * IBS);
        Inv[idx] = (e % IBS) == (e / IBS) ? T(1) : T(0);
    }
}

// Inverses of the nblocks aligned IBS x IBS diagonal blocks of the left-form matrix op'(A)
// (flags as in base_launch), by the blocked solve of each block against the identity, all
// blocks in one launch.  Dinv = the 32 x 32 inverses of the whole diagonal.
template <typename T>
static hipError_t inv_blocks(bool up, bool trA, int nblocks, const T *A, i64 lda, const T *Dinv, T *Inv, hipStream_t s) {
    identity_blocks_kernel<T><<<1024, 256, 0, s>>>(nblocks, Inv);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const i64 a_step = (i64)IBS * (1 + lda), b_step = (i64)IBS * IBS, d_step = (i64)(IBS / DB) * DB * DB;
    if (!up && !trA) return blk_launch_one<T, false, false, false>(IBS, IBS, A, lda, Inv, IBS, Dinv, s, nblocks, a_step, b_step, d_step);
    if (!up && trA) return blk_launch_one<T, false, true, false>(IBS, IBS, A, lda, Inv, IBS, Dinv, s, nblocks, a_step, b_step, d_step);
    if (up && !trA) return blk_launch_one<T, true, false, false>(IBS, IBS, A, lda, Inv, IBS, Dinv, s, nblocks, a_step, b_step, d_step);
    return blk_launch_one<T, true, true, false>(IBS, IBS, A, lda, Inv, IBS, Dinv, s, nblocks, a_step, b_step, d_step);
}

template <typename T>
static hipError_t inv_apply_launch(bool up, bool trB, int m, const T *Inv, T *B, i64 ldb, hipStream_t s) {
    ProfScope prof(PK_TRSM_BASE, s);
    const dim3 grid((unsigned)((m + BCW - 1) / BCW));
    if (!up && !trB) trsm_inv_apply_kernel<T, false, false><<<grid, 512, 0, s>>>(m, Inv, B, ldb);
    else if (!up && trB) trsm_inv_apply_kernel<T, false, true><<<grid, 512, 0, s>>>(m, Inv, B, ldb);
    else if (up && !trB) trsm_inv_apply_kernel<T, true, false><<<grid, 512, 0, s>>>(m, Inv, B, ldb);
    else trsm_inv_apply_kernel<T, true, true><<<grid, 512, 0, s>>>(m, Inv, B, ldb);
    return hipGetLastError();
}

template <typename T>
static hipError_t inv2_launch(int m, const T *Inv1, const T *A21, i64 lda, T *B, i64 ldb, hipStream_t s) {
    ProfScope prof(PK_TRSM_BASE, s);
    const size_t sh = sizeof(T) * 3 * IBS * (BCW + 1);
    static bool attr[2] = {false, false};
    if (sh > 64 * 1024 && !attr[sizeof(T) == 8]) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&trsm_inv2_kernel<T>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
        if (e != hipSuccess) return e;
        attr[sizeof(T) == 8] = true;
    }
    const dim3 grid((unsigned)((m + BCW - 1) / BCW));
    trsm_inv2_kernel<T><<<grid, 512, sh, s>>>(m, Inv1, Inv1 + (i64)IBS * IBS, A21, lda, B, ldb);
    return hipGetLastError();
}

// Left-form flags of the base call for (side, uplo) of op(A), op(A)(i,j) = tr ? A(j,i) : A(i,j):
// X op(A) = B  <=>  op(A)^T X^T = B^T.
static inline void left_form(char side, char uplo, bool tr, bool *up, bool *trA, bool *trB) {
    if (side == 'L') { *up = uplo == 'U'; *trA = tr; *trB = false; }
    else { *up = uplo != 'U'; *trA = !tr; *trB = true; }
}

// Base case of unified_rec.  Dinv (or nullptr) holds the inverted 32x32 diagonal blocks of the
// whole left-form matrix; diag_off is this block's offset along the diagonal.
template <typename T>
static hipError_t base_case(char func, char side, char uplo, bool tr, int n, int m, const T *A, i64 lda, T *B, i64 ldb,
                            const T *Dinv, const T *Inv, i64 diag_off, hipStream_t s) {
    bool up, trA, trB;
    left_form(side, uplo, tr, &up, &trA, &trB);
    // 'S': Inv holds the inverted IBS-blocks; 'M': the blocks themselves (same product)
    if (Inv && n == IBS && (diag_off % IBS) == 0)
        return inv_apply_launch<T>(up, trB, m, Inv + (diag_off / IBS) * (i64)IBS * IBS, B, ldb, s);
    if (func == 'S' && Dinv && (diag_off % DB) == 0 && blk_ok<T>(n))
        return blk_launch<T>(up, trA, trB, n, m, A, lda, B, ldb, Dinv + (diag_off / DB) * DB * DB, s);
    return base_launch<T>(func, up, trA, trB, n, m, A, lda, B, ldb, s);
}

// Pointer to block (i0, j0) of op(A).
template <typename T> static inline const T *opA_block(const T *A, i64 lda, bool tr, i64 i0, i64 j0) {
    return tr ? A + j0 + i0 * lda : A + i0 + j0 * lda;
}

// unified_rec: src/rectrxm.jl:101-198 (same split rule, same forward/backward orders).
template <typename T>
static hipError_t rec(char func, char side, char uplo, bool tr, i64 n, i64 m, const T *A, i64 lda, T *B, i64 ldb,
                      i64 threshold, const T *Dinv, const T *Inv, i64 off, hipStream_t s) {
    if (n <= threshold) return base_case<T>(func, side, uplo, tr, (int)n, (int)m, A, lda, B, ldb, Dinv, Inv, off, s);
    // a lower, left, untransposed solve of two aligned base blocks: one fused launch
    if (func == 'S' && side == 'L' && uplo == 'L' && !tr && Inv && n == 2 * IBS && threshold >= IBS &&
        off % IBS == 0)
        return inv2_launch<T>((int)m, Inv + (off / IBS) * (i64)IBS * IBS, A + IBS, lda, B, ldb, s);
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
        if ((e = rec<T>(func, side, uplo, tr, mid, m, A11, lda, B1, ldb, threshold, Dinv, Inv, off, s)) != hipSuccess) return e;
        if (side == 'L') {
            if (func == 'S') e = gemm<T>(ta, 'N', rem, m, mid, T(-1), A21, lda, B1, ldb, T(1), B2, ldb, s);  // B2 -= A21 B1
            else e = gemm<T>(ta, 'N', mid, m, rem, T(1), A12, lda, B2, ldb, T(1), B1, ldb, s);              // B1 += A12 B2
        } else {
            if (func == 'S') e = gemm<T>('N', ta, m, rem, mid, T(-1), B1, ldb, A12, lda, T(1), B2, ldb, s);  // B2 -= B1 A12
            else e = gemm<T>('N', ta, m, mid, rem, T(1), B2, ldb, A21, lda, T(1), B1, ldb, s);              // B1 += B2 A21
        }
        if (e != hipSuccess) return e;
        return rec<T>(func, side, uplo, tr, rem, m, A22, lda, B2, ldb, threshold, Dinv, Inv, off + mid, s);
    }
    if ((e = rec<T>(func, side, uplo, tr, rem, m, A22, lda, B2, ldb, threshold, Dinv, Inv, off + mid, s)) != hipSuccess)
        return e;
    if (side == 'L') {
        if (func == 'S') e = gemm<T>(ta, 'N', mid, m, rem, T(-1), A12, lda, B2, ldb, T(1), B1, ldb, s);  // B1 -= A12 B2
        else e = gemm<T>(ta, 'N', rem, m, mid, T(1), A21, lda, B1, ldb, T(1), B2, ldb, s);              // B2 += A21 B1
    } else {
        if (func == 'S') e = gemm<T>('N', ta, m, mid, rem, T(-1), B2, ldb, A21, lda, T(1), B1, ldb, s);  // B1 -= B2 A21
        else e = gemm<T>('N', ta, m, rem, mid, T(1), B1, ldb, A12, lda, T(1), B2, ldb, s);              // B2 += B1 A12
    }
    if (e != hipSuccess) return e;
    return rec<T>(func, side, uplo, tr, mid, m, A11, lda, B1, ldb, threshold, Dinv, Inv, off, s);
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

// Device workspace the solve needs beside A and B: the 32 x 32 inverses of every diagonal
// block of op(A) and, with a cut-off in [IBS, 2 IBS), the IBS x IBS inverses as well.
template <typename T> static size_t rectrxm_ws_bytes(char func, i64 n, i64 thr) {
    if (func == 'M') return (thr >= IBS && thr < 2 * IBS && n >= IBS) ? sizeof(T) * (size_t)(n / IBS) * IBS * IBS : 0;
    if (thr < DB) return 0;
    const i64 nd = (n + DB - 1) / DB * DB * DB;
    const i64 ninv = (thr >= IBS && thr < 2 * IBS) ? n / IBS : 0;
    return sizeof(T) * (size_t)(nd + ninv * IBS * IBS);
}

// ws = nullptr: the workspace comes from the per-(device, stream) scratch pool; otherwise
// the caller's buffer of ws_cap bytes is used and a solve that needs more is refused.
template <typename T>
static int rectrxm_entry(char side, char uplo, char trans, T alpha, char func, i64 n, i64 m, const T *A, i64 lda, T *B,
                         i64 ldb, i64 cutoff, void *stream, void *ws = nullptr, size_t ws_cap = 0) {
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
    // 0: this library's default, 256 for both (measured best at n = 8192; the reference's
    // thresholds are 256 for 'S' and 16 for 'M', src/rectrxm.jl:52,63)
    i64 thr = cutoff == 0 ? 256 : cutoff;
    if (thr > 1024) return arg_error(12, "cutoff must be <= 1024 (base kernels hold n <= 1024 rows, src/trsm.jl:10-12)");
    const size_t need = rectrxm_ws_bytes<T>(func, n, thr);
    if (ws && need > ws_cap)
        return arg_error(6, "n = %lld with cut-off %lld needs %zu bytes of device workspace; this entry point holds "
                            "%zu", (long long)n, (long long)thr, need, ws_cap);
    bool tr = false;
    if (trans == 'T' || trans == 'C') { tr = true; uplo = (uplo == 'L') ? 'U' : 'L'; }  // src/rectrxm.jl:56-59
    hipStream_t s = as_stream(stream);
    const i64 rows = side == 'L' ? n : m, cols = side == 'L' ? m : n;
    if (func == 'S') NLA_HIP(scale<T>(rows, cols, alpha, B, ldb, s));
    const T *Dinv = nullptr, *Inv = nullptr;
    if (func == 'S' && thr >= DB) {
        // every base block starts on the 32-grid (see diag_inv_kernel): invert all 32x32 blocks
        // once; with base blocks of IBS rows (cut-off >= IBS) also form the IBS-block inverses
        const i64 nd = (n + DB - 1) / DB * DB * DB;
        const i64 ninv = (thr >= IBS && thr < 2 * IBS) ? n / IBS : 0;
        int st = 0;
        T *D = ws ? (T *)ws : (T *)scratch(SLOT_TRSM, need, s, &st);
        if (!D) return st;
        bool upf, trA, trB;
        left_form(side, uplo, tr, &upf, &trA, &trB);
        NLA_HIP(diag_inv_launch<T>(upf, trA, (int)n, A, lda, D, s));
        Dinv = D;
        if (ninv > 0) {
            ProfScope prof(PK_TRSM_BASE, s);
            NLA_HIP(inv_blocks<T>(upf, trA, (int)ninv, A, lda, D, D + nd, s));
            Inv = D + nd;
        }
    }
    if (func == 'M' && thr >= IBS && thr < 2 * IBS && n >= IBS) {
        // TRMM with 256-row bases: every aligned diagonal block once, for the MFMA base product
        const i64 nblk = n / IBS;
        int st = 0;
        T *D = ws ? (T *)ws : (T *)scratch(SLOT_TRSM, need, s, &st);
        if (!D) return st;
        bool upf, trA, trB;
        left_form(side, uplo, tr, &upf, &trA, &trB);
        ProfScope prof(PK_TRMM_BASE, s);
        NLA_HIP(tri_blocks<T>(upf, trA, (int)nblk, A, lda, D, s));
        Inv = D;
    }
    NLA_HIP(rec<T>(func, side, uplo, tr, n, m, A, lda, B, ldb, thr, Dinv, Inv, 0, s));
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
    hipStream_t s = as_stream(stream);
    const T *Dinv = nullptr;
    if (func == 'S' && blk_ok<T>((int)n)) {
        int st = 0;
        T *D = (T *)scratch(SLOT_TRSM, sizeof(T) * (size_t)((n + DB - 1) / DB) * DB * DB, s, &st);
        if (!D) return st;
        bool upf, trA, trB;
        left_form(side, uplo, false, &upf, &trA, &trB);
        NLA_HIP(diag_inv_launch<T>(upf, trA, (int)n, A, lda, D, s));
        Dinv = D;
    }
    NLA_HIP(base_case<T>(func, side, uplo, false, (int)n, (int)m, A, lda, B, ldb, Dinv, nullptr, 0, s));
    return 0;
}

// Exposed for the mixed-precision and differentiation code.
template <typename T>
int rectrxm_async(char side, char uplo, char trans, T alpha, char func, i64 n, i64 m, const T *A, i64 lda, T *B,
                         i64 ldb, i64 cutoff, hipStream_t s) {
    return rectrxm_entry<T>(side, uplo, trans, alpha, func, n, m, A, lda, B, ldb, cutoff, (void *)s);
}
template int rectrxm_async<double>(char, char, char, double, char, i64, i64, const double *, i64, double *, i64, i64,
                                   hipStream_t);
template int rectrxm_async<float>(char, char, char, float, char, i64, i64, const float *, i64, float *, i64, i64,
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

namespace nla {
// ------------------------------------------------------------------------------------
// Host-pointer variant (NextLAMatrix wraps a host Matrix, src/NextLAMatrix.jl:4-32): A and B
// are copied into the stream's scratch, solved / multiplied there, and B is copied back.
template <typename T>
static int h_rectrxm_entry(char side, char uplo, char trans, T alpha, char func, i64 n, i64 m, const T *A, i64 lda, T *B,
                           i64 ldb, i64 cutoff, void *stream) {
    side = up(side);
    if (side != 'L' && side != 'R') return arg_error(1, "side must be 'L' or 'R', got '%c'", side);
    if (n < 0) return arg_error(6, "n must be non-negative, got %lld", (long long)n);
    if (m < 0) return arg_error(7, "m must be non-negative, got %lld", (long long)m);
    if (n == 0 || m == 0) return 0;
    if (lda < n) return arg_error(9, "lda must be >= n");
    const i64 rows = side == 'L' ? n : m, cols = side == 'L' ? m : n;
    if (ldb < rows) return arg_error(11, "ldb must be >= %lld", (long long)rows);
    if (!A || !B) return arg_error(8, "A and B must be host pointers");
    hipStream_t s = as_stream(stream);
    int st = 0;
    T *dA = (T *)scratch(SLOT_HOST, sizeof(T) * (size_t)(n * n + rows * cols), s, &st);
    if (!dA) return st;
    T *dB = dA + n * n;
    NLA_HIP(hipMemcpy2DAsync(dA, sizeof(T) * n, A, sizeof(T) * lda, sizeof(T) * n, n, hipMemcpyHostToDevice, s));
    NLA_HIP(hipMemcpy2DAsync(dB, sizeof(T) * rows, B, sizeof(T) * ldb, sizeof(T) * rows, cols, hipMemcpyHostToDevice, s));
    NLA_ST(rectrxm_entry<T>(side, uplo, trans, alpha, func, n, m, dA, n, dB, rows, cutoff, stream));
    NLA_HIP(hipMemcpy2DAsync(B, sizeof(T) * ldb, dB, sizeof(T) * rows, sizeof(T) * rows, cols, hipMemcpyDeviceToHost, s));
    NLA_HIP(hipStreamSynchronize(s));
    return 0;
}

}  // namespace nla

NLA_API int nla_hdrectrxm(char side, char uplo, char trans, double alpha, char func, int64_t n, int64_t m,
                          const double *A, int64_t lda, double *B, int64_t ldb, int64_t cutoff, void *stream) {
    return nla::h_rectrxm_entry<double>(side, uplo, trans, alpha, func, n, m, A, lda, B, ldb, cutoff, stream);
}
NLA_API int nla_hsrectrxm(char side, char uplo, char trans, float alpha, char func, int64_t n, int64_t m,
                          const float *A, int64_t lda, float *B, int64_t ldb, int64_t cutoff, void *stream) {
    return nla::h_rectrxm_entry<float>(side, uplo, trans, alpha, func, n, m, A, lda, B, ldb, cutoff, stream);
}

// ------------------------------------------------------------------------------------
// No-allocation variant of nla_drectrxm for callers that cannot allocate device memory
// (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc).  Its
// workspace is this file-scope device array, one per GPU, held until the process ends;
// every call on a GPU shares that GPU's copy, so calls that overlap on one GPU race.
// Nothing in it needs clearing: each solve writes every workspace entry it reads.
constexpr size_t NOALLOC_BYTES = size_t(18) << 20;   // n <= 8192 with the 256-row bases
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

namespace {
constexpr int NOALLOC_MAX_DEV = 64;
std::mutex g_noalloc_mu;
void *g_noalloc_ptr[NOALLOC_MAX_DEV] = {};
}  // namespace

NLA_API int nla_drectrxm_noalloc(char side, char uplo, char trans, double alpha, char func, int64_t n, int64_t m,
                                 const double *A, int64_t lda, double *B, int64_t ldb, int64_t cutoff,
                                 void *stream) {
    // sizes first, on the host: a solve that needs more than the array holds is refused
    {
        const char f = nla::up(func);
        const int64_t thr = cutoff == 0 ? 256 : cutoff;
        if (cutoff >= 0 && thr <= 1024 && n > 0 && m > 0) {
            const size_t need = nla::rectrxm_ws_bytes<double>(f, n, thr);
            if (need > NOALLOC_BYTES)
                return nla::arg_error(6, "n = %lld with cut-off %lld needs %zu bytes of device workspace; the "
                                         "no-allocation entry point holds %zu", (long long)n, (long long)thr, need,
                                      NOALLOC_BYTES);
        }
    }
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return nla::hip_error(e, "hipGetDevice");
    if (dev < 0 || dev >= NOALLOC_MAX_DEV) return nla::hip_error(hipErrorInvalidDevice, "device index");
    void *ptr = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_noalloc_mu);
        if (!g_noalloc_ptr[dev]) {
            // this device's copy of the array (the name itself is a host-side stand-in)
            e = hipGetSymbolAddress(&g_noalloc_ptr[dev], HIP_SYMBOL(scratch));
            if (e != hipSuccess) {
                g_noalloc_ptr[dev] = nullptr;
                return nla::hip_error(e, "hipGetSymbolAddress(scratch)");
            }
        }
        ptr = g_noalloc_ptr[dev];
    }
    nla::prof_suppress(true);
    const int st = nla::rectrxm_entry<double>(side, uplo, trans, alpha, func, n, m, A, lda, B, ldb, cutoff, stream,
                                              ptr, NOALLOC_BYTES);
    nla::prof_suppress(false);
    return st;
}
