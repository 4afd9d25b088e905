// Recursive LU with partial pivoting: getrf2! (src/lu.jl:185-300) and laswp
// (src/lu.jl:470-522), Float64 / Float32.
//
// The recursion is the reference's (LAPACK xgetrf2's): split the columns at n1 = min(m,n)/2,
// factor the left part, swap the right part's rows, A12 = L11^-1 A12, A22 -= A21 A12, factor
// A22, swap the left part's rows.  On the GPU the two updates are this library's recursive
// TRSM (on a unit-diagonal copy of L11) and MFMA GEMM; the recursion stops at leaves of at
// most LEAF columns, factored by one workgroup (a block-wide pivot search per column, the
// row swap, the scaling and the rank-1 update of the leaf's remaining columns), and each
// leaf's interchanges are applied to all other columns right after it, in place of the
// recursion's laswp steps (the same interchanges reach every column in the same order).
// Pivots are stored 1-based and absolute (row indices of the whole matrix).  info: the first
// zero pivot (1-based), 0 if none, written on the device.
#include "nla_common.hpp"
#include "nla_gemm.hpp"

namespace nla {

template <typename T>
int rectrxm_async(char side, char uplo, char trans, T alpha, char func, i64 n, i64 m, const T *A, i64 lda, T *B,
                  i64 ldb, i64 cutoff, hipStream_t s);

constexpr int LU_LEAF = 8;     // widest leaf factored by one workgroup
constexpr int LU_THREADS = 1024;
static_assert(2 * LU_LEAF <= 64, "a leaf's interchanges touch at most one wave's worth of rows");
constexpr unsigned long long LU_NO_ZERO = ~0ull;

template <typename T> __device__ __forceinline__ T lu_sfmin();
template <> __device__ __forceinline__ double lu_sfmin<double>() { return 2.2250738585072014e-308; }
template <> __device__ __forceinline__ float lu_sfmin<float>() { return 1.17549435e-38f; }

// One leaf: columns 0..nb-1 of the m-row block at A (its row 0 is row `base` of the matrix).
template <typename T>
__global__ __launch_bounds__(LU_THREADS) void lu_leaf_kernel(int m, int nb, T *__restrict__ A, i64 lda,
                                                             int64_t *__restrict__ ipiv, i64 base,
                                                             unsigned long long *info) {
    __shared__ T sval[LU_THREADS / 64];
    __shared__ int sidx[LU_THREADS / 64];
    __shared__ T prow[LU_LEAF];
    __shared__ int sp;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int kmax = min(m, nb);
    for (int j = 0; j < kmax; ++j) {
        T *col = A + (i64)j * lda;
        // the first row of largest |a| in rows j..m-1 (ties keep the smaller row, as idamax)
        T best = T(-1);
        int bi = m;
        for (int r = j + t; r < m; r += LU_THREADS) {
            const T v = fabs(col[r]);
            if (v > best) { best = v; bi = r; }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const T ov = __shfl_xor(best, off, 64);
            const int oi = __shfl_xor(bi, off, 64);
            if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
        }
        if (lane == 0) { sval[w] = best; sidx[w] = bi; }
        __syncthreads();
        if (t == 0) {
            T b = sval[0];
            int i0 = sidx[0];
            for (int q = 1; q < LU_THREADS / 64; ++q)
                if (sval[q] > b || (sval[q] == b && sidx[q] < i0)) { b = sval[q]; i0 = sidx[q]; }
            sp = i0;
            ipiv[base + j] = base + i0 + 1;
        }
        __syncthreads();
        const int p = sp;
        const T piv = col[p];
        if (piv == T(0)) {
            // exact zero pivot (the whole column below the diagonal is zero): no swap, no scaling
            if (t == 0) atomicMin(info, (unsigned long long)(base + j + 1));
            __syncthreads();
            continue;
        }
        if (p != j && t < nb) {
            T *a = A + (i64)t * lda;
            const T x = a[j];
            a[j] = a[p];
            a[p] = x;
        }
        __syncthreads();
        if (t < nb) prow[t] = A[j + (i64)t * lda];
        __syncthreads();
        const T d = prow[j];
        const bool recip = fabs(d) >= lu_sfmin<T>();   // src/lu.jl:238-242
        const T rd = T(1) / d;
        for (int r = j + 1 + t; r < m; r += LU_THREADS) {
            const T l = recip ? col[r] * rd : col[r] / d;
            col[r] = l;
            for (int c = j + 1; c < nb; ++c) A[r + (i64)c * lda] -= l * prow[c];
        }
        __syncthreads();
    }
}

// Row interchanges k = k1..k2 (absolute, 1-based; reverse order for inc < 0) with ipiv(k) on
// ncols columns of the block at A whose row 0 is row `base` of the matrix: thread per column.
template <typename T>
__global__ __launch_bounds__(256) void laswp_kernel(int ncols, T *__restrict__ A, i64 lda, i64 k1, i64 k2,
                                                   const int64_t *__restrict__ ipiv, i64 base, int inc) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= ncols) return;
    T *a = A + (i64)c * lda;
    for (i64 s = 0; s <= k2 - k1; ++s) {
        const i64 k = inc > 0 ? k1 + s : k2 - s;
        const i64 ip = ipiv[k - 1];
        if (ip != k) {
            const i64 i0 = k - 1 - base, i1 = ip - 1 - base;
            const T x = a[i0];
            a[i0] = a[i1];
            a[i1] = x;
        }
    }
}

// A leaf's row interchanges applied to the matrix's other columns: rows base..base+kb-1 were
// swapped, in order, with rows ipiv(k) - 1 (absolute).  Every workgroup forms the composed
// permutation of the (at most 2 kb) rows involved in LDS, then each thread gathers one column's
// values of those rows (all loads in flight together) and writes them back permuted, so a leaf
// costs one round trip instead of kb dependent swaps per column.  Columns [skip0, skip1) (the
// leaf's own, swapped by the leaf) are left alone.
template <typename T>
__global__ __launch_bounds__(256) void lu_leaf_swaps_kernel(int ncols, T *__restrict__ A, i64 lda, i64 base, int kb,
                                                            const int64_t *__restrict__ ipiv, int skip0, int skip1) {
    __shared__ i64 rows[2 * LU_LEAF];   // distinct rows involved
    __shared__ int src[2 * LU_LEAF];    // final content of rows[i] = old row rows[src[i]]
    __shared__ int nrows;
    const int t = threadIdx.x;
    if (t < 64) {
        // wave 0 composes the swaps: lane j holds slot j (its row and the slot whose old value it
        // now holds); slots 0..kb-1 are rows base..base+kb-1, later slots the pivot rows below
        i64 row = t < kb ? base + t : -1;
        int content = t;
        int n = kb;
        for (int k = 0; k < kb; ++k) {
            const i64 r = ipiv[base + k] - 1;
            int b;
            if (r < base + kb) {
                b = (int)(r - base);
            } else {
                const unsigned long long hit = __ballot(row == r);
                if (hit) {
                    b = __ffsll((long long)hit) - 1;
                } else {
                    b = n++;
                    if (t == b) row = r;
                }
            }
            const int ca = __shfl(content, k, 64), cb = __shfl(content, b, 64);
            if (t == k) content = cb;
            else if (t == b) content = ca;
        }
        if (t < n) {
            rows[t] = row;
            src[t] = content;
        }
        if (t == 0) nrows = n;
    }
    __syncthreads();
    int c = blockIdx.x * 256 + t;
    if (c >= ncols - (skip1 - skip0)) return;
    if (c >= skip0) c += skip1 - skip0;
    T *a = A + (i64)c * lda;
    const int n = nrows;
    T v[2 * LU_LEAF];
#pragma unroll
    for (int i = 0; i < 2 * LU_LEAF; ++i)
        if (i < n) v[i] = a[rows[src[i]]];   // every load before any store
#pragma unroll
    for (int i = 0; i < 2 * LU_LEAF; ++i)
        if (i < n) a[rows[i]] = v[i];
}

// L (n x n, ld n) = unit lower triangle of A's leading n x n block (zeros above).
template <typename T>
__global__ void unit_lower_copy_kernel(int n, const T *__restrict__ A, i64 lda, T *__restrict__ L) {
    const i64 total = (i64)n * n;
    for (i64 idx = (i64)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (i64)gridDim.x * blockDim.x) {
        const int r = (int)(idx % n), c = (int)(idx / n);
        L[idx] = r > c ? A[r + (i64)c * lda] : (r == c ? T(1) : T(0));
    }
}

__global__ void lu_info_init_kernel(unsigned long long *info) { *info = LU_NO_ZERO; }
__global__ void lu_info_finish_kernel(const unsigned long long *flag, int64_t *info) {
    *info = *flag == LU_NO_ZERO ? 0 : (int64_t)*flag;
}

template <typename T>
static hipError_t laswp_launch(i64 ncols, T *A, i64 lda, i64 k1, i64 k2, const int64_t *ipiv, i64 base, int inc,
                               hipStream_t s) {
    if (ncols <= 0 || k2 < k1) return hipSuccess;
    laswp_kernel<T><<<(unsigned)((ncols + 255) / 256), 256, 0, s>>>((int)ncols, A, lda, k1, k2, ipiv, base, inc);
    return hipGetLastError();
}

// getrf2 on the m x n block at A (rows base.. of the matrix, which is also its first column).
// Each leaf's interchanges are applied to every other column of the whole matrix (A0, ncols0
// columns) as soon as the leaf is factored: every column then sees every interchange once, in
// order, as after the recursion's own laswp steps (src/lu.jl:263-293), which are therefore not
// repeated.
template <typename T>
static int getrf2_rec(i64 m, i64 n, T *A, i64 lda, int64_t *ipiv, i64 base, unsigned long long *flag, T *Lbuf,
                      T *A0, i64 ncols0, hipStream_t s) {
    if (m <= 0 || n <= 0) return 0;
    if (n <= LU_LEAF || m == 1) {
        const int kb = (int)(n < LU_LEAF ? n : LU_LEAF);
        lu_leaf_kernel<T><<<1, LU_THREADS, 0, s>>>((int)m, kb, A, lda, ipiv, base, flag);
        NLA_LAUNCH_CHECK();
        const int kpiv = (int)(m < kb ? m : kb);
        const i64 others = ncols0 - kb;
        if (others > 0 && kpiv > 0) {
            lu_leaf_swaps_kernel<T><<<(unsigned)((others + 255) / 256), 256, 0, s>>>(
                (int)ncols0, A0, lda, base, kpiv, ipiv, (int)base, (int)(base + kb));
            NLA_LAUNCH_CHECK();
        }
        return 0;
    }
    const i64 mn = m < n ? m : n, n1 = mn / 2, n2 = n - n1;
    NLA_ST(getrf2_rec<T>(m, n1, A, lda, ipiv, base, flag, Lbuf, A0, ncols0, s));
    // A12 = L11^-1 A12 (L11 unit lower): the recursive TRSM on a unit-diagonal copy
    {
        const i64 tot = n1 * n1;
        const int blocks = (int)((tot + 255) / 256 < 2048 ? (tot + 255) / 256 : 2048);
        unit_lower_copy_kernel<T><<<blocks, 256, 0, s>>>((int)n1, A, lda, Lbuf);
        NLA_LAUNCH_CHECK();
        NLA_ST(rectrxm_async<T>('L', 'L', 'N', T(1), 'S', n1, n2, Lbuf, n1, A + n1 * lda, lda, 0, s));
    }
    // A22 -= A21 A12
    NLA_HIP(gemm<T>('N', 'N', m - n1, n2, n1, T(-1), A + n1, lda, A + n1 * lda, lda, T(1), A + n1 + n1 * lda, lda, s));
    NLA_ST(getrf2_rec<T>(m - n1, n2, A + n1 + n1 * lda, lda, ipiv, base + n1, flag, Lbuf, A0, ncols0, s));
    return 0;
}

template <typename T>
static int getrf2_entry(i64 m, i64 n, T *A, i64 lda, int64_t *ipiv, int64_t *info, void *stream) {
    if (m < 0) return arg_error(1, "illegal value of m: %lld", (long long)m);            // src/lu.jl:196-199
    if (n < 0) return arg_error(2, "illegal value of n: %lld", (long long)n);            // :200-203
    if (lda < (m > 1 ? m : 1)) return arg_error(4, "lda must be >= max(1,m)");          // :204-207
    hipStream_t s = as_stream(stream);
    if (m == 0 || n == 0) {                                                               // :210-212
        if (info) NLA_HIP(hipMemsetAsync(info, 0, sizeof(int64_t), s));
        return 0;
    }
    if (m > INT32_MAX || n > INT32_MAX) return arg_error(1, "dimension exceeds 2^31-1");
    if (!ipiv || !info) return arg_error(5, "ipiv and info must be device pointers");
    const i64 mn = m < n ? m : n, h = mn / 2;
    int st = 0;
    // the unit-lower copies of L11 (at most (min(m,n)/2)^2) and the zero-pivot flag
    char *base = (char *)scratch(SLOT_LU, sizeof(T) * (size_t)(h * h + 1) + 64, s, &st);
    if (!base) return st;
    unsigned long long *flag = reinterpret_cast<unsigned long long *>(base);
    T *Lbuf = reinterpret_cast<T *>(base + 64);
    lu_info_init_kernel<<<1, 1, 0, s>>>(flag);
    NLA_LAUNCH_CHECK();
    NLA_ST(getrf2_rec<T>(m, n, A, lda, ipiv, 0, flag, Lbuf, A, n, s));
    lu_info_finish_kernel<<<1, 1, 0, s>>>(flag, info);
    NLA_LAUNCH_CHECK();
    return 0;
}

template <typename T>
static int laswp_entry(i64 n, T *A, i64 lda, i64 k1, i64 k2, const int64_t *ipiv, i64 incx, void *stream) {
    if (n < 0) return arg_error(1, "n must be non-negative");
    if (incx == 0 || n == 0 || k2 < k1) return 0;  // src/lu.jl:483-485
    if (k1 < 1) return arg_error(4, "first must be >= 1");
    if (lda < 1) return arg_error(3, "lda must be >= 1");
    NLA_HIP(laswp_launch<T>(n, A, lda, k1, k2, ipiv, 0, incx > 0 ? 1 : -1, as_stream(stream)));
    return 0;
}

}  // namespace nla

NLA_API int nla_dgetrf2(int64_t m, int64_t n, double *A, int64_t lda, int64_t *ipiv, int64_t *info, void *stream) {
    return nla::getrf2_entry<double>(m, n, A, lda, ipiv, info, stream);
}
NLA_API int nla_sgetrf2(int64_t m, int64_t n, float *A, int64_t lda, int64_t *ipiv, int64_t *info, void *stream) {
    return nla::getrf2_entry<float>(m, n, A, lda, ipiv, info, stream);
}
NLA_API int nla_dlaswp(int64_t n, double *A, int64_t lda, int64_t k1, int64_t k2, const int64_t *ipiv, int64_t incx,
                       void *stream) {
    return nla::laswp_entry<double>(n, A, lda, k1, k2, ipiv, incx, stream);
}
NLA_API int nla_slaswp(int64_t n, float *A, int64_t lda, int64_t k1, int64_t k2, const int64_t *ipiv, int64_t incx,
                       void *stream) {
    return nla::laswp_entry<float>(n, A, lda, k1, k2, ipiv, incx, stream);
}
