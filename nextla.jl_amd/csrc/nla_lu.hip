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

constexpr int LU_LEAF = 16;    // widest leaf (the multi-workgroup leaf of tall panels)
constexpr int LU_LEAF_R = 8;   // widest leaf of the one-workgroup kernels (registers: wider spills)
constexpr int LU_THREADS = 1024;
static_assert(2 * LU_LEAF <= 64, "a leaf's interchanges touch at most one wave's worth of rows");
constexpr unsigned long long LU_NO_ZERO = ~0ull;

// f(integral_constant<int, J>) for J = 0, 1, ... in order: a column loop whose register
// indices are constants (the multi-workgroup leaf's sixteen columns)
template <typename F, int... J> __device__ __forceinline__ void lu_for_cols(F &&f, std::integer_sequence<int, J...>) {
    (f(std::integral_constant<int, J>()), ...);
}

template <typename T> __device__ __forceinline__ T lu_sfmin();
template <> __device__ __forceinline__ double lu_sfmin<double>() { return 2.2250738585072014e-308; }
template <> __device__ __forceinline__ float lu_sfmin<float>() { return 1.17549435e-38f; }

// (v, i) over the wave: the largest v, ties to the smaller i, in every lane.  Cross-lane steps
// without the LDS crossbar: v_permlane32_swap / v_permlane16_swap pair lanes l and l ^ 32 /
// l ^ 16, DPP rotations by 8 and 4 within each row of 16 and quad permutations do the rest; the
// selection is a maximum over (v, -i), so the order of the pairings does not change it.
template <typename T> __device__ __forceinline__ void lu_pick(T &v, int &i, T ov, int oi) {
    if (ov > v || (ov == v && oi < i)) { v = ov; i = oi; }
}
template <typename T> __device__ __forceinline__ void lu_wave_argmax(T &v, int &i) {
    {
        T x = v, y = v;
        float fi = __int_as_float(i), fj = fi;
        pl32_swap(x, y);
        pl32_swap(fi, fj);
        v = x; i = __float_as_int(fi);
        lu_pick(v, i, y, __float_as_int(fj));
    }
    {
        T x = v, y = v;
        float fi = __int_as_float(i), fj = fi;
        pl16_swap(x, y);
        pl16_swap(fi, fj);
        v = x; i = __float_as_int(fi);
        lu_pick(v, i, y, __float_as_int(fj));
    }
    lu_pick(v, i, dpp_mov<DPP_ROW_ROR8>(v), __float_as_int(dpp_mov<DPP_ROW_ROR8>(__int_as_float(i))));
    lu_pick(v, i, dpp_mov<DPP_ROW_ROR4>(v), __float_as_int(dpp_mov<DPP_ROW_ROR4>(__int_as_float(i))));
    lu_pick(v, i, dpp_mov<DPP_QUAD_X2>(v), __float_as_int(dpp_mov<DPP_QUAD_X2>(__int_as_float(i))));
    lu_pick(v, i, dpp_mov<DPP_QUAD_X1>(v), __float_as_int(dpp_mov<DPP_QUAD_X1>(__int_as_float(i))));
}

// One leaf: columns 0..nb-1 of the m-row block at A (its row 0 is row `base` of the matrix).
// Per column j: the pivot (the first row of largest |a| in rows j..m-1, ties to the smaller row
// as idamax: per-thread candidates, a wave reduction, then every thread reduces the waves'
// candidates from LDS), the swap of rows j and p by one thread per column (which also puts the
// new row j into LDS), then the scaling and rank-1 update of rows j+1..m-1.  A thread takes its
// rows LU_RB at a time with all their loads in flight before any arithmetic, and notes its
// candidate for column j+1 while it has the updated values (its rows of column j+1 are exactly
// the rows of its next pivot search), so a column costs about three memory round trips.
constexpr int LU_RB = 4;
template <typename T>
__global__ __launch_bounds__(LU_THREADS) void lu_leaf_kernel(int m, int nb, T *__restrict__ A, i64 lda,
                                                             int64_t *__restrict__ ipiv, i64 base,
                                                             unsigned long long *info) {
    __shared__ T sval[LU_THREADS / 64];
    __shared__ int sidx[LU_THREADS / 64];
    __shared__ T prow[LU_LEAF_R];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int kmax = min(m, nb);
    T best = T(-1);
    int bi = m;
    bool have = false;  // best / bi hold this thread's candidate for column j
#pragma unroll
    for (int j = 0; j < LU_LEAF_R; ++j) {
        if (j >= kmax) break;
        if (!have) {
            const T *col = A + (i64)j * lda;
            best = T(-1);
            bi = m;
            for (int r = j + t; r < m; r += LU_THREADS) {
                const T v = fabs(col[r]);
                if (v > best) { best = v; bi = r; }
            }
        }
        lu_wave_argmax(best, bi);
        if (lane == 0) { sval[w] = best; sidx[w] = bi; }
        __syncthreads();
        T b = sval[0];
        int p = sidx[0];
#pragma unroll
        for (int q = 1; q < LU_THREADS / 64; ++q)
            if (sval[q] > b || (sval[q] == b && sidx[q] < p)) { b = sval[q]; p = sidx[q]; }
        if (t == 0) ipiv[base + j] = base + p + 1;
        if (b == T(0)) {
            // exact zero pivot (the whole column below the diagonal is zero): no swap, no scaling
            if (t == 0) atomicMin(info, (unsigned long long)(base + j + 1));
            have = false;
            __syncthreads();  // the waves' candidates are read before the next column's
            continue;
        }
        if (t < nb) {
            T *a = A + (i64)t * lda;
            const T x = a[j], y = a[p];
            if (p != j) {
                a[j] = y;
                a[p] = x;
            }
            prow[t] = y;
        }
        __syncthreads();
        const T d = prow[j];
        const bool recip = fabs(d) >= lu_sfmin<T>();   // src/lu.jl:238-242
        const T rd = T(1) / d;
        T nbest = T(-1);
        int nbi = m;
        for (int r0 = j + 1 + t; r0 < m; r0 += LU_RB * LU_THREADS) {
            T v[LU_RB][LU_LEAF_R];
#pragma unroll
            for (int q = 0; q < LU_RB; ++q) {
                const int r = r0 + q * LU_THREADS;
#pragma unroll
                for (int c = j; c < LU_LEAF_R; ++c)
                    if (r < m && c < nb) v[q][c] = A[r + (i64)c * lda];
            }
#pragma unroll
            for (int q = 0; q < LU_RB; ++q) {
                const int r = r0 + q * LU_THREADS;
                if (r < m) {
                    const T l = recip ? v[q][j] * rd : v[q][j] / d;
                    v[q][j] = l;
#pragma unroll
                    for (int c = j + 1; c < LU_LEAF_R; ++c)
                        if (c < nb) v[q][c] -= l * prow[c];
                    if (j + 1 < LU_LEAF_R && j + 1 < nb) {
                        const T a1 = fabs(v[q][j + 1 < LU_LEAF_R ? j + 1 : j]);
                        if (a1 > nbest) { nbest = a1; nbi = r; }
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < LU_RB; ++q) {
                const int r = r0 + q * LU_THREADS;
#pragma unroll
                for (int c = j; c < LU_LEAF_R; ++c)
                    if (r < m && c < nb) A[r + (i64)c * lda] = v[q][c];
            }
        }
        best = nbest;
        bi = nbi;
        have = true;
        __syncthreads();
    }
}

// The leaf for m <= LU_THREADS * Q rows (m >= nb), its rows held in registers for all its
// columns: thread t owns rows t + LU_THREADS q.  Per column one barrier: each wave's winning
// lane puts its candidate row (all eight columns) into LDS and thread j its row j (row j < 8
// always lives in thread j), so after the barrier every thread has the pivot row and row j
// without a memory round trip; the swap, scaling and rank-1 update are register moves.  The
// LDS slots alternate between even and odd columns, so a fast wave writing column j+1's
// candidates cannot overwrite what a slow one still reads for column j.  Same operations in
// the same order as lu_leaf_kernel, so the factors are identical.
template <typename T, int Q>
__global__ __launch_bounds__(LU_THREADS) void lu_leaf_rows_kernel(int m, int nb, T *__restrict__ A, i64 lda,
                                                                 int64_t *__restrict__ ipiv, i64 base,
                                                                 unsigned long long *info) {
    constexpr int NW = LU_THREADS / 64;
    __shared__ T cval[2][NW];
    __shared__ int cidx[2][NW];
    __shared__ T crow[2][NW][LU_LEAF_R];
    __shared__ T jrow[2][LU_LEAF_R];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    T a[Q][LU_LEAF_R];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int r = t + q * LU_THREADS;
#pragma unroll
        for (int c = 0; c < LU_LEAF_R; ++c) a[q][c] = r < m && c < nb ? A[r + (i64)c * lda] : T(0);
    }
#pragma unroll
    for (int j = 0; j < LU_LEAF_R; ++j) {
        if (j >= nb) break;
        const int par = j & 1;
        T best = T(-1);
        int bi = m, bq = 0;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int r = t + q * LU_THREADS;
            if (r >= j && r < m) {
                const T v = fabs(a[q][j]);
                if (v > best) { best = v; bi = r; bq = q; }
            }
        }
        T wb = best;
        int wi = bi;
        lu_wave_argmax(wb, wi);
        if (wi < m && bi == wi) {  // this wave's winning lane (row indices are distinct)
            cval[par][w] = wb;
            cidx[par][w] = wi;
#pragma unroll
            for (int c = 0; c < LU_LEAF_R; ++c) {
                T x = a[0][c];
#pragma unroll
                for (int q = 1; q < Q; ++q) x = bq == q ? a[q][c] : x;
                crow[par][w][c] = x;
            }
        } else if (wi >= m && lane == 0) {  // no rows of this wave in range
            cval[par][w] = T(-1);
            cidx[par][w] = m;
        }
        if (t == j) {
#pragma unroll
            for (int c = 0; c < LU_LEAF_R; ++c) jrow[par][c] = a[0][c];
        }
        __syncthreads();
        T b = cval[par][0];
        int p = cidx[par][0], pw = 0;
#pragma unroll
        for (int i = 1; i < NW; ++i)
            if (cval[par][i] > b || (cval[par][i] == b && cidx[par][i] < p)) { b = cval[par][i]; p = cidx[par][i]; pw = i; }
        if (t == 0) ipiv[base + j] = base + p + 1;
        if (b == T(0)) {
            // exact zero pivot (the whole column below the diagonal is zero): no swap, no scaling
            if (t == 0) atomicMin(info, (unsigned long long)(base + j + 1));
            continue;
        }
        T prow[LU_LEAF_R];
#pragma unroll
        for (int c = 0; c < LU_LEAF_R; ++c) prow[c] = crow[par][pw][c];
        const T d = prow[j];
        const bool recip = fabs(d) >= lu_sfmin<T>();   // src/lu.jl:238-242
        const T rd = T(1) / d;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int r = t + q * LU_THREADS;
            if (p != j) {
                if (r == j) {
#pragma unroll
                    for (int c = 0; c < LU_LEAF_R; ++c) a[q][c] = prow[c];
                } else if (r == p) {
#pragma unroll
                    for (int c = 0; c < LU_LEAF_R; ++c) a[q][c] = jrow[par][c];
                }
            }
            if (r > j && r < m) {
                const T l = recip ? a[q][j] * rd : a[q][j] / d;
                a[q][j] = l;
#pragma unroll
                for (int c = j + 1; c < LU_LEAF_R; ++c) a[q][c] -= l * prow[c];
            }
        }
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int r = t + q * LU_THREADS;
#pragma unroll
        for (int c = 0; c < LU_LEAF_R; ++c)
            if (r < m && c < nb) A[r + (i64)c * lda] = a[q][c];
    }
}

// The leaf for tall panels (more rows than the one-workgroup register leaf takes): P
// workgroups of 256 threads, workgroup g holding rows g RW .. g RW + RW - 1 (RW = 256 LU_CQ) in
// registers for all the leaf's columns.  Per column: each workgroup takes its candidate as
// lu_leaf_rows_kernel does and publishes it, with the candidate's row (and workgroup 0 row j),
// in a slot; wave 0 of every workgroup polls the P slots (marked words, below) until all are
// complete, and every workgroup takes the same pivot
// (the largest, ties to the smaller row: idamax over the whole column) and applies the swap,
// the scaling and the rank-1 update to its rows, as the one-workgroup leaves do, so the factors
// are the same.  One exchange between the workgroups per column replaces a pass of the panel
// through one CU's memory path.  P <= the CU count, so the P workgroups are resident together;
// the wait is bounded (2 s, then the error flag, after which nothing waits).
// Wave 0 of a workgroup: the kb interchanges of a leaf (row base + k with row piv(k), absolute,
// in order) composed into a permutation of the rows involved: afterwards rows[i] must receive
// the old content of row rows[src[i]], i < *nrows (at most 2 kb rows).  Lane j holds slot j (its
// row and the slot whose old value it now holds); slots 0..kb-1 are rows base..base+kb-1, later
// slots the pivot rows below.
template <typename F>
__device__ __forceinline__ void lu_compose_swaps(int kb, i64 base, F piv, i64 *rows, int *src, int *nrows) {
    const int t = threadIdx.x;
    i64 row = t < kb ? base + t : -1;
    int content = t;
    int n = kb;
    for (int k = 0; k < kb; ++k) {
        const i64 r = piv(k);
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
    if (t == 0) *nrows = n;
}
// One column a of the matrix: the composed interchanges, every load before any store.
template <typename T> __device__ __forceinline__ void lu_swap_column(T *a, const i64 *rows, const int *src, int n) {
    T v[2 * LU_LEAF];
#pragma unroll
    for (int i = 0; i < 2 * LU_LEAF; ++i)
        if (i < n) v[i] = a[rows[src[i]]];
#pragma unroll
    for (int i = 0; i < 2 * LU_LEAF; ++i)
        if (i < n) a[rows[i]] = v[i];
}

// device-coherent (agent scope) loads and stores of the slots and counters: they bypass the
// non-coherent caches, so a value written on one XCD is read correctly on another
template <typename T> __device__ __forceinline__ T lu_ld(const T *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T> __device__ __forceinline__ void lu_st(T *p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
constexpr int LU_CQ = 1;   // rows per thread (8192 rows: 32 workgroups; 2: +2 ms, 4: +6.5 ms per 8192^2 LU)
constexpr int LU_CRW = 256 * LU_CQ;   // rows per workgroup
constexpr int LU_CSL = 2 + LU_LEAF;   // slot: |candidate|, its row, the row's values
// Every slot word starts as a NaN of a fixed payload (arithmetic on numbers yields the default
// quiet NaN, never this one): a slot is complete when none of its words reads as the mark, and
// a word is stored and loaded whole, so the workgroups need no flag, no counter and no wait for
// their stores' completion.  Two slot sets alternate between consecutive leaves: a leaf marks
// the other set, which the previous leaf (finished before this launch) used and the next uses.
template <typename T> struct LuMark;
template <> struct LuMark<double> {
    using U = unsigned long long;
    static constexpr U bits = 0x7ff4a5a5a5a5a5a5ull;
};
template <> struct LuMark<float> {
    using U = unsigned;
    static constexpr U bits = 0x7fa5a5a5u;
};
template <typename T> __device__ __forceinline__ bool lu_unset(T v) {
    return __builtin_bit_cast(typename LuMark<T>::U, v) == LuMark<T>::bits;
}
// words of one slot set for at most pmax workgroups: LU_LEAF columns x pmax slots, then row j
// of each column
__host__ __device__ constexpr i64 lu_slot_set(int pmax) { return (i64)LU_LEAF * pmax * LU_CSL + LU_LEAF * LU_LEAF; }
template <typename T> __global__ void lu_mark_kernel(T *p, i64 n) {
    const T mark = __builtin_bit_cast(T, LuMark<T>::bits);
    for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) p[i] = mark;
}

template <typename T>
__global__ __launch_bounds__(256) void lu_leaf_coop_kernel(int m, int nb, T *__restrict__ A, i64 lda,
                                                           int64_t *__restrict__ ipiv, i64 base,
                                                           unsigned long long *info, T *slots, int pmax, int par,
                                                           int *err) {
    const int P = gridDim.x, g = blockIdx.x, t = threadIdx.x, lane = t & 63, w = t >> 6;
    __shared__ T cval[4];
    __shared__ int cidx[4];
    __shared__ T prow_s[LU_LEAF], jrow_s[LU_LEAF];
    __shared__ T gb_s;
    __shared__ int gp_s;
    const i64 set = lu_slot_set(pmax);
    T *R = slots + par * set;  // this leaf's slots
    {                          // mark the other set for the next leaf
        T *O = slots + (1 - par) * set;
        const T mark = __builtin_bit_cast(T, LuMark<T>::bits);
        for (i64 i = g * 256 + t; i < set; i += (i64)P * 256) O[i] = mark;
    }
    T a[LU_CQ][LU_LEAF];
#pragma unroll
    for (int q = 0; q < LU_CQ; ++q) {
        const int r = g * LU_CRW + q * 256 + t;
#pragma unroll
        for (int c = 0; c < LU_LEAF; ++c) a[q][c] = r < m && c < nb ? A[r + (i64)c * lda] : T(0);
    }
    T *rowj = R + (i64)LU_LEAF * pmax * LU_CSL;  // row j of each column (workgroup 0 holds rows 0..LU_LEAF-1)
    // lane-level poll: *p once it is no longer the mark (bounded: a timeout or another
    // workgroup's error raises the error flag and the value is taken as it is)
    auto fetch = [&](const T *p) -> T {
        T v = lu_ld(p);
        if (!lu_unset(v)) return v;
        const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
        while (lu_unset(v = lu_ld(p))) {
            if (lu_ld(err) || __builtin_amdgcn_s_memrealtime() - t0 > 200000000ull) {
                lu_st(err, 1);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        return v;
    };
    // one column; a lambda over a constant column index so that every register index is constant
    auto column = [&](auto J) __attribute__((always_inline)) {
        constexpr int j = decltype(J)::value;
        if (j >= nb) return;
        T best = T(-1);
        int bi = m;
#pragma unroll
        for (int q = 0; q < LU_CQ; ++q) {
            const int r = g * LU_CRW + q * 256 + t;
            if (r >= j && r < m) {
                const T v = fabs(a[q][j]);
                if (v > best) { best = v; bi = r; }
            }
        }
        T wb = best;
        int wi = bi;
        lu_wave_argmax(wb, wi);
        if (lane == 0) {
            cval[w] = wb;
            cidx[w] = wi;
        }
        __syncthreads();
        T b = cval[0];
        int p = cidx[0];
#pragma unroll
        for (int i = 1; i < 4; ++i) lu_pick(b, p, cval[i], cidx[i]);
        T *S = R + ((i64)j * pmax + g) * LU_CSL;
        if (t == 0) {
            lu_st(&S[0], b);
            lu_st(&S[1], (T)p);
        }
#pragma unroll
        for (int q = 0; q < LU_CQ; ++q)  // the thread holding the workgroup's candidate row
            if (g * LU_CRW + q * 256 + t == p) {
#pragma unroll
                for (int c = 0; c < LU_LEAF; ++c) lu_st(&S[2 + c], a[q][c]);
            }
        if (g == 0 && t == j) {
#pragma unroll
            for (int c = 0; c < LU_LEAF; ++c) lu_st(&rowj[j * LU_LEAF + c], a[0][c]);
        }
        if (w == 0) {  // the pivot over the P candidates (each lane polls its slots), its row and row j
            T cb = T(-1), crow[LU_LEAF];
            int cp = m;
            T jv = lane < LU_LEAF ? lu_ld(&rowj[j * LU_LEAF + lane]) : T(0);  // in flight with the slots
#pragma unroll
            for (int c = 0; c < LU_LEAF; ++c) crow[c] = T(0);
            for (int i = lane; i < P; i += 64) {
                const T *Si = R + ((i64)j * pmax + i) * LU_CSL;
                T sv[LU_CSL];
                unsigned long long t0 = 0;
                for (;;) {  // the whole slot per round trip until no word is the mark (bounded)
#pragma unroll
                    for (int k = 0; k < LU_CSL; ++k) sv[k] = lu_ld(&Si[k]);
                    bool unset = false;
#pragma unroll
                    for (int k = 0; k < LU_CSL; ++k) unset = unset || lu_unset(sv[k]);
                    if (!unset) break;
                    const unsigned long long tn = __builtin_amdgcn_s_memrealtime();
                    if (t0 == 0) t0 = tn;
                    else if (lu_ld(err) || tn - t0 > 200000000ull) {
                        lu_st(err, 1);
                        break;
                    }
                }
                const int si = (int)sv[1];
                if (sv[0] > cb || (sv[0] == cb && si < cp)) {
                    cb = sv[0];
                    cp = si;
#pragma unroll
                    for (int c = 0; c < LU_LEAF; ++c) crow[c] = sv[2 + c];
                }
            }
            T gb = cb;
            int gp = cp;
            lu_wave_argmax(gb, gp);
            if (cp == gp && gp < m) {
#pragma unroll
                for (int c = 0; c < LU_LEAF; ++c) prow_s[c] = crow[c];
            }
            if (lane < LU_LEAF) jrow_s[lane] = lu_unset(jv) ? fetch(&rowj[j * LU_LEAF + lane]) : jv;
            if (lane == 0) {
                gb_s = gb;
                gp_s = gp;
            }
        }
        __syncthreads();
        const T gb = gb_s;
        const int gp = gp_s >= j && gp_s < m ? gp_s : j;  // (out of range only after a timed-out wait)
        if (g == 0 && t == 0) ipiv[base + j] = base + gp + 1;
        if (gb == T(0)) {
            // exact zero pivot (the whole column below the diagonal is zero): no swap, no scaling
            if (g == 0 && t == 0) atomicMin(info, (unsigned long long)(base + j + 1));
            return;
        }
        T prow[LU_LEAF];
#pragma unroll
        for (int c = 0; c < LU_LEAF; ++c) prow[c] = prow_s[c];
        const T d = prow[j];
        const bool recip = fabs(d) >= lu_sfmin<T>();   // src/lu.jl:238-242
        const T rd = T(1) / d;
#pragma unroll
        for (int q = 0; q < LU_CQ; ++q) {
            const int r = g * LU_CRW + q * 256 + t;
            if (gp != j) {
                if (r == j) {
#pragma unroll
                    for (int c = 0; c < LU_LEAF; ++c) a[q][c] = prow[c];
                } else if (r == gp) {
#pragma unroll
                    for (int c = 0; c < LU_LEAF; ++c) a[q][c] = jrow_s[c];
                }
            }
            if (r > j && r < m) {
                const T l = recip ? a[q][j] * rd : a[q][j] / d;
                a[q][j] = l;
#pragma unroll
                for (int c = j + 1; c < LU_LEAF; ++c) a[q][c] -= l * prow[c];
            }
        }
    };
    lu_for_cols(column, std::make_integer_sequence<int, LU_LEAF>());
#pragma unroll
    for (int q = 0; q < LU_CQ; ++q) {
        const int r = g * LU_CRW + q * 256 + t;
#pragma unroll
        for (int c = 0; c < LU_LEAF; ++c)
            if (r < m && c < nb) A[r + (i64)c * lda] = a[q][c];
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
// permutation of the (at most 2 kb) rows involved in LDS (the kb pivots read by one load per
// lane, not one dependent load per interchange), then each thread gathers one column's values
// of those rows (all loads in flight together) and writes them back permuted, so a leaf costs
// one round trip instead of kb dependent swaps per column.  One-wave workgroups spread the
// columns' scattered row accesses over four times as many CUs as 256-thread ones.  Columns
// [skip0, skip1) (the leaf's own, swapped by the leaf) are left alone.
constexpr int LU_SWT = 64;   // threads (columns) per workgroup of the interchanges
template <typename T>
__global__ __launch_bounds__(LU_SWT) void lu_leaf_swaps_kernel(int ncols, T *__restrict__ A, i64 lda, i64 base, int kb,
                                                               const int64_t *__restrict__ ipiv, int skip0, int skip1) {
    __shared__ i64 rows[2 * LU_LEAF];   // distinct rows involved
    __shared__ int src[2 * LU_LEAF];    // final content of rows[i] = old row rows[src[i]]
    __shared__ int nrows;
    const int t = threadIdx.x;
    const i64 mypiv = t < kb ? (i64)(ipiv[base + t] - 1) : 0;
    lu_compose_swaps(kb, base, [&](int k) { return (i64)__shfl((long long)mypiv, k, 64); }, rows, src, &nrows);
    __syncthreads();
    int c = blockIdx.x * LU_SWT + t;
    if (c >= ncols - (skip1 - skip0)) return;
    if (c >= skip0) c += skip1 - skip0;
    lu_swap_column(A + (i64)c * lda, rows, src, nrows);
}

// A12 := L11^-1 A12 for the recursion's small nodes (n1 <= N1 <= 32): L11 the unit lower
// triangle of A's leading n1 x n1 block, read in place; one thread per column of A12 holds the
// column in registers and runs the forward substitution in LAPACK's column order (for k: for
// i > k: x_i -= L_ik x_k), L11 staged in LDS.  Replaces the unit-diagonal copy and the blocked
// recursive solve (three launches) at these sizes.
template <typename T, int N1>
__global__ __launch_bounds__(256) void lu_trsm_small_kernel(int n1, int n2, const T *__restrict__ L, i64 lda,
                                                            T *__restrict__ B, i64 ldb) {
    __shared__ T Ls[N1 * N1];
    for (int idx = threadIdx.x; idx < n1 * n1; idx += 256) {
        const int r = idx % n1, c = idx / n1;
        Ls[c * N1 + r] = L[r + (i64)c * lda];
    }
    __syncthreads();
    const int col = blockIdx.x * 256 + threadIdx.x;
    if (col >= n2) return;
    T *b = B + (i64)col * ldb;
    T x[N1];
#pragma unroll
    for (int i = 0; i < N1; ++i) x[i] = i < n1 ? b[i] : T(0);
#pragma unroll
    for (int k = 0; k < N1; ++k) {
        if (k < n1) {
            const T xk = x[k];
#pragma unroll
            for (int i = k + 1; i < N1; ++i)
                if (i < n1) x[i] -= Ls[k * N1 + i] * xk;
        }
    }
#pragma unroll
    for (int i = 0; i < N1; ++i)
        if (i < n1) b[i] = x[i];
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
__global__ void lu_int_fill_kernel(int *p, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}
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
// Device state of the multi-workgroup leaves: the two candidate slot sets (marked by the entry
// point), the error flag of their bounded wait; ncu caps the workgroups of one leaf; par is the
// slot set of the next such leaf; used is set once one has been launched.
template <typename T> struct LuCoop {
    T *slots;
    int *err;
    int ncu, par;
    bool used;
};

template <typename T>
static int getrf2_rec(i64 m, i64 n, T *A, i64 lda, int64_t *ipiv, i64 base, unsigned long long *flag, T *Lbuf,
                      T *A0, i64 ncols0, LuCoop<T> &co, hipStream_t s) {
    if (m <= 0 || n <= 0) return 0;
    // leaf width: LU_LEAF on the multi-workgroup leaf (panels taller than the one-workgroup
    // register leaves take), LU_LEAF_R elsewhere
    const i64 pw = (m + LU_CRW - 1) / LU_CRW;  // workgroups of a multi-workgroup leaf
    const i64 rmax = (sizeof(T) == 4 ? 4 : 2) * (i64)LU_THREADS;
    const i64 lw = m > rmax && pw <= co.ncu ? LU_LEAF : LU_LEAF_R;
    if (n <= lw || m == 1) {
        const int kb = (int)(n < lw ? n : lw);
        if (m >= kb && m <= LU_THREADS) lu_leaf_rows_kernel<T, 1><<<1, LU_THREADS, 0, s>>>((int)m, kb, A, lda, ipiv, base, flag);
        else if (m >= kb && m <= 2 * LU_THREADS) lu_leaf_rows_kernel<T, 2><<<1, LU_THREADS, 0, s>>>((int)m, kb, A, lda, ipiv, base, flag);
        else if (sizeof(T) == 4 && m >= kb && m <= 4 * LU_THREADS)  // (Float64 at four rows per thread spills)
            lu_leaf_rows_kernel<T, sizeof(T) == 4 ? 4 : 1><<<1, LU_THREADS, 0, s>>>((int)m, kb, A, lda, ipiv, base, flag);
        else if (m >= kb && pw <= co.ncu) {
            lu_leaf_coop_kernel<T><<<(unsigned)pw, 256, 0, s>>>((int)m, kb, A, lda, ipiv, base, flag, co.slots, co.ncu,
                                                               co.par, co.err);
            co.par ^= 1;
            co.used = true;
        } else lu_leaf_kernel<T><<<1, LU_THREADS, 0, s>>>((int)m, kb, A, lda, ipiv, base, flag);
        NLA_LAUNCH_CHECK();
        const int kpiv = (int)(m < kb ? m : kb);
        const i64 others = ncols0 - kb;
        if (others > 0 && kpiv > 0) {
            lu_leaf_swaps_kernel<T><<<(unsigned)((others + LU_SWT - 1) / LU_SWT), LU_SWT, 0, s>>>(
                (int)ncols0, A0, lda, base, kpiv, ipiv, (int)base, (int)(base + kb));
            NLA_LAUNCH_CHECK();
        }
        return 0;
    }
    const i64 mn = m < n ? m : n, n1 = mn / 2, n2 = n - n1;
    NLA_ST(getrf2_rec<T>(m, n1, A, lda, ipiv, base, flag, Lbuf, A0, ncols0, co, s));
    // A12 = L11^-1 A12 (L11 unit lower): in place for small n1, else the recursive TRSM on a
    // unit-diagonal copy
    if (n1 <= 32) {
        const unsigned g = (unsigned)((n2 + 255) / 256);
        if (n1 <= 8) lu_trsm_small_kernel<T, 8><<<g, 256, 0, s>>>((int)n1, (int)n2, A, lda, A + n1 * lda, lda);
        else if (n1 <= 16) lu_trsm_small_kernel<T, 16><<<g, 256, 0, s>>>((int)n1, (int)n2, A, lda, A + n1 * lda, lda);
        else lu_trsm_small_kernel<T, 32><<<g, 256, 0, s>>>((int)n1, (int)n2, A, lda, A + n1 * lda, lda);
        NLA_LAUNCH_CHECK();
    } else {
        const i64 tot = n1 * n1;
        const int blocks = (int)((tot + 255) / 256 < 2048 ? (tot + 255) / 256 : 2048);
        unit_lower_copy_kernel<T><<<blocks, 256, 0, s>>>((int)n1, A, lda, Lbuf);
        NLA_LAUNCH_CHECK();
        NLA_ST(rectrxm_async<T>('L', 'L', 'N', T(1), 'S', n1, n2, Lbuf, n1, A + n1 * lda, lda, 0, s));
    }
    // A22 -= A21 A12
    NLA_HIP(gemm<T>('N', 'N', m - n1, n2, n1, T(-1), A + n1, lda, A + n1 * lda, lda, T(1), A + n1 + n1 * lda, lda, s));
    NLA_ST(getrf2_rec<T>(m - n1, n2, A + n1 + n1 * lda, lda, ipiv, base + n1, flag, Lbuf, A0, ncols0, co, s));
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
    int dev = 0, ncu = 0;
    NLA_HIP(hipGetDevice(&dev));
    NLA_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    int st = 0;
    // the zero-pivot flag, the unit-lower copies of L11 (at most (min(m,n)/2)^2), the
    // multi-workgroup leaves' two slot sets and error flag
    const size_t nslot = 2 * (size_t)lu_slot_set(ncu);
    const size_t lbytes = (sizeof(T) * (size_t)(h * h + 1) + 63) / 64 * 64, sbytes = (sizeof(T) * nslot + 63) / 64 * 64;
    char *base = (char *)scratch(SLOT_LU, 64 + lbytes + sbytes + 64, s, &st);
    if (!base) return st;
    unsigned long long *flag = reinterpret_cast<unsigned long long *>(base);
    T *Lbuf = reinterpret_cast<T *>(base + 64);
    LuCoop<T> co{reinterpret_cast<T *>(base + 64 + lbytes), reinterpret_cast<int *>(base + 64 + lbytes + sbytes), ncu, 0,
                 false};
    lu_info_init_kernel<<<1, 1, 0, s>>>(flag);
    NLA_LAUNCH_CHECK();
    lu_int_fill_kernel<<<1, 64, 0, s>>>(co.err, 1);
    NLA_LAUNCH_CHECK();
    lu_mark_kernel<T><<<(unsigned)((nslot + 255) / 256 < 1024 ? (nslot + 255) / 256 : 1024), 256, 0, s>>>(co.slots, (i64)nslot);
    NLA_LAUNCH_CHECK();
    NLA_ST(getrf2_rec<T>(m, n, A, lda, ipiv, 0, flag, Lbuf, A, n, co, s));
    lu_info_finish_kernel<<<1, 1, 0, s>>>(flag, info);
    NLA_LAUNCH_CHECK();
    if (co.used) {  // the multi-workgroup leaves' bounded wait
        int herr = 0;
        NLA_HIP(hipMemcpyAsync(&herr, co.err, sizeof(int), hipMemcpyDeviceToHost, s));
        NLA_HIP(hipStreamSynchronize(s));
        if (herr) return wait_timeout_error("LU leaf");
    }
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
