// MFMA GEMM for the dense updates of the path: GEMM_ADD!/GEMM_SUB! of the recursive
// TRSM/TRMM (src/matmul.jl:5-81) and the large products of the QR apply routines.
//
//   C[M x N] = alpha * op(A)[M x K] * op(B)[K x N] + beta * C     (column-major)
//
// Design (CDNA4, one 256-thread workgroup = 4 wave64s per BM x BN output tile):
//  - operands are staged through LDS as k-major panels As[BK][BM], Bs[BK][BN]; the next
//    BK slab is fetched into registers while the current one feeds the matrix cores;
//  - each wave owns a (BM/2) x (BN/2) quadrant made of 16x16 MFMA blocks
//    (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32);
//  - the MFMA's A slot takes the op(B) fragment and its B slot the op(A) fragment, so the
//    accumulator's fastest lane index runs down a column of C and stores coalesce;
//  - workgroups are renumbered so that the tiles an XCD (one L2) executes are adjacent.
#include "nla_common.hpp"
#include "nla_gemm.hpp"

#include <type_traits>

namespace nla {

// LDS row padding: rows stay 16-byte aligned (for the vector slab stores) and a wave's MFMA
// operand reads spread over the banks.
template <typename T> struct GemmPad { static constexpr int v = 16 / (int)sizeof(T); };

// EDGE = false: every tile is interior (M % BM == N % BN == K % BK == 0): no bounds checks.
// NSET register sets for the slabs in flight: 1 fetches slab s+1 while slab s multiplies;
// 2 fetches slab s+2, so a load has two slabs of products to arrive (small tiles, whose
// slab is too short to cover a cache miss).
// VL (interior tiles of op = 'N' operands with 16-byte aligned columns): each thread moves 16
// contiguous bytes of a column per global load (A along m, B along k), a quarter (Float32) or
// half (Float64) of the load instructions of one element per load; A's slab goes to LDS as
// 16-byte stores.
template <typename T, int BM, int BN, int BK, bool TA, bool TB, bool EDGE, int NSET = 1, bool VL = false>
__global__ __launch_bounds__(256, 2) void gemm_kernel(int M, int N, int K, T alpha, const T *__restrict__ A,
                                                      i64 lda, const T *__restrict__ B, i64 ldb, T beta,
                                                      T *__restrict__ C, i64 ldc, int tiles_m) {
    constexpr int PAD = GemmPad<T>::v;
    constexpr int LA = (BM * BK) / 256;  // elements of the A slab per thread
    constexpr int LB = (BK * BN) / 256;
    constexpr int FM = BM / 32;          // 16-row blocks per wave (2 waves along M)
    constexpr int FN = BN / 32;
    __shared__ __align__(16) T As[2][BK][BM + PAD];
    __shared__ __align__(16) T Bs[2][BK][BN + PAD];

    // XCD-aware tile numbering: hardware deals workgroups round-robin over 8 XCDs, so the
    // ids are first made contiguous per XCD; within that order tiles go in groups of GM tile
    // rows (column by column inside a group), so the ~64 tiles an XCD runs at once form an
    // 8 x 8 block that shares 8 A row panels and 8 B column panels through the XCD's L2.
    int bid = blockIdx.x;
    const int nb = gridDim.x;
    if ((nb & 7) == 0) bid = (bid & 7) * (nb >> 3) + (bid >> 3);
    constexpr int GM = 8;
    const int tiles_n = nb / tiles_m;
    const int group = bid / (GM * tiles_n), first_m = group * GM;
    const int gsize = tiles_m - first_m < GM ? tiles_m - first_m : GM;
    const int within = bid - group * GM * tiles_n;
    const int tm = first_m + within % gsize, tn = within / gsize;
    const int m0 = tm * BM, n0 = tn * BN;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm = w & 1, wn = w >> 1;

    // Per-thread slab coordinates: element i of this thread is (am + i*ASTEP_M, ak + i*ASTEP_K).
    // !TA: m = t % BM fixed, k = t / BM + i * (256 / BM);   TA: k = t % BK fixed, m = t / BK + i * (256 / BK)
    const int am = TA ? t / BK : t % BM, ak = TA ? t % BK : t / BM;
    const int bk = TB ? t / BN : t % BK, bn = TB ? t % BN : t / BK;
    constexpr int ASM = TA ? 256 / BK : 0, ASK = TA ? 0 : 256 / BM;
    constexpr int BSK = TB ? 256 / BN : 0, BSN = TB ? 0 : 256 / BK;
    const T *pa = TA ? A + (ak) + (i64)(m0 + am) * lda : A + (m0 + am) + (i64)ak * lda;
    const T *pb = TB ? B + (n0 + bn) + (i64)bk * ldb : B + bk + (i64)(n0 + bn) * ldb;
    const i64 a_step_i = TA ? (i64)ASM * lda : (i64)ASK * lda;   // advance per element i
    const i64 b_step_i = TB ? (i64)BSK * ldb : (i64)BSN * ldb;
    const i64 a_step_k = TA ? BK : (i64)BK * lda;               // advance per K slab
    const i64 b_step_k = TB ? (i64)BK * ldb : BK;

    // vector slab layout (VL): V elements per load; A: m = am_v .. am_v + V - 1 of slab column
    // ak_v + i * AKV; B: k = bk_v .. bk_v + V - 1 of column bn_v + i * BNV
    constexpr int V = 16 / (int)sizeof(T);
    constexpr int LAV = VL ? BM * BK / (256 * V) : 1, LBV = VL ? BK * BN / (256 * V) : 1;
    constexpr int AKV = 256 / (BM / V), BNV = 256 / (BK / V);
    static_assert(!VL || (BM % V == 0 && BK % V == 0 && 256 % (BM / V) == 0 && 256 % (BK / V) == 0 &&
                          BK % AKV == 0 && BN % BNV == 0 && LAV >= 1 && LBV >= 1), "vector slab shape");
    typedef T tvec __attribute__((ext_vector_type(V)));
    const int am_v = (t % (BM / V)) * V, ak_v = t / (BM / V);
    const int bk_v = (t % (BK / V)) * V, bn_v = t / (BK / V);
    const T *pav = A + (m0 + am_v) + (i64)ak_v * lda;
    const T *pbv = B + bk_v + (i64)(n0 + bn_v) * ldb;
    tvec va[NSET][LAV], vb[NSET][LBV];

    T ra[NSET][LA], rb[NSET][LB];
    // slabs are loaded in order: the k0 argument only serves the edge checks
    auto load = [&](int k0, int set) {
        if constexpr (VL) {
#pragma unroll
            for (int i = 0; i < LAV; ++i) va[set][i] = *reinterpret_cast<const tvec *>(pav + (i64)i * AKV * lda);
#pragma unroll
            for (int i = 0; i < LBV; ++i) vb[set][i] = *reinterpret_cast<const tvec *>(pbv + (i64)i * BNV * ldb);
            pav += (i64)BK * lda;
            pbv += BK;
            return;
        }
#pragma unroll
        for (int i = 0; i < LA; ++i) {
            if (EDGE) {
                const int gm = m0 + am + i * ASM, gk = k0 + ak + i * ASK;
                ra[set][i] = (gm < M && gk < K) ? pa[i * a_step_i] : T(0);
            } else {
                ra[set][i] = pa[i * a_step_i];
            }
        }
#pragma unroll
        for (int i = 0; i < LB; ++i) {
            if (EDGE) {
                const int gn = n0 + bn + i * BSN, gk = k0 + bk + i * BSK;
                rb[set][i] = (gn < N && gk < K) ? pb[i * b_step_i] : T(0);
            } else {
                rb[set][i] = pb[i * b_step_i];
            }
        }
        pa += a_step_k;
        pb += b_step_k;
    };
    auto store = [&](int buf, int set) {
        if constexpr (VL) {
#pragma unroll
            for (int i = 0; i < LAV; ++i) *reinterpret_cast<tvec *>(&As[buf][ak_v + i * AKV][am_v]) = va[set][i];
#pragma unroll
            for (int i = 0; i < LBV; ++i)
#pragma unroll
                for (int e = 0; e < V; ++e) Bs[buf][bk_v + e][bn_v + i * BNV] = vb[set][i][e];
            return;
        }
#pragma unroll
        for (int i = 0; i < LA; ++i) As[buf][ak + i * ASK][am + i * ASM] = ra[set][i];
#pragma unroll
        for (int i = 0; i < LB; ++i) Bs[buf][bk + i * BSK][bn + i * BSN] = rb[set][i];
    };

    typename Acc<T>::t acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = acc_zero<T>();

    const int lr = lane & 15, lk = lane >> 4;
    auto multiply = [&](int cur) {
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            T a[FM], b[FN];
#pragma unroll
            for (int i = 0; i < FM; ++i) a[i] = As[cur][kk + lk][wm * (BM / 2) + i * 16 + lr];
#pragma unroll
            for (int j = 0; j < FN; ++j) b[j] = Bs[cur][kk + lk][wn * (BN / 2) + j * 16 + lr];
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j) mfma16(acc[i][j], b[j], a[i]);
        }
    };
    load(0, 0);
    store(0, 0);
    if constexpr (NSET == 1) {
        __syncthreads();
        int cur = 0;
        for (int k0 = 0; k0 < K; k0 += BK) {
            const bool more = k0 + BK < K;
            if (more) load(k0 + BK, 0);
            multiply(cur);
            if (more) store(cur ^ 1, 0);
            __syncthreads();
            cur ^= 1;
        }
    } else {
        // slab s multiplies from LDS buffer s&1; register set s&1 is free then and takes
        // slab s+2, set (s+1)&1 holds slab s+1 (fetched one slab earlier) for the buffer
        // the previous slab used
        if (BK < K) load(BK, 1);
        __syncthreads();
        for (int k0 = 0; k0 < K; k0 += 2 * BK) {
            if (k0 + 2 * BK < K) load(k0 + 2 * BK, 0);
            multiply(0);
            if (k0 + BK < K) store(1, 1);
            __syncthreads();
            if (k0 + BK >= K) break;
            if (k0 + 3 * BK < K) load(k0 + 3 * BK, 1);
            multiply(1);
            if (k0 + 2 * BK < K) store(0, 0);
            __syncthreads();
        }
    }

    // acc[i][j] element r of lane l is C[m0 + wm*BM/2 + 16 i + (l%16), n0 + wn*BN/2 + 16 j + acc_row(l/16, r)]
#pragma unroll
    for (int i = 0; i < FM; ++i) {
        const int gm = m0 + wm * (BM / 2) + i * 16 + lr;
        if (EDGE && gm >= M) continue;
#pragma unroll
        for (int j = 0; j < FN; ++j) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gn = n0 + wn * (BN / 2) + j * 16 + acc_row<T>(lk, r);
                if (!EDGE || gn < N) {
                    T *c = &C[gm + (i64)gn * ldc];
                    T v = alpha * acc[i][j][r];
                    if (beta != T(0)) v += beta * *c;
                    *c = v;
                }
            }
        }
    }
}

template <typename T, int BM, int BN, int BK, bool EDGE, int NSET>
static hipError_t launch_tiles(bool ta, bool tb, int M, int N, int K, T alpha, const T *A, i64 lda, const T *B,
                               i64 ldb, T beta, T *C, i64 ldc, hipStream_t s) {
    const int tm = (M + BM - 1) / BM, tn = (N + BN - 1) / BN;
    dim3 grid(tm * tn), block(256);
    constexpr int V = 16 / (int)sizeof(T);
    const bool aligned = !(reinterpret_cast<uintptr_t>(A) & 15) && !(reinterpret_cast<uintptr_t>(B) & 15) &&
                         lda % V == 0 && ldb % V == 0;
    if constexpr (!EDGE) {
        if (!ta && !tb && aligned) {
            gemm_kernel<T, BM, BN, BK, false, false, false, NSET, true><<<grid, block, 0, s>>>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, tm);
            return hipGetLastError();
        }
    }
    if (!ta && !tb) gemm_kernel<T, BM, BN, BK, false, false, EDGE, NSET><<<grid, block, 0, s>>>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, tm);
    else if (ta && !tb) gemm_kernel<T, BM, BN, BK, true, false, EDGE, NSET><<<grid, block, 0, s>>>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, tm);
    else if (!ta && tb) gemm_kernel<T, BM, BN, BK, false, true, EDGE, NSET><<<grid, block, 0, s>>>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, tm);
    else gemm_kernel<T, BM, BN, BK, true, true, EDGE, NSET><<<grid, block, 0, s>>>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, tm);
    return hipGetLastError();
}

template <typename T, int BM, int BN, int BK, int NSET = 1>
static hipError_t launch_sized(bool ta, bool tb, int M, int N, int K, T alpha, const T *A, i64 lda, const T *B,
                               i64 ldb, T beta, T *C, i64 ldc, hipStream_t s) {
    if (M % BM == 0 && N % BN == 0 && K % BK == 0 && K > 0)
        return launch_tiles<T, BM, BN, BK, false, NSET>(ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s);
    return launch_tiles<T, BM, BN, BK, true, NSET>(ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s);
}

// Float32 GEMM with 16-bit operands: the slabs of op(A) and op(B) are rounded to FP16 or BF16
// (round to nearest even) as they are staged in LDS, and v_mfma_f32_16x16x32_{f16,bf16}
// accumulates in Float32.  LDS holds op(A) as As[m][k] and op(B) as Bs[n][k] (k contiguous, 8
// elements per lane and product); as in gemm_kernel the MFMA's A slot takes the op(B) fragment
// so that a lane's accumulator rows run down a column of C.  128 x 128 tiles, BK = 32, bounds
// checked.  Used for the reduced-precision updates of the mixed-precision triangular solve.
template <typename Tl> struct Mfma16x32;
template <> struct Mfma16x32<_Float16> {
    typedef _Float16 v8 __attribute__((ext_vector_type(8)));
    static __device__ __forceinline__ void run(fx4 &c, v8 a, v8 b) { c = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }
};
template <> struct Mfma16x32<__bf16> {
    typedef __bf16 v8 __attribute__((ext_vector_type(8)));
    static __device__ __forceinline__ void run(fx4 &c, v8 a, v8 b) { c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
};

template <typename Tl, bool TA, bool TB>
__global__ __launch_bounds__(256, 2) void gemm_lowp_kernel(int M, int N, int K, float alpha, const float *__restrict__ A,
                                                          i64 lda, const float *__restrict__ B, i64 ldb, float beta,
                                                          float *__restrict__ C, i64 ldc, int tiles_m) {
    constexpr int BM = 128, BN = 128, BK = 32, LD = BK + 8;
    typedef typename Mfma16x32<Tl>::v8 v8;
    __shared__ __align__(16) Tl As[BM * LD];
    __shared__ __align__(16) Tl Bs[BN * LD];
    int bid = blockIdx.x;
    const int nb = gridDim.x;
    if ((nb & 7) == 0) bid = (bid & 7) * (nb >> 3) + (bid >> 3);
    const int tm = bid % tiles_m, tn = bid / tiles_m;
    const int m0 = tm * BM, n0 = tn * BN;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6, wm = w & 1, wn = w >> 1;
    const int lr = lane & 15, lk = lane >> 4;
    // slab element i of this thread: op(A)(m, k) and op(B)(k, n), 16 of each
    const int am = TA ? t / BK : t % BM, ak = TA ? t % BK : t / BM;
    const int bk = TB ? t / BN : t % BK, bn = TB ? t % BN : t / BK;
    constexpr int AMS = TA ? 256 / BK : 0, AKS = TA ? 0 : 256 / BM;
    constexpr int BKS = TB ? 256 / BN : 0, BNS = TB ? 0 : 256 / BK;
    float ra[16], rb[16];
    auto load = [&](int k0) {
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int gm = m0 + am + i * AMS, gk = k0 + ak + i * AKS;
            const bool ok = gm < M && gk < K;
            const i64 off = TA ? (i64)(ok ? gk : 0) + (i64)(ok ? gm : 0) * lda : (i64)(ok ? gm : 0) + (i64)(ok ? gk : 0) * lda;
            const float x = A[off];
            ra[i] = ok ? x : 0.f;
        }
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int gn = n0 + bn + i * BNS, gk = k0 + bk + i * BKS;
            const bool ok = gn < N && gk < K;
            const i64 off = TB ? (i64)(ok ? gn : 0) + (i64)(ok ? gk : 0) * ldb : (i64)(ok ? gk : 0) + (i64)(ok ? gn : 0) * ldb;
            const float x = B[off];
            rb[i] = ok ? x : 0.f;
        }
    };
    auto store = [&]() {
#pragma unroll
        for (int i = 0; i < 16; ++i) As[(am + i * AMS) * LD + ak + i * AKS] = (Tl)ra[i];
#pragma unroll
        for (int i = 0; i < 16; ++i) Bs[(bn + i * BNS) * LD + bk + i * BKS] = (Tl)rb[i];
    };
    fx4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = acc_zero<float>();
    load(0);
    for (int k0 = 0; k0 < K; k0 += BK) {
        __syncthreads();
        store();
        __syncthreads();
        if (k0 + BK < K) load(k0 + BK);   // the next slab's loads overlap this slab's products
        v8 a[4], b[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = *reinterpret_cast<const v8 *>(&As[(wm * 64 + 16 * i + lr) * LD + 8 * lk]);
#pragma unroll
        for (int j = 0; j < 4; ++j) b[j] = *reinterpret_cast<const v8 *>(&Bs[(wn * 64 + 16 * j + lr) * LD + 8 * lk]);
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) Mfma16x32<Tl>::run(acc[i][j], b[j], a[i]);
    }
    // acc[i][j] element r of lane l: C[m0 + wm 64 + 16 i + l % 16][n0 + wn 64 + 16 j + 4 (l / 16) + r]
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int gm = m0 + wm * 64 + 16 * i + lr;
        if (gm >= M) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gn = n0 + wn * 64 + 16 * j + 4 * lk + r;
                if (gn < N) {
                    float *c = &C[gm + (i64)gn * ldc];
                    float v = alpha * acc[i][j][r];
                    if (beta != 0.f) v += beta * *c;
                    *c = v;
                }
            }
    }
}

// Operand precision of the Float32 GEMMs issued by the calling host thread: 'S' (Float32,
// the default), 'H' (FP16 operands) or 'B' (BF16 operands), set for a scope by GemmLowp.
static thread_local char g_gemm_lowp = 'S';
char gemm_lowp_get() { return g_gemm_lowp; }
void gemm_lowp_set(char p) { g_gemm_lowp = p; }

template <typename Tl>
static hipError_t launch_lowp(bool ta, bool tb, int M, int N, int K, float alpha, const float *A, i64 lda,
                              const float *B, i64 ldb, float beta, float *C, i64 ldc, hipStream_t s) {
    const int tm = (M + 127) / 128, tn = (N + 127) / 128;
    dim3 grid(tm * tn), block(256);
    if (!ta && !tb) gemm_lowp_kernel<Tl, false, false><<<grid, block, 0, s>>>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, tm);
    else if (ta && !tb) gemm_lowp_kernel<Tl, true, false><<<grid, block, 0, s>>>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, tm);
    else if (!ta && tb) gemm_lowp_kernel<Tl, false, true><<<grid, block, 0, s>>>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, tm);
    else gemm_lowp_kernel<Tl, true, true><<<grid, block, 0, s>>>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, tm);
    return hipGetLastError();
}

template <typename T>
hipError_t gemm(char ta, char tb, i64 M, i64 N, i64 K, T alpha, const T *A, i64 lda, const T *B, i64 ldb, T beta,
                T *C, i64 ldc, hipStream_t s) {
    if (M <= 0 || N <= 0) return hipSuccess;
    const bool TA = is_trans(ta), TB = is_trans(tb);
    if (K <= 0) {
        if (beta == T(1)) return hipSuccess;
        // C = beta * C through a K=0 product: the kernel reads nothing from A/B.
    }
    const i64 big_tiles = ((M + 127) / 128) * ((N + 127) / 128);
    // "gemm" counts every launch; "gemm128" the 128 x 128-tile launches alone (one kernel)
    ProfScope prof(PK_GEMM, s);
    if constexpr (std::is_same<T, float>::value) {
        if (g_gemm_lowp == 'H' && K > 0)
            return launch_lowp<_Float16>(TA, TB, (int)M, (int)N, (int)K, alpha, A, lda, B, ldb, beta, C, ldc, s);
        if (g_gemm_lowp == 'B' && K > 0)
            return launch_lowp<__bf16>(TA, TB, (int)M, (int)N, (int)K, alpha, A, lda, B, ldb, beta, C, ldc, s);
    }
    constexpr int BK = 16;
    if (big_tiles >= 512) {
        ProfScope prof128(PK_GEMM128, s);
        return launch_sized<T, 128, 128, BK>(TA, TB, (int)M, (int)N, (int)K, alpha, A, lda, B, ldb, beta, C, ldc, s);
    }
    return launch_sized<T, 64, 64, BK>(TA, TB, (int)M, (int)N, (int)K, alpha, A, lda, B, ldb, beta, C, ldc, s);
}

template hipError_t gemm<double>(char, char, i64, i64, i64, double, const double *, i64, const double *, i64, double,
                                 double *, i64, hipStream_t);
template hipError_t gemm<float>(char, char, i64, i64, i64, float, const float *, i64, const float *, i64, float,
                                float *, i64, hipStream_t);

template <typename T>
static int gemm_acc_entry(i64 N, i64 R, i64 M, T alpha, const T *X, i64 ldx, const T *Y, i64 ldy, T *C, i64 ldc,
                          void *stream) {
    if (N < 0) return arg_error(1, "illegal value of N: %lld", (long long)N);
    if (R < 0) return arg_error(2, "illegal value of R: %lld", (long long)R);
    if (M < 0) return arg_error(3, "illegal value of M: %lld", (long long)M);
    if (N == 0 || M == 0 || R == 0) return 0;
    if (ldx < N) return arg_error(6, "ldx must be >= N");
    if (ldy < R) return arg_error(8, "ldy must be >= R");
    if (ldc < N) return arg_error(10, "ldc must be >= N");
    if (N > INT32_MAX || M > INT32_MAX || R > INT32_MAX) return arg_error(1, "dimension exceeds 2^31-1");
    NLA_HIP(gemm<T>('N', 'N', N, M, R, alpha, X, ldx, Y, ldy, T(1), C, ldc, as_stream(stream)));
    return 0;
}

}  // namespace nla

NLA_API int nla_dgemm_acc(int64_t N, int64_t R, int64_t M, double alpha, const double *X, int64_t ldx,
                          const double *Y, int64_t ldy, double *C, int64_t ldc, void *stream) {
    return nla::gemm_acc_entry<double>(N, R, M, alpha, X, ldx, Y, ldy, C, ldc, stream);
}
NLA_API int nla_sgemm_acc(int64_t N, int64_t R, int64_t M, float alpha, const float *X, int64_t ldx,
                          const float *Y, int64_t ldy, float *C, int64_t ldc, void *stream) {
    return nla::gemm_acc_entry<float>(N, R, M, alpha, X, ldx, Y, ldy, C, ldc, stream);
}
