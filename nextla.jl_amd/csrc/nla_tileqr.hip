// Tile-QR driver: flat-tree PLASMA-order composition of geqrt!/unmqr!/tsqrt!/tsmqr!
// (no reference function: the reference leaves this loop to its caller).
//   for k: geqrt(A_kk); unmqr('L','C', A_kk, A_k,j>k); for i>k: tsqrt(A_kk, A_ik);
//          tsmqr('L','C', A_k,j>k, A_i,j>k, V=A_ik, T_ik)
// Each unmqr/tsmqr step is one strip launch over the whole trailing tile row.
#include "nla_common.hpp"
#include "nla_gemm.hpp"
#include "nla_qr.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace nla {

template <typename T>
int tile_geqrf_dag(i64 m, i64 n, i64 nb, i64 ib, T *A, i64 lda, T *Tm, i64 ldt, hipStream_t s, bool *handled);
int tile_geqrf_dag_mixed(i64 m, i64 n, i64 nb, i64 ib, double *A, i64 lda, double *Tm, i64 ldt, char prec, hipStream_t s,
                         bool *handled);
template <typename T>
int tile_apply_qt_dag(i64 m, i64 n, i64 nb, i64 ib, const T *A, i64 lda, const T *Tm, i64 ldt, T *C, i64 ldc,
                      i64 ncols, hipStream_t s, bool *handled, bool qt);
template <typename T>
int rectrxm_async(char side, char uplo, char trans, T alpha, char func, i64 n, i64 m, const T *A, i64 lda, T *B,
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
        if (handled) return r;
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
static int tile_qt_vec(i64 m, i64 n, i64 nb, i64 ib, const T *A, i64 lda, const T *Tm, i64 ldt, T *x, hipStream_t s,
                       bool *handled);

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
    if (nrhs == 1 && !notrans) {  // one column: the single-vector wavefront
        bool handled = false;
        const int r = tile_qt_vec<T>(m, n, nb, ib, A, lda, Tm, ldt, C, s, &handled);
        if (handled) return r;
    }
    {
        // Q^T C / Q C by the dataflow kernel's walkers alone (C padded to a multiple of 64 columns)
        constexpr i64 TWC = 64;
        const i64 ncols = (nrhs + TWC - 1) / TWC * TWC;
        bool handled = false;
        if (ncols == nrhs) {
            const int r = tile_apply_qt_dag<T>(m, n, nb, ib, A, lda, Tm, ldt, C, ldc, ncols, s, &handled, !notrans);
            if (handled) return r;
        } else {
            int st = 0;
            T *P = (T *)scratch(SLOT_RHS, sizeof(T) * (size_t)m * ncols, s, &st);
            if (!P) return st;
            NLA_HIP(hipMemsetAsync(P, 0, sizeof(T) * (size_t)m * ncols, s));
            NLA_HIP(hipMemcpy2DAsync(P, sizeof(T) * m, C, sizeof(T) * ldc, sizeof(T) * m, nrhs, hipMemcpyDeviceToDevice, s));
            const int r = tile_apply_qt_dag<T>(m, n, nb, ib, A, lda, Tm, ldt, P, m, ncols, s, &handled, !notrans);
            if (handled) {
                if (r) return r;
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

// Float64 tile QR with Float32 ('S'), FP16 ('H'), BF16 ('B') or split-BF16 ('X') trailing
// updates, or plain ('D').
static int tile_geqrf_mixed_entry(i64 m, i64 n, i64 nb, i64 ib, double *A, i64 lda, double *Tm, i64 ldt, char prec,
                                  void *stream) {
    prec = up(prec);
    if (prec != 'D' && prec != 'S' && prec != 'H' && prec != 'B' && prec != 'X')
        return arg_error(9, "prec must be 'D', 'S', 'H', 'B' or 'X'");
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
    const int r = tile_geqrf_dag_mixed(m, n, nb, ib, A, lda, Tm, ldt, prec, as_stream(stream), &handled);
    if (!handled) return arg_error(3, "mixed precision needs nb in {64, 128}, ib = 32 and 16-byte aligned columns");
    return r;
}

// ---------------------------------------------------------------------------------
// Triangular solve with one right-hand side: R x = v (trans false) or R^T x = v (trans
// true), R the n x n upper triangle of a column-major matrix (n a multiple of 64), x
// overwriting v.  One persistent launch: workgroups claim 64-row blocks from a counter in
// dependency order (bottom up for R, top down for R^T), accumulate the products with the
// finished blocks as each block's flag appears, solve their diagonal block from LDS and
// publish it.  A block waits only on blocks claimed before it, so the launch cannot
// deadlock; waits are bounded (2 s, then an error flag).  R is read once; the recursive
// triangular solve would instead run log2(n/64) levels of one-column GEMM launches.
constexpr int TRSV_BS = 64;

template <typename T> __device__ __forceinline__ T ld_sc1(const T *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T> __device__ __forceinline__ void st_sc1(T *p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------------
// Q^T x for a single vector x (m entries) with the factors of tile_geqrf, in the flat-tree order
// of the factorization (src/unmqr.jl, src/tsmqr.jl with trans 'C', one column).  One workgroup
// per step k, claimed in order from a counter: unmqr('L','C') of x_k with the blocks of
// geqrt(A_kk), then for every tile row i > k the TS apply of the blocks of tile (i, k) to
// (x_k, x_i), each row published when stored (prog[k] = rows of x final for step k + 1; step k
// waits on prog[k - 1] only, so the lowest unfinished step always runs).  A row is a few
// microseconds of latency, where the dataflow kernel's walker spends a 64-column strip's
// products on it; the wavefront over the kt + mt steps and rows is correspondingly shorter.
//
// Four waves apply the blocks; their only global accesses are the loads of the reflectors and
// T (final before the launch), issued three blocks ahead into registers and staged into the
// second of two LDS buffers while the current block is applied.  A fifth wave does all the
// traffic with the other steps: it waits on the previous step's progress, loads the rows of x
// into LDS (the next row's during the current one, into the second of two row buffers), stores
// each finished row and publishes it.  A wave's memory counter drains in issue order, so this
// keeps the progress waits and the publications from waiting on the block loads in flight, and
// the block loads from being drained by a wait on a row.  Bounded waits as in the dataflow
// kernel.
template <typename T, int NB, int IBV>
__global__ __launch_bounds__(320) void qt_vec_kernel(int mt, int kt, const T *__restrict__ A, i64 lda,
                                                     const T *__restrict__ Tm, i64 ldt, T *x, int *counter,
                                                     int *prog, int *err) {
    constexpr int NBLK = NB / IBV, G = 256 / IBV, RPG = NB / G, XPL = NB / 64, TPR = 256 / NB;
    constexpr int NV = NB * IBV / 256, NT = IBV * IBV / 256;
    static_assert(NB % IBV == 0 && NBLK % 2 == 0 && (G == 8 || G == 16) && (IBV * IBV) % 256 == 0 &&
                      NB % G == 0 && (NB * IBV) % 256 == 0 && NB % 64 == 0,
                  "block shape");
    // LDS row strides: the G lanes of a reflector read G consecutive entries and the next
    // reflector's lanes start G entries further along the 64 banks, so a half wave's reads
    // are free of bank conflicts
    constexpr int BW = 64 * 4 / (int)sizeof(T);  // entries per pass over the banks
    constexpr int VLD = NB + ((G - NB) % BW + BW) % BW, TLD = IBV + ((G - IBV) % BW + BW) % BW;
    __shared__ T xk[NB], xi[2][NB];
    __shared__ T Vs[2][IBV][VLD];
    __shared__ T Ts[2][IBV][TLD];  // Ts[.][j][q] = T(q, j)
    __shared__ T w0[IBV], wv[IBV];
    __shared__ int sh_k;
    const int t = threadIdx.x, jr = t / G, g = t % G, lane = t & 63;
    const bool ctl = t >= 256;  // the fifth wave
    T pv[2][NV], pt[2][NT];     // register buffers: block j in buffer j % 2
    using R0 = std::integral_constant<int, 0>;
    using R1 = std::integral_constant<int, 1>;
    // sum over the G lanes of a reflector (quad swaps, half-row mirror, row rotation by 8)
    auto group_sum = [](T v) {
        v += dpp_mov<DPP_QUAD_X1>(v);
        v += dpp_mov<DPP_QUAD_X2>(v);
        v += dpp_mov<0x141>(v);
        if constexpr (G == 16) v += dpp_mov<DPP_ROW_ROR8>(v);
        return v;
    };
    // item s of a step: s < NBLK the geqrt block s of A_kk, else block (s - NBLK) % NBLK of tile
    // row k + 1 + (s - NBLK) / NBLK; tile row i > k works in row buffer (i - k - 1) % 2
    auto fetch = [&](auto R, int k, int s) {
        constexpr int r = decltype(R)::value;
        const bool ge = s < NBLK;
        const int i = ge ? k : k + 1 + (s - NBLK) / NBLK, b = ge ? s : (s - NBLK) % NBLK;
        const T *Vg = A + (i64)i * NB + ((i64)k * NB + b * IBV) * lda;
        const T *Tg = Tm + (i64)i * IBV + ((i64)k * NB + b * IBV) * ldt;
#pragma unroll
        for (int e = 0; e < NV; ++e) {
            const int idx = t + 256 * e, pr = idx % NB, jj = idx / NB;
            pv[r][e] = Vg[pr + (i64)jj * lda];
        }
#pragma unroll
        for (int e = 0; e < NT; ++e) {
            const int idx = t + 256 * e, q = idx % IBV, jj = idx / IBV;
            pt[r][e] = Tg[q + (i64)jj * ldt];
        }
    };
    // the loaded values are used only here, so the loads stay in flight until the block is
    // staged; a geqrt block gets its unit diagonal and zeros above it here
    auto stage = [&](auto R, int s) {
        constexpr int r = decltype(R)::value;
        const bool ge = s < NBLK;
#pragma unroll
        for (int e = 0; e < NV; ++e) {
            const int idx = t + 256 * e, pr = idx % NB, jj = idx / NB, ja = s * IBV + jj;
            T v = pv[r][e];
            if (ge) v = pr < ja ? T(0) : (pr == ja ? T(1) : v);
            Vs[r][jj][pr] = v;
        }
#pragma unroll
        for (int e = 0; e < NT; ++e) {
            const int idx = t + 256 * e;
            Ts[r][idx / IBV][idx % IBV] = idx % IBV <= idx / IBV ? pt[r][e] : T(0);
        }
    };
    // fifth wave: spin until *p >= v (first: a value of *p read earlier).  On a timeout, or
    // once another workgroup has raised the error flag, it raises the flag and returns at once;
    // the launch then runs to its end without further waits (a result the host discards).  No
    // early exit: the block loop keeps a single shape, which keeps the compiler's waits for the
    // block registers exact.
    auto ctl_wait = [&](const int *p, int v, int first) {
        if (first >= v || ld_sc1(p) >= v) return;
        const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
        while (ld_sc1(p) < v) {
            if (ld_sc1(err) || __builtin_amdgcn_s_memrealtime() - t0 > 200000000ull) {
                if (lane == 0) st_sc1(err, 1);
                return;
            }
            __builtin_amdgcn_s_sleep(2);
        }
    };
    // fifth wave: store tile row i from row buffer c, then publish it
    auto ctl_store = [&](int k, int i, int c) {
#pragma unroll
        for (int e = 0; e < XPL; ++e) st_sc1(&x[(i64)i * NB + lane + 64 * e], xi[c][lane + 64 * e]);
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        if (lane == 0) st_sc1(&prog[k], i + 1);
    };
    // The fifth wave runs its own loop: its loads and stores then never share a code path with
    // the block loads, so the compiler's waits for a block's registers count only block loads.
    // Both loops pass the same barriers: three per block, three per step.
    if (ctl) {
        T xn[XPL];  // the next tile row's x
        int pp = 0;
        for (;;) {
            __syncthreads();
            __syncthreads();
            const int k = sh_k;
            if (k >= kt) break;
            if (k > 0) ctl_wait(&prog[k - 1], k + 1, 0);
#pragma unroll
            for (int e = 0; e < XPL; ++e) xk[lane + 64 * e] = ld_sc1(&x[(i64)k * NB + lane + 64 * e]);
            __syncthreads();
            const int nitems = NBLK * (mt - k);
            for (int s = 0; s < nitems; ++s) {
                const bool ge = s < NBLK;
                const int b = ge ? s : (s - NBLK) % NBLK, i = ge ? k : k + 1 + (s - NBLK) / NBLK;
                const bool more = i + 1 < mt;
                if (!ge && b == 0 && i - 1 > k) ctl_store(k, i - 1, ((i - k - 1) & 1) ^ 1);
                if (b == 0 && more && k > 0) pp = ld_sc1(&prog[k - 1]);
                if (b == NBLK / 2 && more) {  // row i + 1 once the previous step has released it
                    if (k > 0) ctl_wait(&prog[k - 1], i + 2, pp);
#pragma unroll
                    for (int e = 0; e < XPL; ++e) xn[e] = ld_sc1(&x[(i64)(i + 1) * NB + lane + 64 * e]);
                }
                __syncthreads();
                __syncthreads();
                if (b == NBLK - 1 && more) {
#pragma unroll
                    for (int e = 0; e < XPL; ++e) xi[(i - k) & 1][lane + 64 * e] = xn[e];
                }
                __syncthreads();
            }
            if (mt - 1 > k) ctl_store(k, mt - 1, (mt - k) & 1);
#pragma unroll
            for (int e = 0; e < XPL; ++e) st_sc1(&x[(i64)k * NB + lane + 64 * e], xk[lane + 64 * e]);
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        }
        return;
    }
    int k = 0, nitems = 0;
    // H^T of block s (staged in Vs[s % 2]) on S: W0 = V^T S (+ the A1 rows), W = T^T W0,
    // A1 -= W, S -= V W; block s + 1 is staged and block s + 3 fetched in between
    auto item = [&](auto R, int s) {
        constexpr int r = decltype(R)::value, o = 1 - r;
        const bool ge = s < NBLK;
        const int b = ge ? s : (s - NBLK) % NBLK, i = ge ? k : k + 1 + (s - NBLK) / NBLK;
        T *S = ge ? xk : xi[(i - k - 1) & 1], *A1 = ge ? nullptr : xk + b * IBV;
        T acc = 0;
#pragma unroll
        for (int q = 0; q < RPG; ++q) acc += Vs[r][jr][q * G + g] * S[q * G + g];
        acc = group_sum(acc);
        if (g == 0) w0[jr] = A1 ? acc + A1[jr] : acc;
        __syncthreads();
        // W[jr] = sum_{q <= jr} T(q, jr) w0[q]: the G threads of reflector jr take every G-th q
        T w = 0;
#pragma unroll
        for (int q = g; q < IBV; q += G)
            if (q <= jr) w += Ts[r][jr][q] * w0[q];
        w = group_sum(w);
        if (g == 0) {
            wv[jr] = w;
            if (A1) A1[jr] -= w;
        }
        // unconditional (the last items re-load the step's last block), so the number of loads
        // in flight behind each block is the same on every path and the waits for a block's
        // registers stop at that block
        stage(std::integral_constant<int, o>(), s + 1 < nitems ? s + 1 : nitems - 1);
        fetch(std::integral_constant<int, o>(), k, s + 3 < nitems ? s + 3 : nitems - 1);
        __syncthreads();
        {  // S -= V W: the TPR lanes of a row take every TPR-th reflector, summed by quad swaps
            const int row = t / TPR, h = t % TPR;
            T v = 0;
#pragma unroll 8
            for (int jj = h; jj < IBV; jj += TPR) v += Vs[r][jj][row] * wv[jj];
            v += dpp_mov<DPP_QUAD_X1>(v);
            if constexpr (TPR == 4) v += dpp_mov<DPP_QUAD_X2>(v);
            if (h == 0) S[row] -= v;
        }
        __syncthreads();
    };
    for (;;) {
        __syncthreads();
        if (t == 0) sh_k = atomicAdd(counter, 1);
        __syncthreads();
        k = sh_k;
        if (k >= kt) break;
        nitems = NBLK * (mt - k);
        fetch(R0(), k, 0);
        fetch(R1(), k, 1);
        stage(R0(), 0);
        fetch(R0(), k, nitems > 2 ? 2 : nitems - 1);
        __syncthreads();
        for (int s = 0; s < nitems; s += 2) {
            item(R0(), s);
            item(R1(), s + 1);
        }
    }
}

__global__ void qt_vec_clear_kernel(int *p, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

// x := Q^T x (x of m entries, device) with the tile_geqrf factors; *handled = false for the
// shapes it does not take (nb 64 / 128 with ib 32 or 16).
template <typename T>
static int tile_qt_vec(i64 m, i64 n, i64 nb, i64 ib, const T *A, i64 lda, const T *Tm, i64 ldt, T *x, hipStream_t s,
                       bool *handled) {
    *handled = (nb == 128 || nb == 64) && (ib == 32 || ib == 16) && m % nb == 0 && n % nb == 0 && m >= n;
    if (!*handled) return 0;
    const int mt = (int)(m / nb), kt = (int)(n / nb);
    int st = 0;
    int *mem = (int *)scratch(SLOT_SCHED, sizeof(int) * (size_t)(kt + 2), s, &st);
    if (!mem) return st;
    int *counter = mem, *err = mem + 1, *prog = mem + 2;
    qt_vec_clear_kernel<<<(kt + 2 + 255) / 256, 256, 0, s>>>(mem, kt + 2);
    NLA_LAUNCH_CHECK();
    int dev = 0, ncu = 0;
    NLA_HIP(hipGetDevice(&dev));
    NLA_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const int grid = kt < ncu ? kt : ncu;
    if (nb == 128 && ib == 32) qt_vec_kernel<T, 128, 32><<<grid, 320, 0, s>>>(mt, kt, A, lda, Tm, ldt, x, counter, prog, err);
    else if (nb == 128) qt_vec_kernel<T, 128, 16><<<grid, 320, 0, s>>>(mt, kt, A, lda, Tm, ldt, x, counter, prog, err);
    else if (ib == 32) qt_vec_kernel<T, 64, 32><<<grid, 320, 0, s>>>(mt, kt, A, lda, Tm, ldt, x, counter, prog, err);
    else qt_vec_kernel<T, 64, 16><<<grid, 320, 0, s>>>(mt, kt, A, lda, Tm, ldt, x, counter, prog, err);
    NLA_LAUNCH_CHECK();
    int herr = 0;
    NLA_HIP(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, s));
    NLA_HIP(hipStreamSynchronize(s));
    if (herr) return wait_timeout_error("tile Q^T x");
    return 0;
}


// Inverses of the n/64 diagonal 64 x 64 blocks of the upper triangle R (one workgroup per
// block, thread j forms column j by back substitution): Dinv[b][i + 64 j] = inv(R_bb)[i][j].
template <typename T> __global__ __launch_bounds__(64) void diag_block_inv_kernel(const T *R, i64 lda, T *Dinv) {
    constexpr int BS = TRSV_BS;
    __shared__ T U[BS][BS + 1];
    __shared__ T Y[BS][BS + 1];  // Y[i][j] = inv[i][j]
    const int b = blockIdx.x, j = threadIdx.x;
    const i64 r0 = (i64)b * BS;
    for (int i = 0; i < BS; ++i) U[i][j] = R[r0 + i + (r0 + j) * lda];  // row i, column j (coalesced in i? no: lanes = j)
    __syncthreads();
    for (int i = BS - 1; i > j; --i) Y[i][j] = 0;
    Y[j][j] = T(1) / U[j][j];
    for (int i = j - 1; i >= 0; --i) {
        T acc = 0;
        for (int k = i + 1; k <= j; ++k) acc += U[i][k] * Y[k][j];
        Y[i][j] = -acc / U[i][i];
    }
    for (int i = 0; i < BS; ++i) Dinv[(i64)b * BS * BS + i + (i64)j * BS] = Y[i][j];
}

// The solved values are published through a second vector y that starts out filled with a NaN
// of a fixed payload (arithmetic on numbers yields the default quiet NaN, never this one): a
// block's values are ready when none of its words reads as the mark.  A 64-bit (32-bit) word
// is stored and loaded whole, so no flag, no store-completion wait before it and no second
// round trip after it are needed between the workgroup that solves a block and the ones that
// use it.
template <typename T> struct TrsvMark;
template <> struct TrsvMark<double> {
    using U = unsigned long long;
    static constexpr U bits = 0x7ff4a5a5a5a5a5a5ull;
};
template <> struct TrsvMark<float> {
    using U = unsigned;
    static constexpr U bits = 0x7fa5a5a5u;
};
template <typename T> __device__ __forceinline__ bool trsv_unset(T v) {
    return __builtin_bit_cast(typename TrsvMark<T>::U, v) == TrsvMark<T>::bits;
}

// Blocks are claimed in dependency order (each depends on every block claimed before it); a
// workgroup accumulates the products of its block row with the solved blocks as they appear,
// every thread polling the words of y it uses (bounded: a timeout or another workgroup's
// error raises the error flag and the launch runs to its end), then solves its diagonal block
// with the precomputed inverse and writes the block to y (for the later blocks) and to x.
// x_b is read only by block b, so the result overwrites it in place.
template <typename T, bool TRANS>
__global__ __launch_bounds__(256) void trsv_upper_kernel(int n, const T *R, i64 lda, const T *Dinv, T *x, T *y,
                                                          int *counter, int *err) {
    constexpr int BS = TRSV_BS;
    __shared__ T red[4][BS + 1];
    __shared__ T ys[BS];
    __shared__ T tr[TRANS ? 4 * 16 * 64 : 1];  // TRANS: the lanes' partial sums, transposed
    __shared__ int sh_b;
    const int nblk = n / BS, t = threadIdx.x, lane = t & 63, g = t >> 6;
    auto solved = [&](const T *p) -> T {
        T val = ld_sc1(p);
        if (!trsv_unset(val)) return val;
        const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
        while (trsv_unset(val = ld_sc1(p))) {
            if (ld_sc1(err) || __builtin_amdgcn_s_memrealtime() - t0 > 200000000ull) {
                st_sc1(err, 1);
                return T(0);
            }
            __builtin_amdgcn_s_sleep(1);
        }
        return val;
    };
    for (;;) {
        __syncthreads();
        if (t == 0) sh_b = atomicAdd(counter, 1);
        __syncthreads();
        const int claim = sh_b;
        if (claim >= nblk) break;
        const int b = TRANS ? claim : nblk - 1 - claim;
        const i64 r0 = (i64)b * BS;
        T acc[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) acc[q] = 0;
        // TPS tiles per step, all their loads in flight together (past the end a tile is the
        // last one again, with weight 0); !TRANS: lane l < 16 of a wave loads the wave's x value
        // l and the products take it by readlane.  Two tiles: 0.62 / 0.80 ms (plain /
        // transposed) at n = 16384 against 0.78 / 0.92 with one; four were slower again.
        constexpr int TPS = 2;
        auto bcast = [](T v, int l) {
            if constexpr (sizeof(T) == 8) {
                const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
                const unsigned lo = __builtin_amdgcn_readlane((unsigned)u, l), hi = __builtin_amdgcn_readlane((unsigned)(u >> 32), l);
                return __builtin_bit_cast(T, ((unsigned long long)hi << 32) | lo);
            } else {
                return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(unsigned, v), l));
            }
        };
        const int xo = TRANS ? lane : g * 16 + (lane & 15);
        for (int d = 0; d < claim; d += TPS) {
            i64 c0[TPS];
            T rv[TPS][16], xv[TPS];
#pragma unroll
            for (int u = 0; u < TPS; ++u) {
                const int du = d + u < claim ? d + u : claim - 1;
                c0[u] = (i64)(TRANS ? du : nblk - 1 - du) * BS;
                const T *Rc = TRANS ? R + c0[u] + lane + (r0 + g * 16) * lda : R + r0 + lane + (c0[u] + g * 16) * lda;
#pragma unroll
                for (int q = 0; q < 16; ++q) rv[u][q] = Rc[q * lda];
            }
            bool unset = false;
#pragma unroll
            for (int u = 0; u < TPS; ++u) {
                xv[u] = ld_sc1(&y[c0[u] + xo]);
                unset = unset || trsv_unset(xv[u]);
            }
            if (unset) {
#pragma unroll
                for (int u = 0; u < TPS; ++u) xv[u] = solved(&y[c0[u] + xo]);
            }
#pragma unroll
            for (int u = 1; u < TPS; ++u)
                if (d + u >= claim) xv[u] = T(0);
#pragma unroll
            for (int u = 0; u < TPS; ++u) {
                if (!TRANS) {
#pragma unroll
                    for (int q = 0; q < 16; ++q) acc[0] += rv[u][q] * bcast(xv[u], q);
                } else {
#pragma unroll
                    for (int q = 0; q < 16; ++q) acc[q] += rv[u][q] * xv[u];
                }
            }
        }
        // the diagonal block's inverse (independent of the other blocks)
        const T *D = Dinv + (i64)b * BS * BS;
        T dv[16];
#pragma unroll
        for (int q = 0; q < 16; ++q)  // !TRANS: D[lane][g*16 + q];  TRANS: D[g*16 + q][lane] read as (D^T)[lane][..]
            dv[q] = TRANS ? D[(g * 16 + q) + (i64)lane * BS] : D[lane + (i64)(g * 16 + q) * BS];
        if (!TRANS) {
            red[g][lane] = acc[0];
        } else {
            // output g*16 + q is the sum over the wave's 64 lanes of acc[q]: transpose through
            // LDS, then thread t sums 16 lanes of output t/4 and the quad of threads adds up
#pragma unroll
            for (int q = 0; q < 16; ++q) tr[(g * 16 + q) * 64 + lane] = acc[q];
        }
        __syncthreads();
        if (TRANS) {
            const int i = t >> 2, sq = t & 3;
            const T *src = &tr[i * 64 + sq * 16];
            T p0 = 0, p1 = 0, p2 = 0, p3 = 0;
#pragma unroll
            for (int u = 0; u < 16; u += 4) {
                p0 += src[u];
                p1 += src[u + 1];
                p2 += src[u + 2];
                p3 += src[u + 3];
            }
            T p = (p0 + p1) + (p2 + p3);
            p += __shfl_xor(p, 1);
            p += __shfl_xor(p, 2);
            if (sq == 0) red[0][i] = p;
            __syncthreads();
        }
        if (g == 0) {
            // y = v_b - sum of the solved blocks' products
            const T sum = TRANS ? red[0][lane] : (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
            ys[lane] = x[r0 + lane] - sum;
        }
        __syncthreads();
        // x_b = inv(R_bb) y  (or inv(R_bb)^T y): row lane, columns g*16 .. g*16+15, summed over g
        T p = 0;
#pragma unroll
        for (int q = 0; q < 16; ++q) p += dv[q] * ys[g * 16 + q];
        red[g][lane] = p;
        __syncthreads();
        if (g == 0) {
            const T xb = (red[0][lane] + red[1][lane]) + (red[2][lane] + red[3][lane]);
            st_sc1(&y[r0 + lane], xb);
            x[r0 + lane] = xb;
        }
    }
}
// y := the mark (n words), the claim counter := 0
template <typename T> __global__ void trsv_reset_kernel(T *y, int n, int *counter) {
    const T mark = __builtin_bit_cast(T, TrsvMark<T>::bits);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) y[i] = mark;
    if (blockIdx.x == 0 && threadIdx.x == 0) *counter = 0;
}
// x := R^-1 x or R^-T x (R upper n x n at R, lda; Dinv from diag_block_inv_kernel); y: n words
// of device scratch; sync: the claim counter and the error flag (device ints)
template <typename T>
static hipError_t trsv_upper(bool trans, int n, const T *R, i64 lda, const T *Dinv, T *x, T *y, int *sync,
                             hipStream_t s) {
    const int nblk = n / TRSV_BS;
    trsv_reset_kernel<T><<<(n + 255) / 256 < 256 ? (n + 255) / 256 : 256, 256, 0, s>>>(y, n, sync);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    int dev = 0, ncu = 0;
    if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
    if ((e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
    const int grid = std::min(nblk, ncu);
    if (trans) trsv_upper_kernel<T, true><<<grid, 256, 0, s>>>(n, R, lda, Dinv, x, y, sync, sync + 1);
    else trsv_upper_kernel<T, false><<<grid, 256, 0, s>>>(n, R, lda, Dinv, x, y, sync, sync + 1);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// Least squares.  Vector kernels of the refinement (Float64 work for Float64 data).

// part[s][i] = sum over column slice s of A[i, j] p[j]  (two rows per thread, coalesced
// 16-byte column reads; the slices make the grid large enough to stream A at HBM rate)
template <typename T>
__global__ void gemv_n_part_kernel(i64 m, i64 n, const T *A, i64 lda, const T *p, T *part, int nsplit) {
    const i64 i = 2 * ((i64)blockIdx.x * blockDim.x + threadIdx.x);
    const i64 js = (n * blockIdx.y) / nsplit, je = (n * (blockIdx.y + 1)) / nsplit;
    T a0 = 0, a1 = 0;
    if (i + 1 < m) {
        for (i64 j = js; j < je; ++j) {
            const T pj = p[j];
            a0 += A[i + j * lda] * pj;
            a1 += A[i + 1 + j * lda] * pj;
        }
    } else if (i < m) {
        for (i64 j = js; j < je; ++j) a0 += A[i + j * lda] * p[j];
    }
    if (i < m) part[(i64)blockIdx.y * m + i] = a0;
    if (i + 1 < m) part[(i64)blockIdx.y * m + i + 1] = a1;
}
// y[i] = sum_s part[s][i] - beta * y[i];  bsum[block] = sum of y[i]^2 over the block
template <typename T>
__global__ void gemv_n_finish_kernel(i64 m, int nsplit, const T *part, T beta, T *y, double *bsum) {
    __shared__ double red[16];
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    double q = 0;
    if (i < m) {
        T acc = 0;
        for (int k = 0; k < nsplit; ++k) acc += part[(i64)k * m + i];
        const T v = acc - beta * y[i];
        y[i] = v;
        q = (double)v * v;
    }
    q = block_sum(q, red);
    if (threadIdx.x == 0) bsum[blockIdx.x] = q;
}
// z[j] = sum_i A[i, j] u[i]  (one wave per column)
template <typename T> __global__ void gemv_t_kernel(i64 m, i64 n, const T *A, i64 lda, const T *u, T *z) {
    const i64 j = (i64)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (j >= n) return;
    const T *a = A + j * lda;
    T acc = 0;
    for (i64 i = lane; i < m; i += 64) acc += a[i] * u[i];
    acc = wave_sum(acc);
    if (lane == 0) z[j] = acc;
}
// v[i] = z[i] - beta v[i];  bsum[block] = sum of v[i]^2 over the block
template <typename T> __global__ void lincomb_kernel(i64 n, const T *z, T beta, T *v, double *bsum) {
    __shared__ double red[16];
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    double q = 0;
    if (i < n) {
        const T t = z[i] - beta * v[i];
        v[i] = t;
        q = (double)t * t;
    }
    q = block_sum(q, red);
    if (threadIdx.x == 0) bsum[blockIdx.x] = q;
}
// out = sqrt(sum of bsum[0..nb))  (one block; fixed order, so the result is reproducible)
__global__ void norm_finish_kernel(int nb, const double *bsum, double *out) {
    __shared__ double red[16];
    double q = 0;
    for (int k = threadIdx.x; k < nb; k += blockDim.x) q += bsum[k];
    q = block_sum(q, red);
    if (threadIdx.x == 0) *out = sqrt(q);
}
template <typename T> __global__ void scale_kernel(i64 n, T s, T *v) {
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] *= s;
}
// y += a w;  w = v - b w
template <typename T> __global__ void lsqr_update_kernel(i64 n, T a, T b, const T *v, T *w, T *y) {
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const T wi = w[i];
        y[i] += a * wi;
        w[i] = v[i] - b * wi;
    }
}
// y := y + x (n)
template <typename T> __global__ void axpy1_kernel(i64 n, const T *x, T *y) {
    const i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] += x[i];
}

// Least squares min ||A x - b|| (A m x n, m >= n, full rank).
// Tile QR A = Q R (prec 'D': Float64; 'S', 'H', 'B', 'X': Float32, FP16, BF16 or split-BF16
// trailing updates, Float64 data only).  x0 = R^-1 (Q^T b)[0:n], the direct solution (returned for maxit = 0);
// with the Float64 factor and maxit > 0, x0 = R^-1 R^-T (A^T b) instead (the seminormal
// equations: Q^T b = R^-T A^T b for A = QR; one product by A^T and two triangular solves,
// ~1.2 ms at 32768 x 8192, where Q^T b takes 7.4 ms because its walker launch has a single
// strip of work for one right-hand side).  Then LSQR (Paige & Saunders, ACM TOMS 8 (1982),
// algorithm LSQR) on the correction, right-preconditioned by R: min ||(A R^-1) y - (b - A x0)||,
// x = x0 + R^-1 y, every product with A in the working precision on the unmodified A (with the
// seminormal x0: Bjorck's corrected seminormal equations).  With a reduced-precision factor the
// seminormal start (or a zero start) left LSQR short of the direct start's accuracy at kappa
// 1e6 (relative error 1.4e-8 / 1.2e-8 against 1.1e-8 allowed), so those keep Q^T b.  A R^-1 has condition
// number about 1 + kappa(A) times the factorization's backward error, so the iteration
// converges quickly whenever R is a fair preconditioner, also where plain refinement
// (x += R^-1 (Q^T r)[0:n]) diverges: kappa(A) times the factorization error near or above one.
// Stops after maxit iterations or when LSQR's estimate of ||M^T r|| / (||M|| ||r||)
// (M = A R^-1) falls to tol, or the residual estimate to tol ||b - A x0||.
// A 16-bit factor ('H', 'B', 'X') is a usable preconditioner only while kappa(A) times its
// backward error stays well below one.  While refining with one, the contraction of LSQR's
// estimate is watched: from iteration STALL_START on, every STALL_SPAN iterations the
// iterations still needed at the observed rate are projected, and when they exceed what is
// left of maxit (or the estimate has stopped falling) A is factored again with Float32
// updates and LSQR restarts from that factor's direct solution.  prec_used receives the
// factor's precision of the final solve.
constexpr int STALL_START = 12, STALL_SPAN = 6;

template <typename T>
static int tile_gels_entry(i64 m, i64 n, i64 nb, i64 ib, const T *A, i64 lda, const T *b, T *x, char prec,
                           i64 maxit, double tol, i64 *iters, double *last, char *prec_used, void *stream) {
    prec = up(prec);
    if (m < 0) return arg_error(1, "m must be non-negative");
    if (n < 0 || n > m) return arg_error(2, "n must be in 0..m");
    if (nb <= 0) return arg_error(3, "nb must be positive");
    if (ib <= 0 || ib > nb) return arg_error(4, "ib must be in 1..nb");
    if (m % nb || n % nb) return arg_error(1, "m and n must be multiples of nb");
    if (lda < m) return arg_error(6, "lda must be >= m");
    if (prec != 'D' && prec != 'S' && prec != 'H' && prec != 'B' && prec != 'X')
        return arg_error(9, "prec must be 'D', 'S', 'H', 'B' or 'X'");
    if (prec != 'D' && sizeof(T) != 8) return arg_error(9, "prec '%c' needs Float64 data", prec);
    if (maxit < 0) return arg_error(10, "maxit must be non-negative");
    if (iters) *iters = 0;
    if (last) *last = 0;
    if (prec_used) *prec_used = prec;
    if (m == 0 || n == 0) return 0;
    hipStream_t s = as_stream(stream);
    const i64 mt = m / nb, ldt = mt * ib;
    const int rb = (int)((m + 511) / 512);
    const int nsplit = std::max(1, std::min(32, 2048 / rb));
    const int gm = (int)((m + 255) / 256), gn = (int)((n + 255) / 256);
    int st = 0;
    const size_t ne = (size_t)m * n + (size_t)ldt * n + (size_t)m + (size_t)m * (1 + nsplit) + 6 * (size_t)n;
    const size_t nd = (size_t)std::max(gm, gn) + 8;
    const size_t nsync = (size_t)n / TRSV_BS + 2;
    const size_t ndinv = n % TRSV_BS == 0 ? (size_t)n * TRSV_BS : 0;
    char *base = (char *)scratch(SLOT_LS, sizeof(T) * (ne + ndinv) + sizeof(double) * nd + sizeof(int) * (nsync + 2), s, &st);
    if (!base) return st;
    T *AF = (T *)base, *Tm = AF + (size_t)m * n, *Cb = Tm + (size_t)ldt * n, *u = Cb + (size_t)m;
    T *part = u + m, *v = part + (size_t)m * nsplit, *w = v + n, *y = w + n, *t = y + n, *q = t + n, *ty = q + n;
    double *bsum = (double *)(ty + n), *dnorm = bsum + nd - 8;
    int *sync = (int *)(bsum + nd);
    T *Dinv = (T *)(sync + nsync + (nsync & 1));  // 8-byte aligned
    NLA_HIP(hipMemsetAsync(sync, 0, sizeof(int) * nsync, s));
    // x := R^-1 x / R^-T x with R the upper triangle of the factor
    auto rsolve = [&](bool trans, T *vec) -> int {
        if (n % TRSV_BS == 0) {
            const hipError_t e = trsv_upper<T>(trans, (int)n, AF, m, Dinv, vec, ty, sync, s);
            return e == hipSuccess ? 0 : hip_error(e, "trsv_upper");
        }
        return rectrxm_async<T>('L', 'U', trans ? 'T' : 'N', T(1), 'S', n, 1, AF, m, vec, n, 0, s);
    };
    // the factor of a copy of A with trailing updates in precision p, and R's diagonal-block inverses
    auto factor = [&](char p) -> int {
        NLA_HIP(hipMemcpy2DAsync(AF, sizeof(T) * m, A, sizeof(T) * lda, sizeof(T) * m, n, hipMemcpyDeviceToDevice, s));
        NLA_HIP(hipMemsetAsync(Tm, 0, sizeof(T) * (size_t)ldt * n, s));
        if (p != 'D') NLA_ST(tile_geqrf_mixed_entry(m, n, nb, ib, (double *)AF, m, (double *)Tm, ldt, p, s));
        else NLA_ST(tile_geqrf_entry<T>(m, n, nb, ib, AF, m, Tm, ldt, s));
        if (n % TRSV_BS == 0) {
            diag_block_inv_kernel<T><<<(unsigned)(n / TRSV_BS), TRSV_BS, 0, s>>>(AF, m, Dinv);
            NLA_LAUNCH_CHECK();
        }
        return 0;
    };
    // x0 = R^-1 (Q^T b)[0:n]
    auto direct_start = [&]() -> int {
        NLA_HIP(hipMemcpyAsync(Cb, b, sizeof(T) * m, hipMemcpyDeviceToDevice, s));
        NLA_ST(tile_ormqr_entry<T>('T', m, n, 1, nb, ib, AF, m, Tm, ldt, Cb, m, s));
        NLA_ST(rsolve(false, Cb));
        NLA_HIP(hipMemcpyAsync(x, Cb, sizeof(T) * n, hipMemcpyDeviceToDevice, s));
        return 0;
    };
    NLA_ST(factor(prec));
    if (maxit == 0 || prec != 'D') {
        NLA_ST(direct_start());
        if (maxit == 0) {
            NLA_HIP(hipStreamSynchronize(s));
            return 0;
        }
    } else {
        // x0 = R^-1 R^-T (A^T b)
        gemv_t_kernel<T><<<(unsigned)((n + 3) / 4), 256, 0, s>>>(m, n, A, lda, b, x);
        NLA_LAUNCH_CHECK();
        NLA_ST(rsolve(true, x));
        NLA_ST(rsolve(false, x));
    }
    const dim3 gpart((unsigned)rb, (unsigned)nsplit);
    auto norm_read = [&](int nblocks, double *out) -> hipError_t {
        norm_finish_kernel<<<1, 256, 0, s>>>(nblocks, bsum, dnorm);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        if ((e = hipMemcpyAsync(out, dnorm, sizeof(double), hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
        return hipStreamSynchronize(s);
    };
    i64 done = 0;  // LSQR iterations taken over all restarts
    // LSQR from the current x for at most `budget` iterations; x += R^-1 y at the end.
    // watch: give up early (*stalled = true) when the projected iterations exceed the budget.
    auto lsqr = [&](i64 budget, bool watch, bool *stalled) -> int {
        *stalled = false;
        // M = A R^-1.  beta1 u1 = r0 = b - A x;  alpha1 v1 = M^T u1 = R^-T (A^T u1)
        NLA_HIP(hipMemcpyAsync(u, b, sizeof(T) * m, hipMemcpyDeviceToDevice, s));
        gemv_n_part_kernel<T><<<gpart, 256, 0, s>>>(m, n, A, lda, x, part, nsplit);
        NLA_LAUNCH_CHECK();
        gemv_n_finish_kernel<T><<<gm, 256, 0, s>>>(m, nsplit, part, T(1), u, bsum);  // u = A x - b
        NLA_LAUNCH_CHECK();
        double beta = 0, alpha = 0;
        NLA_HIP(norm_read(gm, &beta));
        const double r0norm = beta;
        if (beta == 0) return 0;
        scale_kernel<T><<<gm, 256, 0, s>>>(m, T(-1.0 / beta), u);  // u1 = (b - A x) / beta1
        NLA_LAUNCH_CHECK();
        gemv_t_kernel<T><<<(unsigned)((n + 3) / 4), 256, 0, s>>>(m, n, A, lda, u, q);
        NLA_LAUNCH_CHECK();
        NLA_ST(rsolve(true, q));
        NLA_HIP(hipMemsetAsync(v, 0, sizeof(T) * n, s));
        lincomb_kernel<T><<<gn, 256, 0, s>>>(n, q, T(0), v, bsum);
        NLA_LAUNCH_CHECK();
        NLA_HIP(norm_read(gn, &alpha));
        if (alpha == 0) return 0;
        scale_kernel<T><<<gn, 256, 0, s>>>(n, T(1.0 / alpha), v);
        NLA_LAUNCH_CHECK();
        NLA_HIP(hipMemcpyAsync(w, v, sizeof(T) * n, hipMemcpyDeviceToDevice, s));
        NLA_HIP(hipMemsetAsync(y, 0, sizeof(T) * n, s));
        double phibar = beta, rhobar = alpha, anorm2 = 0, test2 = 1;
        std::vector<double> hist;
        for (i64 it = 0; it < budget; ++it) {
            // beta u = M v - alpha u = A (R^-1 v) - alpha u
            NLA_HIP(hipMemcpyAsync(t, v, sizeof(T) * n, hipMemcpyDeviceToDevice, s));
            NLA_ST(rsolve(false, t));
            gemv_n_part_kernel<T><<<gpart, 256, 0, s>>>(m, n, A, lda, t, part, nsplit);
            NLA_LAUNCH_CHECK();
            gemv_n_finish_kernel<T><<<gm, 256, 0, s>>>(m, nsplit, part, T(alpha), u, bsum);
            NLA_LAUNCH_CHECK();
            NLA_HIP(norm_read(gm, &beta));
            anorm2 += alpha * alpha + beta * beta;
            if (beta > 0) {
                scale_kernel<T><<<gm, 256, 0, s>>>(m, T(1.0 / beta), u);
                NLA_LAUNCH_CHECK();
                // alpha v = M^T u - beta v = R^-T (A^T u) - beta v
                gemv_t_kernel<T><<<(unsigned)((n + 3) / 4), 256, 0, s>>>(m, n, A, lda, u, q);
                NLA_LAUNCH_CHECK();
                NLA_ST(rsolve(true, q));
                lincomb_kernel<T><<<gn, 256, 0, s>>>(n, q, T(beta), v, bsum);
                NLA_LAUNCH_CHECK();
                NLA_HIP(norm_read(gn, &alpha));
                if (alpha > 0) {
                    scale_kernel<T><<<gn, 256, 0, s>>>(n, T(1.0 / alpha), v);
                    NLA_LAUNCH_CHECK();
                }
            } else {
                alpha = 0;
            }
            // plane rotation eliminating beta from the lower bidiagonal
            const double rho = std::hypot(rhobar, beta), c = rhobar / rho, sn = beta / rho;
            const double theta = sn * alpha;
            rhobar = -c * alpha;
            const double phi = c * phibar;
            phibar = sn * phibar;
            lsqr_update_kernel<T><<<gn, 256, 0, s>>>(n, T(phi / rho), T(theta / rho), v, w, y);
            NLA_LAUNCH_CHECK();
            // ||M^T r|| / (||M||_F ||r||) = alpha |c| / ||B_k||_F;  ||r|| = phibar
            test2 = anorm2 > 0 ? alpha * std::fabs(c) / std::sqrt(anorm2) : 0;
            ++done;
            if (iters) *iters = done;
            if (last) *last = test2;
            if (test2 <= tol || phibar <= tol * r0norm || alpha == 0 || beta == 0) break;
            hist.push_back(test2);
            const i64 k = (i64)hist.size();
            if (watch && k >= STALL_START && (k - STALL_START) % STALL_SPAN == 0) {
                const double rate = std::pow(test2 / hist[k - 1 - STALL_SPAN], 1.0 / STALL_SPAN);
                const double need = rate < 1 ? std::log(tol / test2) / std::log(rate) : 1e30;
                if (need > double(budget - it - 1)) {
                    *stalled = true;
                    break;
                }
            }
        }
        // x = x + R^-1 y
        NLA_ST(rsolve(false, y));
        axpy1_kernel<T><<<gn, 256, 0, s>>>(n, y, x);
        NLA_LAUNCH_CHECK();
        return 0;
    };
    bool stalled = false;
    NLA_ST(lsqr(maxit, prec == 'H' || prec == 'B' || prec == 'X', &stalled));
    if (stalled && done < maxit) {
        // the 16-bit factor does not precondition A well enough: Float32 updates instead,
        // from that factor's own direct solution (a restart from the stalled iterate leaves
        // LSQR's attainable accuracy at the size of the correction it still has to make)
        NLA_ST(factor('S'));
        if (prec_used) *prec_used = 'S';
        NLA_ST(direct_start());
        NLA_ST(lsqr(maxit - done, false, &stalled));
    }
    int herr = 0;
    NLA_HIP(hipMemcpyAsync(&herr, sync + 1, sizeof(int), hipMemcpyDeviceToHost, s));
    NLA_HIP(hipStreamSynchronize(s));
    if (herr) return wait_timeout_error("tile_gels triangular solve");
    return 0;
}

}  // namespace nla

NLA_API int nla_dtile_geqrf_mixed(int64_t m, int64_t n, int64_t nb, int64_t ib, double *A, int64_t lda, double *T,
                                  int64_t ldt, char prec, void *stream) {
    return nla::tile_geqrf_mixed_entry(m, n, nb, ib, A, lda, T, ldt, prec, stream);
}
NLA_API int nla_dtile_gels(int64_t m, int64_t n, int64_t nb, int64_t ib, const double *A, int64_t lda, const double *b,
                           double *x, char prec, int64_t maxit, double tol, int64_t *iters, double *last_correction,
                           char *prec_used, void *stream) {
    return nla::tile_gels_entry<double>(m, n, nb, ib, A, lda, b, x, prec, maxit, tol, iters, last_correction, prec_used,
                                        stream);
}
NLA_API int nla_stile_gels(int64_t m, int64_t n, int64_t nb, int64_t ib, const float *A, int64_t lda, const float *b,
                           float *x, char prec, int64_t maxit, double tol, int64_t *iters, double *last_correction,
                           char *prec_used, void *stream) {
    return nla::tile_gels_entry<float>(m, n, nb, ib, A, lda, b, x, prec, maxit, tol, iters, last_correction, prec_used,
                                       stream);
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
