// Persistent dataflow tile QR (flat tree, PLASMA order) for CDNA4.
//
// The composition geqrt!(A_kk) -> unmqr!(A_kj) -> tsqrt!(A_kk, A_ik) -> tsmqr!(A_kj, A_ij)
// (src/geqrt.jl, src/unmqr.jl, src/tsqrt.jl, src/tsmqr.jl) is run as one launch of resident
// workgroups that claim tasks from an ordered list:
//   PANEL(k)    geqrt on A_kk, then tsqrt(A_kk, A_ik) for i = k+1.. (R_kk stays in place)
//   WALK(k, s)  a strip of SW columns of tile column j > k: unmqr('L','C') on A_kj's strip,
//               then tsmqr('L','C') against every A_ij strip, i = k+1.., A_kj's strip held
//               in LDS for the whole walk and each A_ij strip held in MFMA accumulators
//               across all ib-blocks of that tile.
// Columns of C are independent under a left block reflector, so strips never interact.
// Tasks publish progress counters (panel_prog[k] = rows of column k factored,
// strip_prog[k][s] = rows of strip s updated by step k); a task waits on counters of
// tasks listed before it, so the lowest unfinished claimed task can always run and the
// launch cannot deadlock.  Waits are bounded by a wall-clock limit that sets an error
// flag and lets every wave finish, so no input can hang the device.
//
// Arithmetic per tile is that of the reference routines (same blocks, same order); MFMA
// tiles only change the association of the inner sums.
#include "nla_common.hpp"
#include "nla_qr.hpp"

#include <cstdlib>
#include <vector>

namespace nla {
namespace dag {

constexpr int IB = 32;   // inner block (reflectors per T block)
constexpr int SW = 16;   // strip width = one MFMA column tile
constexpr int NTH = 256;

enum : int { TASK_PANEL = 0, TASK_WALK = 1 };
__host__ __device__ inline int enc(int type, int k, int s) { return (type << 30) | (k << 18) | s; }
__device__ inline int task_type(int c) { return c >> 30; }
__device__ inline int task_k(int c) { return (c >> 18) & 0xfff; }
__device__ inline int task_s(int c) { return c & 0x3ffff; }

// ---------------------------------------------------------------------------------
// Coherence.  Each XCD has its own L2; agent-scope acquire/release fences would write back
// (buffer_wbl2) or invalidate (buffer_inv) a whole L2 at every dependency wait.  Instead,
// every access of this kernel to the matrix and T goes through device-coherent (sc1)
// loads/stores, and a flag is published only after the publishing workgroup's stores
// have completed (s_waitcnt vmcnt(0) in every wave, then a barrier).
template <typename T> __device__ __forceinline__ T cld(const T *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T> __device__ __forceinline__ void cst(T *p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int ld_flag(const int *p) { return cld(p); }

// 2 s at the 100 MHz s_memrealtime clock
constexpr unsigned long long SPIN_LIMIT = 200000000ull;

// Workgroup-wide wait until *p >= v (thread 0 polls); on timeout sets *err.
__device__ void wait_ge(const int *p, int v, int *err, int *sh_flag) {
    if (threadIdx.x == 0) {
        if (ld_flag(p) < v) {
            const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
            while (ld_flag(p) < v) {
                if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
                if (__builtin_amdgcn_s_memrealtime() - t0 > SPIN_LIMIT) {
                    __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        *sh_flag = 0;
    }
    __syncthreads();
}

// All threads' coherent stores have completed, then thread 0 publishes v.
__device__ void publish(int *p, int v) {
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (threadIdx.x == 0) cst(p, v);
}

// ---------------------------------------------------------------------------------
// LDS layout (elements of T) shared by all task types.
template <typename T, int NB> struct Smem {
    static constexpr int LDX = NB + 1;                // strip buffers: [SW][LDX] (column c, row r)
    static constexpr int A1 = 0;                      // walker's A_kj strip / panel's R rows
    static constexpr int ST = A1 + SW * LDX;          // staging for global <-> accumulator moves
    static constexpr int P = ST + SW * LDX;           // W0 partials [4][IB/16][16][17]
    static constexpr int W0 = P + 4 * (IB / 16) * 16 * 17;  // [IB][SW+1]
    static constexpr int WF = W0 + IB * (SW + 1);           // [IB][SW+1]
    static constexpr int APPLY_END = WF + IB * (SW + 1);
    // panel block factorization (overlays the apply area; both are never live together)
    static constexpr int F = 0;                       // [IB][LDX] block columns
    static constexpr int RB = F + IB * LDX;           // [IB][IB+1] R block (tsqrt)
    static constexpr int TB = RB + IB * (IB + 1);     // [IB][IB+1] T block
    static constexpr int DS = TB + IB * (IB + 1);     // [IB] dots
    static constexpr int WS = DS + IB;                // [IB] w / z
    static constexpr int PANEL_END = WS + IB;
    static constexpr int TOTAL = APPLY_END > PANEL_END ? APPLY_END : PANEL_END;
};

// ---------------------------------------------------------------------------------
// Strip register file: wave w owns rows [w*NB/4, (w+1)*NB/4) of an NB x SW strip as RT
// 16x16 accumulator tiles; element (lane l, reg r) of tile rt is row
// w*NB/4 + 16 rt + acc_row(l/16, r), column l % 16.
template <typename T, int NB> struct Strip {
    static constexpr int RT = NB / 64;
    typename Acc<T>::t a[RT];
};

template <typename T, int NB>
__device__ void strip_load(Strip<T, NB> &S, const T *G, i64 ld, int ncols, T *sm) {
    constexpr int LDX = Smem<T, NB>::LDX;
    T *st = sm + Smem<T, NB>::ST;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    __syncthreads();
    for (int idx = t; idx < NB * SW; idx += NTH) {
        const int r = idx % NB, c = idx / NB;
        st[c * LDX + r] = c < ncols ? cld(&G[r + (i64)c * ld]) : T(0);
    }
    __syncthreads();
#pragma unroll
    for (int rt = 0; rt < Strip<T, NB>::RT; ++rt)
#pragma unroll
        for (int r = 0; r < 4; ++r)
            S.a[rt][r] = st[(lane & 15) * LDX + w * (NB / 4) + 16 * rt + acc_row<T>(lane >> 4, r)];
}

template <typename T, int NB>
__device__ void strip_store(const Strip<T, NB> &S, T *G, i64 ld, int ncols, T *sm) {
    constexpr int LDX = Smem<T, NB>::LDX;
    T *st = sm + Smem<T, NB>::ST;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    __syncthreads();
#pragma unroll
    for (int rt = 0; rt < Strip<T, NB>::RT; ++rt)
#pragma unroll
        for (int r = 0; r < 4; ++r)
            st[(lane & 15) * LDX + w * (NB / 4) + 16 * rt + acc_row<T>(lane >> 4, r)] = S.a[rt][r];
    __syncthreads();
    for (int idx = t; idx < NB * SW; idx += NTH) {
        const int r = idx % NB, c = idx / NB;
        if (c < ncols) cst(&G[r + (i64)c * ld], st[c * LDX + r]);
    }
}

// Reflector element of block column jrel (absolute tile column jabs) at tile row p.  Vb
// points at the block's first column.  TS: full column; GE: unit lower part of a geqrt
// tile (1 on the diagonal, 0 above, where the tile holds R).
template <typename T, bool GE>
__device__ __forceinline__ T vel(const T *Vb, i64 ldv, int p, int jrel, int jabs) {
    if (GE) {
        if (p < jabs) return T(0);
        if (p == jabs) return T(1);
    }
    return cld(&Vb[p + (i64)jrel * ldv]);
}

// Apply op(H_b) of one inner block (reflectors at tile columns j0.. j0+kb-1, kb <= IB;
// V points at column j0) from the left to
//   [A1s rows j0..j0+kb-1 ; strip]   (TS: V = full NB x kb, A1s = LDS strip)
//   [strip]                          (GE: V unit lower inside the strip rows)
// W0 = V^T strip (+ A1 rows), W = op(T_b) W0, A1 rows -= W, strip -= V W.
// notrans: op(T) = T (apply H); otherwise T^T (apply H^T, the reference's trans 'C').
template <typename T, int NB, bool GE>
__device__ void apply_block(Strip<T, NB> &S, const T *V, i64 ldv, const T *Tb, i64 ldt, int j0, int kb, bool notrans,
                            T *sm) {
    constexpr int LDX = Smem<T, NB>::LDX;
    constexpr int RT = Strip<T, NB>::RT;
    T *P = sm + Smem<T, NB>::P;
    T *W0 = sm + Smem<T, NB>::W0;
    T *WF = sm + Smem<T, NB>::WF;
    T *A1s = sm + Smem<T, NB>::A1;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    const int rbase = w * (NB / 4);

    // W0 partial over this wave's rows: output tile ot = reflectors j0 + 16 ot + (0..15)
#pragma unroll
    for (int ot = 0; ot < IB / 16; ++ot) {
        typename Acc<T>::t acc = acc_zero<T>();
        const int jrel = 16 * ot + lr;
        const bool jok = jrel < kb;
#pragma unroll
        for (int rt = 0; rt < RT; ++rt)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int p = rbase + 16 * rt + acc_row<T>(lk, r);
                const T a = jok ? vel<T, GE>(V, ldv, p, jrel, j0 + jrel) : T(0);
                mfma16(acc, a, S.a[rt][r]);
            }
#pragma unroll
        for (int r = 0; r < 4; ++r) P[((w * (IB / 16) + ot) * 16 + acc_row<T>(lk, r)) * 17 + lr] = acc[r];
    }
    __syncthreads();
    for (int idx = t; idx < IB * SW; idx += NTH) {
        const int j = idx / SW, c = idx % SW, ot = j >> 4, jr = j & 15;
        T s = 0;
#pragma unroll
        for (int ww = 0; ww < 4; ++ww) s += P[((ww * (IB / 16) + ot) * 16 + jr) * 17 + c];
        if (!GE && j < kb) s += A1s[c * LDX + j0 + j];
        W0[j * (SW + 1) + c] = j < kb ? s : T(0);
    }
    __syncthreads();
    // W = op(T_b) W0 (upper triangle of T_b only)
    if (w < IB / 16) {
        typename Acc<T>::t acc = acc_zero<T>();
        const int jj = 16 * w + lr;
#pragma unroll
        for (int k0 = 0; k0 < IB; k0 += 4) {
            const int q = k0 + lk;
            T a = 0;
            if (jj < kb && q < kb) {
                if (notrans) { if (jj <= q) a = cld(&Tb[jj + (i64)q * ldt]); }
                else         { if (q <= jj) a = cld(&Tb[q + (i64)jj * ldt]); }
            }
            mfma16(acc, a, W0[q * (SW + 1) + lr]);
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) WF[(16 * w + acc_row<T>(lk, r)) * (SW + 1) + lr] = acc[r];
    }
    __syncthreads();
    if (!GE)
        for (int idx = t; idx < kb * SW; idx += NTH) {
            const int j = idx % kb, c = idx / kb;
            A1s[c * LDX + j0 + j] -= WF[j * (SW + 1) + c];
        }
    // strip -= V W
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) {
        const int p = rbase + 16 * rt + lr;
#pragma unroll
        for (int k0 = 0; k0 < IB; k0 += 4) {
            const int q = k0 + lk;
            const T a = q < kb ? -vel<T, GE>(V, ldv, p, q, j0 + q) : T(0);
            mfma16(S.a[rt], a, WF[q * (SW + 1) + lr]);
        }
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------------
// Panel block factorizations (LDS resident, one 256-thread workgroup).
// 8 threads per block column compute all column dots with the pivot column in one pass
// (norm included), reduced by shuffles inside the 8-lane group; then every thread derives
// beta/tau/scale itself.  Three barriers per column.

// geqr2! + larft! on block columns [c0, c0+kb) of an NB x NB tile A (rows c0..NB-1).
template <typename T, int NB>
__device__ void ge_block_factor(T *A, i64 lda, T *tau, T *Tm, i64 ldt, int c0, int kb, T *sm) {
    constexpr int LDX = Smem<T, NB>::LDX;
    T *F = sm + Smem<T, NB>::F;
    T *TBs = sm + Smem<T, NB>::TB;
    T *Ds = sm + Smem<T, NB>::DS;
    T *Ws = sm + Smem<T, NB>::WS;
    const int t = threadIdx.x;
    __syncthreads();
    for (int idx = t; idx < kb * NB; idx += NTH) {
        const int r = idx % NB, c = idx / NB;
        F[c * LDX + r] = cld(&A[r + (i64)(c0 + c) * lda]);
    }
    __syncthreads();
    const int gj = t >> 3, part = t & 7;  // 32 columns x 8 threads
    for (int c = 0; c < kb; ++c) {
        const int cc = c0 + c;  // pivot row
        if (gj < kb) {
            T s = 0;
            for (int p = cc + 1 + part; p < NB; p += 8) s += F[gj * LDX + p] * F[c * LDX + p];
            s += __shfl_xor(s, 1, 8);
            s += __shfl_xor(s, 2, 8);
            s += __shfl_xor(s, 4, 8);
            if (part == 0) Ds[gj] = s;
        }
        __syncthreads();
        const T alpha = F[c * LDX + cc], xs = Ds[c];
        T beta = alpha, tv = 0, scale = 0;
        if (xs != T(0)) {
            beta = -copysign(sqrt(alpha * alpha + xs), alpha);
            tv = (beta - alpha) / beta;
            scale = T(1) / (alpha - beta);
        }
        if (t < kb && t != c) Ws[t] = F[t * LDX + cc] + Ds[t] * scale;  // w_j (j > c), z_j (j < c)
        __syncthreads();
        if (tv != T(0)) {
            for (int p = cc + 1 + t; p < NB; p += NTH) {
                const T v = F[c * LDX + p] * scale;
                F[c * LDX + p] = v;
                for (int j = c + 1; j < kb; ++j) F[j * LDX + p] -= tv * v * Ws[j];
            }
            if (t > c && t < kb) F[t * LDX + cc] -= tv * Ws[t];
        }
        if (t < c) {
            T s = 0;
            if (tv != T(0))
                for (int q = t; q < c; ++q) s += TBs[t * (IB + 1) + q] * (-tv * Ws[q]);
            TBs[t * (IB + 1) + c] = s;
        }
        if (t == 0) {
            F[c * LDX + cc] = beta;
            TBs[c * (IB + 1) + c] = tv;
            tau[cc] = tv;
        }
        __syncthreads();
    }
    for (int idx = t; idx < kb * NB; idx += NTH) {
        const int r = idx % NB, c = idx / NB;
        if (r >= c0) cst(&A[r + (i64)(c0 + c) * lda], F[c * LDX + r]);
    }
    for (int idx = t; idx < kb * kb; idx += NTH) {
        const int i = idx % kb, j = idx / kb;
        if (i <= j) cst(&Tm[i + (i64)(c0 + j) * ldt], TBs[i * (IB + 1) + j]);
    }
    __syncthreads();
}

// tsqrt block: [R (NB x NB upper, rows/cols c0..c0+kb-1 used) ; A2 (NB x NB)], block
// columns [c0, c0+kb).  V = A2 columns, T block into Tm[:, c0..].
template <typename T, int NB>
__device__ void ts_block_factor(T *R, i64 ldr, T *A2, i64 lda2, T *tau, T *Tm, i64 ldt, int c0, int kb, T *sm) {
    constexpr int LDX = Smem<T, NB>::LDX;
    T *F = sm + Smem<T, NB>::F;
    T *RBs = sm + Smem<T, NB>::RB;
    T *TBs = sm + Smem<T, NB>::TB;
    T *Ds = sm + Smem<T, NB>::DS;
    T *Ws = sm + Smem<T, NB>::WS;
    const int t = threadIdx.x;
    __syncthreads();
    for (int idx = t; idx < kb * NB; idx += NTH) {
        const int r = idx % NB, c = idx / NB;
        F[c * LDX + r] = cld(&A2[r + (i64)(c0 + c) * lda2]);
    }
    for (int idx = t; idx < kb * kb; idx += NTH) {
        const int i = idx % kb, j = idx / kb;
        RBs[i * (IB + 1) + j] = cld(&R[(c0 + i) + (i64)(c0 + j) * ldr]);
    }
    __syncthreads();
    const int gj = t >> 3, part = t & 7;
    for (int c = 0; c < kb; ++c) {
        if (gj < kb) {
            T s = 0;
            for (int p = part; p < NB; p += 8) s += F[gj * LDX + p] * F[c * LDX + p];
            s += __shfl_xor(s, 1, 8);
            s += __shfl_xor(s, 2, 8);
            s += __shfl_xor(s, 4, 8);
            if (part == 0) Ds[gj] = s;
        }
        __syncthreads();
        const T alpha = RBs[c * (IB + 1) + c], xs = Ds[c];
        T beta = alpha, tv = 0, scale = 0;
        if (xs != T(0)) {
            beta = -copysign(sqrt(alpha * alpha + xs), alpha);
            tv = (beta - alpha) / beta;
            scale = T(1) / (alpha - beta);
        }
        if (t < kb && t != c) Ws[t] = (t > c ? RBs[c * (IB + 1) + t] : T(0)) + Ds[t] * scale;
        __syncthreads();
        if (tv != T(0)) {
            for (int p = t; p < NB; p += NTH) {
                const T v = F[c * LDX + p] * scale;
                F[c * LDX + p] = v;
                for (int j = c + 1; j < kb; ++j) F[j * LDX + p] -= tv * v * Ws[j];
            }
            if (t > c && t < kb) RBs[c * (IB + 1) + t] -= tv * Ws[t];
        }
        if (t < c) {
            T s = 0;
            if (tv != T(0))
                for (int q = t; q < c; ++q) s += TBs[t * (IB + 1) + q] * (-tv * Ws[q]);
            TBs[t * (IB + 1) + c] = s;
        }
        if (t == 0) {
            RBs[c * (IB + 1) + c] = beta;
            TBs[c * (IB + 1) + c] = tv;
            tau[c0 + c] = tv;
        }
        __syncthreads();
    }
    for (int idx = t; idx < kb * NB; idx += NTH) {
        const int r = idx % NB, c = idx / NB;
        cst(&A2[r + (i64)(c0 + c) * lda2], F[c * LDX + r]);
    }
    for (int idx = t; idx < kb * kb; idx += NTH) {
        const int i = idx % kb, j = idx / kb;
        if (i <= j) {
            cst(&R[(c0 + i) + (i64)(c0 + j) * ldr], RBs[i * (IB + 1) + j]);
            cst(&Tm[i + (i64)(c0 + j) * ldt], TBs[i * (IB + 1) + j]);
        }
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------------
// Tile-level operations built from the block routines.

// geqrt!(NB, NB, IB, A, T, tau): src/geqrt.jl:44-67
template <typename T, int NB> __device__ void tile_geqrt(T *A, i64 lda, T *Tm, i64 ldt, T *tau, T *sm) {
    for (int c0 = 0; c0 < NB; c0 += IB) {
        ge_block_factor<T, NB>(A, lda, tau, Tm, ldt, c0, IB, sm);
        // larfb!('L','C') of this block on the tile's trailing columns, strip by strip
        for (int s0 = c0 + IB; s0 < NB; s0 += SW) {
            Strip<T, NB> S;
            strip_load<T, NB>(S, A + (i64)s0 * lda, lda, SW, sm);
            apply_block<T, NB, true>(S, A + (i64)c0 * lda, lda, Tm + (i64)c0 * ldt, ldt, c0, IB, false, sm);
            strip_store<T, NB>(S, A + (i64)s0 * lda, lda, SW, sm);
        }
    }
}

// Load A1 rows [r0, r0+nr) of a strip (columns s0..s0+SW-1 of R) into A1s.
template <typename T, int NB>
__device__ void a1_load(const T *R, i64 ldr, int r0, int nr, int ncols, T *sm) {
    constexpr int LDX = Smem<T, NB>::LDX;
    T *A1s = sm + Smem<T, NB>::A1;
    __syncthreads();
    for (int idx = threadIdx.x; idx < nr * SW; idx += NTH) {
        const int r = idx % nr, c = idx / nr;
        A1s[c * LDX + r0 + r] = c < ncols ? cld(&R[(r0 + r) + (i64)c * ldr]) : T(0);
    }
    __syncthreads();
}
template <typename T, int NB>
__device__ void a1_store(T *R, i64 ldr, int r0, int nr, int ncols, T *sm) {
    constexpr int LDX = Smem<T, NB>::LDX;
    T *A1s = sm + Smem<T, NB>::A1;
    __syncthreads();
    for (int idx = threadIdx.x; idx < nr * SW; idx += NTH) {
        const int r = idx % nr, c = idx / nr;
        if (c < ncols) cst(&R[(r0 + r) + (i64)c * ldr], A1s[c * LDX + r0 + r]);
    }
    __syncthreads();
}

// tsqrt!(NB, NB, IB, R, A2, T, tau): src/tsqrt.jl:146-155
template <typename T, int NB>
__device__ void tile_tsqrt(T *R, i64 ldr, T *A2, i64 lda2, T *Tm, i64 ldt, T *tau, T *sm) {
    for (int c0 = 0; c0 < NB; c0 += IB) {
        ts_block_factor<T, NB>(R, ldr, A2, lda2, tau, Tm, ldt, c0, IB, sm);
        // tsmqr!('L','C', ...) of this block on the trailing columns
        for (int s0 = c0 + IB; s0 < NB; s0 += SW) {
            a1_load<T, NB>(R + (i64)s0 * ldr, ldr, c0, IB, SW, sm);
            Strip<T, NB> S;
            strip_load<T, NB>(S, A2 + (i64)s0 * lda2, lda2, SW, sm);
            apply_block<T, NB, false>(S, A2 + (i64)c0 * lda2, lda2, Tm + (i64)c0 * ldt, ldt, c0, IB, false, sm);
            strip_store<T, NB>(S, A2 + (i64)s0 * lda2, lda2, SW, sm);
            a1_store<T, NB>(R + (i64)s0 * ldr, ldr, c0, IB, SW, sm);
        }
    }
}

// ---------------------------------------------------------------------------------
struct DagArgs {
    int mt, nt, nstrips;      // tile rows/cols, SW-strips of the whole matrix
    int ntasks;
    const int *tasks;
    int *counter;             // next task
    int *panel_prog;          // [kt]
    int *strip_prog;          // [kt][nstrips]
    int *err;
    unsigned long long *trace;  // optional timeline: [ntasks][4] = claim, ready, end, workgroup
    unsigned long long *ptrace; // optional: [kt][mt] panel tile completion times
};

template <typename T, int NB>
__device__ __noinline__ void run_panel(int k, T *A, i64 lda, T *Tm, i64 ldt, T *tau, const DagArgs &g, T *sm,
                                       int *sh_flag) {
    constexpr int SPT = NB / SW;  // strips per tile column
    T *Akk = A + (i64)k * NB + (i64)k * NB * lda;
    T *Tkk = Tm + (i64)k * IB + (i64)k * NB * ldt;
    if (k > 0)
        for (int s = k * SPT; s < (k + 1) * SPT; ++s) wait_ge(&g.strip_prog[(k - 1) * g.nstrips + s], k + 1, g.err, sh_flag);
    tile_geqrt<T, NB>(Akk, lda, Tkk, ldt, tau, sm);
    publish(&g.panel_prog[k], k + 1);
    if (g.ptrace && threadIdx.x == 0) g.ptrace[(size_t)k * g.mt + k] = __builtin_amdgcn_s_memrealtime();
    for (int i = k + 1; i < g.mt; ++i) {
        if (k > 0)
            for (int s = k * SPT; s < (k + 1) * SPT; ++s)
                wait_ge(&g.strip_prog[(k - 1) * g.nstrips + s], i + 1, g.err, sh_flag);
        T *Aik = A + (i64)i * NB + (i64)k * NB * lda;
        T *Tik = Tm + (i64)i * IB + (i64)k * NB * ldt;
        tile_tsqrt<T, NB>(Akk, lda, Aik, lda, Tik, ldt, tau, sm);
        publish(&g.panel_prog[k], i + 1);
        if (g.ptrace && threadIdx.x == 0) g.ptrace[(size_t)k * g.mt + i] = __builtin_amdgcn_s_memrealtime();
    }
}

template <typename T, int NB>
__device__ __noinline__ void run_walk(int k, int s, T *A, i64 lda, T *Tm, i64 ldt, const DagArgs &g, T *sm,
                                      int *sh_flag) {
    const int col0 = s * SW;  // first matrix column of the strip
    int *prev = k > 0 ? &g.strip_prog[(k - 1) * g.nstrips + s] : nullptr;
    int *mine = &g.strip_prog[k * g.nstrips + s];
    const T *Akk = A + (i64)k * NB + (i64)k * NB * lda;
    const T *Tkk = Tm + (i64)k * IB + (i64)k * NB * ldt;
    T *Akj = A + (i64)k * NB + (i64)col0 * lda;
    wait_ge(&g.panel_prog[k], k + 1, g.err, sh_flag);
    if (prev) wait_ge(prev, k + 1, g.err, sh_flag);
    // unmqr!('L','C') with geqrt(A_kk)'s reflectors, forward over blocks
    Strip<T, NB> S;
    strip_load<T, NB>(S, Akj, lda, SW, sm);
    for (int c0 = 0; c0 < NB; c0 += IB)
        apply_block<T, NB, true>(S, Akk + (i64)c0 * lda, lda, Tkk + (i64)c0 * ldt, ldt, c0, IB, false, sm);
    strip_store<T, NB>(S, Akj, lda, SW, sm);
    publish(mine, k + 1);
    a1_load<T, NB>(Akj, lda, 0, NB, SW, sm);
    for (int i = k + 1; i < g.mt; ++i) {
        wait_ge(&g.panel_prog[k], i + 1, g.err, sh_flag);
        if (prev) wait_ge(prev, i + 1, g.err, sh_flag);
        const T *Aik = A + (i64)i * NB + (i64)k * NB * lda;
        const T *Tik = Tm + (i64)i * IB + (i64)k * NB * ldt;
        T *Aij = A + (i64)i * NB + (i64)col0 * lda;
        strip_load<T, NB>(S, Aij, lda, SW, sm);
        for (int c0 = 0; c0 < NB; c0 += IB)
            apply_block<T, NB, false>(S, Aik + (i64)c0 * lda, lda, Tik + (i64)c0 * ldt, ldt, c0, IB, false, sm);
        strip_store<T, NB>(S, Aij, lda, SW, sm);
        if (i + 1 == g.mt) a1_store<T, NB>(Akj, lda, 0, NB, SW, sm);
        publish(mine, i + 1);
    }
}

template <typename T, int NB>
__global__ __launch_bounds__(NTH, 2) void tileqr_dag_kernel(T *A, i64 lda, T *Tm, i64 ldt, T *tau_ws, DagArgs g) {
    extern __shared__ __align__(16) unsigned char dyn_raw[];
    T *sm = reinterpret_cast<T *>(dyn_raw);
    __shared__ int sh_task, sh_flag;
    T *tau = tau_ws + (size_t)blockIdx.x * NB;  // per-workgroup tau scratch (NB entries)
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) sh_task = atomicAdd(g.counter, 1);
        __syncthreads();
        const int task = sh_task;
        if (task >= g.ntasks) break;
        const int code = g.tasks[task];
        if (g.trace && threadIdx.x == 0) {
            g.trace[(size_t)task * 4 + 0] = __builtin_amdgcn_s_memrealtime();
            g.trace[(size_t)task * 4 + 3] = blockIdx.x;
        }
        if (task_type(code) == TASK_PANEL) run_panel<T, NB>(task_k(code), A, lda, Tm, ldt, tau, g, sm, &sh_flag);
        else run_walk<T, NB>(task_k(code), task_s(code), A, lda, Tm, ldt, g, sm, &sh_flag);
        if (g.trace && threadIdx.x == 0) g.trace[(size_t)task * 4 + 2] = __builtin_amdgcn_s_memrealtime();
    }
}

__global__ void fill_int_kernel(int *p, int n, int v) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// Task list in a topological order with one step of lookahead:
// PANEL(0), then for each k: WALK(k, strips of column k+1), PANEL(k+1), WALK(k, other strips).
static std::vector<int> build_tasks(int mt, int nt, int nb) {
    const int kt = mt < nt ? mt : nt, spt = nb / SW;
    std::vector<int> v;
    v.push_back(enc(TASK_PANEL, 0, 0));
    for (int k = 0; k < kt; ++k) {
        if (k + 1 < nt)
            for (int s = (k + 1) * spt; s < (k + 2) * spt; ++s) v.push_back(enc(TASK_WALK, k, s));
        if (k + 1 < kt) v.push_back(enc(TASK_PANEL, k + 1, 0));
        for (int s = (k + 2) * spt; s < nt * spt; ++s) v.push_back(enc(TASK_WALK, k, s));
    }
    return v;
}

}  // namespace dag

// Host driver.  Returns 0, a HIP error code, or -1 (reported as a HIP-side failure by the
// caller) when a dependency wait timed out.
template <typename T, int NB>
static int dag_run(i64 m, i64 n, T *A, i64 lda, T *Tm, i64 ldt, hipStream_t s) {
    using namespace dag;
    const int mt = (int)(m / NB), nt = (int)(n / NB), kt = mt < nt ? mt : nt;
    const int nstrips = (int)(n / SW);
    std::vector<int> tasks = build_tasks(mt, nt, NB);
    const int ntasks = (int)tasks.size();
    int dev = 0, ncu = 0;
    NLA_HIP(hipGetDevice(&dev));
    NLA_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const size_t sh = sizeof(T) * (size_t)Smem<T, NB>::TOTAL;
    if (sh > 64 * 1024)
        NLA_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&tileqr_dag_kernel<T, NB>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
    int per_cu = 0;
    NLA_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, tileqr_dag_kernel<T, NB>, NTH, sh));
    if (per_cu < 1) per_cu = 1;
    int grid = ncu * per_cu;
    if (grid > ntasks) grid = ntasks;
    // scheduler memory: tasks | counter | err | panel_prog[kt] | strip_prog[kt*nstrips]
    const size_t nint = (size_t)ntasks + 2 + kt + (size_t)kt * nstrips;
    int st = 0;
    int *mem = (int *)scratch(SLOT_SCHED, nint * sizeof(int), s, &st);
    if (!mem) return st;
    T *tau_ws = (T *)scratch(SLOT_AUX, sizeof(T) * (size_t)grid * NB, s, &st);
    if (!tau_ws) return st;
    int *d_tasks = mem, *counter = mem + ntasks, *err = counter + 1, *pp = err + 1, *sp = pp + kt;
    NLA_HIP(hipMemcpyAsync(d_tasks, tasks.data(), (size_t)ntasks * sizeof(int), hipMemcpyHostToDevice, s));
    const int nzero = 2 + kt + kt * nstrips;
    fill_int_kernel<<<(nzero + 255) / 256 > 1024 ? 1024 : (nzero + 255) / 256, 256, 0, s>>>(counter, nzero, 0);
    NLA_LAUNCH_CHECK();
    static const bool want_trace = getenv("NLA_TILEQR_TRACE") != nullptr;
    unsigned long long *tr = nullptr, *ptr = nullptr;
    if (want_trace) {
        tr = (unsigned long long *)scratch(SLOT_MIXED, sizeof(unsigned long long) * ((size_t)ntasks * 4 + (size_t)kt * mt),
                                          s, &st);
        if (!tr) return st;
        ptr = tr + (size_t)ntasks * 4;
        NLA_HIP(hipMemsetAsync(tr, 0, sizeof(unsigned long long) * ((size_t)ntasks * 4 + (size_t)kt * mt), s));
    }
    DagArgs g{mt, nt, nstrips, ntasks, d_tasks, counter, pp, sp, err, tr, ptr};
    {
        ProfScope prof(PK_TILE_QR, s);
        tileqr_dag_kernel<T, NB><<<grid, NTH, sh, s>>>(A, lda, Tm, ldt, tau_ws, g);
        NLA_LAUNCH_CHECK();
    }
    int herr = 0;
    NLA_HIP(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, s));
    NLA_HIP(hipStreamSynchronize(s));
    if (want_trace) {
        // "task type k s claim ready end wg" lines, then "panel k i time" lines
        const char *path = getenv("NLA_TILEQR_TRACE");
        std::vector<unsigned long long> h((size_t)ntasks * 4 + (size_t)kt * mt);
        NLA_HIP(hipMemcpy(h.data(), tr, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
        if (FILE *f = fopen(path, "w")) {
            for (int i = 0; i < ntasks; ++i)
                fprintf(f, "task %d %d %d %llu %llu %llu\n", tasks[i] >> 30, (tasks[i] >> 18) & 0xfff, tasks[i] & 0x3ffff,
                        h[i * 4 + 0], h[i * 4 + 2], h[i * 4 + 3]);
            for (int k = 0; k < kt; ++k)
                for (int i = k; i < mt; ++i) fprintf(f, "panel %d %d %llu\n", k, i, h[(size_t)ntasks * 4 + (size_t)k * mt + i]);
            fclose(f);
        }
    }
    if (herr) {
        set_error("tile QR: a dependency wait exceeded its time limit (results are invalid)");
        return -1;
    }
    return 0;
}

template <typename T>
int tile_geqrf_dag(i64 m, i64 n, i64 nb, i64 ib, T *A, i64 lda, T *Tm, i64 ldt, hipStream_t s, bool *handled) {
    *handled = false;
    if (ib != dag::IB) return 0;
    if (m / nb >= 4096 || n / dag::SW >= (1 << 18)) return 0;
    *handled = true;
    if (nb == 64) return dag_run<T, 64>(m, n, A, lda, Tm, ldt, s);
    if (nb == 128) return dag_run<T, 128>(m, n, A, lda, Tm, ldt, s);
    *handled = false;
    return 0;
}

template int tile_geqrf_dag<double>(i64, i64, i64, i64, double *, i64, double *, i64, hipStream_t, bool *);
template int tile_geqrf_dag<float>(i64, i64, i64, i64, float *, i64, float *, i64, hipStream_t, bool *);

}  // namespace nla
