// Persistent dataflow tile QR (flat tree, PLASMA order) for CDNA4.
//
// The composition geqrt!(A_kk) -> unmqr!(A_kj) -> tsqrt!(A_kk, A_ik) -> tsmqr!(A_kj, A_ij)
// (src/geqrt.jl, src/unmqr.jl, src/tsqrt.jl, src/tsmqr.jl) is run as one launch of resident
// workgroups that claim tasks from an ordered list:
//   PANEL(k, b) ib-block b of geqrt on A_kk, then of tsqrt(A_kk, A_ik) for i = k+1.., tile
//               by tile (R_kk stays in place); the NB/IB block stages of one panel form a
//               pipeline over the tiles
//   WALK(k, s)  TW columns of tile column j > k, one 16-column strip per wave: unmqr('L','C')
//               on A_kj's strip, then tsmqr('L','C') against every A_ij strip, i = k+1..,
//               A_kj's strip held in MFMA accumulators for the whole walk.
// Columns of C are independent under a left block reflector, so strips never interact.
// Tasks publish progress counters (panel_prog[k][b] = tiles of column k done by block
// stage b, strip_prog[k][s] = rows of strip group s updated by step k); a task waits only
// on counters of tasks listed before it, so the lowest unfinished claimed task can always
// run and the launch cannot deadlock.  Waits are bounded by a wall-clock limit that sets an
// error flag and lets every wave finish, so no input can hang the device.
//
// Arithmetic per tile is that of the reference routines (same blocks, same order); MFMA
// tiles only change the association of the inner sums.
#include "nla_common.hpp"
#include "nla_qr.hpp"

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <type_traits>
#include <vector>

namespace nla {
namespace dag {

constexpr int IB = 32;   // inner block (reflectors per T block)
constexpr int SW = 16;   // strip width = one MFMA column tile
constexpr int NTH = 256;

// Walker element type of the split-BF16 updates (prec 'X'): every product operand as two BF16
// parts, x = hi + lo with hi = bf16(x), lo = bf16(x - hi) (16 mantissa bits together), and a
// product as hi*hi + hi*lo + lo*hi on the BF16 matrix cores with Float32 accumulation.  The
// dropped lo*lo term and the parts' rounding leave a product good to about 2^-16 (FP16
// operands: 2^-11; Float32: 2^-24) at three 16-bit products, 16/3 times the Float32 MFMA rate.
struct bf16x3 { __bf16 h; };
static_assert(sizeof(bf16x3) == 2, "a 16-bit walker type");
template <typename Tw> constexpr int lp_parts() { return std::is_same<Tw, bf16x3>::value ? 2 : 1; }

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
// Data layout.  Every wave owns a strip of SW = 16 columns and NB rows of the matrix in MFMA
// accumulator registers (NB/16 tiles of 16x16; element (tile rt, lane l, reg r) is row
// 16 rt + acc_row(l/16, r), column l % 16), so a block reflector is applied to a strip by
// one wave with no cross-wave reduction.  The block's reflectors V (NB x IB) and T (IB x IB)
// are staged once in LDS and shared by the four waves (four strips = TW columns).
constexpr int WAVES = NTH / 64;
constexpr int TW = SW * WAVES;  // matrix columns per workgroup task
// Columns per walker task.  The reduced-precision walkers (Tw narrower than T, NB >= 128)
// give each wave two 16-column strips: the block's V/T operands are read once for both, and
// the per-block loads and barrier are paid once per 2 TW columns.
template <typename T, typename Tw, int NB> constexpr int walk_width() {
    return sizeof(Tw) < sizeof(T) && NB >= 128 ? 2 * TW : TW;
}

template <typename T, int NB> struct Lds {
    static constexpr int LDV = NB + 2;                 // Vs[j * LDV + p] = V[p][j]
    static constexpr int V = 0;
    static constexpr int TT = V + IB * LDV;            // Ts[q * (IB+1) + j] = T[q][j] (zero below diag)
    static constexpr int BUF = TT + IB * (IB + 1);     // second V / T pair (walkers double-buffer)
    // panel tasks only:
    static constexpr int RS = BUF + IB * LDV + IB * (IB + 1);  // Rs[i * (IB+1) + j] = R block (tsqrt)
    static constexpr int ZB = RS + IB * (IB + 1);      // block factorization scratch
    static constexpr int TAU = ZB + IB * IB;
    static constexpr int RED = TAU + IB;
    static constexpr int DOT = RED + 16;
    static constexpr int PEND = DOT + 8 * IB;
    // walker tasks only (same space as the panel scratch): the strip staging columns,
    // LDS_S = NB + one 16-byte vector per column, so every column starts 16-byte aligned
    static constexpr int LDS_S = NB + 16 / (int)sizeof(T);
    static constexpr int SSTG = RS;
    static constexpr int WEND = SSTG + TW * LDS_S;
    static constexpr int TOTAL = PEND > WEND ? PEND : WEND;
};

// The Float32-update walker's V/T block image: one LDS buffer of the Float32 walker layout
// (V column-major with LDV, then T with zeros below the diagonal), written once per block by the
// panel; images are 1 KB-aligned in the scratch slot.
template <int NB> struct F32Img {
    static constexpr int BYTES = Lds<float, NB>::BUF * (int)sizeof(float);
    static constexpr int STRIDE = (BYTES + 1023) / 1024 * 1024;
    static_assert(BYTES % 16 == 0, "whole 16-byte pieces");
};

// LDS base with its address space visible at every use (a pointer passed through a call
// would degrade to FLAT accesses).
template <typename T> __device__ __forceinline__ T *lds_base() {
    extern __shared__ __align__(16) unsigned char dyn_raw[];
    return reinterpret_cast<T *>(dyn_raw);
}

template <typename T, int NB> struct WStrip {
    typename Acc<T>::t a[NB / 16];
};

// rows [r0, r0 + 16*NT) of a 16-column strip starting at G (column-major, ld)
template <typename T, int NT>
__device__ __forceinline__ void tiles_load(typename Acc<T>::t *a, const T *G, i64 ld, int r0) {
    const int lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int rt = 0; rt < NT; ++rt)
#pragma unroll
        for (int r = 0; r < 4; ++r) a[rt][r] = cld(&G[(r0 + 16 * rt + acc_row<T>(lk, r)) + (i64)lr * ld]);
}
template <typename T, int NT>
__device__ __forceinline__ void tiles_store(const typename Acc<T>::t *a, T *G, i64 ld, int r0) {
    const int lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int rt = 0; rt < NT; ++rt)
#pragma unroll
        for (int r = 0; r < 4; ++r) cst(&G[(r0 + 16 * rt + acc_row<T>(lk, r)) + (i64)lr * ld], a[rt][r]);
}

// Strip staging.  A wave's 16-column strip moves between HBM and its MFMA registers through
// its own SW columns of LDS (no workgroup synchronisation): loads go straight to LDS with
// global_load_lds, one whole column (NB elements, 16 bytes per lane) per instruction, and
// stores leave as column-contiguous 16-byte device-coherent buffer stores.  In the register
// layout a lane holds 4 rows of one column per 16x16 tile, so direct register loads/stores
// touch 16 partial cache lines per instruction; staged, each instruction moves whole lines.
template <typename T, int NB> struct Stage {
    static constexpr int VEC = 16 / (int)sizeof(T);  // elements per 16-byte access
    static constexpr int LPC = NB / VEC;             // lanes per column
    static constexpr int CPI = 64 / LPC;             // columns per store instruction
    static_assert(LPC <= 64 && 64 % LPC == 0, "a column must be a whole number of lane groups");
};
// The staging columns hold the storage type Ts (the matrix's); they start at the walker
// region of the kernel's LDS layout type Tl (Ts itself, or the Float32 of the mixed walker).
template <typename Ts, typename Tl, int NB> __device__ __forceinline__ Ts *stage_ptr() {
    return reinterpret_cast<Ts *>(lds_base<Tl>() + Lds<Tl, NB>::SSTG) + (threadIdx.x >> 6) * SW * Lds<Ts, NB>::LDS_S;
}
// Issue the loads of the wave's strip (columns G, G + ld, ...) into its staging columns.
// Completion is tracked by vmcnt; the caller waits before stage_to_regs.
template <typename Ts, int NB> __device__ __forceinline__ void stage_load(Ts *stg, const Ts *G, i64 ld) {
    constexpr int LDS_S = Lds<Ts, NB>::LDS_S;
    const int lane = threadIdx.x & 63;
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");  // earlier reads of the staging columns are done
    if (lane < Stage<Ts, NB>::LPC) {
#pragma unroll
        for (int c = 0; c < SW; ++c) {
            const char *src = reinterpret_cast<const char *>(G + (i64)c * ld) + lane * 16;
            auto *d = (__attribute__((address_space(3))) void *)(stg + c * LDS_S);
            __builtin_amdgcn_global_load_lds(src, d, 16, 0, /*sc1*/ 16);
        }
    }
}
// staging columns (Ts) -> accumulator registers (Tr)
template <typename Ts, typename Tr, int NB>
__device__ __forceinline__ void stage_to_regs(const Ts *stg, WStrip<Tr, NB> &S) {
    constexpr int LDS_S = Lds<Ts, NB>::LDS_S;
    const int lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int rt = 0; rt < NB / 16; ++rt)
#pragma unroll
        for (int r = 0; r < 4; ++r) S.a[rt][r] = (Tr)stg[lr * LDS_S + 16 * rt + acc_row<Tr>(lk, r)];
}
// accumulator registers (Tr) -> staging columns (Ts) -> HBM (rows rmin..NB-1 of the wave's 16
// columns at G; rmin a multiple of the 16-byte vector)
template <typename Ts, typename Tr, int NB>
__device__ __forceinline__ void stage_store(Ts *stg, const WStrip<Tr, NB> &S, Ts *G, i64 ld, int rmin = 0) {
    using St = Stage<Ts, NB>;
    typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
    typedef Ts tvec __attribute__((ext_vector_type(St::VEC)));
    constexpr int LDS_S = Lds<Ts, NB>::LDS_S;
    const int lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int rt = 0; rt < NB / 16; ++rt)
#pragma unroll
        for (int r = 0; r < 4; ++r) stg[lr * LDS_S + 16 * rt + acc_row<Tr>(lk, r)] = (Ts)S.a[rt][r];
    // the wave's LDS accesses execute in order; this keeps the compiler from moving the
    // column reads above the writes of other lanes' elements
    asm volatile("" ::: "memory");
    const int cl = lane / St::LPC, e = (lane % St::LPC) * St::VEC;
    const unsigned bytes = (unsigned)((((i64)SW - 1) * ld + NB) * (i64)sizeof(Ts));
    // G is the same in every lane of the wave; say so, so the descriptor is built in SGPRs
    const unsigned long long gb = reinterpret_cast<unsigned long long>(G);
    const unsigned glo = __builtin_amdgcn_readfirstlane((unsigned)gb), ghi = __builtin_amdgcn_readfirstlane((unsigned)(gb >> 32));
    Ts *Gu = reinterpret_cast<Ts *>(((unsigned long long)ghi << 32) | glo);
    const __amdgpu_buffer_rsrc_t rsrc =
        __builtin_amdgcn_make_buffer_rsrc(Gu, 0, __builtin_amdgcn_readfirstlane(bytes), 0x00020000);
#pragma unroll
    for (int c0 = 0; c0 < SW; c0 += St::CPI) {
        const int c = c0 + cl;
        const Ts *p = stg + c * LDS_S + e;
        if (e < rmin) continue;  // rows above rmin are left as they are in memory
        tvec v;
#pragma unroll
        for (int q = 0; q < St::VEC; ++q) v[q] = p[q];
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, v), rsrc,
                                               (int)(((i64)c * ld + e) * (i64)sizeof(Ts)), 0, /*sc1*/ 16);
    }
}

// Reflector block prefetch: columns j0..j0+IB-1 of a panel tile (Vg = first of them) and
// its T block go global -> registers (vload, issued one block ahead) -> LDS buffer (vstore).
// GE: unit lower part of a geqrt tile (1 on the diagonal, 0 above); TS: full columns.
template <typename T, int NB> struct VPre {
    static constexpr int NV = NB * IB / NTH, NT = IB * IB / NTH;
    T v[NV], t[NT];
};
// Element e of vload below: e < NV a V element, then the T elements.
template <typename T, int NB>
__device__ __forceinline__ void vload_one(VPre<T, NB> &P, const T *Vg, i64 ldv, const T *Tg, i64 ldt, int e) {
    constexpr int NV = VPre<T, NB>::NV;
    if (e < NV) {
        const int idx = threadIdx.x + e * NTH, p = idx % NB, j = idx / NB;
        P.v[e] = cld(&Vg[p + (i64)j * ldv]);
    } else {
        const int idx = threadIdx.x + (e - NV) * NTH, q = idx % IB, j = idx / IB;
        P.t[e - NV] = cld(&Tg[q + (i64)j * ldt]);
    }
}
template <typename T, int NB>
__device__ __forceinline__ void vload(VPre<T, NB> &P, const T *Vg, i64 ldv, const T *Tg, i64 ldt) {
#pragma unroll
    for (int e = 0; e < VPre<T, NB>::NV + VPre<T, NB>::NT; ++e) vload_one<T, NB>(P, Vg, ldv, Tg, ldt, e);
}
template <typename T, int NB, bool GE>
__device__ __forceinline__ void vstore(const VPre<T, NB> &P, int buf, int j0) {
    T *sm = lds_base<T>() + buf * Lds<T, NB>::BUF;
    constexpr int LDV = Lds<T, NB>::LDV;
    T *Vs = sm + Lds<T, NB>::V;
    T *Ts = sm + Lds<T, NB>::TT;
#pragma unroll
    for (int e = 0; e < VPre<T, NB>::NV; ++e) {
        const int idx = threadIdx.x + e * NTH, p = idx % NB, j = idx / NB;
        T v = P.v[e];
        if (GE) {
            const int ja = j0 + j;
            v = p < ja ? T(0) : (p == ja ? T(1) : v);
        }
        Vs[j * LDV + p] = v;
    }
#pragma unroll
    for (int e = 0; e < VPre<T, NB>::NT; ++e) {
        const int idx = threadIdx.x + e * NTH, q = idx % IB, j = idx / IB;
        Ts[q * (IB + 1) + j] = q <= j ? P.t[e] : T(0);
    }
}

// tsqrt V block (full columns j = 0..IB-1 of Vg) straight into LDS buffer `buf` with
// global_load_lds (no registers; completion is tracked by vmcnt like any vector load).  Each
// instruction moves 64 lanes x SZ bytes, a contiguous run of one column of Vs.  Column bases
// must be 16-byte aligned when SZ is 16 (the caller checks the leading dimension).
// Column jj (of this wave's IB / WAVES) of vload_lds below.
template <typename T, int NB>
__device__ __forceinline__ void vload_lds_col(const T *Vg, i64 ldv, int buf, int jj) {
    constexpr int LDV = Lds<T, NB>::LDV;
    constexpr int CB = NB * (int)sizeof(T);
    constexpr int SZ = (CB % 1024 == 0) ? 16 : 4;
    constexpr int CHUNK = 64 * SZ, NCH = CB / CHUNK;
    static_assert(CB % CHUNK == 0, "column bytes must be a multiple of the chunk");
    T *Vs = lds_base<T>() + buf * Lds<T, NB>::BUF + Lds<T, NB>::V;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int j = w + WAVES * jj;
    const char *src = reinterpret_cast<const char *>(Vg + (i64)j * ldv) + lane * SZ;
    char *dst = reinterpret_cast<char *>(Vs + j * LDV);
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        auto *d = (__attribute__((address_space(3))) void *)(dst + c * CHUNK);
        if constexpr (SZ == 16) __builtin_amdgcn_global_load_lds(src + c * CHUNK, d, 16, 0, /*sc1*/ 16);
        else __builtin_amdgcn_global_load_lds(src + c * CHUNK, d, 4, 0, /*sc1*/ 16);
    }
}
template <typename T, int NB>
__device__ __forceinline__ void vload_lds(const T *Vg, i64 ldv, int buf) {
#pragma unroll
    for (int jj = 0; jj < IB / WAVES; ++jj) vload_lds_col<T, NB>(Vg, ldv, buf, jj);
}
template <typename T, int NB> struct TPre {
    static constexpr int NT = IB * IB / NTH;
    T t[NT];
};
template <typename T, int NB> __device__ __forceinline__ void tload_one(TPre<T, NB> &P, const T *Tg, i64 ldt, int e) {
    const int idx = threadIdx.x + e * NTH, q = idx % IB, j = idx / IB;
    P.t[e] = cld(&Tg[q + (i64)j * ldt]);
}
template <typename T, int NB> __device__ __forceinline__ void tload(TPre<T, NB> &P, const T *Tg, i64 ldt) {
#pragma unroll
    for (int e = 0; e < TPre<T, NB>::NT; ++e) tload_one<T, NB>(P, Tg, ldt, e);
}
template <typename T, int NB> __device__ __forceinline__ void tstore(const TPre<T, NB> &P, int buf) {
    T *Ts = lds_base<T>() + buf * Lds<T, NB>::BUF + Lds<T, NB>::TT;
#pragma unroll
    for (int e = 0; e < TPre<T, NB>::NT; ++e) {
        const int idx = threadIdx.x + e * NTH, q = idx % IB, j = idx / IB;
        Ts[q * (IB + 1) + j] = q <= j ? P.t[e] : T(0);
    }
}

// op(H) strip with H = I - V T V^T of the staged block (one wave, registers only):
//   W0 = V^T S (+ A1 rows),  W = op(T) W0,  A1 rows -= W,  S -= V W.
// notrans: op(T) = T (apply H); otherwise T^T (apply H^T: the reference's trans 'C').
struct NoIssue {
    __device__ void operator()(int) const {}
};
// issue(part): the caller's loads for the next block, one part per product pair of the first
// phase (parts 0, 1, ... from its third pair on), so that the load instructions go out spread
// over the time the matrix pipe is busy instead of in one burst that the memory path throttles.
template <typename T, int NB, bool GE, typename F = NoIssue>
__device__ __forceinline__ void apply_wave(WStrip<T, NB> &S, typename Acc<T>::t *a1, bool notrans, int buf = 0,
                                           F &&issue = F()) {
    const T *sm = lds_base<T>() + buf * Lds<T, NB>::BUF;
    constexpr int LDV = Lds<T, NB>::LDV;
    constexpr int RT = NB / 16;
    const T *Vs = sm + Lds<T, NB>::V;
    const T *Ts = sm + Lds<T, NB>::TT;
    const int lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
    // Every MFMA's LDS operand is read two products ahead of its use (the reads of one
    // product then overlap the previous two in the matrix pipe instead of waiting before it).
    typename Acc<T>::t w0[2] = {acc_zero<T>(), acc_zero<T>()};
    {
        auto vop = [&](int idx, int ot) {
            const int rt = idx >> 2, r = idx & 3;
            return Vs[(16 * ot + lr) * LDV + 16 * rt + acc_row<T>(lk, r)];
        };
        constexpr int N = RT * 4;
        T va[3][2];
#pragma unroll
        for (int d = 0; d < 2; ++d)
#pragma unroll
            for (int ot = 0; ot < 2; ++ot) va[d][ot] = vop(d, ot);
#pragma unroll
        for (int idx = 0; idx < N; ++idx) {
            if (idx + 2 < N) {
#pragma unroll
                for (int ot = 0; ot < 2; ++ot) va[(idx + 2) % 3][ot] = vop(idx + 2, ot);
            }
#pragma unroll
            for (int ot = 0; ot < 2; ++ot) mfma16(w0[ot], va[idx % 3][ot], S.a[idx >> 2][idx & 3]);
            if (idx >= 2) issue(idx - 2);
        }
    }
    if (!GE) {
#pragma unroll
        for (int ot = 0; ot < 2; ++ot)
#pragma unroll
            for (int r = 0; r < 4; ++r) w0[ot][r] += a1[ot][r];
    }
    typename Acc<T>::t wv[2] = {acc_zero<T>(), acc_zero<T>()};
    {
        auto top = [&](int idx, int ot) {
            const int q = 16 * (idx >> 2) + acc_row<T>(lk, idx & 3), j = 16 * ot + lr;
            return notrans ? Ts[j * (IB + 1) + q] : Ts[q * (IB + 1) + j];
        };
        T ta[3][2];
#pragma unroll
        for (int d = 0; d < 2; ++d)
#pragma unroll
            for (int ot = 0; ot < 2; ++ot) ta[d][ot] = top(d, ot);
#pragma unroll
        for (int idx = 0; idx < 8; ++idx) {
            if (idx + 2 < 8) {
#pragma unroll
                for (int ot = 0; ot < 2; ++ot) ta[(idx + 2) % 3][ot] = top(idx + 2, ot);
            }
#pragma unroll
            for (int ot = 0; ot < 2; ++ot) mfma16(wv[ot], ta[idx % 3][ot], w0[idx >> 2][idx & 3]);
        }
    }
    // S -= V W as S += V (-W): W negated once (eight values) rather than every V operand
    T wn[2][4];
#pragma unroll
    for (int tq = 0; tq < 2; ++tq)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            if (!GE) a1[tq][r] -= wv[tq][r];
            wn[tq][r] = -wv[tq][r];
        }
    {
        auto vop = [&](int idx) {
            const int tq = idx / (4 * RT), r = (idx / RT) & 3, rt = idx % RT;
            return Vs[(16 * tq + acc_row<T>(lk, r)) * LDV + 16 * rt + lr];
        };
        constexpr int N = 2 * 4 * RT;
        T va[3];
        va[0] = vop(0);
        va[1] = vop(1);
#pragma unroll
        for (int idx = 0; idx < N; ++idx) {
            if (idx + 2 < N) va[(idx + 2) % 3] = vop(idx + 2);
            mfma16(S.a[idx % RT], va[idx % 3], wn[idx / (4 * RT)][(idx / RT) & 3]);
        }
    }
}

// apply_wave (H^T only) for the wave's NS strips of the Float32-update walker: the same
// products, every LDS operand read once for the NS strips; A1: the NS strips' A1 rows.
template <typename T, int NB, bool GE, int NS, typename F = NoIssue>
__device__ __forceinline__ void apply_wave_ns(WStrip<T, NB> *S, WStrip<T, NB> *A1, int B, int buf, F &&issue = F()) {
    const T *sm = lds_base<T>() + buf * Lds<T, NB>::BUF;
    constexpr int LDV = Lds<T, NB>::LDV;
    constexpr int RT = NB / 16;
    const T *Vs = sm + Lds<T, NB>::V;
    const T *Ts = sm + Lds<T, NB>::TT;
    const int lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
    typename Acc<T>::t w0[NS][2];
#pragma unroll
    for (int st = 0; st < NS; ++st) w0[st][0] = w0[st][1] = acc_zero<T>();
    {
        auto vop = [&](int idx, int ot) {
            const int rt = idx >> 2, r = idx & 3;
            return Vs[(16 * ot + lr) * LDV + 16 * rt + acc_row<T>(lk, r)];
        };
        constexpr int N = RT * 4;
        T va[3][2];
#pragma unroll
        for (int d = 0; d < 2; ++d)
#pragma unroll
            for (int ot = 0; ot < 2; ++ot) va[d][ot] = vop(d, ot);
#pragma unroll
        for (int idx = 0; idx < N; ++idx) {
            if (idx + 2 < N) {
#pragma unroll
                for (int ot = 0; ot < 2; ++ot) va[(idx + 2) % 3][ot] = vop(idx + 2, ot);
            }
#pragma unroll
            for (int st = 0; st < NS; ++st)
#pragma unroll
                for (int ot = 0; ot < 2; ++ot) mfma16(w0[st][ot], va[idx % 3][ot], S[st].a[idx >> 2][idx & 3]);
            if (idx >= 2) issue(idx - 2);
        }
    }
    if (!GE) {
#pragma unroll
        for (int st = 0; st < NS; ++st)
#pragma unroll
            for (int ot = 0; ot < 2; ++ot)
#pragma unroll
                for (int r = 0; r < 4; ++r) w0[st][ot][r] += A1[st].a[2 * B + ot][r];
    }
    typename Acc<T>::t wv[NS][2];
#pragma unroll
    for (int st = 0; st < NS; ++st) wv[st][0] = wv[st][1] = acc_zero<T>();
    {
        auto top = [&](int idx, int ot) {
            const int q = 16 * (idx >> 2) + acc_row<T>(lk, idx & 3), j = 16 * ot + lr;
            return Ts[q * (IB + 1) + j];
        };
        T ta[3][2];
#pragma unroll
        for (int d = 0; d < 2; ++d)
#pragma unroll
            for (int ot = 0; ot < 2; ++ot) ta[d][ot] = top(d, ot);
#pragma unroll
        for (int idx = 0; idx < 8; ++idx) {
            if (idx + 2 < 8) {
#pragma unroll
                for (int ot = 0; ot < 2; ++ot) ta[(idx + 2) % 3][ot] = top(idx + 2, ot);
            }
#pragma unroll
            for (int st = 0; st < NS; ++st)
#pragma unroll
                for (int ot = 0; ot < 2; ++ot) mfma16(wv[st][ot], ta[idx % 3][ot], w0[st][idx >> 2][idx & 3]);
        }
    }
    T wn[NS][2][4];
#pragma unroll
    for (int st = 0; st < NS; ++st)
#pragma unroll
        for (int tq = 0; tq < 2; ++tq)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                if (!GE) A1[st].a[2 * B + tq][r] -= wv[st][tq][r];
                wn[st][tq][r] = -wv[st][tq][r];
            }
    {
        auto vop = [&](int idx) {
            const int tq = idx / (4 * RT), r = (idx / RT) & 3, rt = idx % RT;
            return Vs[(16 * tq + acc_row<T>(lk, r)) * LDV + 16 * rt + lr];
        };
        constexpr int N = 2 * 4 * RT;
        T va[3];
        va[0] = vop(0);
        va[1] = vop(1);
#pragma unroll
        for (int idx = 0; idx < N; ++idx) {
            if (idx + 2 < N) va[(idx + 2) % 3] = vop(idx + 2);
#pragma unroll
            for (int st = 0; st < NS; ++st) mfma16(S[st].a[idx % RT], va[idx % 3], wn[st][idx / (4 * RT)][(idx / RT) & 3]);
        }
    }
}

// ---------------------------------------------------------------------------------
// Block factorization by the whole workgroup.  Wave w owns the eight columns j = 8w .. 8w+7
// of the IB-column block: lane l holds column 8w + l / 8 on the RPT = NB/8 consecutive rows
// from row0 = RPT * (l % 8), in registers, so the eight lanes holding one column are
// neighbours in one wave and its reductions are three DPP exchanges (every lane of the column
// gets the same sum).  Per column c one workgroup barrier:
//   the wave owning column c forms its norm below the pivot, beta, tau and the scale
//     (identically in its lanes), scales the column into the reflector v (unit pivot, zeros
//     above for geqrt) and stores v, tau (and beta into R for tsqrt) in LDS;
//   after the barrier every wave forms the dots of v with its columns (v read on its own
//     rows), updates each column j > c with w_j = tau (v^T x_j + R[c][j]) and records
//     z_j = v_j^T v for j < c.
// The next column's owner may start while other waves still update: what it writes for
// column c + 1 (the reflector, tau, R's diagonal entry) is not read for column c.
// The T factor (larft forward/columnwise: T[0:c, c] = -tau_c T[0:c, 0:c] z) is formed after
// the loop by wave 0, one row per lane, while the other waves return to the caller; the
// caller synchronises before it reads Ts.
// Sum over each aligned group of eight lanes, the same bits in all eight: quad_perm [1,0,3,2]
// (xor 1), quad_perm [2,3,0,1] (xor 2), then row_half_mirror (lane i <-> 7 - i of its eight,
// which pairs each quad with the other); every pairing is of two equal-shaped sums, so the
// additions are commutative images of each other.
template <int CTRL, typename T> __device__ __forceinline__ T dpp_mov(T v) {
    static_assert(sizeof(T) % 4 == 0, "whole dwords");
    struct W { int d[sizeof(T) / 4]; };
    W a = __builtin_bit_cast(W, v);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); ++i) a.d[i] = __builtin_amdgcn_mov_dpp(a.d[i], CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(T, a);
}
template <typename T> __device__ __forceinline__ T col_sum(T v) {
    v += dpp_mov<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp_mov<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dpp_mov<0x141>(v);  // row_half_mirror
    return v;
}

template <typename T, int NB, bool GE>
__device__ __forceinline__ void block_factor(int c0, T *Ablk, i64 lda) {
    T *sm = lds_base<T>();
    constexpr int LDV = Lds<T, NB>::LDV, RPT = NB / 8;
    static_assert(RPT % 4 == 0, "rows per thread must be a multiple of 4");
    static_assert(IB == 8 * WAVES, "eight columns per wave");
    T *Vs = sm + Lds<T, NB>::V;
    T *Ts = sm + Lds<T, NB>::TT;
    T *Rs = sm + Lds<T, NB>::RS;
    T *Zs = sm + Lds<T, NB>::ZB;      // Zs[c * IB + q] = z_q of column c (q < c)
    T *taus = sm + Lds<T, NB>::TAU;   // [IB]
    const int t = threadIdx.x, lane = t & 63;
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);
    const int j = 8 * w + (lane >> 3), grp = lane & 7, row0 = grp * RPT;
    T x[RPT];
#pragma unroll
    for (int r = 0; r < RPT; ++r) x[r] = Vs[j * LDV + row0 + r];
    __syncthreads();  // every column is in registers before Vs is overwritten
#pragma unroll 1
    for (int c = 0; c < IB; ++c) {
        const int piv = c0 + c;
        if (w == (c >> 3)) {  // the wave owning column c (uniform branch)
            const bool own = j == c;
            T q0 = 0, q1 = 0, q2 = 0, q3 = 0, ap = 0;
            if (own) {
#pragma unroll
                for (int r = 0; r < RPT; r += 4) {
                    const T y0 = (!GE || row0 + r > piv) ? x[r] : T(0);
                    const T y1 = (!GE || row0 + r + 1 > piv) ? x[r + 1] : T(0);
                    const T y2 = (!GE || row0 + r + 2 > piv) ? x[r + 2] : T(0);
                    const T y3 = (!GE || row0 + r + 3 > piv) ? x[r + 3] : T(0);
                    q0 = fma(y0, y0, q0);
                    q1 = fma(y1, y1, q1);
                    q2 = fma(y2, y2, q2);
                    q3 = fma(y3, y3, q3);
                }
                if (GE && piv >= row0 && piv < row0 + RPT) {
#pragma unroll
                    for (int r = 0; r < RPT; ++r) ap = (row0 + r == piv) ? x[r] : ap;
                }
            }
            const T ssq = col_sum((q0 + q1) + (q2 + q3));
            const T alpha = GE ? col_sum(ap) : Rs[c * (IB + 1) + c];
            T beta = alpha, tau = 0, scale = 1;
            if (ssq != T(0)) {
                beta = -copysign(sqrt(alpha * alpha + ssq), alpha);
                tau = (beta - alpha) / beta;
                scale = T(1) / (alpha - beta);
            }
            if (own) {
#pragma unroll
                for (int r = 0; r < RPT; ++r) {
                    const int p = row0 + r;
                    const bool below = !GE || p > piv, pv = GE && p == piv;
                    const T v = x[r] * scale;
                    Vs[c * LDV + p] = below ? v : (pv ? T(1) : T(0));
                    x[r] = below ? v : (pv ? beta : x[r]);
                }
                if (grp == 0) {
                    taus[c] = tau;
                    if (!GE) Rs[c * (IB + 1) + c] = beta;
                }
            }
        }
        __syncthreads();
        // ---- dots with every column, updates
        T v[RPT];
        T q0 = 0, q1 = 0, q2 = 0, q3 = 0;
#pragma unroll
        for (int r = 0; r < RPT; r += 4) {
            v[r] = Vs[c * LDV + row0 + r];
            v[r + 1] = Vs[c * LDV + row0 + r + 1];
            v[r + 2] = Vs[c * LDV + row0 + r + 2];
            v[r + 3] = Vs[c * LDV + row0 + r + 3];
            q0 = fma(v[r], x[r], q0);
            q1 = fma(v[r + 1], x[r + 1], q1);
            q2 = fma(v[r + 2], x[r + 2], q2);
            q3 = fma(v[r + 3], x[r + 3], q3);
        }
        const T gsum = col_sum((q0 + q1) + (q2 + q3));
        if (j > c) {
            const T rc = GE ? T(0) : Rs[c * (IB + 1) + j];
            const T cw = taus[c] * (gsum + rc);
#pragma unroll
            for (int r = 0; r < RPT; ++r) x[r] = fma(-cw, v[r], x[r]);
            if (!GE && grp == 0) Rs[c * (IB + 1) + j] = rc - cw;
        } else if (j < c && grp == 0) {
            Zs[c * IB + j] = gsum;
        }
    }
    if (GE) {
        // Vs holds the unit-lower reflectors for the trailing update; the tile gets R, beta
        // and V (rows >= c0; the rows above are untouched) from the registers.
#pragma unroll
        for (int r = 0; r < RPT; ++r)
            if (row0 + r >= c0) cst(&Ablk[(row0 + r) + (i64)j * lda], x[r]);
    }
    __syncthreads();
    if (w == 0) {
        const int i = lane;
        T trow[IB];
#pragma unroll
        for (int c = 0; c < IB; ++c) {
            T acc = 0;
#pragma unroll
            for (int q = 0; q < c; ++q) acc += trow[q] * Zs[c * IB + q];
            const T tc = taus[c];
            trow[c] = (i < c) ? -tc * acc : (i == c ? tc : T(0));
        }
        if (i < IB) {
#pragma unroll
            for (int c = 0; c < IB; ++c) Ts[i * (IB + 1) + c] = trow[c];
        }
    }
}

// ---------------------------------------------------------------------------------
// Tile operations (one workgroup).

// One ib-block of geqrt!(NB, NB, IB, A, T) (src/geqrt.jl:44-67): geqr2+larft of columns
// c0..c0+IB-1, then larfb!('L','C') of that block on the trailing columns of the tile.
template <typename T, int NB> __device__ __forceinline__ void geqrt_block(T *A, i64 lda, T *Tm, i64 ldt, int c0) {
    T *sm = lds_base<T>();
    constexpr int LDV = Lds<T, NB>::LDV;
    T *Vs = sm + Lds<T, NB>::V;
    T *Ts = sm + Lds<T, NB>::TT;
    const int t = threadIdx.x, w = t >> 6;
    __syncthreads();
    for (int idx = t; idx < NB * IB; idx += NTH) {
        const int p = idx % NB, j = idx / NB;
        Vs[j * LDV + p] = cld(&A[p + (i64)(c0 + j) * lda]);
    }
    __syncthreads();
    block_factor<T, NB, true>(c0, A + (i64)c0 * lda, lda);
    __syncthreads();
    for (int idx = t; idx < IB * IB; idx += NTH) {
        const int q = idx % IB, j = idx / IB;
        if (q <= j) cst(&Tm[q + (i64)(c0 + j) * ldt], Ts[q * (IB + 1) + j]);
    }
    // trailing strips through the wave's staging columns (free once T is formed), as in
    // tsqrt_block below
    const int nst = (NB - c0 - IB) / SW;
    T *stg = stage_ptr<T, T, NB>();
    for (int st = w; st < nst; st += WAVES) {
        T *G = A + (i64)(c0 + IB + st * SW) * lda;
        stage_load<T, NB>(stg, G, lda);
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        WStrip<T, NB> S;
        stage_to_regs<T, T, NB>(stg, S);
        apply_wave<T, NB, true>(S, nullptr, false);
        // rows above c0 are unchanged by these reflectors (zero there) and may be in flight
        // in another block stage (R rows of earlier blocks updated by tsqrt): store only rows
        // >= c0 (the rows above may have been read stale; they do not enter the products)
        stage_store<T, T, NB>(stg, S, G, lda, c0);
    }
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
}

// One ib-block of tsqrt!(NB, NB, IB, R, A2, T) (src/tsqrt.jl:146-155): the TS block QR of
// columns c0..c0+IB-1 of [R; A2], then tsmqr!('L','C') of that block on the trailing columns
// (A1 = rows c0..c0+IB-1 of R, A2).
template <typename T, int NB>
__device__ __forceinline__ void tsqrt_block(T *R, i64 ldr, T *A2, i64 lda2, T *Tm, i64 ldt, int c0) {
    T *sm = lds_base<T>();
    constexpr int LDV = Lds<T, NB>::LDV;
    T *Vs = sm + Lds<T, NB>::V;
    T *Ts = sm + Lds<T, NB>::TT;
    T *Rs = sm + Lds<T, NB>::RS;
    const int t = threadIdx.x, w = t >> 6;
    __syncthreads();
    for (int idx = t; idx < IB * IB; idx += NTH) {
        const int i = idx % IB, j = idx / IB;
        Rs[i * (IB + 1) + j] = cld(&R[(c0 + i) + (i64)(c0 + j) * ldr]);
    }
    for (int idx = t; idx < NB * IB; idx += NTH) {
        const int p = idx % NB, j = idx / NB;
        Vs[j * LDV + p] = cld(&A2[p + (i64)(c0 + j) * lda2]);
    }
    __syncthreads();
    block_factor<T, NB, false>(c0, nullptr, 0);
    for (int idx = t; idx < NB * IB; idx += NTH) {
        const int p = idx % NB, j = idx / NB;
        cst(&A2[p + (i64)(c0 + j) * lda2], Vs[j * LDV + p]);
    }
    for (int idx = t; idx < IB * IB; idx += NTH) {
        const int i = idx % IB, j = idx / IB;
        if (i <= j) cst(&R[(c0 + i) + (i64)(c0 + j) * ldr], Rs[i * (IB + 1) + j]);
    }
    __syncthreads();
    for (int idx = t; idx < IB * IB; idx += NTH) {
        const int i = idx % IB, j = idx / IB;
        if (i <= j) cst(&Tm[i + (i64)(c0 + j) * ldt], Ts[i * (IB + 1) + j]);
    }
    // The trailing columns of A2 move through the wave's strip staging columns (whole-column
    // LDS-DMA loads, column-contiguous stores), which overlay the R / scratch regions, free
    // once the block's R and T are stored.  No other block stage touches these columns of
    // A2 meanwhile, so whole columns are stored back.
    const int nst = (NB - c0 - IB) / SW;
    T *stg = stage_ptr<T, T, NB>();
    for (int st = w; st < nst; st += WAVES) {
        const i64 col = c0 + IB + st * SW;
        stage_load<T, NB>(stg, A2 + col * lda2, lda2);
        typename Acc<T>::t a1[2];
        tiles_load<T, 2>(a1, R + col * ldr, ldr, c0);
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        WStrip<T, NB> S;
        stage_to_regs<T, T, NB>(stg, S);
        apply_wave<T, NB, false>(S, a1, false);
        stage_store<T, T, NB>(stg, S, A2 + col * lda2, lda2);
        tiles_store<T, 2>(a1, R + col * ldr, ldr, c0);
    }
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
}

// ---------------------------------------------------------------------------------
struct DagArgs {
    int mt, nt, nwalk;        // tile rows/cols, TW-column walker groups of the matrix
    int ntasks;
    const int *tasks;
    int *counter;             // next task
    int *panel_prog;          // [kt][NB/IB]: tiles of column k done by block stage b
    int *strip_prog;          // [kt][nwalk]
    int *err;
    void *C;                    // matrix the walkers update (A itself when factoring)
    i64 ldc;
    int qmode;                  // walkers apply Q (reverse order) instead of Q^T
    // Reduced-precision factorizations only (null otherwise): per (panel tile, block) the
    // V/T image the panel writes and the walkers read, and for the 16-bit walkers the
    // trailing matrix in Float32 between steps (same positions and leading dimension).
    unsigned char *lpimg;
    float *shadow;
};

// 16-bit walker support, defined with that walker below
template <typename Tw, int NB> struct LdsLp;
template <typename T, typename Tw, int NB> __device__ void write_lp_image(unsigned char *dst);
__host__ __device__ inline i64 lp_tile_index(int k, int i, int mt);

// Block stage b of panel k: ib-block b of geqrt(A_kk), then of tsqrt(A_kk, A_ik) for
// i = k+1.., tile by tile.  Block b of tile i needs block b of tile i-1 (rows c0.. of R_kk,
// this task's own previous step) and block b-1 of tile i (the trailing update of A_ik's
// columns c0.. and the factorization of R_kk's rows above), so the stages of one panel
// run as a pipeline over the tiles.  Stage 0 waits for the previous step's walkers instead.
// Tw: the walkers' type; a Tw narrower than T makes the panel also write each block's image
// for them (16-bit: write_lp_image; Float32 with a Float64 matrix: write_f32_image).  The
// Float64 and Float32 factorizations share one instance (PanelTw maps their Tw to T): that
// out-of-line panel compiles as it did before the image-reading variants existed.
template <typename T, typename Tw> struct PanelTwT { using type = T; };
template <typename T> struct PanelTwT<T, _Float16> { using type = _Float16; };
template <typename T> struct PanelTwT<T, __bf16> { using type = __bf16; };
template <typename T> struct PanelTwT<T, bf16x3> { using type = bf16x3; };
template <> struct PanelTwT<double, float> { using type = float; };
template <typename T, typename Tw> using PanelTw = typename PanelTwT<T, Tw>::type;
// Panel side of F32Img: the block just factored (V, T of the panel's LDS, V unit lower for a
// geqrt block) rounded to Float32 and written to dst with device-coherent stores.
template <typename T, int NB> __device__ void write_f32_image(unsigned char *dst) {
    using L = Lds<float, NB>;
    constexpr int LDV = Lds<T, NB>::LDV;
    static_assert(Lds<T, NB>::LDV == L::LDV, "one V column stride");
    const T *Vs = lds_base<T>() + Lds<T, NB>::V;
    const T *Ts = lds_base<T>() + Lds<T, NB>::TT;
    float *d = reinterpret_cast<float *>(dst);
    for (int e = threadIdx.x; e < L::BUF; e += NTH) {
        float v = 0.0f;
        if (e < L::TT) {
            const int j = e / LDV, p = e % LDV;
            if (p < NB) v = (float)Vs[j * LDV + p];
        } else {
            const int f = e - L::TT, q = f / (IB + 1), j = f % (IB + 1);
            if (j < IB && q <= j) v = (float)Ts[q * (IB + 1) + j];
        }
        cst(&d[e], v);
    }
}
template <typename T, int NB, typename Tw>
__device__ __noinline__ void run_panel(int k, int b, T *A, i64 lda, T *Tm, i64 ldt, DagArgs g, int *sh_flag) {
    // walker groups per tile column (NB / the kernel's walk_width; from the launch, so that the
    // Float64 and Float32-update kernels share this panel), stages
    const int WPT = g.nwalk / g.nt;
    constexpr int NBLK = NB / IB;
    const int c0 = b * IB;
    int *mine = &g.panel_prog[k * NBLK + b];
    const int *prev = b > 0 ? &g.panel_prog[k * NBLK + b - 1] : nullptr;
    T *Akk = A + (i64)k * NB + (i64)k * NB * lda;
    for (int i = k; i < g.mt; ++i) {
        if (prev) {
            wait_ge(prev, i + 1, g.err, sh_flag);
        } else if (k > 0) {
            for (int s = k * WPT; s < (k + 1) * WPT; ++s)
                wait_ge(&g.strip_prog[(k - 1) * g.nwalk + s], i + 1, g.err, sh_flag);
        }
        T *Tik = Tm + (i64)i * IB + (i64)k * NB * ldt;
        if (i == k) geqrt_block<T, NB>(Akk, lda, Tik, ldt, c0);
        else tsqrt_block<T, NB>(Akk, lda, A + (i64)i * NB + (i64)k * NB * lda, lda, Tik, ldt, c0);
        // the block's V and T for the reduced-precision walkers
        if constexpr (sizeof(Tw) == 2)
            write_lp_image<T, Tw, NB>(g.lpimg + (lp_tile_index(k, i, g.mt) * NBLK + b) * (i64)LdsLp<Tw, NB>::IMG_BYTES);
        else if constexpr (sizeof(Tw) < sizeof(T))
            write_f32_image<T, NB>(g.lpimg + (lp_tile_index(k, i, g.mt) * NBLK + b) * (i64)F32Img<NB>::STRIDE);
        publish(mine, i + 1);
    }
}

// Non-blocking: has panel stage counter *p reached v?  (workgroup-uniform result)
__device__ __forceinline__ bool ready_ge(const int *p, int v, int *sh) {
    if (threadIdx.x == 0) *sh = ld_flag(p) >= v;
    __syncthreads();
    return *sh != 0;
}

// The walker: unmqr('L','C') of A_kj's strip with geqrt(A_kk)'s blocks, then tsmqr('L','C')
// of (A_kj, A_ij) for every tile row i > k, A_kj's strip (A1) in MFMA accumulators
// throughout.  The reflector blocks are a pipeline over two LDS buffers: block B+1 is loaded
// (V straight into LDS, T through registers) while block B is applied; block 0 of the next
// tile row is fetched the same way when its panel tile is already done (a walker never
// waits for the next tile before finishing the current one).
template <typename T, int NB>
__device__ __forceinline__ void run_walk(int k, int s, T *A, i64 lda, T *Tm, i64 ldt, DagArgs g, int *sh_flag) {
    constexpr int NBLK = NB / IB;
    static_assert(NBLK % 2 == 0, "block count must be even for the buffer parity");
    const int w = threadIdx.x >> 6;
    const i64 col = (i64)s * TW + w * SW;  // this wave's first matrix column
    const int *prev = k > 0 ? &g.strip_prog[(k - 1) * g.nwalk + s] : nullptr;
    int *mine = &g.strip_prog[k * g.nwalk + s];
    const int *pdone = &g.panel_prog[k * NBLK + NBLK - 1];  // last block stage of panel k
    const T *Akk = A + (i64)k * NB + (i64)k * NB * lda;
    const T *Tkk = Tm + (i64)k * IB + (i64)k * NB * ldt;
    T *Cm = reinterpret_cast<T *>(g.C);
    const i64 ldc = g.ldc;
    T *Akj = Cm + (i64)k * NB + col * ldc;
    T *stg = stage_ptr<T, T, NB>();
    wait_ge(pdone, k + 1, g.err, sh_flag);
    if (prev) wait_ge(prev, k + 1, g.err, sh_flag);
    WStrip<T, NB> A1;
    tiles_load<T, NB / 16>(A1.a, Akj, ldc, 0);
    // unmqr!('L','C'): geqrt(A_kk)'s blocks in forward order (unit-lower V through registers)
    {
        VPre<T, NB> P;
        vload<T, NB>(P, Akk, lda, Tkk, ldt);
        vstore<T, NB, true>(P, 0, 0);
#pragma unroll
        for (int B = 0; B < NBLK; ++B) {
            if (B + 1 < NBLK) vload<T, NB>(P, Akk + (i64)(B + 1) * IB * lda, lda, Tkk + (i64)(B + 1) * IB * ldt, ldt);
            __syncthreads();
            apply_wave<T, NB, true>(A1, nullptr, false, B & 1);
            if (B + 1 < NBLK) vstore<T, NB, true>(P, (B + 1) & 1, (B + 1) * IB);
        }
    }
    // tsmqr!('L','C') against every tile below.  Row i's strip is stored at the end of its
    // last block and published at the first wait of row i+1 (whose strip and reflector loads
    // are then in flight together with those stores); the next strip is loaded early when the
    // previous step has already released it.
    TPre<T, NB> P;
    WStrip<T, NB> S;
    bool have0 = false;  // block 0 of the next tile row is already in (or on its way to) buffer 0
    bool haveS = false;  // the next strip is already on its way to S
    int pending = 0;     // a finished row whose stores are in flight, not yet published (row + 1)
    // Thread 0 reads the two flags that decide the next row's prefetches (block 0 of the next
    // panel tile, the next strip) a block before the decisions and shares the outcomes at
    // the last block's barrier, so their round trips stay off the row's critical path.  An
    // early read can only say "not yet", and the next row then waits for it.
    __shared__ int sh_nx[2];
    int fl_p = 0, fl_s = 0;
    for (int i = k + 1; i < g.mt; ++i) {
        const T *Aik = A + (i64)i * NB + (i64)k * NB * lda;
        const T *Tik = Tm + (i64)i * IB + (i64)k * NB * ldt;
        T *Aij = Cm + (i64)i * NB + col * ldc;
        if (pending && (!have0 || !haveS)) {  // about to block: release the finished row first
            publish(mine, pending);
            pending = 0;
        }
        if (!have0) {
            wait_ge(pdone, i + 1, g.err, sh_flag);
            vload_lds<T, NB>(Aik, lda, 0);
            tload<T, NB>(P, Tik, ldt);
            tstore<T, NB>(P, 0);
        }
        if (!haveS) {
            if (prev) wait_ge(prev, i + 1, g.err, sh_flag);
            stage_load<T, NB>(stg, Aij, ldc);
        }
        bool next0 = false;
#pragma unroll
        for (int B = 0; B < NBLK; ++B) {
            if (B == NBLK - 1 && threadIdx.x == 0) {
                sh_nx[0] = fl_p >= i + 2;
                sh_nx[1] = fl_s >= i + 2;
            }
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // this block's V, the strip, earlier stores
            __syncthreads();
            if (B == 0) {
                if (pending) {
                    if (threadIdx.x == 0) cst(mine, pending);
                    pending = 0;
                }
                stage_to_regs<T, T, NB>(stg, S);
            }
            if (B == NBLK - 2 && threadIdx.x == 0) {
                fl_p = i + 1 < g.mt ? ld_flag(pdone) : 0;
                fl_s = i + 1 < g.mt && prev ? ld_flag(prev) : 0;
            }
            // the next block's V (LDS DMA) and T, or block 0 of the next tile row
            bool load_next = B + 1 < NBLK;
            if (!load_next && i + 1 < g.mt) load_next = next0 = sh_nx[0] != 0;
            const T *Vn = B + 1 < NBLK ? Aik + (i64)(B + 1) * IB * lda : Aik + NB;
            const T *Tn = B + 1 < NBLK ? Tik + (i64)(B + 1) * IB * ldt : Tik + IB;
            static_assert(IB / WAVES + TPre<T, NB>::NT <= NB / 16 * 4 - 2, "every load part gets a product pair");
            apply_wave<T, NB, false>(S, &A1.a[2 * B], false, B & 1, [&](int part) {
                constexpr int NV = IB / WAVES, NTL = TPre<T, NB>::NT;
                if (load_next) {
                    if (part < NV) vload_lds_col<T, NB>(Vn, lda, (B + 1) & 1, part);
                    else if (part < NV + NTL) tload_one<T, NB>(P, Tn, ldt, part - NV);
                }
            });
            if (B + 1 < NBLK || next0) tstore<T, NB>(P, (B + 1) & 1);
        }
        have0 = next0;
        stage_store<T, T, NB>(stg, S, Aij, ldc);
        pending = i + 1;
        haveS = i + 1 < g.mt && (!prev || sh_nx[1] != 0);
        if (haveS) stage_load<T, NB>(stg, Aij + NB, ldc);
    }
    if (pending) publish(mine, pending);
    tiles_store<T, NB / 16>(A1.a, Akj, ldc, 0);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
}

// Reverse walker: Q C for the strip, step k of Q = Q_0 Q_1 ... (steps run k = kt-1 down to 0).
// Q_k = G_k TS_{k,k+1} ... TS_{k,mt-1}, so the rows go mt-1 down to k+1 (tsmqr('L','N'),
// blocks in reverse order) and then row k (unmqr('L','N')).  Row k is final input for step
// k at once; row i > k must be done by step k+1, whose counter counts rows finished from the
// bottom (its own row k+1 last).
template <typename T, int NB>
__device__ __forceinline__ void run_walk_q(int k, int s, T *A, i64 lda, T *Tm, i64 ldt, DagArgs g, int *sh_flag) {
    constexpr int NBLK = NB / IB;
    const int kt = g.mt < g.nt ? g.mt : g.nt;
    const int w = threadIdx.x >> 6;
    const i64 col = (i64)s * TW + w * SW;
    const int *prev = k + 1 < kt ? &g.strip_prog[(k + 1) * g.nwalk + s] : nullptr;
    int *mine = &g.strip_prog[k * g.nwalk + s];
    T *Cm = reinterpret_cast<T *>(g.C);
    const i64 ldc = g.ldc;
    T *Ckj = Cm + (i64)k * NB + col * ldc;
    WStrip<T, NB> A1;
    tiles_load<T, NB / 16>(A1.a, Ckj, ldc, 0);
    VPre<T, NB> P;
    for (int i = g.mt - 1; i > k; --i) {
        if (prev) wait_ge(prev, g.mt - i, g.err, sh_flag);
        const T *Aik = A + (i64)i * NB + (i64)k * NB * lda;
        const T *Tik = Tm + (i64)i * IB + (i64)k * NB * ldt;
        T *Cij = Cm + (i64)i * NB + col * ldc;
        WStrip<T, NB> S;
        tiles_load<T, NB / 16>(S.a, Cij, ldc, 0);
#pragma unroll
        for (int B = NBLK - 1; B >= 0; --B) {
            vload<T, NB>(P, Aik + (i64)B * IB * lda, lda, Tik + (i64)B * IB * ldt, ldt);
            __syncthreads();
            vstore<T, NB, false>(P, 0, 0);
            __syncthreads();
            apply_wave<T, NB, false>(S, &A1.a[2 * B], true, 0);
        }
        tiles_store<T, NB / 16>(S.a, Cij, ldc, 0);
        publish(mine, g.mt - i);
    }
    const T *Akk = A + (i64)k * NB + (i64)k * NB * lda;
    const T *Tkk = Tm + (i64)k * IB + (i64)k * NB * ldt;
#pragma unroll
    for (int B = NBLK - 1; B >= 0; --B) {
        vload<T, NB>(P, Akk + (i64)B * IB * lda, lda, Tkk + (i64)B * IB * ldt, ldt);
        __syncthreads();
        vstore<T, NB, true>(P, 0, B * IB);
        __syncthreads();
        apply_wave<T, NB, true>(A1, nullptr, true, 0);
    }
    tiles_store<T, NB / 16>(A1.a, Ckj, ldc, 0);
    publish(mine, g.mt - k);
}

// ---------------------------------------------------------------------------------
// Mixed-precision walkers: the panel (geqrt/tsqrt blocks and T) stays in T = Float64; the
// trailing update of the strips (unmqr/tsmqr) runs with Float32 accumulators (Float32, FP16
// or BF16 products), the strips kept in a Float32 shadow between steps.

// Float32 accumulator strips <-> memory without staging.  In the Float32-accumulator 16x16
// MFMA layout (v_mfma_f32_16x16x4f32, _16x16x16_f16 / _bf16) a lane holds rows
// 4(l/16)..4(l/16)+3 of column l%16 of each 16-row tile: 16 contiguous bytes of a
// column-major Float32 matrix (32 of a Float64 one), so a tile moves as one device-coherent
// 16-byte buffer access per lane (two for Float64).  G: the strip's first element (the same
// in every lane), ld its leading dimension; a 16-column strip of NB rows.
template <typename Ts, int NB> __device__ __forceinline__ __amdgpu_buffer_rsrc_t strip_rsrc(const Ts *G, i64 ld) {
    const unsigned long long gb = reinterpret_cast<unsigned long long>(G);
    const unsigned glo = __builtin_amdgcn_readfirstlane((unsigned)gb), ghi = __builtin_amdgcn_readfirstlane((unsigned)(gb >> 32));
    const unsigned bytes = (unsigned)((((i64)SW - 1) * ld + NB) * (i64)sizeof(Ts));
    return __builtin_amdgcn_make_buffer_rsrc(reinterpret_cast<void *>(((unsigned long long)ghi << 32) | glo), 0,
                                             __builtin_amdgcn_readfirstlane(bytes), 0x00020000);
}
template <typename Ts, int NB> __device__ __forceinline__ void strip_load(fx4 *a, const Ts *G, i64 ld) {
    typedef double f64x2 __attribute__((ext_vector_type(2)));
    const __amdgpu_buffer_rsrc_t rsrc = strip_rsrc<Ts, NB>(G, ld);
    const int lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int rt = 0; rt < NB / 16; ++rt) {
        const int off = (int)(((i64)lr * ld + 16 * rt + 4 * lk) * (i64)sizeof(Ts));
        if constexpr (sizeof(Ts) == 4) {
            a[rt] = __builtin_bit_cast(fx4, __builtin_amdgcn_raw_buffer_load_b128(rsrc, off, 0, /*sc1*/ 16));
        } else {
            const f64x2 d0 = __builtin_bit_cast(f64x2, __builtin_amdgcn_raw_buffer_load_b128(rsrc, off, 0, 16));
            const f64x2 d1 = __builtin_bit_cast(f64x2, __builtin_amdgcn_raw_buffer_load_b128(rsrc, off + 16, 0, 16));
            a[rt] = fx4{(float)d0[0], (float)d0[1], (float)d1[0], (float)d1[1]};
        }
    }
}
// AUX: the store's cache policy, device-coherent (sc1) with or without the non-temporal hint
// (measured: it helps the 16-bit walkers' Float32 shadow rows, 1-2 %, and costs the Float32
// walker's about as much).
constexpr int ST_SC1 = 16, ST_SC1_NT = 18;
template <typename Ts, int NB, int AUX = ST_SC1> __device__ __forceinline__ void strip_store(const fx4 *a, Ts *G, i64 ld) {
    typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
    typedef double f64x2 __attribute__((ext_vector_type(2)));
    const __amdgpu_buffer_rsrc_t rsrc = strip_rsrc<Ts, NB>(G, ld);
    const int lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int rt = 0; rt < NB / 16; ++rt) {
        const int off = (int)(((i64)lr * ld + 16 * rt + 4 * lk) * (i64)sizeof(Ts));
        if constexpr (sizeof(Ts) == 4) {
            __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, a[rt]), rsrc, off, 0, AUX);
        } else {
            const f64x2 d0{(double)a[rt][0], (double)a[rt][1]}, d1{(double)a[rt][2], (double)a[rt][3]};
            __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, d0), rsrc, off, 0, AUX);
            __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, d1), rsrc, off + 16, 0, AUX);
        }
    }
}

// The next row's Float32 strips of the reduced-precision walkers arrive by LDS-DMA (no
// registers held while the current row is applied).  A wave's NS * SW columns are staged at
// LDS byte offset BASE in groups of 256 / NB columns, one 1 KB DMA instruction each (64 lanes
// x 16 bytes), the groups 16 bytes apart against bank conflicts; inside the Float64 staging
// region of the kernel's layout.
template <int BASE, int NB, int NS> struct StripStage {
    static constexpr int CPD = 256 / NB, LPC = NB / 4, GRP = 1024 + 16, NG = NS * SW / CPD;
    static_assert(NS * SW % CPD == 0, "whole groups per wave");
    __device__ static unsigned char *base() {
        return lds_base<unsigned char>() + BASE + (threadIdx.x >> 6) * NG * GRP;
    }
    // issue the loads of the wave's columns G, G + ld, ... (rows 0..NB-1)
    __device__ static void load(const float *G, i64 ld) {
        const int lane = threadIdx.x & 63;
        unsigned char *b = base();
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");  // earlier reads of the staging are done
#pragma unroll
        for (int gi = 0; gi < NG; ++gi) {
            const float *src = G + (i64)(gi * CPD + lane / LPC) * ld + (lane % LPC) * 4;
            auto *d = (__attribute__((address_space(3))) void *)(b + gi * GRP);
            __builtin_amdgcn_global_load_lds(src, d, 16, 0, /*sc1*/ 16);
        }
    }
    // staged columns -> the wave's NS accumulator strips (after the loads' vmcnt wait)
    __device__ static void read(WStrip<float, NB> *S) {
        const int lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
        const unsigned char *b = base();
#pragma unroll
        for (int q = 0; q < NS; ++q) {
            const int c = q * SW + lr;
            const unsigned char *cb = b + (c / CPD) * GRP + (c % CPD) * NB * 4 + 16 * lk;
#pragma unroll
            for (int rt = 0; rt < NB / 16; ++rt) S[q].a[rt] = *reinterpret_cast<const fx4 *>(cb + 64 * rt);
        }
    }
};

// An image of BYTES (a multiple of 16) from global src into LDS dst with 1 KB LDS-DMA
// instructions (waves take every fourth KB; the last one only its lanes inside BYTES);
// completion is tracked by vmcnt like any vector load.
template <int BYTES> __device__ __forceinline__ void img_dma(const unsigned char *src, unsigned char *dst) {
    constexpr int NI = (BYTES + 1023) / 1024;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int c = 0; c < NI; c += WAVES) {
        const int ci = c + w;
        if (ci < NI && ci * 1024 + lane * 16 < BYTES) {
            auto *d = (__attribute__((address_space(3))) void *)(dst + ci * 1024);
            __builtin_amdgcn_global_load_lds(src + ci * 1024 + lane * 16, d, 16, 0, /*sc1*/ 16);
        }
    }
}

// The Float32-update walker (Tw = Float32 over a Float64 matrix): the row pipeline of the
// 16-bit walker with v_mfma_f32_16x16x4 products: each block's V/T image (F32Img, written by
// the panel) copied into an LDS buffer by DMA one block ahead, two 16-column strips per wave.
template <typename T, typename Tw, int NB>
__device__ __forceinline__ void run_walk_mixed(int k, int s, T * /*A*/, i64 /*lda*/, T * /*Tm*/, i64 /*ldt*/, DagArgs g,
                                               int *sh_flag) {
    constexpr int NBLK = NB / IB;
    static_assert(NBLK % 2 == 0, "block count must be even for the buffer parity");
    static_assert(sizeof(Tw) == 4, "Float32 accumulator strips");
    __shared__ int sh_ready;
    constexpr int TWK = walk_width<T, Tw, NB>(), NS = TWK / TW;  // columns per task, strips per wave
    const int w = threadIdx.x >> 6;
    const i64 col = (i64)s * TWK + w * NS * SW;  // the wave's strip q starts at column col + 16 q
    const int *prev = k > 0 ? &g.strip_prog[(k - 1) * g.nwalk + s] : nullptr;
    int *mine = &g.strip_prog[k * g.nwalk + s];
    const int *pdone = &g.panel_prog[k * NBLK + NBLK - 1];
    T *Cm = reinterpret_cast<T *>(g.C);
    const i64 ldc = g.ldc;
    // Strips between steps live in the Float32 shadow, as in the 16-bit walker: step 0 reads
    // the matrix (16-byte loads at the row start), later steps the shadow (prefetched through
    // StripStage, behind the two image buffers); rows go to the matrix only for the next
    // panel's tile column and the final R row (A1).  Every row leaves straight from the
    // registers.
    using Stg = StripStage<Lds<Tw, NB>::SSTG * (int)sizeof(Tw), NB, NS>;
    static_assert(2 * Lds<Tw, NB>::BUF <= Lds<Tw, NB>::SSTG, "the image buffers end before the staging");
    static_assert(WAVES * Stg::NG * Stg::GRP <= TW * Lds<T, NB>::LDS_S * (int)sizeof(T), "staging fits the strip region");
    float *const sh32 = g.shadow;
    const bool src32 = k > 0, dst64 = s / (NB / TWK) == k + 1;
    auto load_strips = [&](WStrip<Tw, NB> *X, int row_tile) {
#pragma unroll
        for (int q = 0; q < NS; ++q) {
            const i64 off = (i64)row_tile * NB + (col + q * SW) * ldc;
            if (src32) strip_load<float, NB>(X[q].a, sh32 + off, ldc);
            else strip_load<T, NB>(X[q].a, Cm + off, ldc);
        }
    };
    auto store_strips = [&](const WStrip<Tw, NB> *X, int row_tile, bool f64) {
#pragma unroll
        for (int q = 0; q < NS; ++q) {
            const i64 off = (i64)row_tile * NB + (col + q * SW) * ldc;
            if (f64) strip_store<T, NB>(X[q].a, Cm + off, ldc);
            else strip_store<float, NB>(X[q].a, sh32 + off, ldc);
        }
    };
    unsigned char *const imgs = g.lpimg;
    auto img_load = [&](int i, int B, int buf) {
        img_dma<F32Img<NB>::BYTES>(imgs + (lp_tile_index(k, i, g.mt) * NBLK + B) * (i64)F32Img<NB>::STRIDE,
                                   lds_base<unsigned char>() + buf * Lds<Tw, NB>::BUF * (int)sizeof(Tw));
    };
    wait_ge(pdone, k + 1, g.err, sh_flag);
    if (prev) wait_ge(prev, k + 1, g.err, sh_flag);
    WStrip<Tw, NB> A1[NS];
    load_strips(A1, k);
    // unmqr!('L','C') with geqrt(A_kk)'s blocks (their images carry the unit-lower V)
    img_load(k, 0, 0);
    bool have0 = false;
#pragma unroll
    for (int B = 0; B < NBLK; ++B) {
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
        if (B + 1 < NBLK) {
            img_load(k, B + 1, (B + 1) & 1);
        } else if (k + 1 < g.mt) {
            have0 = ready_ge(pdone, k + 2, &sh_ready);
            if (have0) img_load(k + 1, 0, 0);
        }
        apply_wave_ns<Tw, NB, true, NS>(A1, nullptr, B, B & 1);
    }
    WStrip<Tw, NB> S[NS];
    bool haveS = false;  // this row's strips are staged (or being staged)
    int pending = 0;
    // the next-row prefetch flags, read a block early (as in the Float64 walker)
    __shared__ int sh_nx[2];
    int fl_p = 0, fl_s = 0;
    for (int i = k + 1; i < g.mt; ++i) {
        if (pending && (!have0 || !haveS)) {
            publish(mine, pending);
            pending = 0;
        }
        if (!have0) {
            wait_ge(pdone, i + 1, g.err, sh_flag);
            img_load(i, 0, 0);
        }
        if (!haveS && src32) {
            if (prev) wait_ge(prev, i + 1, g.err, sh_flag);
            Stg::load(sh32 + (i64)i * NB + col * ldc, ldc);
        }
        bool next0 = false;
#pragma unroll
        for (int B = 0; B < NBLK; ++B) {
            if (B == NBLK - 1 && threadIdx.x == 0) {
                sh_nx[0] = fl_p >= i + 2;
                sh_nx[1] = fl_s >= i + 2;
            }
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // this block's image, the strips, earlier stores
            __syncthreads();
            if (B == 0) {
                if (pending) {
                    if (threadIdx.x == 0) cst(mine, pending);
                    pending = 0;
                }
                if (src32) Stg::read(S);
                else load_strips(S, i);  // step 0 (no previous step to wait for)
            }
            if (B == NBLK - 2 && threadIdx.x == 0) {
                fl_p = i + 1 < g.mt ? ld_flag(pdone) : 0;
                fl_s = i + 1 < g.mt && prev ? ld_flag(prev) : 0;
            }
            if (B + 1 < NBLK) {
                img_load(i, B + 1, (B + 1) & 1);
            } else if (i + 1 < g.mt) {
                next0 = sh_nx[0] != 0;
                if (next0) img_load(i + 1, 0, 0);
            }
            apply_wave_ns<Tw, NB, false, NS>(S, A1, B, B & 1);
        }
        have0 = next0;
        // Straight from the registers: the prefetch region of one wave lies inside another
        // wave's Float64 staging columns, so this walker stages nothing in Float64.
        store_strips(S, i, dst64);
        pending = i + 1;
        haveS = i + 1 < g.mt && (!prev || sh_nx[1] != 0);
        if (haveS && src32) Stg::load(sh32 + (i64)(i + 1) * NB + col * ldc, ldc);
    }
    if (pending) publish(mine, pending);
    store_strips(A1, k, true);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
}

// ---------------------------------------------------------------------------------
// 16-bit walker (Tw = _Float16 or __bf16, Float32 accumulation).  The strips stay Float32 in
// MFMA accumulators; every product rounds its operands to Tw: V and T when they are staged,
// the strip and the intermediate W from the accumulators right before each
// v_mfma_f32_16x16x16.  The Float32 accumulator layout of a 16 x 16 tile (lane l holds rows
// 4 (l/16) .. 4 (l/16) + 3 of column l % 16) is exactly the B operand of the 16x16x16 MFMA for
// a 16-deep slice of K, so the strip and W feed the products without data movement.  V is
// staged twice, column-major for V^T S and row-major for V W, so each lane's four A-operand
// elements are one 8-byte LDS read in both products.
template <typename Tw> struct LpMfma;
// v_mfma_f32_16x16x32_{f16,bf16}: lane l holds A[l % 16][8 (l/16) + i] and B[8 (l/16) + i][l % 16],
// i = 0..7; mma8 takes each as two halves of four (the products only need the same K order
// in both operands).
template <> struct LpMfma<_Float16> {
    typedef _Float16 v4 __attribute__((ext_vector_type(4)));
    typedef _Float16 v8 __attribute__((ext_vector_type(8)));
    static __device__ __forceinline__ void mma8(fx4 &c, v4 alo, v4 ahi, v4 blo, v4 bhi) {
        const v8 a = __builtin_shufflevector(alo, ahi, 0, 1, 2, 3, 4, 5, 6, 7);
        const v8 b = __builtin_shufflevector(blo, bhi, 0, 1, 2, 3, 4, 5, 6, 7);
        c = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
};
template <> struct LpMfma<__bf16> {
    typedef __bf16 v4 __attribute__((ext_vector_type(4)));
    typedef __bf16 v8 __attribute__((ext_vector_type(8)));
    static __device__ __forceinline__ void mma8(fx4 &c, v4 alo, v4 ahi, v4 blo, v4 bhi) {
        const v8 a = __builtin_shufflevector(alo, ahi, 0, 1, 2, 3, 4, 5, 6, 7);
        const v8 b = __builtin_shufflevector(blo, bhi, 0, 1, 2, 3, 4, 5, 6, 7);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
};
template <typename Tw> __device__ __forceinline__ typename LpMfma<Tw>::v4 lp_cvt(const fx4 &x, float sign) {
    typename LpMfma<Tw>::v4 r;
#pragma unroll
    for (int q = 0; q < 4; ++q) r[q] = (Tw)(sign * x[q]);
    return r;
}
// LDS of the 16-bit walker, in Tw elements: per buffer V column-major (Vc[j * LDC + p]),
// V row-major (Vr[p * LDR + j]) and T column-major (Tc[j * LDR + q] = T[q][j]); the next
// row's strip prefetch (StripStage) follows the two buffers.
// Each buffer is padded to whole KB: the panel writes exactly this image (IMG_BYTES) to HBM
// once per (tile, block) and walkers copy it into a buffer with 1 KB LDS-DMA instructions.
// Split BF16 ('X') images hold the hi parts and then, LO elements further, the lo parts.
template <typename Tw, int NB> struct LdsLp {
    static constexpr int LDC = NB + 8, LDR = IB + 8;
    static constexpr int VC = 0, VR = VC + IB * LDC, TC = VR + NB * LDR, BUF = TC + IB * LDR;
    static constexpr int PART_BYTES = (BUF * 2 + 1023) / 1024 * 1024;
    static constexpr int LO = PART_BYTES / 2;  // the lo part's offset in 16-bit elements
    static constexpr int IMG_BYTES = lp_parts<Tw>() * PART_BYTES;
    static constexpr int BUFS = IMG_BYTES / 2;  // buffer stride in 16-bit elements
    static constexpr int SSTG_BYTES = 2 * IMG_BYTES;
    static_assert(SSTG_BYTES % 16 == 0, "staging columns must start 16-byte aligned");
};
static_assert(LdsLp<_Float16, 128>::IMG_BYTES == LdsLp<__bf16, 128>::IMG_BYTES &&
                  LdsLp<_Float16, 64>::IMG_BYTES == LdsLp<__bf16, 64>::IMG_BYTES,
              "the FP16 and BF16 block images have one size");
// position of panel tile (i, k), i >= k, among the factorization's panel tiles
__host__ __device__ inline i64 lp_tile_index(int k, int i, int mt) {
    return (i64)k * mt - (i64)k * (k - 1) / 2 + (i - k);
}
// Float64 -> Tw through Float32: the hardware converts Float32 -> FP16 / BF16 in one
// instruction, while a direct Float64 -> FP16 rounding is a long integer sequence (and the
// compiler folds a plain two-step cast back into it; the empty asm keeps the Float32 value).
// The extra rounding is below 2^-24 relative, far under the 16-bit formats' 2^-11 / 2^-8.
template <typename Tw> __device__ __forceinline__ Tw to_lp(double v) {
    float f = (float)v;
    asm volatile("" : "+v"(f));
    return (Tw)f;
}
// Panel side: the block just factored (V in Vs, T in Ts of the panel's LDS layout, V with the
// unit diagonal and zeros above for a geqrt block) converted to Tw in the 16-bit walker's
// buffer layout and written to dst (IMG_BYTES), two elements per coherent 32-bit store.
template <typename T, typename Tw, int NB> __device__ void write_lp_image(unsigned char *dst) {
    using L = LdsLp<Tw, NB>;
    constexpr int LDV = Lds<T, NB>::LDV;
    const T *Vs = lds_base<T>() + Lds<T, NB>::V;
    const T *Ts = lds_base<T>() + Lds<T, NB>::TT;
    unsigned *d32 = reinterpret_cast<unsigned *>(dst);
    for (int wd = threadIdx.x; wd < L::PART_BYTES / 4; wd += NTH) {
        unsigned short h[2], l[2];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int e = 2 * wd + q;
            T v = T(0);
            if (e < L::VR) {
                const int j = e / L::LDC, p = e % L::LDC;
                if (p < NB) v = Vs[j * LDV + p];
            } else if (e < L::TC) {
                const int f = e - L::VR, p = f / L::LDR, j = f % L::LDR;
                if (j < IB) v = Vs[j * LDV + p];
            } else if (e < L::BUF) {
                const int f = e - L::TC, j = f / L::LDR, r = f % L::LDR;
                if (r < IB && r <= j) v = Ts[r * (IB + 1) + j];
            }
            if constexpr (lp_parts<Tw>() == 2) {
                float f = (float)v;
                asm volatile("" : "+v"(f));
                const __bf16 hi = (__bf16)f;
                h[q] = __builtin_bit_cast(unsigned short, hi);
                l[q] = __builtin_bit_cast(unsigned short, (__bf16)(f - (float)hi));  // f - hi is exact
            } else {
                h[q] = __builtin_bit_cast(unsigned short, to_lp<Tw>(v));
            }
        }
        cst(&d32[wd], (unsigned)h[0] | ((unsigned)h[1] << 16));
        if constexpr (lp_parts<Tw>() == 2) cst(&d32[wd + L::PART_BYTES / 4], (unsigned)l[0] | ((unsigned)l[1] << 16));
    }
}
// Walker side: one block image into LDS buffer buf with 1 KB LDS-DMA instructions (waves
// take every fourth KB); completion is tracked by vmcnt like any vector load.
template <typename Tw, int NB> __device__ __forceinline__ void lp_img_load(const unsigned char *src, int buf) {
    constexpr int BYTES = LdsLp<Tw, NB>::IMG_BYTES, NI = BYTES / 1024;
    unsigned char *dst = lds_base<unsigned char>() + buf * BYTES;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int c = 0; c < NI; c += WAVES) {
        if (c + w < NI) {
            auto *d = (__attribute__((address_space(3))) void *)(dst + (c + w) * 1024);
            __builtin_amdgcn_global_load_lds(src + (c + w) * 1024 + lane * 16, d, 16, 0, /*sc1*/ 16);
        }
    }
}
// H^T (the reference's trans 'C') of the staged block applied to the wave's NS strips:
//   W0 = V^T S (+ A1 rows),  W = T^T W0,  A1 rows -= W,  S -= V W.
// Every V / T operand read from LDS serves the NS strips' products.  Each product has a
// 32-deep K: a lane's eight K values are the four rows 4 (l/16)..+3 of two accumulator tiles
// 16 rows apart (the Float32 strip / W0 / W registers as they are), and the A operands are
// read at the same rows (two 8-byte LDS reads).
template <typename Tw, int NB, bool GE, int NS>
__device__ __forceinline__ void apply_lp(WStrip<float, NB> *S, WStrip<float, NB> *A1, int B, int buf) {
    using L = LdsLp<Tw, NB>;
    using M = LpMfma<Tw>;
    typedef typename M::v4 v4;
    const Tw *sm = lds_base<Tw>() + buf * L::BUFS;
    const int lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
    auto rd = [&](int off) { return *reinterpret_cast<const v4 *>(sm + off); };
    fx4 w0[NS][2];
#pragma unroll
    for (int q = 0; q < NS; ++q) w0[q][0] = w0[q][1] = acc_zero<float>();
#pragma unroll
    for (int t = 0; t < NB / 32; ++t) {  // rows 32 t .. 32 t + 31 of the strip
        v4 alo[2], ahi[2];
#pragma unroll
        for (int ot = 0; ot < 2; ++ot) {
            const int o = L::VC + (16 * ot + lr) * L::LDC + 32 * t + 4 * lk;
            alo[ot] = rd(o);
            ahi[ot] = rd(o + 16);
        }
#pragma unroll
        for (int q = 0; q < NS; ++q) {
            const v4 blo = lp_cvt<Tw>(S[q].a[2 * t], 1.0f), bhi = lp_cvt<Tw>(S[q].a[2 * t + 1], 1.0f);
#pragma unroll
            for (int ot = 0; ot < 2; ++ot) M::mma8(w0[q][ot], alo[ot], ahi[ot], blo, bhi);
        }
    }
    if (!GE) {
#pragma unroll
        for (int q = 0; q < NS; ++q)
#pragma unroll
            for (int ot = 0; ot < 2; ++ot) w0[q][ot] += A1[q].a[2 * B + ot];
    }
    fx4 wv[NS][2];
#pragma unroll
    for (int q = 0; q < NS; ++q) wv[q][0] = wv[q][1] = acc_zero<float>();
    {
        v4 alo[2], ahi[2];
#pragma unroll
        for (int ot = 0; ot < 2; ++ot) {
            const int o = L::TC + (16 * ot + lr) * L::LDR + 4 * lk;
            alo[ot] = rd(o);
            ahi[ot] = rd(o + 16);
        }
#pragma unroll
        for (int q = 0; q < NS; ++q) {
            const v4 blo = lp_cvt<Tw>(w0[q][0], 1.0f), bhi = lp_cvt<Tw>(w0[q][1], 1.0f);
#pragma unroll
            for (int ot = 0; ot < 2; ++ot) M::mma8(wv[q][ot], alo[ot], ahi[ot], blo, bhi);
        }
    }
    if (!GE) {
#pragma unroll
        for (int q = 0; q < NS; ++q)
#pragma unroll
            for (int ot = 0; ot < 2; ++ot) A1[q].a[2 * B + ot] -= wv[q][ot];
    }
    v4 blo[NS], bhi[NS];
#pragma unroll
    for (int q = 0; q < NS; ++q) {
        blo[q] = lp_cvt<Tw>(wv[q][0], -1.0f);
        bhi[q] = lp_cvt<Tw>(wv[q][1], -1.0f);
    }
#pragma unroll
    for (int rt = 0; rt < NB / 16; ++rt) {
        const int o = L::VR + (16 * rt + lr) * L::LDR + 4 * lk;
        const v4 alo = rd(o), ahi = rd(o + 16);
#pragma unroll
        for (int q = 0; q < NS; ++q) M::mma8(S[q].a[rt], alo, ahi, blo[q], bhi[q]);
    }
}

// The split-BF16 apply: apply_lp's products with both operands in two parts.  A operands (V, T)
// come from the image's hi and lo parts; B operands (the Float32 strips, W0, W) are split in
// registers right before their products.
__device__ __forceinline__ void bf16_split(const fx4 &x, float sign, LpMfma<__bf16>::v4 &hi, LpMfma<__bf16>::v4 &lo) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const float f = sign * x[q];
        const __bf16 h = (__bf16)f;
        hi[q] = h;
        lo[q] = (__bf16)(f - (float)h);
    }
}
// c += A B as lo*hi + hi*lo + hi*hi (the small terms first)
__device__ __forceinline__ void mma8x3(fx4 &c, const LpMfma<__bf16>::v4 *a, const LpMfma<__bf16>::v4 *b) {
    // a / b: [0] hi of K 0..3, [1] hi of K 4..7, [2] lo of K 0..3, [3] lo of K 4..7
    LpMfma<__bf16>::mma8(c, a[2], a[3], b[0], b[1]);
    LpMfma<__bf16>::mma8(c, a[0], a[1], b[2], b[3]);
    LpMfma<__bf16>::mma8(c, a[0], a[1], b[0], b[1]);
}
template <int NB, bool GE, int NS>
__device__ __forceinline__ void apply_lp3(WStrip<float, NB> *S, WStrip<float, NB> *A1, int B, int buf) {
    using L = LdsLp<bf16x3, NB>;
    typedef LpMfma<__bf16>::v4 v4;
    const __bf16 *sm = lds_base<__bf16>() + buf * L::BUFS;
    const int lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
    auto rd = [&](int off) { return *reinterpret_cast<const v4 *>(sm + off); };
    auto rd4 = [&](v4 *a, int o) {  // hi and lo parts of a lane's eight K values at o, o + 16
        a[0] = rd(o);
        a[1] = rd(o + 16);
        a[2] = rd(o + L::LO);
        a[3] = rd(o + L::LO + 16);
    };
    fx4 w0[NS][2];
#pragma unroll
    for (int q = 0; q < NS; ++q) w0[q][0] = w0[q][1] = acc_zero<float>();
#pragma unroll
    for (int t = 0; t < NB / 32; ++t) {
        v4 a[2][4];
#pragma unroll
        for (int ot = 0; ot < 2; ++ot) rd4(a[ot], L::VC + (16 * ot + lr) * L::LDC + 32 * t + 4 * lk);
#pragma unroll
        for (int q = 0; q < NS; ++q) {
            v4 b[4];
            bf16_split(S[q].a[2 * t], 1.0f, b[0], b[2]);
            bf16_split(S[q].a[2 * t + 1], 1.0f, b[1], b[3]);
#pragma unroll
            for (int ot = 0; ot < 2; ++ot) mma8x3(w0[q][ot], a[ot], b);
        }
    }
    if (!GE) {
#pragma unroll
        for (int q = 0; q < NS; ++q)
#pragma unroll
            for (int ot = 0; ot < 2; ++ot) w0[q][ot] += A1[q].a[2 * B + ot];
    }
    fx4 wv[NS][2];
#pragma unroll
    for (int q = 0; q < NS; ++q) wv[q][0] = wv[q][1] = acc_zero<float>();
    {
        v4 a[2][4];
#pragma unroll
        for (int ot = 0; ot < 2; ++ot) rd4(a[ot], L::TC + (16 * ot + lr) * L::LDR + 4 * lk);
#pragma unroll
        for (int q = 0; q < NS; ++q) {
            v4 b[4];
            bf16_split(w0[q][0], 1.0f, b[0], b[2]);
            bf16_split(w0[q][1], 1.0f, b[1], b[3]);
#pragma unroll
            for (int ot = 0; ot < 2; ++ot) mma8x3(wv[q][ot], a[ot], b);
        }
    }
    if (!GE) {
#pragma unroll
        for (int q = 0; q < NS; ++q)
#pragma unroll
            for (int ot = 0; ot < 2; ++ot) A1[q].a[2 * B + ot] -= wv[q][ot];
    }
    v4 b[NS][4];
#pragma unroll
    for (int q = 0; q < NS; ++q) {
        bf16_split(wv[q][0], -1.0f, b[q][0], b[q][2]);
        bf16_split(wv[q][1], -1.0f, b[q][1], b[q][3]);
    }
#pragma unroll
    for (int rt = 0; rt < NB / 16; ++rt) {
        v4 a[4];
        rd4(a, L::VR + (16 * rt + lr) * L::LDR + 4 * lk);
#pragma unroll
        for (int q = 0; q < NS; ++q) mma8x3(S[q].a[rt], a, b[q]);
    }
}

template <typename T, typename Tw, int NB>
__device__ __forceinline__ void run_walk_lp(int k, int s, T * /*A*/, i64 /*lda*/, T * /*Tm*/, i64 /*ldt*/, DagArgs g,
                                            int *sh_flag) {
    constexpr int NBLK = NB / IB;
    constexpr i64 IMG = LdsLp<Tw, NB>::IMG_BYTES;
    static_assert(NBLK % 2 == 0, "block count must be even for the buffer parity");
    __shared__ int sh_ready;
    constexpr int TWK = walk_width<T, Tw, NB>(), NS = TWK / TW;  // columns per task, strips per wave
    const int w = threadIdx.x >> 6;
    const i64 col = (i64)s * TWK + w * NS * SW;  // the wave's strip q starts at column col + 16 q
    const int *prev = k > 0 ? &g.strip_prog[(k - 1) * g.nwalk + s] : nullptr;
    int *mine = &g.strip_prog[k * g.nwalk + s];
    const int *pdone = &g.panel_prog[k * NBLK + NBLK - 1];
    T *Cm = reinterpret_cast<T *>(g.C);
    const i64 ldc = g.ldc;
    // Strips between steps live in the Float32 shadow: step 0 reads the matrix, later steps
    // the shadow; a step writes its rows into the matrix (Float64) only for the strips of the
    // next panel's tile column, which that panel reads, and the final R row (A1) always.
    // Strips are stored straight from registers (strip_store); the next row's Float32 strips
    // are prefetched into LDS (LpStage), step 0's Float64 ones loaded at the row start.
    float *const sh32 = g.shadow;
    const bool src32 = k > 0, dst64 = s / (NB / TWK) == k + 1;
    auto load_strips = [&](WStrip<float, NB> *X, int row_tile) {
#pragma unroll
        for (int q = 0; q < NS; ++q) {
            const i64 off = (i64)row_tile * NB + (col + q * SW) * ldc;
            if (src32) strip_load<float, NB>(X[q].a, sh32 + off, ldc);
            else strip_load<T, NB>(X[q].a, Cm + off, ldc);
        }
    };
    auto store_strips = [&](const WStrip<float, NB> *X, int row_tile, bool f64) {
#pragma unroll
        for (int q = 0; q < NS; ++q) {
            const i64 off = (i64)row_tile * NB + (col + q * SW) * ldc;
            if (f64) strip_store<T, NB>(X[q].a, Cm + off, ldc);
            else strip_store<float, NB, ST_SC1_NT>(X[q].a, sh32 + off, ldc);
        }
    };
    unsigned char *const lpimg = g.lpimg;
    auto img = [&](int i, int B) { return lpimg + (lp_tile_index(k, i, g.mt) * NBLK + B) * IMG; };
    wait_ge(pdone, k + 1, g.err, sh_flag);
    if (prev) wait_ge(prev, k + 1, g.err, sh_flag);
    WStrip<float, NB> A1[NS];
    load_strips(A1, k);
    // unmqr!('L','C') with geqrt(A_kk)'s blocks (their images carry the unit-lower V)
    lp_img_load<Tw, NB>(img(k, 0), 0);
    bool have0 = false;
#pragma unroll
    for (int B = 0; B < NBLK; ++B) {
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
        if (B + 1 < NBLK) {
            lp_img_load<Tw, NB>(img(k, B + 1), (B + 1) & 1);
        } else if (k + 1 < g.mt) {
            have0 = ready_ge(pdone, k + 2, &sh_ready);
            if (have0) lp_img_load<Tw, NB>(img(k + 1, 0), 0);
        }
        if constexpr (lp_parts<Tw>() == 2) apply_lp3<NB, true, NS>(A1, nullptr, B, B & 1);
        else apply_lp<Tw, NB, true, NS>(A1, nullptr, B, B & 1);
    }
    using Stg = StripStage<LdsLp<Tw, NB>::SSTG_BYTES, NB, NS>;
    static_assert(WAVES * Stg::NG * Stg::GRP <= TW * Lds<T, NB>::LDS_S * (int)sizeof(T), "staging fits the strip region");
    WStrip<float, NB> S[NS];
    bool haveS = false;  // this row's strips are staged (or being staged)
    int pending = 0;
    // the next-row prefetch flags, read a block early (as in the Float64 walker)
    __shared__ int sh_nx[2];
    int fl_p = 0, fl_s = 0;
    for (int i = k + 1; i < g.mt; ++i) {
        if (pending && (!have0 || !haveS)) {
            publish(mine, pending);
            pending = 0;
        }
        if (!have0) {
            wait_ge(pdone, i + 1, g.err, sh_flag);
            lp_img_load<Tw, NB>(img(i, 0), 0);
        }
        if (!haveS && src32) {
            if (prev) wait_ge(prev, i + 1, g.err, sh_flag);
            Stg::load(sh32 + (i64)i * NB + col * ldc, ldc);
        }
        bool next0 = false;
#pragma unroll
        for (int B = 0; B < NBLK; ++B) {
            if (B == NBLK - 1 && threadIdx.x == 0) {
                sh_nx[0] = fl_p >= i + 2;
                sh_nx[1] = fl_s >= i + 2;
            }
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // this block's image, the strip, earlier stores
            __syncthreads();
            if (B == 0) {
                if (pending) {
                    if (threadIdx.x == 0) cst(mine, pending);
                    pending = 0;
                }
                if (src32) Stg::read(S);
                else load_strips(S, i);  // step 0 (no previous step to wait for)
            }
            if (B == NBLK - 2 && threadIdx.x == 0) {
                fl_p = i + 1 < g.mt ? ld_flag(pdone) : 0;
                fl_s = i + 1 < g.mt && prev ? ld_flag(prev) : 0;
            }
            // (the image DMA goes out before the block's products: spread over them, as in the
            // Float64 walker, it delays the short 16-bit apply more than it hides)
            if (B + 1 < NBLK) {
                lp_img_load<Tw, NB>(img(i, B + 1), (B + 1) & 1);
            } else if (i + 1 < g.mt) {
                next0 = sh_nx[0] != 0;
                if (next0) lp_img_load<Tw, NB>(img(i + 1, 0), 0);
            }
            if (B == NBLK - 1) {  // the next row's strip, when its previous step has released it
                haveS = i + 1 < g.mt && (!prev || sh_nx[1] != 0);
                if (haveS && src32) Stg::load(sh32 + (i64)(i + 1) * NB + col * ldc, ldc);
            }
            if constexpr (lp_parts<Tw>() == 2) apply_lp3<NB, false, NS>(S, A1, B, B & 1);
            else apply_lp<Tw, NB, false, NS>(S, A1, B, B & 1);
        }
        have0 = next0;
        store_strips(S, i, dst64);
        pending = i + 1;
    }
    if (pending) publish(mine, pending);
    store_strips(A1, k, true);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
}

// Dynamic LDS of a launch: the larger of the T and Tw layouts, and for the 16-bit walkers the
// two block-image buffers with the next row's strip staging after them (split-BF16 images are
// twice as large, which takes the launch past the Float64 layout's 150 KB).
template <typename T, typename Tw, int NB> constexpr size_t dag_lds_bytes() {
    const size_t t = sizeof(T) * (size_t)Lds<T, NB>::TOTAL, w = sizeof(Tw) * (size_t)Lds<Tw, NB>::TOTAL;
    size_t b = t > w ? t : w;
    if constexpr (sizeof(Tw) == 2) {
        const size_t lp = (size_t)LdsLp<Tw, NB>::SSTG_BYTES + TW * Lds<T, NB>::LDS_S * sizeof(T);
        if (lp > b) b = lp;
    }
    return b;
}

template <typename T, int NB, typename Tw>
__global__ __launch_bounds__(NTH) void tileqr_dag_kernel(T *A, i64 lda, T *Tm, i64 ldt, DagArgs g) {
    static_assert((sizeof(Tw) == 2 ? (size_t)LdsLp<Tw, NB>::SSTG_BYTES : Lds<Tw, NB>::SSTG * sizeof(Tw)) +
                          TW * Lds<T, NB>::LDS_S * sizeof(T) <=
                      dag_lds_bytes<T, Tw, NB>(),
                  "the walker's staging columns must fit the launch's LDS");
    static_assert(dag_lds_bytes<T, Tw, NB>() <= 160 * 1024, "a CU's LDS");
    __shared__ int sh_task, sh_flag;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) sh_task = atomicAdd(g.counter, 1);
        __syncthreads();
        const int task = sh_task;
        if (task >= g.ntasks) break;
        const int code = g.tasks[task];
        if (task_type(code) == TASK_PANEL) run_panel<T, NB, PanelTw<T, Tw>>(task_k(code), task_s(code), A, lda, Tm, ldt, g, &sh_flag);
        else if constexpr (sizeof(Tw) == sizeof(T)) {
            if (g.qmode) run_walk_q<T, NB>(task_k(code), task_s(code), A, lda, Tm, ldt, g, &sh_flag);
            else run_walk<T, NB>(task_k(code), task_s(code), A, lda, Tm, ldt, g, &sh_flag);
        } else if constexpr (sizeof(Tw) == 2) {
            run_walk_lp<T, Tw, NB>(task_k(code), task_s(code), A, lda, Tm, ldt, g, &sh_flag);
        } else {
            run_walk_mixed<T, Tw, NB>(task_k(code), task_s(code), A, lda, Tm, ldt, g, &sh_flag);
        }
    }
}

__global__ void fill_int_kernel(int *p, int n, int v) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// Task list, sorted by urgency: WALK(k, s) on tile column j has key 5 j + k, PANEL(k, b) has
// key 6 k; ties go to the lower step, then column, then stage/strip.  The key is a
// topological order of the waits (WALK(k-1, column k) = 6k - 1 < PANEL(k) = 6k < WALK(k, j > k);
// WALK(k, j) < WALK(k+1, j)), and it lets the panels run several steps ahead of the far
// columns' walkers: a walker holds its workgroup for its whole column walk, so with a
// one-step lookahead order the walkers of step k fill the machine and the panel of step k+2
// cannot be claimed before they finish.  The weight 5 : 1 (a panel about 0.2 steps of column
// distance ahead per step) is the best of a timed simulation of claim order and row
// dependencies at 4096..16384 squared and 32768 x 8192; it is not critical between 0.1 and 0.3.
static std::vector<int> build_tasks(int mt, int nt, int nb, int tw) {
    const int kt = mt < nt ? mt : nt, spt = nb / tw, nblk = nb / IB;
    struct Item {
        long long key;
        int k, j, sub, code;
    };
    std::vector<Item> items;
    for (int k = 0; k < kt; ++k) {
        for (int b = 0; b < nblk; ++b) items.push_back({6ll * k, k, k, b, enc(TASK_PANEL, k, b)});
        for (int s = (k + 1) * spt; s < nt * spt; ++s) {
            const int j = s / spt;
            items.push_back({5ll * j + k, k, j, nblk + s, enc(TASK_WALK, k, s)});
        }
    }
    std::sort(items.begin(), items.end(), [](const Item &a, const Item &b) {
        if (a.key != b.key) return a.key < b.key;
        if (a.k != b.k) return a.k < b.k;
        if (a.j != b.j) return a.j < b.j;
        return a.sub < b.sub;
    });
    std::vector<int> v;
    v.reserve(items.size());
    for (const Item &it : items) v.push_back(it.code);
    return v;
}

}  // namespace dag

// Host driver.  FACTOR: the whole tile QR of A (walkers update A itself).  APPLY: Q^T of
// an already factored A applied to C (m x ncols, ncols a multiple of TW) by the walkers
// alone, every panel marked done up front.  Returns 0, a HIP error code, or -1 (reported as
// a HIP-side failure by the caller) when a dependency wait timed out.
enum DagMode { DAG_FACTOR = 0, DAG_APPLY = 1, DAG_APPLY_Q = 2 };

template <typename T, int NB, typename Tw>
static int dag_run(DagMode mode, i64 m, i64 n, T *A, i64 lda, T *Tm, i64 ldt, T *C, i64 ldc, i64 ncols,
                   hipStream_t s) {
    using namespace dag;
    const int mt = (int)(m / NB), nt = (int)(n / NB), kt = mt < nt ? mt : nt;
    const int nblk = NB / IB;
    constexpr int TWK = walk_width<T, Tw, NB>();
    const int nwalk = (int)((mode == DAG_FACTOR ? n : ncols) / TWK);
    std::vector<int> tasks;
    if (mode == DAG_FACTOR) {
        tasks = build_tasks(mt, nt, NB, TWK);
    } else if (mode == DAG_APPLY) {
        for (int k = 0; k < kt; ++k)
            for (int w = 0; w < nwalk; ++w) tasks.push_back(enc(TASK_WALK, k, w));
    } else {
        for (int k = kt - 1; k >= 0; --k)
            for (int w = 0; w < nwalk; ++w) tasks.push_back(enc(TASK_WALK, k, w));
    }
    const int ntasks = (int)tasks.size();
    int dev = 0, ncu = 0;
    NLA_HIP(hipGetDevice(&dev));
    NLA_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    constexpr size_t sh = dag_lds_bytes<T, Tw, NB>();
    if (sh > 64 * 1024)
        NLA_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&tileqr_dag_kernel<T, NB, Tw>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
    int per_cu = 0;
    NLA_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, tileqr_dag_kernel<T, NB, Tw>, NTH, sh));
    if (per_cu < 1) per_cu = 1;
    int grid = ncu * per_cu;
    if (grid > ntasks) grid = ntasks;
    // scheduler memory: tasks | counter | err | panel_prog[kt*NB/IB] | strip_prog[kt*nwalk]
    const int npp = kt * nblk;
    const size_t nint = (size_t)ntasks + 2 + npp + (size_t)kt * nwalk;
    int st = 0;
    int *mem = (int *)scratch(SLOT_SCHED, nint * sizeof(int), s, &st);
    if (!mem) return st;
    int *d_tasks = mem, *counter = mem + ntasks, *err = counter + 1, *pp = err + 1, *sp = pp + npp;
    NLA_HIP(hipMemcpyAsync(d_tasks, tasks.data(), (size_t)ntasks * sizeof(int), hipMemcpyHostToDevice, s));
    const int nzero = 2 + npp + kt * nwalk;
    fill_int_kernel<<<(nzero + 255) / 256 > 1024 ? 1024 : (nzero + 255) / 256, 256, 0, s>>>(counter, nzero, 0);
    NLA_LAUNCH_CHECK();
    if (mode != DAG_FACTOR) {  // every panel tile is factored already
        fill_int_kernel<<<(npp + 255) / 256, 256, 0, s>>>(pp, npp, mt);
        NLA_LAUNCH_CHECK();
    }
    unsigned char *lpimg = nullptr;
    if constexpr (sizeof(Tw) < sizeof(T)) {
        // every panel tile's NB/IB block images, written by the panels, read by the walkers
        const size_t ntiles = (size_t)kt * mt - (size_t)kt * (kt - 1) / 2;
        const size_t img_bytes = sizeof(Tw) == 2 ? (size_t)LdsLp<Tw, NB>::IMG_BYTES : (size_t)F32Img<NB>::STRIDE;
        lpimg = (unsigned char *)scratch(SLOT_LPIMG, ntiles * nblk * img_bytes, s, &st);
        if (!lpimg) return st;
    }
    float *shadow = nullptr;
    if constexpr (sizeof(Tw) < sizeof(T)) {
        // the trailing matrix's Float32 shadow (same positions and leading dimension as C)
        shadow = (float *)scratch(SLOT_SHADOW, sizeof(float) * (size_t)ldc * (size_t)ncols, s, &st);
        if (!shadow) return st;
    }
    DagArgs g{mt, nt, nwalk, ntasks, d_tasks, counter, pp, sp, err, (void *)C, ldc, mode == DAG_APPLY_Q ? 1 : 0,
              lpimg, shadow};
    {
        ProfScope prof(PK_TILE_QR, s);
        tileqr_dag_kernel<T, NB, Tw><<<grid, NTH, sh, s>>>(A, lda, Tm, ldt, g);
        NLA_LAUNCH_CHECK();
    }
    int herr = 0;
    NLA_HIP(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, s));
    NLA_HIP(hipStreamSynchronize(s));
    if (herr) return wait_timeout_error("tile QR");
    return 0;
}

// Shapes the dataflow kernel handles: ib = 32, nb in {64, 128}, 16-byte aligned columns.
template <typename T> static bool dag_shape_ok(i64 m, i64 n, i64 nb, i64 ib, const T *A, i64 lda) {
    if (ib != dag::IB || (nb != 64 && nb != 128)) return false;
    if (m / nb >= 4096 || n / dag::TW >= (1 << 18)) return false;
    return !((reinterpret_cast<uintptr_t>(A) & 15) || ((lda * (i64)sizeof(T)) & 15));
}

template <typename T>
int tile_geqrf_dag(i64 m, i64 n, i64 nb, i64 ib, T *A, i64 lda, T *Tm, i64 ldt, hipStream_t s, bool *handled) {
    *handled = dag_shape_ok<T>(m, n, nb, ib, A, lda);
    if (!*handled) return 0;
    if (nb == 64) return dag_run<T, 64, T>(DAG_FACTOR, m, n, A, lda, Tm, ldt, A, lda, n, s);
    return dag_run<T, 128, T>(DAG_FACTOR, m, n, A, lda, Tm, ldt, A, lda, n, s);
}

// Trailing updates in Float32 (prec 'S'), FP16 ('H'), BF16 ('B') or split BF16 ('X'); the panels
// in Float64.
int tile_geqrf_dag_mixed(i64 m, i64 n, i64 nb, i64 ib, double *A, i64 lda, double *Tm, i64 ldt, char prec,
                         hipStream_t s, bool *handled) {
    *handled = dag_shape_ok<double>(m, n, nb, ib, A, lda);
    if (!*handled) return 0;
    if (prec == 'X') {
        if (nb == 64) return dag_run<double, 64, dag::bf16x3>(DAG_FACTOR, m, n, A, lda, Tm, ldt, A, lda, n, s);
        return dag_run<double, 128, dag::bf16x3>(DAG_FACTOR, m, n, A, lda, Tm, ldt, A, lda, n, s);
    }
    if (prec == 'H') {
        if (nb == 64) return dag_run<double, 64, _Float16>(DAG_FACTOR, m, n, A, lda, Tm, ldt, A, lda, n, s);
        return dag_run<double, 128, _Float16>(DAG_FACTOR, m, n, A, lda, Tm, ldt, A, lda, n, s);
    }
    if (prec == 'B') {
        if (nb == 64) return dag_run<double, 64, __bf16>(DAG_FACTOR, m, n, A, lda, Tm, ldt, A, lda, n, s);
        return dag_run<double, 128, __bf16>(DAG_FACTOR, m, n, A, lda, Tm, ldt, A, lda, n, s);
    }
    if (nb == 64) return dag_run<double, 64, float>(DAG_FACTOR, m, n, A, lda, Tm, ldt, A, lda, n, s);
    return dag_run<double, 128, float>(DAG_FACTOR, m, n, A, lda, Tm, ldt, A, lda, n, s);
}

// C (m x ncols, ncols a multiple of 64) := Q^T C (qt) or Q C with the factors of tile_geqrf.
template <typename T>
int tile_apply_qt_dag(i64 m, i64 n, i64 nb, i64 ib, const T *A, i64 lda, const T *Tm, i64 ldt, T *C, i64 ldc,
                      i64 ncols, hipStream_t s, bool *handled, bool qt) {
    *handled = dag_shape_ok<T>(m, n, nb, ib, A, lda) && ncols % dag::TW == 0 &&
               !((reinterpret_cast<uintptr_t>(C) & 15) || ((ldc * (i64)sizeof(T)) & 15));
    if (!*handled) return 0;
    T *Am = const_cast<T *>(A), *Tmm = const_cast<T *>(Tm);
    const DagMode md = qt ? DAG_APPLY : DAG_APPLY_Q;
    if (nb == 64) return dag_run<T, 64, T>(md, m, n, Am, lda, Tmm, ldt, C, ldc, ncols, s);
    return dag_run<T, 128, T>(md, m, n, Am, lda, Tmm, ldt, C, ldc, ncols, s);
}

template int tile_geqrf_dag<double>(i64, i64, i64, i64, double *, i64, double *, i64, hipStream_t, bool *);
template int tile_geqrf_dag<float>(i64, i64, i64, i64, float *, i64, float *, i64, hipStream_t, bool *);
template int tile_apply_qt_dag<double>(i64, i64, i64, i64, const double *, i64, const double *, i64, double *, i64, i64,
                                       hipStream_t, bool *, bool);
template int tile_apply_qt_dag<float>(i64, i64, i64, i64, const float *, i64, const float *, i64, float *, i64, i64,
                                      hipStream_t, bool *, bool);

}  // namespace nla

// The factorization's task list (the claim order of the persistent launch), for inspection:
// code = type << 30 | k << 18 | s with type 0 = PANEL(k, block stage s), 1 = WALK(k, strip
// group s of TW = 64 columns; the reduced-precision factorizations' (Float32 / FP16 / BF16
// updates) list is the same with groups of 128 columns when nb = 128).  Returns the number of tasks (codes written up to cap), or -1
// for a shape the dataflow kernel does not take.
NLA_API int64_t nla_tile_qr_schedule(int64_t m, int64_t n, int64_t nb, int32_t *codes, int64_t cap) {
    if ((nb != 64 && nb != 128) || m < nb || n < nb || m % nb || n % nb || m / nb >= 4096) {
        nla::set_error("tile QR schedule: nb must be 64 or 128 and m, n positive multiples of nb");
        return -1;
    }
    const std::vector<int> v = nla::dag::build_tasks((int)(m / nb), (int)(n / nb), (int)nb, nla::dag::TW);
    for (size_t i = 0; i < v.size() && (int64_t)i < cap; ++i) codes[i] = v[i];
    return (int64_t)v.size();
}
