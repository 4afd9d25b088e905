// Tile-QR kernels of NextLA.jl on CDNA4: larfg!, geqr2!, larft!, larfb!, geqrt!, unmqr!,
// tsqrt!, tsmqr!, parfb! (src/larfg.jl, src/geqr2.jl, src/larft.jl, src/larfb.jl,
// src/geqrt.jl, src/unmqr.jl, src/tsqrt.jl, src/tsmqr.jl, src/parfb.jl).
//
// Two kernel families:
//  * Panel kernels (one 256-thread workgroup): the column-by-column Householder work of
//    geqr2!/tsqrt! with block reductions for the norms, the trailing in-panel update and
//    the T column fused into one pass per column (the reference runs these on a single
//    work-item, src/geqr2.jl:119, src/tsqrt.jl:148).
//  * Strip kernels: a block reflector applied from the left acts on every column of C
//    independently, so a workgroup owns a strip of CW columns and runs the *whole*
//    unmqr!/tsmqr! block loop on it in one launch (the reference issues one larfb!/parfb!
//    per ib-block, src/unmqr.jl:127, src/tsmqr.jl:123).  Right-side calls are the left
//    form on the transposed view with trans flipped: C op(H) = (op(H)^T C^T)^T.
//  Blocks wider than 64 reflectors are applied as consecutive 64-wide sub-blocks, whose
//  T factors are the diagonal sub-blocks of the block's T (larft's column recurrence makes
//  every leading/trailing diagonal block of T the T factor of those reflectors).
#include "nla_common.hpp"
#include "nla_qr.hpp"

#include <utility>
#include <vector>

namespace nla {

// ------------------------------------------------------------------------------------
// larfg!: src/larfg.jl:43-106 with its complex branch (safe rescale loop kept; beta via
// hypot for reals as LAPACK dlarfg, via a scaled 3-term norm for complex elements).
template <typename R> __device__ __forceinline__ R lamch_sfmin();
template <> __device__ __forceinline__ double lamch_sfmin<double>() { return 2.2250738585072014e-308; }
template <> __device__ __forceinline__ float lamch_sfmin<float>() { return 1.17549435e-38f; }
template <typename R> __device__ __forceinline__ R lamch_eps();
template <> __device__ __forceinline__ double lamch_eps<double>() { return 1.1102230246251565e-16; }
template <> __device__ __forceinline__ float lamch_eps<float>() { return 5.96046448e-08f; }

// sqrt(a^2 + b^2 + c^2) without overflow or underflow of the squares
template <typename R> __device__ __forceinline__ R hyp3(R a, R b, R c) {
    a = fabs(a); b = fabs(b); c = fabs(c);
    const R mx = fmax(a, fmax(b, c));
    if (mx == R(0)) return R(0);
    a /= mx; b /= mx; c /= mx;
    return mx * sqrt(a * a + b * b + c * c);
}
template <typename T> __device__ __forceinline__ typename RealOf<T>::t beta_of(typename RealOf<T>::t alphr,
                                                                              typename RealOf<T>::t alphi,
                                                                              typename RealOf<T>::t xnorm) {
    if constexpr (IsCplx<T>::v) return -copysign(hyp3(alphr, alphi, xnorm), alphr);
    else return -copysign(hypot(alphr, xnorm), alphr);
}

template <typename T>
__global__ __launch_bounds__(256) void larfg_kernel(int n, T *alpha, T *x, i64 incx, T *tau) {
    typedef typename RealOf<T>::t R;
    __shared__ R red[16];
    __shared__ T sh_scale;
    const int t = threadIdx.x;
    const int nx = n - 1;
    // two-pass overflow-safe norm: max |component|, then sum |x/max|^2
    R mx = 0;
    for (int i = t; i < nx; i += 256) mx = fmax(mx, absmax_(x[i * incx]));
    mx = wave_max(mx);
    __syncthreads();
    if ((t & 63) == 0) red[t >> 6] = mx;
    __syncthreads();
    mx = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    R ss = 0;
    if (mx > 0)
        for (int i = t; i < nx; i += 256) ss += abs2_(x[i * incx] / mx);
    ss = block_sum(ss, red);
    R xnorm = mx * sqrt(ss);
    const T a0 = *alpha;
    R alphr = re_(a0), alphi = im_(a0);
    if (xnorm == R(0) && alphi == R(0)) {
        if (t == 0) *tau = T(0);
        return;
    }
    R beta = beta_of<T>(alphr, alphi, xnorm);
    const R safmin = lamch_sfmin<R>() / lamch_eps<R>();
    const R rsafmn = R(1) / safmin;
    int knt = 0;
    if (fabs(beta) < safmin) {
        while (fabs(beta) < safmin && knt < 20) {
            ++knt;
            beta *= rsafmn; alphr *= rsafmn; alphi *= rsafmn;
        }
        R f = 1;
        for (int j = 0; j < knt; ++j) f *= rsafmn;
        for (int i = t; i < nx; i += 256) x[i * incx] = x[i * incx] * f;
        __syncthreads();
        mx = 0;
        for (int i = t; i < nx; i += 256) mx = fmax(mx, absmax_(x[i * incx]));
        mx = wave_max(mx);
        __syncthreads();
        if ((t & 63) == 0) red[t >> 6] = mx;
        __syncthreads();
        mx = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
        ss = 0;
        for (int i = t; i < nx; i += 256) ss += abs2_(x[i * incx] / mx);
        ss = block_sum(ss, red);
        xnorm = mx * sqrt(ss);
        beta = beta_of<T>(alphr, alphi, xnorm);
    }
    if (t == 0) sh_scale = T(1) / (mkc_<T>(alphr, alphi) - T(beta));
    __syncthreads();
    const T s = sh_scale;
    for (int i = t; i < nx; i += 256) x[i * incx] *= s;
    if (t == 0) {
        *tau = mkc_<T>((beta - alphr) / beta, -alphi / beta);
        for (int j = 0; j < knt; ++j) beta *= safmin;
        *alpha = T(beta);
    }
}

// Householder vector of one column in the panel kernels: alpha = the diagonal entry, xs =
// sum |x|^2 of the rest.  The reference's inline larfg (src/geqr2.jl:38-73,
// src/tsqrt.jl:53-75): no reflector (tau = 0, beta = alpha) when xs = 0 and alpha is real.
template <typename T>
__device__ __forceinline__ void reflector(T alpha, typename RealOf<T>::t xs, bool make, T *sh) {
    typedef typename RealOf<T>::t R;
    const R alphr = re_(alpha), alphi = im_(alpha);
    T tv = T(0), beta = alpha, scale = T(0);
    if (make && !(xs == R(0) && alphi == R(0))) {
        const R b = -copysign(sqrt(alphr * alphr + alphi * alphi + xs), alphr);
        tv = mkc_<T>((b - alphr) / b, -alphi / b);
        scale = T(1) / (alpha - T(b));
        beta = T(b);
    }
    sh[0] = tv; sh[1] = beta; sh[2] = scale;
}

// ------------------------------------------------------------------------------------
// Panel QR (geqr2! + larft!('F','C') fused).  geqr2_kernel! semantics (src/geqr2.jl:23-105):
// plain sum-of-squares norm, beta = -sign(re alpha) sqrt(|alpha|^2 + |x|^2), tau = (beta -
// alpha)/beta (conjugated imaginary part), x /= (alpha - beta); a column of length 1 gets
// tau = 0; the in-panel update applies H^H: A -= conj(tau) v (v^H A), v(1) = 1.
// If Tm != nullptr also writes T (larft_kernel! F/C, src/larft.jl:55-126): for column i,
// T[0:i, i] = -tau_i * T[0:i,0:i] * (V^H v_i), T[i,i] = tau_i.  LAPACK's lastv trimming
// only skips exact-zero products, so the dense sums here give the same T.
// red: 16 entries of T (used as real scratch).
template <typename T>
__device__ void panel_qr_dev(int m, int n, T *__restrict__ A, i64 lda, T *__restrict__ tau, T *__restrict__ Tm,
                             i64 ldt, T *z, T *tcol, T *red, T *sh) {
    typedef typename RealOf<T>::t R;
    // z[n]: v^H A[:, j] for every panel column j; tcol[n] (only with Tm)
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int k = min(m, n);
    for (int i = 0; i < k; ++i) {
        const int len = m - i;
        T *col = A + i + (i64)i * lda;  // column i from row i
        R xs = 0;
        for (int p = 1 + t; p < len; p += 256) xs += abs2_(col[p]);
        xs = block_sum(xs, reinterpret_cast<R *>(red));
        if (t == 0) reflector<T>(col[0], xs, len > 1, sh);
        __syncthreads();
        const T tv = sh[0], beta = sh[1], scale = sh[2];
        if (tv != T(0)) {
            for (int p = 1 + t; p < len; p += 256) col[p] *= scale;
        }
        __syncthreads();
        if (t == 0) { col[0] = beta; tau[i] = tv; }
        // z_j = A[i, j] + sum_{p>=1} conj(v_p) A[i+p, j] = v^H a_j for all j != i (j > i: the
        // update; j < i: conj(z_j) = v_j^H v_i feeds T)
        const int jlo = Tm ? 0 : i + 1;
        if (tv != T(0)) {
            for (int j = jlo + w; j < n; j += 4) {
                if (j == i) continue;
                const T *cj = A + i + (i64)j * lda;
                T s = T(0);
                for (int p = 1 + lane; p < len; p += 64) s += conj_(col[p]) * cj[p];
                s = wave_sum(s);
                if (lane == 0) z[j] = s + cj[0];
            }
        }
        __syncthreads();
        if (tv != T(0)) {
            // trailing in-panel update A[i+p, j] -= conj(tau) v_p z_j (thread per row: no index
            // division, the row's factor formed once)
            const T ct = conj_(tv);
            for (int p = t; p < len; p += 256) {
                const T cv = ct * (p == 0 ? T(1) : col[p]);
                T *row = A + i + p;
                for (int j = i + 1; j < n; ++j) row[(i64)j * lda] -= cv * z[j];
            }
        }
        if (Tm) {
            // T[0:i, i] = -tau * T[0:i, 0:i] * conj(z[0:i])
            if (t < i) tcol[t] = tv != T(0) ? -tv * conj_(z[t]) : T(0);
            __syncthreads();
            if (t < i) {
                T s = T(0);
                if (tv != T(0))
                    for (int q = t; q < i; ++q) s += Tm[t + (i64)q * ldt] * tcol[q];
                Tm[t + (i64)i * ldt] = s;
            }
            if (t == 0) Tm[i + (i64)i * ldt] = tv;
        }
        __syncthreads();
    }
    // rows/columns beyond k: geqr2 leaves tau[k..] untouched (k = min(m, n)).
}

template <typename T>
__global__ __launch_bounds__(256) void panel_qr_kernel(int m, int n, T *__restrict__ A, i64 lda, T *__restrict__ tau,
                                                       T *__restrict__ Tm, i64 ldt) {
    __shared__ T red[16];
    __shared__ T sh[4];
    extern __shared__ __align__(16) unsigned char dyn_raw[];
    T *z = reinterpret_cast<T *>(dyn_raw);
    panel_qr_dev<T>(m, n, A, lda, tau, Tm, ldt, z, z + n, red, sh);
}

// ------------------------------------------------------------------------------------
// TS / TT panel: one ib-block of tsqrt_block_kernel! (src/tsqrt.jl:46-126) or
// ttqrt_block_kernel! (src/ttqrt.jl:2-79).  Column r = ii + i of [A1 (upper); A2]:
// alpha = A1[r,r], x = A2[0:mi, r] with mi = m (TS) or min(r + 1, m) (TT, A2 upper
// triangular); the block's later columns get H^H; the T column uses the first mj = m (TS)
// or min(r, m) (TT) rows of A2: T[p, r] = -tau sum_q conj(A2[q, ii+p]) A2[q, r].
template <typename T, bool TT>
__global__ __launch_bounds__(256) void ts_panel_kernel(int m, int ii, int sb, T *__restrict__ A1, i64 lda1,
                                                       T *__restrict__ A2, i64 lda2, T *__restrict__ Tm, i64 ldt,
                                                       T *__restrict__ tau) {
    typedef typename RealOf<T>::t R;
    __shared__ R red[16];
    __shared__ T z[256];     // block width sb <= 256
    __shared__ T tcol[256];
    __shared__ T sh[4];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    for (int i = 0; i < sb; ++i) {
        const int r = ii + i;
        const int mi = TT ? min(r + 1, m) : m, mj = TT ? min(r, m) : m;
        T *x = A2 + (i64)r * lda2;
        R xs = 0;
        for (int p = t; p < mi; p += 256) xs += abs2_(x[p]);
        xs = block_sum(xs, red);
        if (t == 0) reflector<T>(A1[r + (i64)r * lda1], xs, true, sh);
        __syncthreads();
        const T tv = sh[0], beta = sh[1], scale = sh[2];
        if (tv != T(0))
            for (int p = t; p < mi; p += 256) x[p] *= scale;
        __syncthreads();
        if (t == 0) { A1[r + (i64)r * lda1] = beta; tau[r] = tv; }
        // j > i: z_j = sum_{q < mi} conj(x_q) A2[q, ii+j] (the update's v^H a_j without A1's row);
        // j < i: z_j = sum_{q < mj} conj(A2[q, ii+j]) x_q (v_j^H v_r for the T column)
        if (tv != T(0)) {
            for (int j = w; j < sb; j += 4) {
                if (j == i) continue;
                const T *cj = A2 + (i64)(ii + j) * lda2;
                T s = T(0);
                if (j > i)
                    for (int p = lane; p < mi; p += 64) s += conj_(x[p]) * cj[p];
                else
                    for (int p = lane; p < mj; p += 64) s += conj_(cj[p]) * x[p];
                s = wave_sum(s);
                if (lane == 0) z[j] = s;
            }
        }
        __syncthreads();
        if (tv != T(0)) {
            const T ct = conj_(tv);
            // w_j = A1[r, r+1+j] + z_j ;  A2[0:mi, r+1+j] -= conj(tau) x w ;  A1[r, r+1+j] -= conj(tau) w
            for (int j = i + 1 + t; j < sb; j += 256) z[j] += A1[r + (i64)(ii + j) * lda1];
            __syncthreads();
            for (int p = t; p < mi; p += 256) {
                const T cx = ct * x[p];
                T *row = A2 + p;
                for (int j = i + 1; j < sb; ++j) row[(i64)(ii + j) * lda2] -= cx * z[j];
            }
            for (int j = i + 1 + t; j < sb; j += 256) A1[r + (i64)(ii + j) * lda1] -= ct * z[j];
        }
        // T column: T[p, r] = -tau * sum_q T[p, ii+q] z_q  (p < i), T[i, r] = tau
        if (t < i) tcol[t] = tv != T(0) ? -tv * z[t] : T(0);
        __syncthreads();
        if (t < i) {
            T s = T(0);
            if (tv != T(0))
                for (int q = t; q < i; ++q) s += Tm[t + (i64)(ii + q) * ldt] * tcol[q];
            Tm[t + (i64)r * ldt] = s;
        }
        if (t == 0) Tm[i + (i64)r * ldt] = tv;
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------------
// larft!('F', storev) standalone (src/larft.jl:55-126): T upper k x k with
// T[0:i,i] = -tau_i T[0:i,0:i] (V^H v_i) (storev 'C', V n x k) or -tau_i T[0:i,0:i] (V v_i^H)
// over the rows of V (storev 'R', V k x n).
template <typename T, bool ROW>
__global__ __launch_bounds__(256) void larft_kernel(int n, int k, const T *__restrict__ V, i64 ldv,
                                                    const T *__restrict__ tau, T *__restrict__ Tm, i64 ldt) {
    extern __shared__ __align__(16) unsigned char dyn_raw[];
    T *z = reinterpret_cast<T *>(dyn_raw);  // z[k]
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    for (int i = 0; i < k; ++i) {
        const T ti = tau[i];
        // C: z_p = conj(V[i, p]) + sum_{r > i} conj(V[r, p]) V[r, i];  R: z_p = V[p, i] + sum_{r > i} V[p, r] conj(V[i, r])
        for (int p = w; p < i; p += 4) {
            T s = T(0);
            for (int r = i + 1 + lane; r < n; r += 64) {
                if (ROW) s += V[p + (i64)r * ldv] * conj_(V[i + (i64)r * ldv]);
                else s += conj_(V[r + (i64)p * ldv]) * V[r + (i64)i * ldv];
            }
            s = wave_sum(s);
            const T head = ROW ? V[p + (i64)i * ldv] : conj_(V[i + (i64)p * ldv]);
            if (lane == 0) z[p] = ti == T(0) ? T(0) : -ti * (head + s);
        }
        __syncthreads();
        for (int p = t; p < i; p += 256) {
            T s = T(0);
            if (ti != T(0))
                for (int q = p; q < i; ++q) s += Tm[p + (i64)q * ldt] * z[q];
            Tm[p + (i64)i * ldt] = s;
        }
        if (t == 0) Tm[i + (i64)i * ldt] = ti;
        __syncthreads();
    }
}


// ------------------------------------------------------------------------------------
// Strip kernels.  View element (r, c) of a matrix stored at p with leading dimension ld:
// TR ? conj(p[c + r*ld]) : p[r + c*ld]  (the conjugate-transposed view for right-side calls).
template <typename T, bool TR> __device__ __forceinline__ T vget(const T *p, i64 ld, int r, int c) {
    return TR ? conj_(p[c + (i64)r * ld]) : p[r + (i64)c * ld];
}
template <typename T, bool TR> __device__ __forceinline__ void vsub(T *p, i64 ld, int r, int c, T s) {
    if (TR) p[c + (i64)r * ld] -= conj_(s);
    else p[r + (i64)c * ld] -= s;
}

constexpr int CW = 16;    // strip width (columns of the left-form view)
constexpr int KBM = 64;   // reflectors per applied sub-block
constexpr int RC = 64;    // rows per staged chunk

// Apply op(H) = I - Vh op(T) Vh^H of kb reflectors from the left to rows [Top (kb); Bot (mb)]
// of the strip c0..c0+cw-1 (left-form views, TRC = conjugate-transposed storage).
//   GE: Vh = [V1 unit lower (kb x kb) ; V2 (mb x kb)]   (larfb!, src/larfb.jl:57-98)
//   TS: Vh = [I ; V2 (mb x kb)]                         (parfb!, src/parfb.jl:121-129,189)
// W0 = Vh^H C, W = op(T) W0 (op(T) = T for notrans, T^H otherwise), C -= Vh W.
// TS with l > 0 (ttmqr!, parfb! with a triangular V): V2 is pentagonal, its last l rows the
// first l rows of a upper triangle; column j here is column jofs + j of that block, and
// V2[r, j] = 0 for r >= mb - l and r - (mb - l) > jofs + j (src/pamm.jl:104-127,306-325).
template <typename T, bool GE, bool TRC>
__device__ void apply_strip(int kb, int mb, const T *__restrict__ V1, const T *__restrict__ V2, i64 ldv,
                            const T *__restrict__ Tb, i64 ldt, bool notrans, T *Top, i64 ldtop, T *Bot, i64 ldbot,
                            int c0, int cw, T *sV, T *sC, T *sW0, T *sW, int l = 0, int jofs = 0) {
    const int t = threadIdx.x;
    const int oc = t % CW, oj = t / CW;  // thread -> W0 entries (oj + 16 q, oc)
    const int rz = mb - l;               // first row of the triangular part (l > 0)
    T acc[KBM / 16];
#pragma unroll
    for (int q = 0; q < KBM / 16; ++q) acc[q] = T(0);
    for (int r0 = 0; r0 < mb; r0 += RC) {
        const int rc = min(RC, mb - r0);
        __syncthreads();
        for (int idx = t; idx < rc * kb; idx += 256) {
            const int r = idx % rc, j = idx / rc;
            const bool zero = !GE && l > 0 && r0 + r >= rz && r0 + r - rz > jofs + j;
            sV[j * RC + r] = zero ? T(0) : conj_(V2[(r0 + r) + (i64)j * ldv]);
        }
        for (int idx = t; idx < rc * CW; idx += 256) {
            int r, c;
            if (TRC) { c = idx % CW; r = idx / CW; } else { r = idx % rc; c = idx / rc; }
            sC[c * RC + r] = (c < cw) ? vget<T, TRC>(Bot, ldbot, r0 + r, c0 + c) : T(0);
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < KBM / 16; ++q) {
            const int j = oj + 16 * q;
            if (j < kb) {
                T s = T(0);
                for (int r = 0; r < rc; ++r) s += sV[j * RC + r] * sC[oc * RC + r];
                acc[q] += s;
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < KBM / 16; ++q) {
        const int j = oj + 16 * q;
        if (j < kb) {
            T s = acc[q];
            if (oc < cw) {
                s += vget<T, TRC>(Top, ldtop, j, c0 + oc);
                if (GE)
                    for (int r = j + 1; r < kb; ++r)
                        s += conj_(V1[r + (i64)j * ldv]) * vget<T, TRC>(Top, ldtop, r, c0 + oc);
            }
            sW0[j * CW + oc] = s;
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < KBM / 16; ++q) {
        const int j = oj + 16 * q;
        if (j < kb) {
            T s = T(0);
            if (notrans) for (int p = j; p < kb; ++p) s += Tb[j + (i64)p * ldt] * sW0[p * CW + oc];
            else         for (int p = 0; p <= j; ++p) s += conj_(Tb[p + (i64)j * ldt]) * sW0[p * CW + oc];
            sW[j * CW + oc] = s;
        }
    }
    __syncthreads();
    for (int idx = t; idx < kb * cw; idx += 256) {
        int r, c;
        if (TRC) { c = idx % cw; r = idx / cw; } else { r = idx % kb; c = idx / kb; }
        T s = sW[r * CW + c];
        if (GE)
            for (int j = 0; j < r; ++j) s += V1[r + (i64)j * ldv] * sW[j * CW + c];
        vsub<T, TRC>(Top, ldtop, r, c0 + c, s);
    }
    for (int idx = t; idx < mb * cw; idx += 256) {
        int r, c;
        if (TRC) { c = idx % cw; r = idx / cw; } else { r = idx % mb; c = idx / mb; }
        T s = T(0);
        // pentagonal V2: row r >= rz has entries only in columns jofs + j >= r - rz
        const int jlo = (!GE && l > 0 && r >= rz) ? max(0, r - rz - jofs) : 0;
        for (int j = jlo; j < kb; ++j) s += V2[r + (i64)j * ldv] * sW[j * CW + c];
        vsub<T, TRC>(Bot, ldbot, r, c0 + c, s);
    }
    __syncthreads();
}

// unmqr!/larfb!/geqrt! trailing update (left form): k reflectors of V (unit lower, ldv) in
// blocks of ib; block i acts on rows i..m-1 of C (m x ncols view).
template <typename T, bool TRC>
__global__ __launch_bounds__(256) void ge_apply_kernel(int m, int ncols, int k, int ib, const T *__restrict__ V,
                                                       i64 ldv, const T *__restrict__ Tm, i64 ldt, bool notrans,
                                                       bool forward, T *C, i64 ldc) {
    __shared__ T sV[KBM * RC];
    __shared__ T sC[CW * RC];
    __shared__ T sW0[KBM * CW];
    __shared__ T sW[KBM * CW];
    const int c0 = blockIdx.x * CW;
    const int cw = min(CW, ncols - c0);
    const int nblk = (k + ib - 1) / ib;
    for (int bb = 0; bb < nblk; ++bb) {
        const int b = forward ? bb : nblk - 1 - bb;
        const int i = b * ib;
        const int kb = min(ib, k - i);
        const int nsub = (kb + KBM - 1) / KBM;
        for (int ss = 0; ss < nsub; ++ss) {
            const int s = forward ? ss : nsub - 1 - ss;
            const int j0 = s * KBM, kk = min(KBM, kb - j0);
            const int row = i + j0;  // first row touched by this sub-block
            const T *V1 = V + row + (i64)row * ldv;
            const T *Tb = Tm + j0 + (i64)(i + j0) * ldt;
            T *Top = TRC ? C + (i64)row * ldc : C + row;
            T *Bot = TRC ? C + (i64)(row + kk) * ldc : C + row + kk;
            apply_strip<T, true, TRC>(kk, m - row - kk, V1, V1 + kk, ldv, Tb, ldt, notrans, Top, ldc, Bot, ldc, c0, cw,
                                      sV, sC, sW0, sW);
        }
    }
}

// tsmqr! / ttmqr! (left form): k reflectors of V (m2 x k) in blocks of ib; block i updates
// rows i..i+kb-1 of A1 and the first mb rows of A2 (parfb!, src/tsmqr.jl:123-145,
// src/ttmqr.jl:64-92).  mode 0: mb = m2 and pentagon order l = lfix (tsmqr!: 0; one
// parfb! block of order lfix); mode 1 (ttmqr!): mb = min(i + kb, m2), l = min(kb, m2 - i - 1).
template <typename T, bool TRC>
__global__ __launch_bounds__(256) void ts_apply_kernel(int m2, int ncols, int k, int ib, const T *__restrict__ V,
                                                       i64 ldv, const T *__restrict__ Tm, i64 ldt, bool notrans,
                                                       bool forward, T *A1, i64 lda1, T *A2, i64 lda2, int mode,
                                                       int lfix) {
    __shared__ T sV[KBM * RC];
    __shared__ T sC[CW * RC];
    __shared__ T sW0[KBM * CW];
    __shared__ T sW[KBM * CW];
    const int c0 = blockIdx.x * CW;
    const int cw = min(CW, ncols - c0);
    const int nblk = (k + ib - 1) / ib;
    for (int bb = 0; bb < nblk; ++bb) {
        const int b = forward ? bb : nblk - 1 - bb;
        const int i = b * ib;
        const int kb = min(ib, k - i);
        const int mb = mode == 1 ? min(i + kb, m2) : m2;
        const int l = mode == 1 ? min(kb, max(0, m2 - i - 1)) : lfix;
        const int nsub = (kb + KBM - 1) / KBM;
        for (int ss = 0; ss < nsub; ++ss) {
            const int s = forward ? ss : nsub - 1 - ss;
            const int j0 = s * KBM, kk = min(KBM, kb - j0);
            const int row = i + j0;
            const T *Vb = V + (i64)row * ldv;
            const T *Tb = Tm + j0 + (i64)(i + j0) * ldt;
            T *Top = TRC ? A1 + (i64)row * lda1 : A1 + row;
            apply_strip<T, false, TRC>(kk, mb, nullptr, Vb, ldv, Tb, ldt, notrans, Top, lda1, A2, lda2, c0, cw, sV, sC,
                                       sW0, sW, l, j0);
        }
    }
}

// ------------------------------------------------------------------------------------
// General parfb! / pamm! (src/parfb.jl:44-191, src/pamm.jl:49-473): one block of k
// reflectors in any direct / storev with pentagon order l and an arbitrary T (parfb! applies
// I - Vc T Vc^H exactly; no sub-block splitting, so T need not come from larft!).  Left
// form: A1 (k x ncols), A2 (rows x ncols); right-side calls use conjugate-transposed views.
// Column form of V: Vc(r, j) = V(r, j) (storev 'C') or conj(V(j, r)) (storev 'R'), with the
// structural zeros of direct 'F' (last l rows: first l rows of an upper triangle) or 'B'
// (first l rows: last l rows of a lower triangle).  T upper ('F') or lower ('B').
// Phases (flags): PW: W = A1 + Vc^H A2;  PT: W = op(T) W, A1 -= W;  PA: A2 -= Vc W.  W lives in
// LDS, or in the caller's W (pamm!'s 'W' output / 'A' input; Wg != nullptr).
constexpr int PCW = 8;       // strip width of the general kernel
constexpr int PKMAX = 256;   // largest k of one general block

template <typename T> __device__ __forceinline__ T pent_vc(const T *V, i64 ldv, bool row, int r, int j) {
    return row ? conj_(V[j + (i64)r * ldv]) : V[r + (i64)j * ldv];
}
__device__ __forceinline__ bool pent_nz(bool fwd, int rows, int k, int l, int r, int j) {
    return fwd ? !(r >= rows - l && r - (rows - l) > j) : !(r < l && j > k - l + r);
}

enum { PENT_W = 1, PENT_T = 2, PENT_A = 4 };

template <typename T, bool TRC>
__global__ __launch_bounds__(256) void pent_apply_kernel(int rows, int ncols, int k, int l, bool fwd, bool row,
                                                         bool notrans, int phases, const T *__restrict__ V, i64 ldv,
                                                         const T *__restrict__ Tm, i64 ldt, T *A1, i64 lda1, T *A2,
                                                         i64 lda2, T *Wg, i64 ldw) {
    __shared__ T sW0[PKMAX * PCW];
    __shared__ T sW[PKMAX * PCW];
    const int t = threadIdx.x;
    const int c0 = blockIdx.x * PCW;
    const int cw = min(PCW, ncols - c0);
    // W0 = A1 + Vc^H A2, or the caller's W
    for (int idx = t; idx < k * PCW; idx += 256) {
        const int j = idx / PCW, c = idx % PCW;
        T s = T(0);
        if (c < cw) {
            if (phases & PENT_W) {
                s = vget<T, TRC>(A1, lda1, j, c0 + c);
                for (int r = 0; r < rows; ++r)
                    if (pent_nz(fwd, rows, k, l, r, j)) s += conj_(pent_vc(V, ldv, row, r, j)) * vget<T, TRC>(A2, lda2, r, c0 + c);
            } else {
                s = vget<T, TRC>(Wg, ldw, j, c0 + c);
            }
        }
        sW0[j * PCW + c] = s;
    }
    __syncthreads();
    if (phases & PENT_T) {
        for (int idx = t; idx < k * PCW; idx += 256) {
            const int i = idx / PCW, c = idx % PCW;
            T s = T(0);
            for (int j = 0; j < k; ++j) {
                // element (i, j) of op(T) within T's triangle
                const int a = notrans ? i : j, b = notrans ? j : i;
                if (fwd ? a > b : a < b) continue;
                const T tv = Tm[a + (i64)b * ldt];
                s += (notrans ? tv : conj_(tv)) * sW0[j * PCW + c];
            }
            sW[i * PCW + c] = s;
        }
        __syncthreads();
        for (int idx = t; idx < k * cw; idx += 256) {
            const int i = idx / cw, c = idx % cw;
            vsub<T, TRC>(A1, lda1, i, c0 + c, sW[i * PCW + c]);
        }
    } else {
        for (int idx = t; idx < k * PCW; idx += 256) sW[idx] = sW0[idx];
        __syncthreads();
    }
    if ((phases & PENT_W) && !(phases & PENT_T) && Wg) {
        // pamm! 'W': the product is the output
        for (int idx = t; idx < k * cw; idx += 256) {
            const int j = idx / cw, c = idx % cw;
            const T v = sW[j * PCW + c];
            if (TRC) Wg[c0 + c + (i64)j * ldw] = conj_(v);
            else Wg[j + (i64)(c0 + c) * ldw] = v;
        }
    }
    if (phases & PENT_A) {
        for (int idx = t; idx < rows * cw; idx += 256) {
            int r, c;
            if (TRC) { c = idx % cw; r = idx / cw; } else { r = idx % rows; c = idx / rows; }
            T s = T(0);
            for (int j = 0; j < k; ++j)
                if (pent_nz(fwd, rows, k, l, r, j)) s += pent_vc(V, ldv, row, r, j) * sW[j * PCW + c];
            vsub<T, TRC>(A2, lda2, r, c0 + c, s);
        }
    }
}

// ------------------------------------------------------------------------------------
// pemv! (src/pemv.jl:39-181): y = alpha op(A) x + beta y over a pentagonal A (m x n).
// COLS (storev 'C', trans 'T'/'C'): y has n entries, column j's rows r >= m - l with
// r - (m - l) > j are zero (j < l); !COLS (storev 'R', trans 'N'): y has m entries, row r's
// columns c >= n - l with c - (n - l) > r are zero (r < l).  beta = 0 does not read y.
template <typename T, bool COLS>
__global__ __launch_bounds__(256) void pemv_kernel(int m, int n, int l, bool conjA, T alpha, const T *__restrict__ A,
                                                   i64 lda, const T *__restrict__ x, T beta, T *__restrict__ y) {
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    if (COLS) {
        // one wave per output entry: a column of A against x (coalesced down the column)
        const int j = blockIdx.x * 4 + w;
        if (j >= n) return;
        const int rlim = (j < l) ? m - l + j + 1 : m;
        T s = T(0);
        for (int r = lane; r < rlim; r += 64) {
            const T a = A[r + (i64)j * lda];
            s += (conjA ? conj_(a) : a) * x[r];
        }
        s = wave_sum(s);
        if (lane == 0) y[j] = alpha * s + (beta == T(0) ? T(0) : beta * y[j]);
    } else {
        // one thread per output entry: threads of a wave read consecutive rows of a column
        const int r = blockIdx.x * 256 + t;
        if (r >= m) return;
        const int clim = (r < l) ? n - l + r + 1 : n;
        T s = T(0);
        for (int c = 0; c < clim; ++c) s += A[r + (i64)c * lda] * x[c];
        y[r] = alpha * s + (beta == T(0) ? T(0) : beta * y[r]);
    }
}

// ------------------------------------------------------------------------------------
// Host launchers (no validation; used by the entry points below and by the tile driver).
// ------------------------------------------------------------------------------------
// Register-resident panels of the per-call geqrt!/tsqrt! (real element types, panel rows
// mr <= 256, block width sb <= SB in {16, 32}): thread t holds row t of the block's columns.
// A column step is two workgroup reductions (the norm; the products of the new reflector with
// the block's later columns, folded across the wave with v_permlane / DPP and across the four
// waves through LDS); the reflector is formed identically in every thread.  GE form (geqr2!
// inside geqrt!, src/geqr2.jl:23-105): the block's own rows, v_i = 1 on row i.  TS form
// (tsqrt_block_kernel!, src/tsqrt.jl:46-126): alpha = A1[i, i] and the block's A1 row i is held
// by thread i, the reflector's lower part is all of A2's rows.  T (larft!('F','C'),
// src/larft.jl:55-126) from the Gram matrix of the block's V (one MFMA product) after the last
// column: T[t, i] = sum_{q=t}^{i-1} T[t, q] (-tau_i G[q, i]), T[i, i] = tau_i.

// Transposing butterfly over the wave: afterwards lane l holds the wave's sum of p[idx(l)],
// idx(l) = (l >> 2) & 15 for SB = 16, (l >> 1) & 31 for SB = 32.
template <typename T, int SB> __device__ __forceinline__ T fold_values(T (&p)[SB], int lane) {
    if constexpr (SB == 32) {
#pragma unroll
        for (int k = 0; k < 16; ++k) p[k] = fold32(p[k], p[k + 16]);
#pragma unroll
        for (int k = 0; k < 8; ++k) p[k] = fold16(p[k], p[k + 8]);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const bool hi = lane & 8;
            const T send = hi ? p[k] : p[k + 4];
            p[k] = (hi ? p[k + 4] : p[k]) + dpp_mov<DPP_ROW_ROR8>(send);
        }
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const bool hi = lane & 4;
            const T send = hi ? p[k] : p[k + 2];
            const T from_up = dpp_mov<DPP_ROW_SHL4>(send), from_down = dpp_mov<DPP_ROW_SHR4>(send);
            p[k] = (hi ? p[k + 2] : p[k]) + (hi ? from_down : from_up);
        }
        {
            const bool hi = lane & 2;
            const T send = hi ? p[0] : p[1];
            p[0] = (hi ? p[1] : p[0]) + dpp_mov<DPP_QUAD_X2>(send);
        }
        return p[0] + dpp_mov<DPP_QUAD_X1>(p[0]);
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k) p[k] = fold32(p[k], p[k + 8]);
#pragma unroll
        for (int k = 0; k < 4; ++k) p[k] = fold16(p[k], p[k + 4]);
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const bool hi = lane & 8;
            const T send = hi ? p[k] : p[k + 2];
            p[k] = (hi ? p[k + 2] : p[k]) + dpp_mov<DPP_ROW_ROR8>(send);
        }
        {
            const bool hi = lane & 4;
            const T send = hi ? p[0] : p[1];
            const T from_up = dpp_mov<DPP_ROW_SHL4>(send), from_down = dpp_mov<DPP_ROW_SHR4>(send);
            p[0] = (hi ? p[1] : p[0]) + (hi ? from_down : from_up);
        }
        p[0] += dpp_mov<DPP_QUAD_X2>(p[0]);
        return p[0] + dpp_mov<DPP_QUAD_X1>(p[0]);
    }
}

template <typename T, bool TS, int SB, int I>
__device__ __forceinline__ void rpc_column(T (&a)[SB], T (&a1)[SB], T (&taus)[SB], int mr, int sb, T *nred, T *zred) {
    constexpr int i = I;
    if (i >= sb) return;
    const int t = threadIdx.x, lane = t & 63, w = __builtin_amdgcn_readfirstlane(t >> 6);
    // sum of squares of the reflector's lower part (rows >= mr hold zeros) and alpha
    T xs = (TS || t > i) ? a[i] * a[i] : T(0);
    xs = wave_sum_fast(xs);
    T *nr = nred + 8 * (i & 1);
    if (lane == 0) nr[w] = xs;
    if (t == i) nr[4] = TS ? a1[i] : a[i];
    lds_barrier();
    const T xsum = (nr[0] + nr[1]) + (nr[2] + nr[3]);
    const T alpha = nr[4];
    T tau = T(0), beta = alpha, scale = T(0);
    if ((TS || mr - i > 1) && xsum != T(0)) {
        beta = -copysign(sqrt(alpha * alpha + xsum), alpha);
        tau = (beta - alpha) / beta;
        scale = T(1) / (alpha - beta);
    }
    taus[i] = tau;
    if (tau != T(0)) {
        T v;
        if (TS) {
            v = a[i] * scale;
            a[i] = v;
        } else {
            v = t > i ? a[i] * scale : (t == i ? T(1) : T(0));
            if (t > i) a[i] = v;
        }
        T p[SB];
#pragma unroll
        for (int j = 0; j < SB; ++j) p[j] = T(0);
#pragma unroll
        for (int j = i + 1; j < SB; ++j) p[j] = v * a[j] + ((TS && t == i) ? a1[j] : T(0));
        const T zsum = fold_values<T, SB>(p, lane);
        constexpr int SH = SB == 32 ? 1 : 2;
        if ((lane & ((1 << SH) - 1)) == 0) zred[((lane >> SH) & (SB - 1)) * 4 + w] = zsum;
        lds_barrier();
        T u[SB];
#pragma unroll
        for (int j = i + 1; j < SB; ++j)
            u[j] = tau * ((zred[4 * j] + zred[4 * j + 1]) + (zred[4 * j + 2] + zred[4 * j + 3]));
#pragma unroll
        for (int j = i + 1; j < SB; ++j) a[j] -= v * u[j];
        if (TS && t == i) {
#pragma unroll
            for (int j = i + 1; j < SB; ++j) a1[j] -= u[j];
        }
    }
    if (t == i) {
        if (TS) a1[i] = beta;
        else a[i] = beta;
    }
}

template <typename T, bool TS, int SB, int... I>
__device__ __forceinline__ void rpc_columns(T (&a)[SB], T (&a1)[SB], T (&taus)[SB], int mr, int sb, T *nred, T *zred,
                                            std::integer_sequence<int, I...>) {
    (rpc_column<T, TS, SB, I>(a, a1, taus, mr, sb, nred, zred), ...);
}

// One ib-block panel: GE form on P (mr x sb, the block's rows), TS form on P = the block's
// columns of A2 (mr x sb) with A1 = the block's diagonal block of A1 (upper triangle read and
// written).  Writes the factors back, tau[0:sb] and T (sb x sb, zeros below the diagonal).
template <typename T, bool TS, int SB>
__global__ __launch_bounds__(256) void rpc_panel_kernel(int mr, int sb, T *P_, i64 ldp, T *A1_, i64 lda1, T *Tg_,
                                                        i64 ldt, T *tau_) {
    static_assert(!IsCplx<T>::v, "real element types only");
    constexpr int VLD = SB + 1;
    __shared__ T zred[SB * 4];
    __shared__ T nred[16];
    __shared__ T Gs[SB * SB];
    __shared__ T Vs[256 * VLD];
    const GPtr<T> P = to_global(P_), A1 = to_global(A1_), Tg = to_global(Tg_), tau = to_global(tau_);
    const int t = threadIdx.x, lane = t & 63, lr = lane & 15, lk = lane >> 4;
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);
    T a[SB], a1[SB], taus[SB];
#pragma unroll
    for (int j = 0; j < SB; ++j) {
        const T x = P[min(t, mr - 1) + (i64)min(j, sb - 1) * ldp];
        a[j] = (t < mr && j < sb) ? x : T(0);
        if (TS) {
            const T y = A1[min(t, sb - 1) + (i64)min(j, sb - 1) * lda1];
            a1[j] = (t < sb && j < sb && j >= t) ? y : T(0);
        }
    }
    rpc_columns<T, TS, SB>(a, a1, taus, mr, sb, nred, zred, std::make_integer_sequence<int, SB>{});
    if (t < mr) {
#pragma unroll
        for (int j = 0; j < SB; ++j)
            if (j < sb) P[t + (i64)j * ldp] = a[j];
    }
    if (TS && t < sb) {
#pragma unroll
        for (int j = 0; j < SB; ++j)
            if (j < sb && j >= t) A1[t + (i64)j * lda1] = a1[j];
    }
    if (t < sb) {
        T tv = T(0);
#pragma unroll
        for (int j = 0; j < SB; ++j)
            if (j == t) tv = taus[j];
        tau[t] = tv;
    }
    // V (GE: unit lower trapezoid; TS: A2's part) to LDS, then G = V^T V on MFMA
#pragma unroll
    for (int j = 0; j < SB; ++j)
        Vs[t * VLD + j] = (t < mr && j < sb) ? (TS ? a[j] : (t > j ? a[j] : (t == j ? T(1) : T(0)))) : T(0);
    lds_barrier();
    constexpr int NGR = SB / 16;
    if (w < NGR * NGR) {
        typedef typename Acc<T>::t AT;
        const int g1 = w / NGR, g2 = w % NGR;
        AT acc = acc_zero<T>();
        const int nkt = (mr + 15) >> 4;
        for (int s = 0; s < nkt; ++s) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int r = 16 * s + acc_row<T>(lk, q);
                mfma16(acc, Vs[r * VLD + 16 * g1 + lr], Vs[r * VLD + 16 * g2 + lr]);
            }
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) Gs[(16 * g1 + acc_row<T>(lk, e)) * SB + 16 * g2 + lr] = acc[e];
    }
    lds_barrier();
    if (t < sb) {
        T Trow[SB];
#pragma unroll
        for (int i = 0; i < SB; ++i) {
            T s = T(0);
            if (i < sb && taus[i] != T(0)) {
                if (t == i) s = taus[i];
#pragma unroll
                for (int q = 0; q < i; ++q)
                    if (q >= t) s += Trow[q] * (-taus[i] * Gs[q * SB + i]);
            }
            Trow[i] = s;
        }
#pragma unroll
        for (int j = 0; j < SB; ++j)
            if (j < sb) Tg[t + (i64)j * ldt] = j >= t ? Trow[j] : T(0);
    }
}

// The register panel takes real element types with at most 256 rows and 32 columns; it returns
// false (nothing launched) otherwise.
template <typename T>
static bool launch_rpc_panel(bool ts, i64 mr, i64 sb, T *P, i64 ldp, T *A1, i64 lda1, T *Tg, i64 ldt, T *tau,
                             hipStream_t s, hipError_t *err) {
    if constexpr (IsCplx<T>::v) {
        return false;
    } else {
        if (mr <= 0 || mr > 256 || sb <= 0 || sb > 32 || !Tg) return false;
        if (ts) {
            if (sb <= 16) rpc_panel_kernel<T, true, 16><<<1, 256, 0, s>>>((int)mr, (int)sb, P, ldp, A1, lda1, Tg, ldt, tau);
            else rpc_panel_kernel<T, true, 32><<<1, 256, 0, s>>>((int)mr, (int)sb, P, ldp, A1, lda1, Tg, ldt, tau);
        } else {
            if (sb <= 16) rpc_panel_kernel<T, false, 16><<<1, 256, 0, s>>>((int)mr, (int)sb, P, ldp, A1, lda1, Tg, ldt, tau);
            else rpc_panel_kernel<T, false, 32><<<1, 256, 0, s>>>((int)mr, (int)sb, P, ldp, A1, lda1, Tg, ldt, tau);
        }
        *err = hipGetLastError();
        return true;
    }
}

template <typename T>
hipError_t launch_panel_qr(i64 m, i64 n, T *A, i64 lda, T *tau, T *Tm, i64 ldt, hipStream_t s) {
    if (m <= 0 || n <= 0) return hipSuccess;
    size_t sh = (size_t)n * sizeof(T) * (Tm ? 2 : 1);
    ProfScope prof(PK_PANEL, s);
    hipError_t e = hipSuccess;
    if (launch_rpc_panel<T>(false, m, n, A, lda, (T *)nullptr, 0, Tm, ldt, tau, s, &e)) return e;
    panel_qr_kernel<T><<<1, 256, sh, s>>>((int)m, (int)n, A, lda, tau, Tm, ldt);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_ts_panel(bool tt, i64 m, i64 ii, i64 sb, T *A1, i64 lda1, T *A2, i64 lda2, T *Tm, i64 ldt, T *tau,
                           hipStream_t s) {
    ProfScope prof(PK_TS_PANEL, s);
    hipError_t e = hipSuccess;
    if (!tt && launch_rpc_panel<T>(true, m, sb, A2 + ii * lda2, lda2, A1 + ii + ii * lda1, lda1, Tm + ii * ldt, ldt,
                                   tau + ii, s, &e))
        return e;
    if (tt) ts_panel_kernel<T, true><<<1, 256, 0, s>>>((int)m, (int)ii, (int)sb, A1, lda1, A2, lda2, Tm, ldt, tau);
    else ts_panel_kernel<T, false><<<1, 256, 0, s>>>((int)m, (int)ii, (int)sb, A1, lda1, A2, lda2, Tm, ldt, tau);
    return hipGetLastError();
}

// Left-form GE apply on an m x ncols view (trc: C stored conjugate-transposed).
// ------------------------------------------------------------------------------------
__device__ __forceinline__ int cc_of(int c, int ncols) { return c < ncols ? c : ncols - 1; }

// MFMA block-reflector apply of the per-call path (left side, real types): for each ib-block
// (forward or backward), W = V_b^T C (+ the block's rows of A1 for the TS form), W = op(T_b) W,
// A1 rows -= W (TS form), C -= V_b W: larfb!/unmqr! (GE: V_b the unit lower trapezoid stored in
// the tile, src/larfb.jl:57-98, src/unmqr.jl:127-148) and tsmqr!/parfb! with l = 0 (TS: V_b full,
// src/tsmqr.jl:104-145, src/parfb.jl:121-129, src/pamm.jl:104-127,306-325).  A workgroup owns 16
// columns of C for all blocks; wave w keeps rows 16 s + acc_row(lk, q) (s = w, w + 4, ...) of its
// columns in registers from the first block to the last, in the B layout of V^T C and the D
// layout of V W.  Rows <= 64 NU, blocks of at most 16 NG reflectors.  Global memory is read in
// the prologue (C, the A1 rows, T) and, one block ahead, for V_b, so a block's phases wait only on
// LDS, MFMA and barriers.  Every product is unconditional (rows past `rows` meet zero rows of V_b,
// reflectors past kb zero columns of V_b and zero rows of T_b), so the accumulators never cross a
// branch.
constexpr int MA_KB = 64;      // widest block
constexpr int MA_VCAP = 8448;  // V_b elements staged in LDS (row stride kb + 1)
constexpr int MA_VX = 16;      // V_b elements per thread prefetched a block ahead (rows x kb <= 4096)
constexpr int MA_TCAP = 4096;  // T elements staged in LDS at the start (ib x k)
constexpr int MA_AX = 8;       // A1 elements per thread prefetched at the start (k <= 128)
template <typename T, bool TS, int NU, int NG>
__global__ __launch_bounds__(256) void mfma_apply_kernel(int rows, int ncols, int k, int ib, const T *__restrict__ V_,
                                                         i64 ldv, const T *__restrict__ Tm_, i64 ldt, bool notrans,
                                                         bool forward, T *__restrict__ A1_, i64 lda1, T *__restrict__ C_,
                                                         i64 ldc) {
    typedef typename Acc<T>::t AT;
    __shared__ T Wp[4 * MA_KB * 16];
    __shared__ T Ws[MA_KB * 16];
    __shared__ T Wt[MA_KB * 16];
    __shared__ T Vl[MA_VCAP];
    __shared__ T Tl[MA_TCAP];
    const GPtr<const T> V = to_global(V_), Tm = to_global(Tm_);
    const GPtr<T> A1 = to_global(A1_), C = to_global(C_);
    const int t = threadIdx.x, lane = t & 63, lr = lane & 15, lk = lane >> 4;
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);
    const int c0 = blockIdx.x * 16;
    const int nkt = (rows + 15) >> 4;
    const int cc = min(c0 + lr, ncols - 1);
    const int nblk = (k + ib - 1) / ib, kbmax = min(ib, k);
    // V_b[r][j] (block origin column i, width kb): zero outside 0 <= r < rows, 0 <= j < kb; the GE
    // form is unit lower trapezoidal below row i + j.  The load is unconditional at a clamped address.
    auto vb = [&](int r, int j, int i, int kb) -> T {
        const T x = V[min(r, rows - 1) + (i64)(i + min(j, kb - 1)) * ldv];
        if (r >= rows || j >= kb) return T(0);
        if (TS) return x;
        return r > i + j ? x : (r == i + j ? T(1) : T(0));
    };
    auto blk_i = [&](int bb) { return (forward ? bb : nblk - 1 - bb) * ib; };
    // prologue loads, all in flight together: this workgroup's columns of C
    T cv[NU][4];
#pragma unroll
    for (int u = 0; u < NU; ++u) {
        const int s = w + 4 * u;
#pragma unroll
        for (int q = 0; q < 4; ++q) cv[u][q] = C[min(16 * s + acc_row<T>(lk, q), rows - 1) + (i64)cc * ldc];
    }
    // the A1 rows of every block (TS form), element (row (t >> 4) + 16 x, column t & 15)
    const bool a1pre = TS && k <= 16 * MA_AX && ib % 16 == 0;   // block rows start on a multiple of 16
    T a1r[MA_AX];
    if (a1pre) {
#pragma unroll
        for (int x = 0; x < MA_AX; ++x)
            a1r[x] = A1[min((t >> 4) + 16 * x, k - 1) + (i64)cc_of(c0 + (t & 15), ncols) * lda1];
    }
    // T (ib x k) to LDS
    const bool tpre = kbmax * k <= MA_TCAP;
    if (tpre)
        for (int col = t; col < k; col += 256)
            for (int p0 = 0; p0 < kbmax; p0 += 16) {
                T x[16];   // 16 loads in flight, then the stores
#pragma unroll
                for (int u = 0; u < 16; ++u) x[u] = Tm[min(p0 + u, kbmax - 1) + (i64)col * ldt];
#pragma unroll
                for (int u = 0; u < 16; ++u)
                    if (p0 + u < kbmax) Tl[(p0 + u) * k + col] = x[u];
            }
    // V_b of the first block into registers: element e = t + 256 x of the block, row e % rows,
    // column e / rows
    const bool vpre = NU <= 4 && rows * kbmax <= 256 * MA_VX && rows * (kbmax + 1) <= MA_VCAP;
    int vr[MA_VX], vj[MA_VX];
    T vx[MA_VX];
    if (vpre) {
#pragma unroll
        for (int x = 0; x < MA_VX; ++x) {
            const int e = t + 256 * x;
            vj[x] = e / rows;
            vr[x] = e - vj[x] * rows;
            vx[x] = vb(vr[x], vj[x], blk_i(0), kbmax);
        }
    }
    for (int bb = 0; bb < nblk; ++bb) {
        const int i = blk_i(bb);
        const int kb = min(ib, k - i);
        const int vld = kb + 1;
        const bool staged = rows * vld <= MA_VCAP;   // uniform
        if (bb > 0) lds_barrier();   // the previous block's readers of Vl and Wt are done
        if (vpre) {
#pragma unroll
            for (int x = 0; x < MA_VX; ++x)
                if (vj[x] < kb) Vl[vr[x] * vld + vj[x]] = vx[x];
            if (bb + 1 < nblk) {
                const int i1 = blk_i(bb + 1), kb1 = min(ib, k - i1);
#pragma unroll
                for (int x = 0; x < MA_VX; ++x) vx[x] = vb(vr[x], vj[x], i1, kb1);
            }
        } else if (staged) {
            for (int r = t; r < rows; r += 256)
                for (int j0 = 0; j0 < kb; j0 += 16) {
                    T x[16];   // 16 loads in flight, then the stores
#pragma unroll
                    for (int u = 0; u < 16; ++u) x[u] = vb(r, j0 + u, i, kb);
#pragma unroll
                    for (int u = 0; u < 16; ++u)
                        if (j0 + u < kb) Vl[r * vld + j0 + u] = x[u];
                }
        }
        lds_barrier();
        auto vv = [&](int r, int j) -> T {
            if (staged) return (r < rows && j < kb) ? Vl[min(r, rows - 1) * vld + min(j, kb - 1)] : T(0);
            return vb(r, j, i, kb);
        };
        // W partials of this wave's rows
        // one accumulator chain per (group, step pair) so that consecutive products are independent
        constexpr int NC = NU >= 2 ? 2 : 1;
        AT acc[NG][NC];
#pragma unroll
        for (int g = 0; g < NG; ++g)
#pragma unroll
            for (int c = 0; c < NC; ++c) acc[g][c] = acc_zero<T>();
#pragma unroll
        for (int u = 0; u < NU; ++u) {
            const int s = w + 4 * u;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int r = 16 * s + acc_row<T>(lk, q);
#pragma unroll
                for (int g = 0; g < NG; ++g) mfma16(acc[g][u % NC], vv(r, 16 * g + lr), cv[u][q]);
            }
        }
#pragma unroll
        for (int g = 0; g < NG; ++g)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                T sum = acc[g][0][e];
#pragma unroll
                for (int c = 1; c < NC; ++c) sum += acc[g][c][e];
                Wp[(w * MA_KB + 16 * g + acc_row<T>(lk, e)) * 16 + lr] = sum;
            }
        lds_barrier();
        // W = the partials (fixed order) + the block's A1 rows; element (j, c) = (t >> 4 + 16 x, t & 15)
        T a1v[NG];
#pragma unroll
        for (int x = 0; x < NG; ++x) {
            const int j = (t >> 4) + 16 * x, c = t & 15;
            {
                T s = T(0);
                if (j < kb) {
                    s = (Wp[j * 16 + c] + Wp[(MA_KB + j) * 16 + c]) + (Wp[(2 * MA_KB + j) * 16 + c] + Wp[(3 * MA_KB + j) * 16 + c]);
                    if (TS) {
                        T a;
                        if (a1pre) {
                            a = T(0);
#pragma unroll
                            for (int y = 0; y < MA_AX; ++y)
                                if (16 * y == i + 16 * x) a = a1r[y];   // row i + j = (t >> 4) + 16 y
                        } else {
                            a = A1[i + j + (i64)cc_of(c0 + c, ncols) * lda1];
                        }
                        a1v[x] = a;
                        s += a;
                    }
                }
                Ws[j * 16 + c] = s;
            }
        }
        lds_barrier();
        // W := op(T_b) W on MFMA, wave w for reflector group w
        if (w < NG) {
            AT d = acc_zero<T>();
            const int j = 16 * w + lr;
            T ta[4 * NG];   // op(T_b)[j][4 q + lk]
#pragma unroll
            for (int q = 0; q < 4 * NG; ++q) {
                const int p = 4 * q + lk;
                const int jc = min(j, kb - 1), pc = min(p, kb - 1);
                const int rr = notrans ? jc : pc, col = i + (notrans ? pc : jc);
                const T x = tpre ? Tl[rr * k + col] : Tm[rr + (i64)col * ldt];
                ta[q] = (j < kb && p < kb && (notrans ? p >= j : p <= j)) ? x : T(0);
            }
#pragma unroll
            for (int q = 0; q < 4 * NG; ++q) mfma16(d, ta[q], Ws[(4 * q + lk) * 16 + lr]);
#pragma unroll
            for (int e = 0; e < 4; ++e) Wt[(16 * w + acc_row<T>(lk, e)) * 16 + lr] = d[e];
        }
        lds_barrier();
        // the TS form's A1 rows -= W, by the threads that read them
        if (TS) {
#pragma unroll
            for (int x = 0; x < NG; ++x) {
                const int j = (t >> 4) + 16 * x, c = t & 15;
                if (j < kb && c0 + c < ncols) A1[i + j + (i64)(c0 + c) * lda1] = a1v[x] - Wt[j * 16 + c];
            }
        }
        // C -= V_b W: A = V_b rows of the step, B = W
#pragma unroll
        for (int u = 0; u < NU; ++u) {
            const int s = w + 4 * u;
            AT d0 = acc_zero<T>(), d1 = acc_zero<T>();
#pragma unroll
            for (int q = 0; q < 4 * NG; q += 2) {
                mfma16(d0, vv(16 * s + lr, 4 * q + lk), Wt[(4 * q + lk) * 16 + lr]);
                mfma16(d1, vv(16 * s + lr, 4 * q + 4 + lk), Wt[(4 * q + 4 + lk) * 16 + lr]);
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) cv[u][e] -= d0[e] + d1[e];
        }
    }
#pragma unroll
    for (int u = 0; u < NU; ++u) {
        const int s = w + 4 * u;
        if (s < nkt) {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int r = 16 * s + acc_row<T>(lk, e);
                if (r < rows && c0 + lr < ncols) C[r + (i64)(c0 + lr) * ldc] = cv[u][e];
            }
        }
    }
}

// The MFMA apply takes the left form of real element types with blocks of at most 64
// reflectors and at most 512 rows; it returns false (nothing launched) otherwise.
template <typename T>
static bool launch_mfma_apply(bool ts, int rows, int ncols, int k, int ib, const T *V, i64 ldv, const T *Tm, i64 ldt,
                              bool notrans, bool forward, T *A1, i64 lda1, T *C, i64 ldc, hipStream_t s,
                              hipError_t *err) {
    if constexpr (IsCplx<T>::v) {
        return false;
    } else {
        if (ib > MA_KB || rows > 512 || rows <= 0 || ncols <= 0 || k <= 0) return false;
        dim3 grid((unsigned)((ncols + 15) / 16)), block(256);
        const int nu = rows <= 128 ? 2 : rows <= 256 ? 4 : 8;
        const int kbm = ib < k ? ib : k;
        const int ng = kbm <= 16 ? 1 : kbm <= 32 ? 2 : 4;
#define NLA_MA(TSV, NUV, NGV)                                                                                         \
    mfma_apply_kernel<T, TSV, NUV, NGV><<<grid, block, 0, s>>>(rows, ncols, k, ib, V, ldv, Tm, ldt, notrans, forward, \
                                                              A1, lda1, C, ldc)
#define NLA_MA_NU(TSV, NGV)                                                                                           \
    do {                                                                                                              \
        if (nu == 2) NLA_MA(TSV, 2, NGV); else if (nu == 4) NLA_MA(TSV, 4, NGV); else NLA_MA(TSV, 8, NGV);            \
    } while (0)
        if (ts) {
            if (ng == 1) NLA_MA_NU(true, 1); else if (ng == 2) NLA_MA_NU(true, 2); else NLA_MA_NU(true, 4);
        } else {
            if (ng == 1) NLA_MA_NU(false, 1); else if (ng == 2) NLA_MA_NU(false, 2); else NLA_MA_NU(false, 4);
        }
#undef NLA_MA_NU
#undef NLA_MA
        *err = hipGetLastError();
        return true;
    }
}

template <typename T>
hipError_t launch_ge_apply(bool trc, i64 m, i64 ncols, i64 k, i64 ib, const T *V, i64 ldv, const T *Tm, i64 ldt,
                           bool notrans, bool forward, T *C, i64 ldc, hipStream_t s) {
    if (m <= 0 || ncols <= 0 || k <= 0) return hipSuccess;
    dim3 grid((unsigned)((ncols + CW - 1) / CW)), block(256);
    ProfScope prof(PK_GE_APPLY, s);
    hipError_t e = hipSuccess;
    if (!trc && launch_mfma_apply<T>(false, (int)m, (int)ncols, (int)k, (int)ib, V, ldv, Tm, ldt, notrans, forward,
                                     (T *)nullptr, 0, C, ldc, s, &e))
        return e;
    if (trc) ge_apply_kernel<T, true><<<grid, block, 0, s>>>((int)m, (int)ncols, (int)k, (int)ib, V, ldv, Tm, ldt, notrans, forward, C, ldc);
    else ge_apply_kernel<T, false><<<grid, block, 0, s>>>((int)m, (int)ncols, (int)k, (int)ib, V, ldv, Tm, ldt, notrans, forward, C, ldc);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_ts_apply_mode(bool trc, i64 m2, i64 ncols, i64 k, i64 ib, const T *V, i64 ldv, const T *Tm, i64 ldt,
                                bool notrans, bool forward, T *A1, i64 lda1, T *A2, i64 lda2, int mode, int lfix,
                                hipStream_t s) {
    if (m2 <= 0 || ncols <= 0 || k <= 0) return hipSuccess;
    dim3 grid((unsigned)((ncols + CW - 1) / CW)), block(256);
    ProfScope prof(PK_TS_APPLY, s);
    hipError_t e = hipSuccess;
    if (!trc && mode == 0 && lfix == 0 && launch_mfma_apply<T>(true, (int)m2, (int)ncols, (int)k, (int)ib, V, ldv, Tm, ldt, notrans,
                                                  forward, A1, lda1, A2, lda2, s, &e))
        return e;
    if (trc) ts_apply_kernel<T, true><<<grid, block, 0, s>>>((int)m2, (int)ncols, (int)k, (int)ib, V, ldv, Tm, ldt, notrans, forward, A1, lda1, A2, lda2, mode, lfix);
    else ts_apply_kernel<T, false><<<grid, block, 0, s>>>((int)m2, (int)ncols, (int)k, (int)ib, V, ldv, Tm, ldt, notrans, forward, A1, lda1, A2, lda2, mode, lfix);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_ts_apply(bool trc, i64 m2, i64 ncols, i64 k, i64 ib, const T *V, i64 ldv, const T *Tm, i64 ldt,
                           bool notrans, bool forward, T *A1, i64 lda1, T *A2, i64 lda2, hipStream_t s) {
    return launch_ts_apply_mode<T>(trc, m2, ncols, k, ib, V, ldv, Tm, ldt, notrans, forward, A1, lda1, A2, lda2, 0, 0, s);
}

template <typename T>
hipError_t launch_pent(bool trc, int rows, int ncols, int k, int l, bool fwd, bool row, bool notrans, int phases,
                       const T *V, i64 ldv, const T *Tm, i64 ldt, T *A1, i64 lda1, T *A2, i64 lda2, T *W, i64 ldw,
                       hipStream_t s) {
    if (rows < 0 || ncols <= 0 || k <= 0) return hipSuccess;
    dim3 grid((unsigned)((ncols + PCW - 1) / PCW)), block(256);
    ProfScope prof(PK_TS_APPLY, s);
    if (trc) pent_apply_kernel<T, true><<<grid, block, 0, s>>>(rows, ncols, k, l, fwd, row, notrans, phases, V, ldv, Tm, ldt, A1, lda1, A2, lda2, W, ldw);
    else pent_apply_kernel<T, false><<<grid, block, 0, s>>>(rows, ncols, k, l, fwd, row, notrans, phases, V, ldv, Tm, ldt, A1, lda1, A2, lda2, W, ldw);
    return hipGetLastError();
}

// geqrt! body: src/geqrt.jl:44-67 (geqr2 + larft fused in the panel kernel, then larfb).
template <typename T>
hipError_t geqrt_async(i64 m, i64 n, i64 ib, T *A, i64 lda, T *Tm, i64 ldt, T *tau, hipStream_t s) {
    const i64 k = m < n ? m : n;
    for (i64 i = 0; i < k; i += ib) {
        const i64 sb = ib < k - i ? ib : k - i;
        T *panel = A + i + i * lda;
        hipError_t e = launch_panel_qr<T>(m - i, sb, panel, lda, tau + i, Tm + i * ldt, ldt, s);
        if (e != hipSuccess) return e;
        if (n > i + sb) {
            // larfb!('L','C','F','C', m-i, n-i-sb, sb, ...)
            e = launch_ge_apply<T>(false, m - i, n - i - sb, sb, sb, panel, lda, Tm + i * ldt, ldt, false, true,
                                   A + i + (i + sb) * lda, lda, s);
            if (e != hipSuccess) return e;
        }
    }
    return hipSuccess;
}

// tsqrt! body: src/tsqrt.jl:146-155.
template <typename T>
hipError_t tsqrt_async(i64 m, i64 n, i64 ib, i64 k, T *A1, i64 lda1, T *A2, i64 lda2, T *Tm, i64 ldt, T *tau,
                       hipStream_t s) {
    for (i64 ii = 0; ii < k; ii += ib) {
        const i64 sb = ib < k - ii ? ib : k - ii;
        hipError_t e = launch_ts_panel<T>(false, m, ii, sb, A1, lda1, A2, lda2, Tm, ldt, tau, s);
        if (e != hipSuccess) return e;
        if (n > ii + sb) {
            // tsmqr!('L','C', sb, n-ii-sb, m, n-ii-sb, sb, ib, A1[ii:, ii+sb:], A2[:, ii+sb:], A2[:, ii:], T[:, ii:])
            e = launch_ts_apply<T>(false, m, n - ii - sb, sb, ib, A2 + ii * lda2, lda2, Tm + ii * ldt, ldt, false, true,
                                   A1 + ii + (ii + sb) * lda1, lda1, A2 + (ii + sb) * lda2, lda2, s);
            if (e != hipSuccess) return e;
        }
    }
    return hipSuccess;
}

// ttqrt! body: src/ttqrt.jl:101-117 -- per ib-block the TT panel, then parfb!('L','C','F','C')
// of the block on the trailing columns with the first mi = min(ii + sb, m) rows of A2 and
// pentagon order l = min(sb, mi - ii).
template <typename T>
hipError_t ttqrt_async(i64 m, i64 n, i64 ib, T *A1, i64 lda1, T *A2, i64 lda2, T *Tm, i64 ldt, T *tau, hipStream_t s) {
    const i64 k = m < n ? m : n;
    for (i64 ii = 0; ii < k; ii += ib) {
        const i64 sb = ib < k - ii ? ib : k - ii;
        hipError_t e = launch_ts_panel<T>(true, m, ii, sb, A1, lda1, A2, lda2, Tm, ldt, tau, s);
        if (e != hipSuccess) return e;
        if (n > ii + sb) {
            const i64 mi = ii + sb < m ? ii + sb : m;
            const i64 l = mi - ii > 0 ? (sb < mi - ii ? sb : mi - ii) : 0;
            e = launch_ts_apply_mode<T>(false, mi, n - ii - sb, sb, sb, A2 + ii * lda2, lda2, Tm + ii * ldt, ldt, false,
                                        true, A1 + ii + (ii + sb) * lda1, lda1, A2 + (ii + sb) * lda2, lda2, 0, (int)l, s);
            if (e != hipSuccess) return e;
        }
    }
    return hipSuccess;
}

#define NLA_QR_INST(T)                                                                                                 \
    template hipError_t launch_ge_apply<T>(bool, i64, i64, i64, i64, const T *, i64, const T *, i64, bool, bool, T *, \
                                           i64, hipStream_t);                                                          \
    template hipError_t launch_ts_apply<T>(bool, i64, i64, i64, i64, const T *, i64, const T *, i64, bool, bool, T *, \
                                           i64, T *, i64, hipStream_t);                                                \
    template hipError_t geqrt_async<T>(i64, i64, i64, T *, i64, T *, i64, T *, hipStream_t);                          \
    template hipError_t tsqrt_async<T>(i64, i64, i64, i64, T *, i64, T *, i64, T *, i64, T *, hipStream_t);
NLA_QR_INST(double)
NLA_QR_INST(float)
NLA_QR_INST(cplx<double>)
NLA_QR_INST(cplx<float>)
#undef NLA_QR_INST

// ------------------------------------------------------------------------------------
// Batched uneven problems: one persistent launch over a device work list.  Task t < count is
// geqrt!(m, n, ib, A, T) of problem t (src/geqrt.jl:22-68: per ib-block the fused
// geqr2+larft panel, then larfb!('L','C') of the block on the trailing columns, 16-column
// strips); task count + t is the lower-triangular solve X = L^-1 B of problem t (the
// LeftLowerTRSM! semantics of src/trsm.jl:5-33 with any number of right-hand sides).  A
// workgroup owns a whole task, so problems of any shape share the launch; the caller orders
// the work list largest first.  The launch allocates nothing and uses no memset (the task
// counter is cleared by a fill kernel), so it can be recorded in a hipGraph and replayed.
template <typename T> struct BatchProblem {
    i64 m, n, ib;
    T *A;
    i64 lda;
    T *Tm;
    i64 ldt;
    i64 nt, nrhs;
    const T *L;
    i64 ldl;
    T *B;
    i64 ldb;
};

constexpr int BATCH_RHS = 32;    // right-hand sides staged per LDS pass
constexpr int BATCH_MAXN = 512;  // largest triangular order / panel width handled per problem

// Register-resident blocks of the batched geqrt!.  An ib-block of width sb <= 16 over
// mr <= 1024 rows is factored with its columns in registers: thread t holds rows t, t + 256,
// t + 512 and t + 768 of the block, so a column step costs two workgroup reductions (the
// norm below the diagonal; the products v^T a_j of the new reflector with every other
// column, geqr2!, src/geqr2.jl:23-105) and no LDS traffic for the panel itself.  T
// (larft!('F','C'), src/larft.jl:55-126) is formed after the last column from the Gram matrix
// V^T V, one MFMA product over the block's rows.  The trailing columns then take W = V^T C, W = T^T W,
// C -= V W (larfb!('L','T'), src/larfb.jl:57-98) on MFMA 16x16x4 tiles, V staged in LDS and
// each wave's rows of C held in registers between the two products.  Real element types only
// (the batched entry points are s / d).
constexpr int RP_SB = 16;            // widest block of the register path
constexpr int RP_NQ = 4;             // rows per thread: problems of up to 1024 rows
constexpr int RP_VLD = RP_SB + 1;    // odd row stride of V in LDS (spreads the MFMA reads over banks)
// trailing columns per update chunk: two column tiles in Float32, one in Float64 (a wave keeps
// 16 RP_NQ x tiles x 4 elements of C in registers)
template <typename T> struct RpWc { static constexpr int v = sizeof(T) == 4 ? 32 : 16; };

__device__ __forceinline__ bool rp_fits(int mr, int sb, i64 lda, i64 n) {
    return sb <= RP_SB && mr <= 256 * RP_NQ && lda * n < ((i64)1 << 31);
}
template <typename T> static constexpr size_t rp_lds_elems() {
    // zred, two norm slots, T, W partials (4 waves x tiles; first the Gram partials), W (first
    // the Gram matrix), V.  Float32: 81216 bytes, so two workgroups share a CU.
    return 64 + 16 + 256 + 4 * 16 * RpWc<T>::v + 16 * RpWc<T>::v + (size_t)256 * RP_NQ * RP_VLD;
}

// C (mr x nc, nc <= RpWc) -= V (T^T (V^T C)) for mr <= 256 NQ; Vs: V (rows padded with zeros
// to 256 NQ, columns sb..15 zero) row-major with stride RP_VLD; Tl: T row-major 16 x 16.  Wave
// w owns the 16-row steps w, w + 4, ..., 16 NQ - 4 + w; lane (lr, lk) holds rows
// 16 s + acc_row(lk, q) of column 16 ct + lr, which is the B operand of V^T C (the row order
// inside a step is free) and the D layout of V W, so C is read once, all loads in flight
// together, and written once.  Offsets are 32-bit (the entry point checks lda * n < 2^31).
template <typename T, int NQ>
__device__ __forceinline__ void rp_apply(int mr, int sb, int nc, const T *Vs, const T *Tl, T *Wp, T *Ws, GPtr<T> C,
                                         int ldc) {
    typedef typename Acc<T>::t AT;
    constexpr int WC = RpWc<T>::v, NCT = WC / 16, NU = 4 * NQ;
    const int t = threadIdx.x, lane = t & 63, lr = lane & 15, lk = lane >> 4;
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);
    const int nct = (nc + 15) >> 4;           // column tiles, 1..NCT
    // Unconditional loads at addresses clamped into the block, so no branch splits them: rows
    // >= mr meet zero rows of V in V^T C and are not stored back, columns >= nc only reach
    // columns of W that are never used.
    T cv[NU][NCT][4];
#pragma unroll
    for (int u = 0; u < NU; ++u) {
        const int s = w + 4 * u;
#pragma unroll
        for (int ct = 0; ct < NCT; ++ct) {
            const int c = min(16 * ct + lr, nc - 1);
#pragma unroll
            for (int q = 0; q < 4; ++q) cv[u][ct][q] = C[min(16 * s + acc_row<T>(lk, q), mr - 1) + c * ldc];
        }
    }
    {
        AT acc[NCT];
#pragma unroll
        for (int ct = 0; ct < NCT; ++ct) acc[ct] = acc_zero<T>();
#pragma unroll
        for (int u = 0; u < NU; ++u) {
            const int s = w + 4 * u;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const T a = Vs[(16 * s + acc_row<T>(lk, q)) * RP_VLD + lr];
#pragma unroll
                for (int ct = 0; ct < NCT; ++ct) mfma16(acc[ct], a, cv[u][ct][q]);
            }
        }
#pragma unroll
        for (int ct = 0; ct < NCT; ++ct)
#pragma unroll
            for (int e = 0; e < 4; ++e) Wp[(w * NCT + ct) * 256 + acc_row<T>(lk, e) * 16 + lr] = acc[ct][e];
    }
    lds_barrier();
    // W = the four waves' partials summed in a fixed order, then W := T^T W into the partials' space
    for (int idx = t; idx < 16 * nc; idx += 256) {
        const int p = idx & 15, c = idx >> 4;
        const int o = (c >> 4) * 256 + p * 16 + (c & 15);
        Ws[p * WC + c] = (Wp[o] + Wp[NCT * 256 + o]) + (Wp[2 * NCT * 256 + o] + Wp[3 * NCT * 256 + o]);
    }
    lds_barrier();
    for (int idx = t; idx < 16 * nc; idx += 256) {
        const int j = idx & 15, c = idx >> 4;
        T s = T(0);
        if (j < sb)
            for (int p = 0; p <= j; ++p) s += Tl[p * 16 + j] * Ws[p * WC + c];
        Wp[j * WC + c] = s;
    }
    lds_barrier();
    // C -= V W: A = V rows of the step (lane: V[16 s + lr][k0 + lk]), B = W (lane: W[k0 + lk][c]);
    // W columns past nc are whatever the partials left there, and their results are not stored.
    T wf[NCT][4];
#pragma unroll
    for (int ct = 0; ct < NCT; ++ct)
#pragma unroll
        for (int k = 0; k < 4; ++k) wf[ct][k] = Wp[(4 * k + lk) * WC + 16 * ct + lr];
#pragma unroll
    for (int u = 0; u < NU; ++u) {
        const int s = w + 4 * u;
        if (16 * s >= mr) break;   // wave-uniform
        T va[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) va[k] = Vs[(16 * s + lr) * RP_VLD + 4 * k + lk];
#pragma unroll
        for (int ct = 0; ct < NCT; ++ct) {
            AT acc = acc_zero<T>();
#pragma unroll
            for (int k = 0; k < 4; ++k) mfma16(acc, va[k], wf[ct][k]);
            const int c = 16 * ct + lr;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int r = 16 * s + acc_row<T>(lk, e);
                if (r < mr && c < nc) C[r + c * ldc] = cv[u][ct][e] - acc[e];
            }
        }
    }
    lds_barrier();   // Wp is rewritten by the next chunk
}

// Column I of a register-resident block (see rp_block); a template so that every register
// index is a constant.  taus[I] receives the reflector's tau.
template <typename T, int I>
__device__ __forceinline__ void rp_column(T (&a)[RP_NQ][RP_SB], T (&taus)[RP_SB], int mr, int sb, int nq, T *nred,
                                          T *zred) {
    constexpr int i = I;
    if (i >= sb) return;
    const int t = threadIdx.x, lane = t & 63, w = __builtin_amdgcn_readfirstlane(t >> 6);
    // sum of squares below the diagonal (rows >= mr hold zeros) and alpha = A[i, i]
    T xs = T(0);
#pragma unroll
    for (int q = 0; q < RP_NQ; ++q)
        if (q > 0 || t > i) xs += a[q][i] * a[q][i];
    xs = wave_sum_fast(xs);
    T *nr = nred + 8 * (i & 1);   // alternating slots: a step without a reflector has one barrier
    if (lane == 0) nr[w] = xs;
    if (t == i) nr[4] = a[0][i];
    lds_barrier();
    const T xsum = (nr[0] + nr[1]) + (nr[2] + nr[3]);
    const T alpha = nr[4];
    // the reflector, identically in every thread (src/geqr2.jl:38-73)
    T tau = T(0), beta = alpha, scale = T(0);
    if (mr - i > 1 && xsum != T(0)) {
        beta = -copysign(sqrt(alpha * alpha + xsum), alpha);
        tau = (beta - alpha) / beta;
        scale = T(1) / (alpha - beta);
    }
    taus[i] = tau;
    if (tau != T(0)) {
        T v[RP_NQ];
#pragma unroll
        for (int q = 0; q < RP_NQ; ++q) {
            if (q == 0) {
                v[0] = t > i ? a[0][i] * scale : (t == i ? T(1) : T(0));
                if (t > i) a[0][i] = v[0];
            } else {
                v[q] = a[q][i] * scale;
                a[q][i] = v[q];
            }
        }
        // p_j = v^T a_j for j > i (rows past mr hold zeros)
        T p[RP_SB];
#pragma unroll
        for (int j = 0; j < RP_SB; ++j) p[j] = T(0);
#pragma unroll
        for (int q = 0; q < RP_NQ; ++q)
            if (q < nq) {
#pragma unroll
                for (int j = i + 1; j < RP_SB; ++j) p[j] += v[q] * a[q][j];
            }
        // transposing butterfly: afterwards lane l holds the wave's sum of p_{(l >> 2) & 15}
#pragma unroll
        for (int k = 0; k < 8; ++k) p[k] = k + 8 <= i ? T(0) : fold32(p[k], p[k + 8]);
#pragma unroll
        for (int k = 0; k < 4; ++k) p[k] = fold16(p[k], p[k + 4]);
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const bool hi = lane & 8;
            const T send = hi ? p[k] : p[k + 2];
            p[k] = (hi ? p[k + 2] : p[k]) + dpp_mov<DPP_ROW_ROR8>(send);
        }
        {
            const bool hi = lane & 4;
            const T send = hi ? p[0] : p[1];
            const T from_up = dpp_mov<DPP_ROW_SHL4>(send), from_down = dpp_mov<DPP_ROW_SHR4>(send);
            p[0] = (hi ? p[1] : p[0]) + (hi ? from_down : from_up);
        }
        p[0] += dpp_mov<DPP_QUAD_X2>(p[0]);
        p[0] += dpp_mov<DPP_QUAD_X1>(p[0]);
        if ((lane & 3) == 0) zred[((lane >> 2) & 15) * 4 + w] = p[0];
        lds_barrier();
        // A[:, j] -= tau v z_j for j > i (H^T applied to the block's later columns)
        T u[RP_SB];
#pragma unroll
        for (int j = i + 1; j < RP_SB; ++j)
            u[j] = tau * ((zred[4 * j] + zred[4 * j + 1]) + (zred[4 * j + 2] + zred[4 * j + 3]));
#pragma unroll
        for (int q = 0; q < RP_NQ; ++q)
            if (q < nq) {
#pragma unroll
                for (int j = i + 1; j < RP_SB; ++j) a[q][j] -= v[q] * u[j];
            }
    }
    if (t == i) a[0][i] = beta;
}

template <typename T, int... I>
__device__ __forceinline__ void rp_columns(T (&a)[RP_NQ][RP_SB], T (&taus)[RP_SB], int mr, int sb, int nq, T *nred,
                                           T *zred, std::integer_sequence<int, I...>) {
    (rp_column<T, I>(a, taus, mr, sb, nq, nred, zred), ...);
}

// One ib-block (mr x sb at `panel`, ncols trailing columns to its right) on the register path.
template <typename T>
__device__ __forceinline__ void rp_block(int mr, int sb, int ncols, T *panel_, i64 lda, T *Tg_, i64 ldt, T *smem) {
    static_assert(!IsCplx<T>::v, "real element types only");
    const GPtr<T> panel = to_global(panel_), Tg = to_global(Tg_);
    constexpr int WC = RpWc<T>::v;
    T *zred = smem, *nred = zred + 64, *Tl = nred + 16, *Wp = Tl + 256, *Ws = Wp + 4 * 16 * WC;
    T *Vs = Ws + 16 * WC;
    static_assert(4 * 16 * WC >= 4 * 256 && 16 * WC >= 256, "the Gram partials and sum fit the W space");
    const int t = threadIdx.x, lane = t & 63, lr = lane & 15, lk = lane >> 4;
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);
    const int nq = (mr + 255) >> 8;
    T a[RP_NQ][RP_SB];
#pragma unroll
    for (int q = 0; q < RP_NQ; ++q) {
        const int r = t + 256 * q;
#pragma unroll
        for (int j = 0; j < RP_SB; ++j) {
            // unconditional load at a clamped address, then the zero padding (no divergent branch)
            const T x = panel[(r < mr ? r : mr - 1) + (i64)(j < sb ? j : sb - 1) * lda];
            a[q][j] = (r < mr && j < sb) ? x : T(0);
        }
    }
    T taus[RP_SB];
    rp_columns<T>(a, taus, mr, sb, nq, nred, zred, std::make_integer_sequence<int, RP_SB>{});

    // R and V back to memory; V (unit lower trapezoid, zero columns past sb, zero rows up to
    // 256 nq) to LDS
#pragma unroll
    for (int q = 0; q < RP_NQ; ++q) {
        const int r = t + 256 * q;
        if (r < mr) {
#pragma unroll
            for (int j = 0; j < RP_SB; ++j)
                if (j < sb) panel[r + (i64)j * lda] = a[q][j];
        }
    }
#pragma unroll
    for (int q = 0; q < RP_NQ; ++q) {
        if (q < nq) {
            const int r = t + 256 * q;
#pragma unroll
            for (int j = 0; j < RP_SB; ++j)
                Vs[r * RP_VLD + j] = j < sb ? (r > j ? a[q][j] : (r == j ? T(1) : T(0))) : T(0);
        }
    }
    lds_barrier();
    // G = V^T V on MFMA (both operands are the same V fragment), four partials summed in order
    {
        typedef typename Acc<T>::t AT;
        AT acc = acc_zero<T>();
        const int nkt = (mr + 15) >> 4;
        for (int s = w; s < nkt; s += 4) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const T x = Vs[(16 * s + acc_row<T>(lk, q)) * RP_VLD + lr];
                mfma16(acc, x, x);
            }
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) Wp[w * 256 + acc_row<T>(lk, e) * 16 + lr] = acc[e];
    }
    lds_barrier();
    Ws[t] = (Wp[t] + Wp[256 + t]) + (Wp[512 + t] + Wp[768 + t]);
    lds_barrier();
    // T (larft!('F','C')), row t by thread t < 16: T[t, i] = sum_{q=t}^{i-1} T[t, q] (-tau_i G[q, i]),
    // T[i, i] = tau_i; to memory (zeros below the diagonal) and to LDS for the update
    if (t < 16) {
        T Trow[RP_SB];
#pragma unroll
        for (int i = 0; i < RP_SB; ++i) {
            T s = T(0);
            if (i < sb && taus[i] != T(0)) {
                if (t == i) s = taus[i];
#pragma unroll
                for (int q = 0; q < i; ++q)
                    if (q >= t) s += Trow[q] * (-taus[i] * Ws[q * 16 + i]);
            }
            Trow[i] = s;
        }
#pragma unroll
        for (int j = 0; j < RP_SB; ++j) {
            const T v = (j >= t && j < sb && t < sb) ? Trow[j] : T(0);
            if (t < sb && j < sb) Tg[t + (i64)j * ldt] = v;
            Tl[t * 16 + j] = v;
        }
    }
    // (Tl is read after the first barrier of the update, which wave 0 reaches after writing it)
    const GPtr<T> C = panel + (i64)sb * lda;
    for (int cc = 0; cc < ncols; cc += WC) {
        const int nc = ncols - cc < WC ? ncols - cc : WC;
        const GPtr<T> Cc = C + (i64)cc * lda;
        switch (nq) {
        case 1: rp_apply<T, 1>(mr, sb, nc, Vs, Tl, Wp, Ws, Cc, (int)lda); break;
        case 2: rp_apply<T, 2>(mr, sb, nc, Vs, Tl, Wp, Ws, Cc, (int)lda); break;
        case 3: rp_apply<T, 3>(mr, sb, nc, Vs, Tl, Wp, Ws, Cc, (int)lda); break;
        default: rp_apply<T, 4>(mr, sb, nc, Vs, Tl, Wp, Ws, Cc, (int)lda); break;
        }
    }
}

template <typename T> __device__ __forceinline__ void batched_geqrt(const BatchProblem<T> &P, T *smem) {
    T *sV = smem, *sC = sV + KBM * RC, *sW0 = sC + CW * RC, *sW = sW0 + KBM * CW;
    T *z = sW + KBM * CW, *tcol = z + 64, *red = tcol + 64, *sh = red + 16, *tau = sh + 4;
    const int m = (int)P.m, n = (int)P.n, ib = (int)P.ib;
    const int k = m < n ? m : n;
    for (int i = 0; i < k; i += ib) {
        const int sb = ib < k - i ? ib : k - i;
        T *panel = P.A + i + (i64)i * P.lda;
        if (rp_fits(m - i, sb, P.lda, n)) {
            rp_block<T>(m - i, sb, n - i - sb, panel, P.lda, P.Tm + (i64)i * P.ldt, P.ldt, smem);
            __syncthreads();
            continue;
        }
        panel_qr_dev<T>(m - i, sb, panel, P.lda, tau, P.Tm + (i64)i * P.ldt, P.ldt, z, tcol, red, sh);
        __syncthreads();
        const int ncols = n - i - sb;
        for (int c0 = 0; c0 < ncols; c0 += CW) {
            const int cw = ncols - c0 < CW ? ncols - c0 : CW;
            T *C = P.A + i + (i64)(i + sb) * P.lda;
            apply_strip<T, true, false>(sb, m - i - sb, panel, panel + sb, P.lda, P.Tm + (i64)i * P.ldt, P.ldt, false,
                                        C, P.lda, C + sb, P.lda, c0, cw, sV, sC, sW0, sW);
        }
        __syncthreads();
    }
}

// X = L^-1 B (L nt x nt lower, B nt x nrhs), blocked by 16 columns of L.  A block of
// BATCH_RHS right-hand sides lives in LDS (row-major, stride BATCH_RHS + 1); for each 16-column
// block of L its rows are staged in LDS TR_CH at a time, the 16 x 16 diagonal block is solved by
// one lane per right-hand side (forward substitution), and the rows below take
// X[p, :] -= L[p, block] X[block, :] with thread t on right-hand side t % 32 and rows t / 32 + 8 k.
constexpr int TR_XLD = BATCH_RHS + 1;
constexpr int TR_CH = 128;

template <typename T> __device__ __forceinline__ void batched_trsm(const BatchProblem<T> &P, T *smem) {
    const int n = (int)P.nt, t = threadIdx.x, c = t & (BATCH_RHS - 1), g = t / BATCH_RHS;
    constexpr int NG = 256 / BATCH_RHS;
    T *X = smem, *Lp = smem + BATCH_MAXN * TR_XLD;
    const GPtr<const T> L = to_global(P.L);
    for (i64 r0 = 0; r0 < P.nrhs; r0 += BATCH_RHS) {
        const int nr = P.nrhs - r0 < BATCH_RHS ? (int)(P.nrhs - r0) : BATCH_RHS;
        const GPtr<T> Bb = to_global(P.B + r0 * P.ldb);
        lds_barrier();
        for (int cc = 0; cc < BATCH_RHS; ++cc)
            for (int p = t; p < n; p += 256) X[p * TR_XLD + cc] = cc < nr ? Bb[p + (i64)cc * P.ldb] : T(0);
        for (int k0 = 0; k0 < n; k0 += 16) {
            const int kw = n - k0 < 16 ? n - k0 : 16;
            for (int q0 = k0; q0 < n; q0 += TR_CH) {
                const int q1 = n - q0 < TR_CH ? n : q0 + TR_CH;
                lds_barrier();
                for (int idx = t; idx < TR_CH * 16; idx += 256) {
                    const int rr = idx & (TR_CH - 1), j = idx / TR_CH;
                    if (q0 + rr < q1 && j < kw) Lp[rr * 16 + j] = L[q0 + rr + (i64)(k0 + j) * P.ldl];
                }
                lds_barrier();
                if (q0 == k0) {
                    if (t < BATCH_RHS) {
                        T x[16];
#pragma unroll
                        for (int r = 0; r < 16; ++r) x[r] = r < kw ? X[(k0 + r) * TR_XLD + c] : T(0);
#pragma unroll
                        for (int r = 0; r < 16; ++r) {
                            if (r < kw) {
                                T sum = x[r];
#pragma unroll
                                for (int q = 0; q < r; ++q) sum -= Lp[r * 16 + q] * x[q];
                                x[r] = sum / Lp[r * 16 + r];
                            }
                        }
#pragma unroll
                        for (int r = 0; r < 16; ++r)
                            if (r < kw) X[(k0 + r) * TR_XLD + c] = x[r];
                    }
                    lds_barrier();
                }
                const int u0 = q0 > k0 + kw ? q0 : k0 + kw;
                if (u0 < q1) {
                    T xb[16];
#pragma unroll
                    for (int j = 0; j < 16; ++j) xb[j] = j < kw ? X[(k0 + j) * TR_XLD + c] : T(0);
                    for (int p = u0 + g; p < q1; p += NG) {
                        T sum = X[p * TR_XLD + c];
                        const T *Lr = Lp + (p - q0) * 16;
#pragma unroll
                        for (int j = 0; j < 16; ++j)
                            if (j < kw) sum -= Lr[j] * xb[j];
                        X[p * TR_XLD + c] = sum;
                    }
                }
            }
        }
        lds_barrier();
        for (int cc = 0; cc < nr; ++cc)
            for (int p = t; p < n; p += 256) Bb[p + (i64)cc * P.ldb] = X[p * TR_XLD + cc];
    }
}

static_assert(2 * (sizeof(float) * rp_lds_elems<float>() + 64) <= 160 * 1024, "two Float32 workgroups per CU");
static_assert((size_t)BATCH_MAXN * TR_XLD + TR_CH * 16 <= rp_lds_elems<float>(), "the solve fits the QR's LDS");

template <typename T> static size_t batched_lds_bytes() {
    const size_t geq = (size_t)KBM * RC + (size_t)CW * RC + 2 * (size_t)KBM * CW + 64 + 64 + 16 + 4 + 64;
    const size_t trs = (size_t)BATCH_MAXN * TR_XLD + TR_CH * 16;
    size_t e = geq > trs ? geq : trs;
    if (rp_lds_elems<T>() > e) e = rp_lds_elems<T>();
    return sizeof(T) * e;
}

// Float32 runs two workgroups per CU (LDS and registers sized for it); Float64 runs one.
template <typename T> struct BatchOcc { static constexpr int waves_per_eu = sizeof(T) == 4 ? 2 : 1; };

template <typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(BatchOcc<T>::waves_per_eu))) void batched_kernel(int count, const BatchProblem<T> *__restrict__ probs, int *counter) {
    extern __shared__ __align__(16) unsigned char dyn_raw[];
    T *smem = reinterpret_cast<T *>(dyn_raw);
    __shared__ int sh_task;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) sh_task = atomicAdd(counter, 1);
        __syncthreads();
        const int task = __builtin_amdgcn_readfirstlane(sh_task);   // uniform: scalar loads of P
        if (task >= 2 * count) break;
        // task 2i: the QR of problem i, task 2i + 1: its solve (independent of each other), so
        // latency-bound solves share the CUs with the QRs instead of forming a tail
        const BatchProblem<T> P = probs[task >> 1];
        if ((task & 1) == 0) {
            if (P.m > 0 && P.n > 0) batched_geqrt<T>(P, smem);
        } else if (P.nt > 0 && P.nrhs > 0) {
            batched_trsm<T>(P, smem);
        }
    }
}

__global__ void batched_reset_kernel(int *counter) { *counter = 0; }

template <typename T>
static hipError_t batched_launch(int count, const BatchProblem<T> *probs, int *counter, hipStream_t s) {
    // device queries and the LDS attribute once per device (nothing but the two launches runs
    // when the call is being recorded into a graph after a first eager call)
    static std::mutex mu;
    static int cached_dev = -1, cached_grid = 0;
    const size_t sh = batched_lds_bytes<T>();
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    int gridmax;
    {
        std::lock_guard<std::mutex> lk(mu);
        if (cached_dev != dev) {
            int ncu = 0, per = 0;
            if ((e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
            if (sh > 64 * 1024 &&
                (e = hipFuncSetAttribute(reinterpret_cast<const void *>(&batched_kernel<T>),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh)) != hipSuccess)
                return e;
            if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, batched_kernel<T>, 256, sh)) != hipSuccess)
                return e;
            cached_grid = ncu * (per > 0 ? per : 1);
            cached_dev = dev;
        }
        gridmax = cached_grid;
    }
    int grid = gridmax;
    if (grid > 2 * count) grid = 2 * count;
    batched_reset_kernel<<<1, 1, 0, s>>>(counter);
    batched_kernel<T><<<grid, 256, sh, s>>>(count, probs, counter);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------
// Entry points: validation mirrors the reference's ArgumentError checks (file:line on each).
static bool valid_trans(char t) { return t == 'N' || t == 'C' || t == 'T'; }

template <typename T> static int larfg_entry(i64 n, T *alpha, T *x, i64 incx, T *tau, void *stream) {
    hipStream_t s = as_stream(stream);
    if (n <= 1) {  // src/larfg.jl:47-50
        NLA_HIP(hipMemsetAsync(tau, 0, sizeof(T), s));
        return 0;
    }
    if (incx <= 0) return arg_error(4, "incx must be positive, got %lld", (long long)incx);
    if (n > INT32_MAX) return arg_error(1, "n exceeds 2^31-1");
    larfg_kernel<T><<<1, 256, 0, s>>>((int)n, alpha, x, incx, tau);
    NLA_LAUNCH_CHECK();
    return 0;
}

template <typename T> static int geqr2_entry(i64 m, i64 n, T *A, i64 lda, T *tau, void *stream) {
    if (m < 0) return arg_error(1, "illegal value of m: %lld", (long long)m);   // src/geqr2.jl:108-110
    if (n < 0) return arg_error(2, "illegal value of n: %lld", (long long)n);   // src/geqr2.jl:111-113
    if (m == 0 || n == 0) return 0;
    if (lda < m) return arg_error(4, "lda must be >= max(1,m)");
    if (n > 4096) return arg_error(2, "geqr2 panel holds at most 4096 columns, got %lld", (long long)n);
    if (m > INT32_MAX) return arg_error(1, "m exceeds 2^31-1");
    NLA_HIP(launch_panel_qr<T>(m, n, A, lda, tau, (T *)nullptr, 0, as_stream(stream)));
    return 0;
}

template <typename T>
static int larft_entry(char direct, char storev, i64 n, i64 k, const T *V, i64 ldv, const T *tau, T *Tm, i64 ldt,
                       void *stream) {
    direct = up(direct); storev = up(storev);
    if (n == 0) return 0;  // src/larft.jl:206-208
    if (direct != 'F' && direct != 'B') return arg_error(1, "direct must be 'F' or 'B', got '%c'", direct);
    if (storev != 'C' && storev != 'R') return arg_error(2, "storev must be 'C' or 'R', got '%c'", storev);
    if (direct != 'F') return arg_error(1, "direct='B' is not implemented (the QR routines use 'F')");
    if (n < 0) return arg_error(3, "n must be non-negative");
    if (k < 0) return arg_error(4, "k must be non-negative");
    if (k == 0) return 0;
    if (ldv < (storev == 'C' ? n : k)) return arg_error(6, "ldv must be >= %lld", (long long)(storev == 'C' ? n : k));
    if (ldt < k) return arg_error(9, "ldt must be >= k");
    if (k > 4096) return arg_error(4, "k must be <= 4096");
    hipStream_t s = as_stream(stream);
    const size_t sh = (size_t)k * sizeof(T);
    if (storev == 'C') larft_kernel<T, false><<<1, 256, sh, s>>>((int)n, (int)k, V, ldv, tau, Tm, ldt);
    else larft_kernel<T, true><<<1, 256, sh, s>>>((int)n, (int)k, V, ldv, tau, Tm, ldt);
    NLA_LAUNCH_CHECK();
    return 0;
}

template <typename T>
static int larfb_entry(char side, char trans, char direct, char storev, i64 m, i64 n, i64 k, const T *V, i64 ldv,
                       const T *Tm, i64 ldt, T *C, i64 ldc, void *stream) {
    side = up(side); trans = up(trans); direct = up(direct); storev = up(storev);
    if (m <= 0 || n <= 0 || k <= 0) return 0;  // src/larfb.jl:38-40
    if (storev != 'C' || direct != 'F')         // src/larfb.jl:42-44
        return arg_error(storev != 'C' ? 4 : 3, "larfb! currently supports only forward, columnwise reflectors");
    if (side != 'L' && side != 'R') return arg_error(1, "side must be 'L' or 'R', got '%c'", side);
    if (!valid_trans(trans)) return arg_error(2, "trans must be 'N', 'C', or 'T', got '%c'", trans);
    if (IsCplx<T>::v && trans == 'T') return arg_error(2, "trans='T' is not valid for complex element types");
    const i64 q = side == 'L' ? m : n;
    if (k > q) return arg_error(7, "k must be <= %lld", (long long)q);
    if (ldv < q) return arg_error(9, "ldv must be >= %lld", (long long)q);
    if (ldt < k) return arg_error(11, "ldt must be >= k");
    if (ldc < m) return arg_error(13, "ldc must be >= m");
    const bool notrans = trans == 'N';
    hipStream_t s = as_stream(stream);
    if (side == 'L') NLA_HIP(launch_ge_apply<T>(false, m, n, k, k, V, ldv, Tm, ldt, notrans, true, C, ldc, s));
    else NLA_HIP(launch_ge_apply<T>(true, n, m, k, k, V, ldv, Tm, ldt, !notrans, true, C, ldc, s));
    return 0;
}

template <typename T>
static int geqrt_entry(i64 m, i64 n, i64 ib, T *A, i64 lda, T *Tm, i64 ldt, T *tau, void *stream) {
    if (m < 0) return arg_error(1, "illegal value of m: %lld", (long long)m);   // src/geqrt.jl:24-26
    if (n < 0) return arg_error(2, "illegal value of n: %lld", (long long)n);   // src/geqrt.jl:28-30
    if (ib < 0 || (ib == 0 && m > 0 && n > 0))                                 // src/geqrt.jl:32-34
        return arg_error(3, "illegal value of ib: %lld", (long long)ib);
    if (m == 0 || n == 0 || ib == 0) return 0;
    if (lda < m) return arg_error(5, "lda must be >= max(1,m)");
    const i64 k = m < n ? m : n;
    if (ldt < (ib < k ? ib : k)) return arg_error(7, "ldt must be >= min(ib, min(m,n))");
    if (ib > 256) return arg_error(3, "ib must be <= 256 (one panel workgroup), got %lld", (long long)ib);
    if (m > INT32_MAX || n > INT32_MAX) return arg_error(1, "dimension exceeds 2^31-1");
    NLA_HIP(geqrt_async<T>(m, n, ib, A, lda, Tm, ldt, tau, as_stream(stream)));
    return 0;
}

template <typename T>
static int unmqr_entry(char side, char trans, i64 m, i64 n, i64 k, i64 ib, const T *A, i64 lda, i64 a_cols,
                       const T *Tm, i64 ldt, T *C, i64 ldc, void *stream) {
    side = up(side); trans = up(trans);
    if (side != 'L' && side != 'R') return arg_error(1, "side must be 'L' or 'R', got '%c'", side);  // src/unmqr.jl:54-56
    const i64 nq = side == 'L' ? m : n;
    if (!valid_trans(trans)) return arg_error(2, "trans must be 'N', 'C', or 'T', got '%c'", trans);  // :66-68
    if (IsCplx<T>::v && trans == 'T') return arg_error(2, "trans='T' is not valid for complex element types");
    if (m < 0) return arg_error(3, "m must be non-negative, got %lld", (long long)m);                 // :70-72
    if (n < 0) return arg_error(4, "n must be non-negative, got %lld", (long long)n);                 // :74-76
    if (k < 0) return arg_error(5, "k must be non-negative, got %lld", (long long)k);                 // :78-80
    if (k > nq) k = nq;                                                                                  // :83
    if (k > a_cols) k = a_cols;
    if (ib < 0) return arg_error(6, "ib must be non-negative, got %lld", (long long)ib);               // :85-87
    if (lda < (nq > 1 ? nq : 1) && nq > 0)                                                               // :89-91
        return arg_error(8, "lda must be >= max(1,%lld), got %lld", (long long)nq, (long long)lda);
    if (m == 0 || n == 0 || k == 0) return 0;                                                            // :94-96
    if (ib == 0) return arg_error(6, "step cannot be zero");  // the block loop at src/unmqr.jl:127 has step ib
    if (ldt < (ib < k ? ib : k)) return arg_error(11, "ldt must be >= min(ib, k)");
    if (ldc < m) return arg_error(13, "ldc must be >= m");
    const bool notrans = trans == 'N';
    const bool fwd = (side == 'L' && !notrans) || (side == 'R' && notrans);  // src/unmqr.jl:99-109
    hipStream_t s = as_stream(stream);
    if (side == 'L') NLA_HIP(launch_ge_apply<T>(false, m, n, k, ib, A, lda, Tm, ldt, notrans, fwd, C, ldc, s));
    else NLA_HIP(launch_ge_apply<T>(true, n, m, k, ib, A, lda, Tm, ldt, !notrans, fwd, C, ldc, s));
    return 0;
}

template <typename T>
static int tsqrt_entry(i64 m, i64 n, i64 ib, T *A1, i64 lda1, i64 m1, T *A2, i64 lda2, T *Tm, i64 ldt, T *tau,
                       void *stream) {
    if (m < 0) return arg_error(1, "m must be non-negative, got %lld", (long long)m);    // src/tsqrt.jl:129-131
    if (n < 0) return arg_error(2, "n must be non-negative, got %lld", (long long)n);    // :132-134
    if (ib < 0) return arg_error(3, "ib must be non-negative, got %lld", (long long)ib); // :135-137
    const i64 k = m1 < n ? m1 : n;                                                         // :139-140
    if (m == 0 || n == 0 || ib == 0 || k <= 0) return 0;                                   // :141-143
    if (lda1 < m1) return arg_error(5, "lda1 must be >= m1");
    if (lda2 < m) return arg_error(8, "lda2 must be >= m");
    if (ldt < (ib < k ? ib : k)) return arg_error(10, "ldt must be >= min(ib, k)");
    if (ib > 256) return arg_error(3, "ib must be <= 256 (one panel workgroup), got %lld", (long long)ib);
    if (m > INT32_MAX || n > INT32_MAX) return arg_error(1, "dimension exceeds 2^31-1");
    NLA_HIP(tsqrt_async<T>(m, n, ib, k, A1, lda1, A2, lda2, Tm, ldt, tau, as_stream(stream)));
    return 0;
}

// ttqrt!(m, n, ib, A1, A2, T, tau, work): src/ttqrt.jl:81-118.
template <typename T>
static int ttqrt_entry(i64 m, i64 n, i64 ib, T *A1, i64 lda1, T *A2, i64 lda2, T *Tm, i64 ldt, T *tau, void *stream) {
    if (m < 0) return arg_error(1, "illegal value of m");    // src/ttqrt.jl:82-84
    if (n < 0) return arg_error(2, "illegal value of n");    // :86-88
    if (ib < 0) return arg_error(3, "illegal value of ib");  // :90-92
    if (m == 0 || n == 0 || ib == 0) return 0;               // :95-97
    const i64 k = m < n ? m : n;
    if (lda1 < m) return arg_error(5, "lda1 must be >= m");
    if (lda2 < m) return arg_error(7, "lda2 must be >= m");
    if (ldt < (ib < k ? ib : k)) return arg_error(9, "ldt must be >= min(ib, k)");
    if (ib > 256) return arg_error(3, "ib must be <= 256 (one panel workgroup), got %lld", (long long)ib);
    if (m > INT32_MAX || n > INT32_MAX) return arg_error(1, "dimension exceeds 2^31-1");
    NLA_HIP(ttqrt_async<T>(m, n, ib, A1, lda1, A2, lda2, Tm, ldt, tau, as_stream(stream)));
    return 0;
}

// The ib-block loop of tsmqr!/ttmqr! on any T: blocks of up to 64 reflectors run in the strip
// kernel (one launch for every block); wider blocks are applied one block per launch by the
// general kernel, which applies I - V T V^H exactly without splitting the block.
template <typename T>
static int ts_blocks(bool tt, char side, char trans, i64 m1, i64 n1, i64 m2, i64 n2, i64 k, i64 ib, T *A1, i64 lda1,
                     T *A2, i64 lda2, const T *V, i64 ldv, const T *Tm, i64 ldt, hipStream_t s) {
    const bool notrans = trans == 'N';
    const bool fwd = (side == 'L' && !notrans) || (side == 'R' && notrans);  // src/tsmqr.jl:104-114
    const bool L = side == 'L';
    const i64 rows = L ? m2 : n2, ncols = L ? n1 : m1;   // left form
    if (ib <= KBM) {
        NLA_HIP(launch_ts_apply_mode<T>(!L, rows, ncols, k, ib, V, ldv, Tm, ldt, L ? notrans : !notrans, fwd, A1, lda1,
                                        A2, lda2, tt ? 1 : 0, 0, s));
        return 0;
    }
    const i64 nblk = (k + ib - 1) / ib;
    for (i64 bb = 0; bb < nblk; ++bb) {
        const i64 i = (fwd ? bb : nblk - 1 - bb) * ib;
        const i64 kb = ib < k - i ? ib : k - i;
        const i64 mb = tt ? (i + kb < rows ? i + kb : rows) : rows;
        const i64 l = tt ? (rows - i - 1 > 0 ? (kb < rows - i - 1 ? kb : rows - i - 1) : 0) : 0;
        T *a1 = L ? A1 + i : A1 + i * lda1;
        NLA_HIP(launch_pent<T>(!L, (int)mb, (int)ncols, (int)kb, (int)l, true, false, L ? notrans : !notrans,
                               PENT_W | PENT_T | PENT_A, V + i * ldv, ldv, Tm + i * ldt, ldt, a1, lda1, A2, lda2,
                               (T *)nullptr, 0, s));
    }
    return 0;
}

template <typename T>
static int tsmqr_entry(char side, char trans, i64 m1, i64 n1, i64 m2, i64 n2, i64 k, i64 ib, T *A1, i64 lda1, T *A2,
                       i64 lda2, const T *V, i64 ldv, i64 v_cols, const T *Tm, i64 ldt, i64 t_cols, void *stream) {
    side = up(side); trans = up(trans);
    if (side != 'L' && side != 'R') return arg_error(1, "side must be 'L' or 'R', got '%c'", side);  // src/tsmqr.jl:51-53
    if (!valid_trans(trans)) return arg_error(2, "trans must be 'N', 'C', or 'T', got '%c'", trans);  // :55-57
    if (IsCplx<T>::v && trans == 'T') return arg_error(2, "trans='T' is not valid for complex element types");
    if (m1 < 0) return arg_error(3, "m1 must be non-negative, got %lld", (long long)m1);              // :59-61
    if (n1 < 0) return arg_error(4, "n1 must be non-negative, got %lld", (long long)n1);              // :63-65
    if (m2 < 0 || (m2 != m1 && side == 'R')) {                                                        // :67-73
        if (side == 'R')
            return arg_error(5, "For side='R', m2 must equal m1. Got m1=%lld, m2=%lld", (long long)m1, (long long)m2);
        return arg_error(5, "m2 must be non-negative, got %lld", (long long)m2);
    }
    if (n2 < 0 || (n2 != n1 && side == 'L')) {                                                        // :75-81
        if (side == 'L')
            return arg_error(6, "For side='L', n2 must equal n1. Got n1=%lld, n2=%lld", (long long)n1, (long long)n2);
        return arg_error(6, "n2 must be non-negative, got %lld", (long long)n2);
    }
    if (k < 0) return arg_error(7, "k must be non-negative, got %lld", (long long)k);                 // :83-85
    if (ib < 0) return arg_error(8, "ib must be non-negative, got %lld", (long long)ib);              // :87-89
    const i64 kmax = side == 'L' ? m1 : n1;                                                           // :93-96
    const i64 stored = v_cols < t_cols ? v_cols : t_cols;
    if (k > kmax) k = kmax;
    if (k > stored) k = stored;
    if (m1 == 0 || n1 == 0 || m2 == 0 || n2 == 0 || k <= 0 || ib == 0) return 0;                      // :99-101
    if (lda1 < m1) return arg_error(10, "lda1 must be >= m1");
    if (lda2 < m2) return arg_error(12, "lda2 must be >= m2");
    const i64 vr = side == 'L' ? m2 : n2;
    if (ldv < vr) return arg_error(14, "ldv must be >= %lld", (long long)vr);
    if (ldt < (ib < k ? ib : k)) return arg_error(17, "ldt must be >= min(ib, k)");
    if (ib > PKMAX) return arg_error(8, "ib must be <= %d", PKMAX);
    return ts_blocks<T>(false, side, trans, m1, n1, m2, n2, k, ib, A1, lda1, A2, lda2, V, ldv, Tm, ldt,
                        as_stream(stream));
}

// The reference's tile-QR call sequence (its callers' flat-tree loop over geqrt!/unmqr!/tsqrt!/
// tsmqr!, one call per tile or tile pair), issued through the same per-call entry points a
// drop-in caller reaches.  Same output layout as nla_?tile_geqrf: T tile (i,k) at
// T[i*ib, k*nb].  Mirrors the caller loop; the library's own tile_geqrf runs the same
// factorization as one persistent launch.
template <typename T>
static int tile_geqrf_calls_entry(i64 m, i64 n, i64 nb, i64 ib, T *A, i64 lda, T *Tm, i64 ldt, void *stream) {
    if (m < 0) return arg_error(1, "m must be non-negative");
    if (n < 0) return arg_error(2, "n must be non-negative");
    if (nb <= 0) return arg_error(3, "nb must be positive");
    if (ib <= 0 || ib > nb) return arg_error(4, "ib must be in 1..nb");
    if (m % nb || n % nb) return arg_error(1, "m and n must be multiples of nb");
    if (m == 0 || n == 0) return 0;
    if (lda < m) return arg_error(6, "lda must be >= m");
    const i64 mt = m / nb, nt = n / nb, kt = mt < nt ? mt : nt;
    if (ldt < mt * ib) return arg_error(8, "ldt must be >= (m/nb)*ib");
    hipStream_t s = as_stream(stream);
    int st = 0;
    T *tau = (T *)scratch(SLOT_AUX, sizeof(T) * (size_t)nb, s, &st);
    if (!tau) return st;
    auto tile = [&](i64 i, i64 j) { return A + i * nb + j * nb * lda; };
    auto ttile = [&](i64 i, i64 j) { return Tm + i * ib + j * nb * ldt; };
    for (i64 k = 0; k < kt; ++k) {
        if ((st = geqrt_entry<T>(nb, nb, ib, tile(k, k), lda, ttile(k, k), ldt, tau, s))) return st;
        for (i64 j = k + 1; j < nt; ++j)
            if ((st = unmqr_entry<T>('L', 'C', nb, nb, nb, ib, tile(k, k), lda, nb, ttile(k, k), ldt, tile(k, j), lda, s)))
                return st;
        for (i64 i = k + 1; i < mt; ++i) {
            if ((st = tsqrt_entry<T>(nb, nb, ib, tile(k, k), lda, nb, tile(i, k), lda, ttile(i, k), ldt, tau, s)))
                return st;
            for (i64 j = k + 1; j < nt; ++j)
                if ((st = tsmqr_entry<T>('L', 'C', nb, nb, nb, nb, nb, ib, tile(k, j), lda, tile(i, j), lda, tile(i, k),
                                         lda, nb, ttile(i, k), ldt, nb, s)))
                    return st;
        }
    }
    return 0;
}

// ttmqr!(side, trans, m1, n1, m2, n2, k, ib, A1, A2, V, T, work): src/ttmqr.jl:1-95.
template <typename T>
static int ttmqr_entry(char side, char trans, i64 m1, i64 n1, i64 m2, i64 n2, i64 k, i64 ib, T *A1, i64 lda1, T *A2,
                       i64 lda2, const T *V, i64 ldv, const T *Tm, i64 ldt, void *stream) {
    side = up(side); trans = up(trans);
    if (side != 'L' && side != 'R') return arg_error(1, "illegal value of side");                  // src/ttmqr.jl:3-5
    if (!valid_trans(trans)) return arg_error(2, "illegal value of trans");                        // :7-9
    if (IsCplx<T>::v && trans == 'T') return arg_error(2, "trans='T' is not valid for complex element types");
    if (m1 < 0) return arg_error(3, "illegal value of m1");                                         // :11-13
    if (n1 < 0) return arg_error(4, "illegal value of n1");                                         // :15-17
    if (m2 < 0 || (m2 != m1 && side == 'R')) return arg_error(5, "illegal value of m2");            // :19-21
    if (n2 < 0 || (n2 != n1 && side == 'L')) return arg_error(6, "illegal value of n2");            // :23-25
    if (k < 0) return arg_error(7, "illegal value of k");                                           // :27-29
    const i64 kmax = side == 'L' ? m1 : n1;                                                         // :31-32
    if (k > kmax) k = kmax;
    if (ib < 0) return arg_error(8, "illegal value of ib");                                         // :34-36
    if (m1 == 0 || n1 == 0 || m2 == 0 || n2 == 0 || k == 0 || ib == 0) return 0;                    // :39-41
    if (lda1 < m1) return arg_error(10, "lda1 must be >= m1");
    if (lda2 < m2) return arg_error(12, "lda2 must be >= m2");
    const i64 vr = side == 'L' ? m2 : n2;
    if (ldv < vr) return arg_error(14, "ldv must be >= %lld", (long long)vr);
    if (ldt < (ib < k ? ib : k)) return arg_error(16, "ldt must be >= min(ib, k)");
    if (ib > PKMAX) return arg_error(8, "ib must be <= %d", PKMAX);
    return ts_blocks<T>(true, side, trans, m1, n1, m2, n2, k, ib, A1, lda1, A2, lda2, V, ldv, Tm, ldt,
                        as_stream(stream));
}

// parfb!(side, trans, direct, storev, m1, n1, m2, n2, k, l, A1, A2, V, T, work): src/parfb.jl:44-191.
template <typename T>
static int parfb_entry(char side, char trans, char direct, char storev, i64 m1, i64 n1, i64 m2, i64 n2, i64 k, i64 l,
                       T *A1, i64 lda1, T *A2, i64 lda2, const T *V, i64 ldv, const T *Tm, i64 ldt, void *stream) {
    side = up(side); trans = up(trans); direct = up(direct); storev = up(storev);
    if (side != 'L' && side != 'R') return arg_error(1, "illegal value of side");          // src/parfb.jl:47-49
    if (!valid_trans(trans)) return arg_error(2, "illegal value of trans");                 // :51-53
    if (direct != 'F' && direct != 'B') return arg_error(3, "illegal value of direct");     // :55-57
    if (storev != 'C' && storev != 'R') return arg_error(4, "illegal value of storev");     // :59-61
    if (m1 < 0) return arg_error(5, "illegal value of m1");                                  // :63-65
    if (n1 < 0) return arg_error(6, "illegal value of n1");                                  // :67-69
    if (m2 < 0 || (side == 'R' && m1 != m2)) return arg_error(7, "illegal value of m2");     // :71-73
    if (n2 < 0 || (side == 'L' && n1 != n2)) return arg_error(8, "illegal value of n2");     // :75-77
    if (k < 0) return arg_error(9, "illegal value of k");                                    // :79-81
    if (l < 0 || l > k) return arg_error(10, "illegal value of l");                          // :83-85
    if (m1 == 0 || n1 == 0 || n2 == 0 || k == 0) return 0;                                   // :87-89
    if (IsCplx<T>::v && trans == 'T') return arg_error(2, "trans='T' is not valid for complex element types");
    if (k > PKMAX) return arg_error(9, "k must be <= %d", PKMAX);
    const bool L = side == 'L';
    const i64 rows = L ? m2 : n2, ncols = L ? n2 : m2;   // left form
    if (rows == 0 || ncols == 0) return 0;   // no rows of A2 to combine with
    if (lda1 < (L ? k : m2)) return arg_error(12, "lda1 too small");
    if (rows > 0 && lda2 < m2) return arg_error(14, "lda2 must be >= m2");
    if (rows > 0 && ldv < (storev == 'C' ? rows : k)) return arg_error(16, "ldv too small");
    if (ldt < k) return arg_error(18, "ldt must be >= k");
    const bool notrans = (trans == 'N') == L;   // right-side calls flip op(T) in the left form
    hipStream_t s = as_stream(stream);
    if (direct == 'F' && storev == 'C' && k <= KBM) {
        NLA_HIP(launch_ts_apply_mode<T>(!L, rows, ncols, k, k, V, ldv, Tm, ldt, notrans, true, A1, lda1, A2, lda2, 0,
                                        (int)l, s));
        return 0;
    }
    NLA_HIP(launch_pent<T>(!L, (int)rows, (int)ncols, (int)k, (int)l, direct == 'F', storev == 'R', notrans,
                           PENT_W | PENT_T | PENT_A, V, ldv, Tm, ldt, A1, lda1, A2, lda2, (T *)nullptr, 0, s));
    return 0;
}

// pamm!(op, side, storev, direct, m, n, k, l, A1, A2, V, W): src/pamm.jl:49-100.
template <typename T>
static int pamm_entry(char op, char side, char storev, char direct, i64 m, i64 n, i64 k, i64 l, const T *A1, i64 lda1,
                      T *A2, i64 lda2, const T *V, i64 ldv, T *W, i64 ldw, void *stream) {
    op = up(op); side = up(side); storev = up(storev); direct = up(direct);
    if (op != 'W' && op != 'A') return arg_error(1, "op must be 'W' or 'A', got '%c'", op);                // src/pamm.jl:51-53
    if (side != 'L' && side != 'R') return arg_error(2, "side must be 'L' or 'R', got '%c'", side);        // :54-56
    if (storev != 'C' && storev != 'R') return arg_error(3, "storev must be 'C' or 'R', got '%c'", storev); // :57-59
    if (direct != 'F' && direct != 'B') return arg_error(4, "direct must be 'F' or 'B', got '%c'", direct); // :60-62
    if (m < 0) return arg_error(5, "m must be non-negative, got %lld", (long long)m);                      // :65-67
    if (n < 0) return arg_error(6, "n must be non-negative, got %lld", (long long)n);                      // :68-70
    if (k < 0) return arg_error(7, "k must be non-negative, got %lld", (long long)k);                      // :71-73
    if (l < 0) return arg_error(8, "l must be non-negative, got %lld", (long long)l);                      // :74-76
    if (m == 0 || n == 0 || k == 0) return 0;                                                               // :80-82
    if (l > k) return arg_error(8, "l must be <= k");
    if (k > PKMAX) return arg_error(7, "k must be <= %d", PKMAX);
    const bool L = side == 'L';
    const i64 rows = L ? m : n, ncols = L ? n : m;
    if (lda2 < m) return arg_error(12, "lda2 must be >= m");
    if (ldv < (storev == 'C' ? rows : k)) return arg_error(14, "ldv too small");
    if (ldw < (L ? k : m)) return arg_error(16, "ldw too small");
    if (op == 'W' && lda1 < (L ? k : m)) return arg_error(10, "lda1 too small");
    NLA_HIP(launch_pent<T>(!L, (int)rows, (int)ncols, (int)k, (int)l, direct == 'F', storev == 'R', true,
                           op == 'W' ? PENT_W : PENT_A, V, ldv, (const T *)nullptr, 0, const_cast<T *>(A1), lda1, A2,
                           lda2, W, ldw, as_stream(stream)));
    return 0;
}

// pemv!(trans, storev, m, n, l, alpha, A, X, beta, Y, work): src/pemv.jl:39-181.
template <typename T>
static int pemv_entry(char trans, char storev, i64 m, i64 n, i64 l, T alpha, const T *A, i64 lda, const T *x, T beta,
                      T *y, void *stream) {
    trans = up(trans); storev = up(storev);
    if (trans != 'N' && trans != 'T' && trans != 'C') return arg_error(1, "illegal value of trans");   // src/pemv.jl:40-42
    if (storev != 'C' && storev != 'R') return arg_error(2, "illegal value of storev");               // :44-46
    if (!((storev == 'C' && trans != 'N') || (storev == 'R' && trans == 'N')))                         // :48-50
        return arg_error(1, "illegal values of trans/storev");
    if (m < 0) return arg_error(3, "illegal value of m");                                               // :52-54
    if (n < 0) return arg_error(4, "illegal value of n");                                               // :56-58
    if (l > (m < n ? m : n) || l < 0) return arg_error(5, "illegal value of l");                       // :60-62
    if (m == 0 || n == 0) return 0;                                                                      // :64-66
    if (alpha == T(0) && beta == T(0)) return 0;                                                         // :68-70
    if (l == 1) l = 0;                                                                                   // :72-74
    if (lda < m) return arg_error(8, "lda must be >= m");
    if (m > INT32_MAX || n > INT32_MAX) return arg_error(3, "dimension exceeds 2^31-1");
    hipStream_t s = as_stream(stream);
    if (storev == 'C') {
        pemv_kernel<T, true><<<(unsigned)((n + 3) / 4), 256, 0, s>>>((int)m, (int)n, (int)l, trans == 'C', alpha, A, lda,
                                                                       x, beta, y);
    } else {
        pemv_kernel<T, false><<<(unsigned)((m + 255) / 256), 256, 0, s>>>((int)m, (int)n, (int)l, false, alpha, A, lda,
                                                                           x, beta, y);
    }
    NLA_LAUNCH_CHECK();
    return 0;
}

}  // namespace nla

using namespace nla;

// Exported entry points for one element type (prefix p, element type T).
#define NLA_QR_EXPORTS(p, T)                                                                                           \
    NLA_API int nla_##p##larfg(int64_t n, T *alpha, T *x, int64_t incx, T *tau, void *stream) {                       \
        return larfg_entry<T>(n, alpha, x, incx, tau, stream);                                                         \
    }                                                                                                                  \
    NLA_API int nla_##p##geqr2(int64_t m, int64_t n, T *A, int64_t lda, T *tau, T *work, void *stream) {              \
        (void)work;                                                                                                    \
        return geqr2_entry<T>(m, n, A, lda, tau, stream);                                                              \
    }                                                                                                                  \
    NLA_API int nla_##p##larft(char direct, char storev, int64_t n, int64_t k, const T *V, int64_t ldv, const T *tau, \
                               T *Tm, int64_t ldt, void *stream) {                                                     \
        return larft_entry<T>(direct, storev, n, k, V, ldv, tau, Tm, ldt, stream);                                     \
    }                                                                                                                  \
    NLA_API int nla_##p##geqrt(int64_t m, int64_t n, int64_t ib, T *A, int64_t lda, T *Tm, int64_t ldt, T *tau,       \
                               T *work, void *stream) {                                                                \
        (void)work;                                                                                                    \
        return geqrt_entry<T>(m, n, ib, A, lda, Tm, ldt, tau, stream);                                                 \
    }                                                                                                                  \
    NLA_API int nla_##p##larfb(char side, char trans, char direct, char storev, int64_t m, int64_t n, int64_t k,      \
                               const T *V, int64_t ldv, const T *Tm, int64_t ldt, T *C, int64_t ldc, T *work,          \
                               int64_t ldwork, void *stream) {                                                         \
        (void)work; (void)ldwork;                                                                                      \
        return larfb_entry<T>(side, trans, direct, storev, m, n, k, V, ldv, Tm, ldt, C, ldc, stream);                  \
    }                                                                                                                  \
    NLA_API int nla_##p##unmqr(char side, char trans, int64_t m, int64_t n, int64_t k, int64_t ib, const T *A,        \
                               int64_t lda, int64_t a_cols, const T *Tm, int64_t ldt, T *C, int64_t ldc, T *work,      \
                               void *stream) {                                                                         \
        (void)work;                                                                                                    \
        return unmqr_entry<T>(side, trans, m, n, k, ib, A, lda, a_cols, Tm, ldt, C, ldc, stream);                      \
    }                                                                                                                  \
    NLA_API int nla_##p##tsqrt(int64_t m, int64_t n, int64_t ib, T *A1, int64_t lda1, int64_t m1, T *A2,              \
                               int64_t lda2, T *Tm, int64_t ldt, T *tau, T *work, void *stream) {                      \
        (void)work;                                                                                                    \
        return tsqrt_entry<T>(m, n, ib, A1, lda1, m1, A2, lda2, Tm, ldt, tau, stream);                                 \
    }                                                                                                                  \
    NLA_API int nla_##p##tsmqr(char side, char trans, int64_t m1, int64_t n1, int64_t m2, int64_t n2, int64_t k,      \
                               int64_t ib, T *A1, int64_t lda1, T *A2, int64_t lda2, const T *V, int64_t ldv,          \
                               int64_t v_cols, const T *Tm, int64_t ldt, int64_t t_cols, T *work, void *stream) {      \
        (void)work;                                                                                                    \
        return tsmqr_entry<T>(side, trans, m1, n1, m2, n2, k, ib, A1, lda1, A2, lda2, V, ldv, v_cols, Tm, ldt, t_cols, \
                              stream);                                                                                 \
    }                                                                                                                  \
    NLA_API int nla_##p##tile_geqrf_calls(int64_t m, int64_t n, int64_t nb, int64_t ib, T *A, int64_t lda, T *Tm,    \
                                          int64_t ldt, void *stream) {                                                 \
        return tile_geqrf_calls_entry<T>(m, n, nb, ib, A, lda, Tm, ldt, stream);                                       \
    }                                                                                                                  \
    NLA_API int nla_##p##ttqrt(int64_t m, int64_t n, int64_t ib, T *A1, int64_t lda1, T *A2, int64_t lda2, T *Tm,     \
                               int64_t ldt, T *tau, T *work, void *stream) {                                           \
        (void)work;                                                                                                    \
        return ttqrt_entry<T>(m, n, ib, A1, lda1, A2, lda2, Tm, ldt, tau, stream);                                     \
    }                                                                                                                  \
    NLA_API int nla_##p##ttmqr(char side, char trans, int64_t m1, int64_t n1, int64_t m2, int64_t n2, int64_t k,      \
                               int64_t ib, T *A1, int64_t lda1, T *A2, int64_t lda2, const T *V, int64_t ldv,          \
                               const T *Tm, int64_t ldt, T *work, void *stream) {                                      \
        (void)work;                                                                                                    \
        return ttmqr_entry<T>(side, trans, m1, n1, m2, n2, k, ib, A1, lda1, A2, lda2, V, ldv, Tm, ldt, stream);        \
    }                                                                                                                  \
    NLA_API int nla_##p##parfb(char side, char trans, char direct, char storev, int64_t m1, int64_t n1, int64_t m2,   \
                               int64_t n2, int64_t k, int64_t l, T *A1, int64_t lda1, T *A2, int64_t lda2, const T *V, \
                               int64_t ldv, const T *Tm, int64_t ldt, T *work, int64_t ldwork, void *stream) {         \
        (void)work; (void)ldwork;                                                                                      \
        return parfb_entry<T>(side, trans, direct, storev, m1, n1, m2, n2, k, l, A1, lda1, A2, lda2, V, ldv, Tm, ldt,  \
                              stream);                                                                                 \
    }                                                                                                                  \
    NLA_API int nla_##p##pamm(char op, char side, char storev, char direct, int64_t m, int64_t n, int64_t k,          \
                              int64_t l, const T *A1, int64_t lda1, T *A2, int64_t lda2, const T *V, int64_t ldv,      \
                              T *W, int64_t ldw, void *stream) {                                                       \
        return pamm_entry<T>(op, side, storev, direct, m, n, k, l, A1, lda1, A2, lda2, V, ldv, W, ldw, stream);        \
    }                                                                                                                  \
    NLA_API int nla_##p##pemv(char trans, char storev, int64_t m, int64_t n, int64_t l, T alpha, const T *A,          \
                              int64_t lda, const T *x, T beta, T *y, T *work, void *stream) {                          \
        (void)work;                                                                                                    \
        return pemv_entry<T>(trans, storev, m, n, l, alpha, A, lda, x, beta, y, stream);                               \
    }
NLA_QR_EXPORTS(d, double)
NLA_QR_EXPORTS(s, float)
NLA_QR_EXPORTS(z, cplx<double>)
NLA_QR_EXPORTS(c, cplx<float>)
#undef NLA_QR_EXPORTS

namespace nla {
// Host-side validation of a work list (copied to the host once; a device list is the
// interface so that a graph replay reads the list the caller may have refreshed).
template <typename T>
static int batched_entry(i64 count, const BatchProblem<T> *probs, int *counter, int validate, void *stream) {
    if (count < 0) return arg_error(1, "count must be non-negative");
    if (count == 0) return 0;
    if (count > (1 << 29)) return arg_error(1, "count too large");
    if (!probs || !counter) return arg_error(2, "problems and counter must be device pointers");
    hipStream_t s = as_stream(stream);
    if (validate) {
        std::vector<BatchProblem<T>> h((size_t)count);
        NLA_HIP(hipMemcpyAsync(h.data(), probs, sizeof(BatchProblem<T>) * (size_t)count, hipMemcpyDeviceToHost, s));
        NLA_HIP(hipStreamSynchronize(s));
        for (i64 i = 0; i < count; ++i) {
            const BatchProblem<T> &P = h[i];
            if (P.m < 0 || P.n < 0 || P.n > P.m) return arg_error(2, "problem %lld: need 0 <= n <= m", (long long)i);
            if (P.m > 0 && P.n > 0) {
                if (P.ib <= 0 || P.ib > 64) return arg_error(2, "problem %lld: ib must be in 1..64", (long long)i);
                if (P.lda < P.m) return arg_error(2, "problem %lld: lda < m", (long long)i);
                if (P.ldt < P.ib) return arg_error(2, "problem %lld: ldt < ib", (long long)i);
                if (!P.A || !P.Tm) return arg_error(2, "problem %lld: null A or T", (long long)i);
            }
            if (P.nt < 0 || P.nrhs < 0) return arg_error(2, "problem %lld: negative solve size", (long long)i);
            if (P.nt > 0 && P.nrhs > 0) {
                if (P.nt > BATCH_MAXN) return arg_error(2, "problem %lld: triangular order must be <= %d", (long long)i, BATCH_MAXN);
                if (P.ldl < P.nt || P.ldb < P.nt) return arg_error(2, "problem %lld: ldl/ldb < nt", (long long)i);
                if (!P.L || !P.B) return arg_error(2, "problem %lld: null L or B", (long long)i);
            }
        }
    }
    NLA_HIP(batched_launch<T>((int)count, probs, counter, s));
    return 0;
}
}  // namespace nla

NLA_API int nla_dbatched_qr_trsm(int64_t count, const void *problems, int *counter, int validate, void *stream) {
    return nla::batched_entry<double>(count, (const nla::BatchProblem<double> *)problems, counter, validate, stream);
}
NLA_API int nla_sbatched_qr_trsm(int64_t count, const void *problems, int *counter, int validate, void *stream) {
    return nla::batched_entry<float>(count, (const nla::BatchProblem<float> *)problems, counter, validate, stream);
}

namespace nla {
// fill with a deterministic pseudo-random N(0,1)-like sequence (sum of 4 uniforms)
template <typename T> __global__ void fill_rand_kernel(T *p, i64 n, unsigned seed) {
    for (i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
        unsigned h = (unsigned)i * 2654435761u ^ seed;
        T acc = 0;
        for (int k = 0; k < 4; ++k) {
            h ^= h << 13; h ^= h >> 17; h ^= h << 5;
            acc += T(h & 0xffffff) * T(1.0 / 16777216.0);
        }
        p[i] = (acc - T(2)) * T(1.7320508);
    }
}

// Autotuner of the batched QR's inner block ib: for each (m, n) class, time the batched
// launch on `samples` problems of that shape for every candidate and keep the fastest.
template <typename T>
static int batched_autotune_entry(i64 nclass, const i64 *m_values, const i64 *n_values, i64 samples,
                                  const i64 *cand, i64 ncand, i64 *ib_out, double *ms_out, void *stream) {
    if (nclass < 0) return arg_error(1, "nclass must be non-negative, got %lld", (long long)nclass);
    if (samples <= 0) return arg_error(4, "samples must be positive, got %lld", (long long)samples);
    if (ncand <= 0 || !cand) return arg_error(6, "need at least one candidate ib");
    // the batched kernel holds ib <= 64 (its LDS panel buffers and the 64-row T blocks)
    for (i64 q = 0; q < ncand; ++q)
        if (cand[q] < 1 || cand[q] > 64)
            return arg_error(5, "candidate ib must be in 1..64, got %lld", (long long)cand[q]);
    for (i64 c = 0; c < nclass; ++c)
        if (m_values[c] <= 0 || n_values[c] <= 0 || n_values[c] > m_values[c])
            return arg_error(2, "class %lld: need 0 < n <= m", (long long)c);
    if (nclass == 0) return 0;
    hipStream_t s = as_stream(stream);
    int *counter = nullptr;
    BatchProblem<T> *dprob = nullptr;
    NLA_HIP(hipMalloc(&counter, sizeof(int)));
    NLA_HIP(hipMalloc(&dprob, sizeof(BatchProblem<T>) * (size_t)samples));
    hipEvent_t e0, e1;
    NLA_HIP(hipEventCreate(&e0));
    NLA_HIP(hipEventCreate(&e1));
    int rc = 0;
    for (i64 c = 0; c < nclass && rc == 0; ++c) {
        const i64 m = m_values[c], n = n_values[c];
        if (m <= 0 || n <= 0 || n > m) { rc = arg_error(2, "class %lld: need 0 < n <= m", (long long)c); break; }
        T *A = nullptr, *A0 = nullptr, *Tm = nullptr;
        const size_t na = (size_t)m * n * samples;
        if (hipMalloc(&A, sizeof(T) * na) != hipSuccess || hipMalloc(&A0, sizeof(T) * na) != hipSuccess ||
            hipMalloc(&Tm, sizeof(T) * 64 * (size_t)n * samples) != hipSuccess) {
            rc = hip_error(hipErrorOutOfMemory, "autotune buffers");
        } else {
            fill_rand_kernel<T><<<1024, 256, 0, s>>>(A0, (i64)na, 0x9e3779b9u + (unsigned)c);
            double best = 1e300;
            i64 bestib = cand[0];
            for (i64 q = 0; q < ncand && rc == 0; ++q) {
                const i64 ib = cand[q] < n ? cand[q] : n;
                std::vector<BatchProblem<T>> h((size_t)samples);
                for (i64 i = 0; i < samples; ++i)
                    h[i] = BatchProblem<T>{m, n, ib, A + (size_t)i * m * n, m, Tm + (size_t)i * 64 * n, 64, 0, 0,
                                           nullptr, 1, nullptr, 1};
                if (hipMemcpyAsync(dprob, h.data(), sizeof(BatchProblem<T>) * samples, hipMemcpyHostToDevice, s) ||
                    hipMemcpyAsync(A, A0, sizeof(T) * na, hipMemcpyDeviceToDevice, s)) {
                    rc = hip_error(hipErrorUnknown, "autotune copy");
                    break;
                }
                float ms = 0, tot = 0;
                for (int rep = 0; rep < 3 && rc == 0; ++rep) {
                    (void)hipMemcpyAsync(A, A0, sizeof(T) * na, hipMemcpyDeviceToDevice, s);
                    (void)hipEventRecord(e0, s);
                    if (batched_launch<T>((int)samples, dprob, counter, s) != hipSuccess) rc = hip_error(hipGetLastError(), "autotune launch");
                    (void)hipEventRecord(e1, s);
                    (void)hipEventSynchronize(e1);
                    (void)hipEventElapsedTime(&ms, e0, e1);
                    if (rep > 0) tot += ms;  // first run warms up
                }
                if (tot / 2 < best) best = tot / 2, bestib = ib;
            }
            ib_out[c] = bestib;
            if (ms_out) ms_out[c] = best;
        }
        (void)hipFree(A);
        (void)hipFree(A0);
        (void)hipFree(Tm);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(dprob);
    (void)hipFree(counter);
    return rc;
}
}  // namespace nla

NLA_API int nla_sbatched_autotune(int64_t nclass, const int64_t *m_values, const int64_t *n_values, int64_t samples,
                                  const int64_t *candidates, int64_t ncand, int64_t *ib_out, double *ms_out,
                                  void *stream) {
    return nla::batched_autotune_entry<float>(nclass, m_values, n_values, samples, candidates, ncand, ib_out, ms_out,
                                              stream);
}
NLA_API int nla_dbatched_autotune(int64_t nclass, const int64_t *m_values, const int64_t *n_values, int64_t samples,
                                  const int64_t *candidates, int64_t ncand, int64_t *ib_out, double *ms_out,
                                  void *stream) {
    return nla::batched_autotune_entry<double>(nclass, m_values, n_values, samples, candidates, ncand, ib_out, ms_out,
                                               stream);
}
