/*
 * CPU restatement of NextLA.jl's tile-QR and recursive TRSM/TRMM algorithms.
 * TEST INFRASTRUCTURE ONLY: the checker used by tests/, __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py.  Nothing in the shipped library calls it.
 *
 * This file is a type template: nla_oracle.c includes it twice, once with
 * R=double / P(x)=nlao_d##x and once with R=float / P(x)=nlao_s##x.
 * All matrices are column-major with explicit leading dimensions; indices are
 * 0-based here where the reference is 1-based.
 */

#define AT(A, ld, i, j) (A)[(size_t)(i) + (size_t)(j) * (size_t)(ld)]

/* lamch: src/NextLA.jl:30-53 (eps is halved because rounding is to nearest). */
R P(lamch)(char cmach) {
    R ep = (sizeof(R) == 8) ? (R)DBL_EPSILON : (R)FLT_EPSILON;
    ep *= (R)0.5;
    if (cmach == 'E') return ep;
    if (cmach == 'S') {
        R sfmin = (sizeof(R) == 8) ? (R)DBL_MIN : (R)FLT_MIN;
        R small = (R)1 / ((sizeof(R) == 8) ? (R)DBL_MAX : (R)FLT_MAX);
        if (small >= sfmin) sfmin = small * ((R)1 + ep);
        return sfmin;
    }
    return ep * 2;
}

/* Overflow-safe 2-norm (the BLAS nrm2 that Julia's norm(x,2) reaches). */
static R P(nrm2)(int64_t n, const R *x, int64_t incx) {
    R scale = 0, ssq = 1;
    for (int64_t i = 0; i < n; ++i) {
        R v = x[i * incx];
        if (v != 0) {
            R a = (R)fabs((double)v);
            if (scale < a) { ssq = 1 + ssq * (scale / a) * (scale / a); scale = a; }
            else ssq += (a / scale) * (a / scale);
        }
    }
    return scale * (R)sqrt((double)ssq);
}

/* larfg!: src/larfg.jl:43-106, including the safmin rescale loop.
 * On return *alpha holds beta and x holds v[2:n].
 * Deviation (LAPACK dlarfg semantics): beta uses hypot instead of
 * sqrt(alpha^2 + xnorm^2), and the rescale loop stops after 20 passes.  With the
 * reference's form, |alpha|, ||x|| ~ 1e-300 underflow beta to 0 and the loop at
 * src/larfg.jl:70-77 never ends. */
void P(larfg)(int64_t n, R *alpha, R *x, int64_t incx, R *tau) {
    if (n <= 1) { *tau = 0; return; }
    R xnorm = P(nrm2)(n - 1, x, incx);
    R alphr = *alpha;
    if (xnorm == 0) { *tau = 0; return; }
    R beta = -(R)copysign(hypot((double)alphr, (double)xnorm), (double)alphr);
    R safmin = P(lamch)('S') / P(lamch)('E');
    R rsafmn = (R)1 / safmin;
    int knt = 0;
    R a = *alpha;
    if (fabs((double)beta) < safmin) {
        while (fabs((double)beta) < safmin && knt < 20) {
            knt++;
            for (int64_t i = 0; i < n - 1; ++i) x[i * incx] *= rsafmn;
            beta *= rsafmn; alphr *= rsafmn; a *= rsafmn;
        }
        xnorm = P(nrm2)(n - 1, x, incx);
        beta = -(R)copysign(hypot((double)alphr, (double)xnorm), (double)alphr);
    }
    *tau = (beta - alphr) / beta;
    R s = (R)1 / (a - beta);
    for (int64_t i = 0; i < n - 1; ++i) x[i * incx] *= s;
    for (int j = 0; j < knt; ++j) beta *= safmin;
    *alpha = beta;
}

/* geqr2!: kernel src/geqr2.jl:23-105 (plain sum-of-squares norm, no rescale). */
void P(geqr2)(int64_t m, int64_t n, R *A, int64_t lda, R *tau, R *work) {
    if (m <= 0 || n <= 0) return;
    int64_t k = m < n ? m : n;
    for (int64_t i = 0; i < k; ++i) {
        int64_t len = m - i;
        if (len <= 1) { tau[i] = 0; continue; }
        R alpha = AT(A, lda, i, i);
        R xs = 0;
        for (int64_t p = 1; p < len; ++p) { R v = AT(A, lda, i + p, i); xs += v * v; }
        if (xs == 0) { tau[i] = 0; continue; }
        R beta = -(R)copysign(sqrt((double)(alpha * alpha + xs)), (double)alpha);
        R t = (beta - alpha) / beta;
        R scale = (R)1 / (alpha - beta);
        for (int64_t p = 1; p < len; ++p) AT(A, lda, i + p, i) *= scale;
        AT(A, lda, i, i) = beta;
        tau[i] = t;
        if (i < n - 1) {
            AT(A, lda, i, i) = 1;
            int64_t nc = n - i - 1;
            for (int64_t j = 0; j < nc; ++j) {
                R w = 0;
                for (int64_t p = 0; p < len; ++p) w += AT(A, lda, i + p, i + 1 + j) * AT(A, lda, i + p, i);
                work[j] = w;
            }
            for (int64_t j = 0; j < nc; ++j)
                for (int64_t p = 0; p < len; ++p)
                    AT(A, lda, i + p, i + 1 + j) -= t * AT(A, lda, i + p, i) * work[j];
            AT(A, lda, i, i) = beta;
        }
    }
}

/* larft! direct='F', storev='C': kernel src/larft.jl:48-126 (LAPACK lastv trimming). */
void P(larft_fc)(int64_t n, int64_t k, const R *V, int64_t ldv, const R *tau, R *T, int64_t ldt) {
    if (n <= 0) return;
    R work[4096];
    int64_t prevlastv = n - 1;
    for (int64_t i = 0; i < k; ++i) {
        if (prevlastv < i) prevlastv = i;
        R ti = tau[i];
        if (ti == 0) {
            for (int64_t j = 0; j <= i; ++j) AT(T, ldt, j, i) = 0;
            continue;
        }
        int64_t lastv = n - 1;
        while (lastv >= i + 1 && AT(V, ldv, lastv, i) == 0) lastv--;
        int64_t j = lastv < prevlastv ? lastv : prevlastv;
        for (int64_t p = 0; p < i; ++p) AT(T, ldt, p, i) = -ti * AT(V, ldv, i, p);
        for (int64_t p = 0; p < i; ++p) {
            R acc = 0;
            for (int64_t r = i + 1; r <= j; ++r) acc += AT(V, ldv, r, p) * AT(V, ldv, r, i);
            AT(T, ldt, p, i) += -ti * acc;
        }
        prevlastv = (i > 0) ? (prevlastv > lastv ? prevlastv : lastv) : lastv;
        for (int64_t p = 0; p < i; ++p) work[p] = AT(T, ldt, p, i);
        for (int64_t p = 0; p < i; ++p) {
            R acc = 0;
            for (int64_t q = p; q < i; ++q) acc += AT(T, ldt, p, q) * work[q];
            AT(T, ldt, p, i) = acc;
        }
        AT(T, ldt, i, i) = ti;
    }
}

/* larfb! direct='F', storev='C': src/larfb.jl:37-142.
 * work is n x k (side L) or m x k (side R), leading dimension ldw. */
void P(larfb)(char side, char trans, int64_t m, int64_t n, int64_t k, const R *V, int64_t ldv,
              const R *T, int64_t ldt, R *C, int64_t ldc, R *W, int64_t ldw) {
    if (m <= 0 || n <= 0 || k <= 0) return;
    int notrans = (trans == 'N');
    if (side == 'L') {
        /* W = C1^T (n x k) */
        for (int64_t j = 0; j < k; ++j)
            for (int64_t c = 0; c < n; ++c) AT(W, ldw, c, j) = AT(C, ldc, j, c);
        /* W = W * V1 (unit lower k x k): column j uses rows j..k-1 of V1 */
        for (int64_t j = 0; j < k; ++j)
            for (int64_t c = 0; c < n; ++c) {
                R acc = AT(W, ldw, c, j);
                for (int64_t r = j + 1; r < k; ++r) acc += AT(W, ldw, c, r) * AT(V, ldv, r, j);
                AT(W, ldw, c, j) = acc;
            }
        /* W += C2^T V2 */
        if (m > k)
            for (int64_t j = 0; j < k; ++j)
                for (int64_t c = 0; c < n; ++c) {
                    R acc = 0;
                    for (int64_t r = k; r < m; ++r) acc += AT(C, ldc, r, c) * AT(V, ldv, r, j);
                    AT(W, ldw, c, j) += acc;
                }
        /* W = W * T^T ('N') or W * T ('C'/'T'); T upper */
        for (int64_t c = 0; c < n; ++c) {
            R row[4096];
            for (int64_t j = 0; j < k; ++j) row[j] = AT(W, ldw, c, j);
            for (int64_t j = 0; j < k; ++j) {
                R acc = 0;
                if (notrans) for (int64_t q = j; q < k; ++q) acc += row[q] * AT(T, ldt, j, q);
                else         for (int64_t q = 0; q <= j; ++q) acc += row[q] * AT(T, ldt, q, j);
                AT(W, ldw, c, j) = acc;
            }
        }
        /* C2 -= V2 W^T */
        if (m > k)
            for (int64_t c = 0; c < n; ++c)
                for (int64_t r = k; r < m; ++r) {
                    R acc = 0;
                    for (int64_t j = 0; j < k; ++j) acc += AT(V, ldv, r, j) * AT(W, ldw, c, j);
                    AT(C, ldc, r, c) -= acc;
                }
        /* W = W V1^T ; C1 -= W^T */
        for (int64_t c = 0; c < n; ++c) {
            R row[4096];
            for (int64_t j = 0; j < k; ++j) row[j] = AT(W, ldw, c, j);
            for (int64_t j = 0; j < k; ++j) {
                R acc = row[j];
                for (int64_t q = 0; q < j; ++q) acc += row[q] * AT(V, ldv, j, q);
                AT(C, ldc, j, c) -= acc;
            }
        }
    } else {
        /* W = C1 (m x k) */
        for (int64_t j = 0; j < k; ++j)
            for (int64_t r = 0; r < m; ++r) AT(W, ldw, r, j) = AT(C, ldc, r, j);
        for (int64_t j = 0; j < k; ++j)
            for (int64_t r = 0; r < m; ++r) {
                R acc = AT(W, ldw, r, j);
                for (int64_t q = j + 1; q < k; ++q) acc += AT(W, ldw, r, q) * AT(V, ldv, q, j);
                AT(W, ldw, r, j) = acc;
            }
        if (n > k)
            for (int64_t j = 0; j < k; ++j)
                for (int64_t r = 0; r < m; ++r) {
                    R acc = 0;
                    for (int64_t q = k; q < n; ++q) acc += AT(C, ldc, r, q) * AT(V, ldv, q, j);
                    AT(W, ldw, r, j) += acc;
                }
        /* W = W T ('N') or W T^T ('C') */
        for (int64_t r = 0; r < m; ++r) {
            R row[4096];
            for (int64_t j = 0; j < k; ++j) row[j] = AT(W, ldw, r, j);
            for (int64_t j = 0; j < k; ++j) {
                R acc = 0;
                if (notrans) for (int64_t q = 0; q <= j; ++q) acc += row[q] * AT(T, ldt, q, j);
                else         for (int64_t q = j; q < k; ++q) acc += row[q] * AT(T, ldt, j, q);
                AT(W, ldw, r, j) = acc;
            }
        }
        if (n > k)
            for (int64_t q = k; q < n; ++q)
                for (int64_t r = 0; r < m; ++r) {
                    R acc = 0;
                    for (int64_t j = 0; j < k; ++j) acc += AT(W, ldw, r, j) * AT(V, ldv, q, j);
                    AT(C, ldc, r, q) -= acc;
                }
        for (int64_t r = 0; r < m; ++r) {
            R row[4096];
            for (int64_t j = 0; j < k; ++j) row[j] = AT(W, ldw, r, j);
            for (int64_t j = 0; j < k; ++j) {
                R acc = row[j];
                for (int64_t q = 0; q < j; ++q) acc += row[q] * AT(V, ldv, j, q);
                AT(C, ldc, r, j) -= acc;
            }
        }
    }
}

/* geqrt!: src/geqrt.jl:22-68.  T is ib x n (block j's sb x sb triangle at T[0:sb, i:i+sb]). */
void P(geqrt)(int64_t m, int64_t n, int64_t ib, R *A, int64_t lda, R *T, int64_t ldt, R *tau, R *work) {
    if (m <= 0 || n <= 0 || ib <= 0) return;
    int64_t k = m < n ? m : n;
    for (int64_t i = 0; i < k; i += ib) {
        int64_t sb = (ib < k - i) ? ib : k - i;
        P(geqr2)(m - i, sb, &AT(A, lda, i, i), lda, tau + i, work);
        P(larft_fc)(m - i, sb, &AT(A, lda, i, i), lda, tau + i, &AT(T, ldt, 0, i), ldt);
        if (n > i + sb)
            P(larfb)('L', 'C', m - i, n - i - sb, sb, &AT(A, lda, i, i), lda, &AT(T, ldt, 0, i), ldt,
                     &AT(A, lda, i, i + sb), lda, work, n - i - sb);
    }
}

/* unmqr!: src/unmqr.jl:52-149.  a_cols = size(A,2) used by the k clamp (src/unmqr.jl:83).
 * work is n x ib (side L) or m x ib (side R). */
void P(unmqr)(char side, char trans, int64_t m, int64_t n, int64_t k, int64_t ib, const R *A, int64_t lda,
              int64_t a_cols, const R *T, int64_t ldt, R *C, int64_t ldc, R *work) {
    int64_t nq = side == 'L' ? m : n;
    if (k > nq) k = nq;
    if (k > a_cols) k = a_cols;
    if (m <= 0 || n <= 0 || k <= 0 || ib <= 0) return;
    int fwd = (side == 'L' && trans != 'N') || (side == 'R' && trans == 'N');
    int64_t nblk = (k + ib - 1) / ib;
    for (int64_t b = 0; b < nblk; ++b) {
        int64_t i = fwd ? b * ib : (nblk - 1 - b) * ib;
        int64_t kb = (ib < k - i) ? ib : k - i;
        if (side == 'L')
            P(larfb)('L', trans, m - i, n, kb, &AT(A, lda, i, i), lda, &AT(T, ldt, 0, i), ldt,
                     &AT(C, ldc, i, 0), ldc, work, n);
        else
            P(larfb)('R', trans, m, n - i, kb, &AT(A, lda, i, i), lda, &AT(T, ldt, 0, i), ldt,
                     &AT(C, ldc, 0, i), ldc, work, m);
    }
}

/* parfb! with direct='F', storev='C', l=0: src/parfb.jl:44-191 and the
 * pamm_w!/pamm_a! F/C branches (src/pamm.jl:104-176, 306-368).
 * Side L: W (k x n2, ld ldw) = A1[0:k,:] + V^T A2; W = op(T) W; A1[0:k,:] -= W; A2 -= V W.
 * Side R: W (m2 x k) = A1[:,0:k] + A2 V; W = W op(T); A1[:,0:k] -= W; A2 -= W V^T. */
void P(parfb)(char side, char trans, int64_t m1, int64_t n1, int64_t m2, int64_t n2, int64_t k,
              R *A1, int64_t lda1, R *A2, int64_t lda2, const R *V, int64_t ldv, const R *T, int64_t ldt,
              R *W, int64_t ldw) {
    if (m1 <= 0 || n1 <= 0 || n2 <= 0 || k <= 0) return;
    int notrans = (trans == 'N');
    if (side == 'L') {
        for (int64_t c = 0; c < n2; ++c)
            for (int64_t i = 0; i < k; ++i) {
                R acc = 0;
                for (int64_t r = 0; r < m2; ++r) acc += AT(V, ldv, r, i) * AT(A2, lda2, r, c);
                AT(W, ldw, i, c) = acc + AT(A1, lda1, i, c);
            }
        for (int64_t c = 0; c < n2; ++c) {
            R col[4096];
            for (int64_t i = 0; i < k; ++i) col[i] = AT(W, ldw, i, c);
            for (int64_t i = 0; i < k; ++i) {
                R acc = 0;
                if (notrans) for (int64_t q = i; q < k; ++q) acc += AT(T, ldt, i, q) * col[q];
                else         for (int64_t q = 0; q <= i; ++q) acc += AT(T, ldt, q, i) * col[q];
                AT(W, ldw, i, c) = acc;
            }
        }
        for (int64_t c = 0; c < n2; ++c)
            for (int64_t i = 0; i < k; ++i) AT(A1, lda1, i, c) -= AT(W, ldw, i, c);
        for (int64_t c = 0; c < n2; ++c)
            for (int64_t r = 0; r < m2; ++r) {
                R acc = 0;
                for (int64_t i = 0; i < k; ++i) acc += AT(V, ldv, r, i) * AT(W, ldw, i, c);
                AT(A2, lda2, r, c) -= acc;
            }
    } else {
        for (int64_t j = 0; j < k; ++j)
            for (int64_t r = 0; r < m2; ++r) {
                R acc = 0;
                for (int64_t q = 0; q < n2; ++q) acc += AT(A2, lda2, r, q) * AT(V, ldv, q, j);
                AT(W, ldw, r, j) = acc + AT(A1, lda1, r, j);
            }
        for (int64_t r = 0; r < m2; ++r) {
            R row[4096];
            for (int64_t j = 0; j < k; ++j) row[j] = AT(W, ldw, r, j);
            for (int64_t j = 0; j < k; ++j) {
                R acc = 0;
                if (notrans) for (int64_t q = 0; q <= j; ++q) acc += row[q] * AT(T, ldt, q, j);
                else         for (int64_t q = j; q < k; ++q) acc += row[q] * AT(T, ldt, j, q);
                AT(W, ldw, r, j) = acc;
            }
        }
        for (int64_t j = 0; j < k; ++j)
            for (int64_t r = 0; r < m2; ++r) AT(A1, lda1, r, j) -= AT(W, ldw, r, j);
        for (int64_t q = 0; q < n2; ++q)
            for (int64_t r = 0; r < m2; ++r) {
                R acc = 0;
                for (int64_t j = 0; j < k; ++j) acc += AT(W, ldw, r, j) * AT(V, ldv, q, j);
                AT(A2, lda2, r, q) -= acc;
            }
    }
}

/* tsmqr!: src/tsmqr.jl:47-146.  v_cols / t_cols are size(V,2) / size(T,2) for the k clamp.
 * work holds ib*n1 (side L) or m1*ib (side R) elements. */
void P(tsmqr)(char side, char trans, int64_t m1, int64_t n1, int64_t m2, int64_t n2, int64_t k, int64_t ib,
              R *A1, int64_t lda1, R *A2, int64_t lda2, const R *V, int64_t ldv, int64_t v_cols,
              const R *T, int64_t ldt, int64_t t_cols, R *work) {
    int64_t kmax = side == 'L' ? m1 : n1;
    int64_t stored = v_cols < t_cols ? v_cols : t_cols;
    if (k > kmax) k = kmax;
    if (k > stored) k = stored;
    if (m1 <= 0 || n1 <= 0 || m2 <= 0 || n2 <= 0 || k <= 0 || ib <= 0) return;
    int fwd = (side == 'L' && trans != 'N') || (side == 'R' && trans == 'N');
    int64_t nblk = (k + ib - 1) / ib;
    for (int64_t b = 0; b < nblk; ++b) {
        int64_t i = fwd ? b * ib : (nblk - 1 - b) * ib;
        int64_t kb = (ib < k - i) ? ib : k - i;
        if (side == 'L')
            P(parfb)('L', trans, m1 - i, n1, m2, n2, kb, &AT(A1, lda1, i, 0), lda1, A2, lda2,
                     &AT(V, ldv, 0, i), ldv, &AT(T, ldt, 0, i), ldt, work, kb);
        else
            P(parfb)('R', trans, m1, n1 - i, m2, n2, kb, &AT(A1, lda1, 0, i), lda1, A2, lda2,
                     &AT(V, ldv, 0, i), ldv, &AT(T, ldt, 0, i), ldt, work, m1);
    }
}

/* tsqrt!: src/tsqrt.jl:46-156.  QR of [A1 (m1 x n, upper); A2 (m x n)], k = min(m1, n).
 * A column of A2 that is already zero gets tau = 0 and a zero T column (LAPACK dtpqrt2
 * semantics; the reference leaves that T column unwritten). */
void P(tsqrt)(int64_t m, int64_t n, int64_t ib, R *A1, int64_t lda1, int64_t m1, R *A2, int64_t lda2,
              R *T, int64_t ldt, R *tau, R *work) {
    int64_t k = m1 < n ? m1 : n;
    if (m <= 0 || n <= 0 || ib <= 0 || k <= 0) return;
    for (int64_t ii = 0; ii < k; ii += ib) {
        int64_t sb = (ib < k - ii) ? ib : k - ii;
        for (int64_t i = 0; i < sb; ++i) {
            int64_t r = ii + i;
            R alpha = AT(A1, lda1, r, r);
            R xs = 0;
            for (int64_t p = 0; p < m; ++p) { R v = AT(A2, lda2, p, r); xs += v * v; }
            if (xs == 0) {
                tau[r] = 0;
                for (int64_t p = 0; p <= i; ++p) AT(T, ldt, p, r) = 0;
                continue;
            }
            R beta = -(R)copysign(sqrt((double)(alpha * alpha + xs)), (double)alpha);
            R t = (beta - alpha) / beta;
            R scale = (R)1 / (alpha - beta);
            for (int64_t p = 0; p < m; ++p) AT(A2, lda2, p, r) *= scale;
            AT(A1, lda1, r, r) = beta;
            tau[r] = t;
            for (int64_t j = 0; j < sb - i - 1; ++j) {
                R w = AT(A1, lda1, r, r + 1 + j);
                for (int64_t p = 0; p < m; ++p) w += AT(A2, lda2, p, r + 1 + j) * AT(A2, lda2, p, r);
                work[j] = w;
            }
            for (int64_t j = 0; j < sb - i - 1; ++j) AT(A1, lda1, r, r + 1 + j) -= t * work[j];
            for (int64_t j = 0; j < sb - i - 1; ++j)
                for (int64_t p = 0; p < m; ++p) AT(A2, lda2, p, r + 1 + j) -= t * AT(A2, lda2, p, r) * work[j];
            if (i > 0) {
                for (int64_t p = 0; p < i; ++p) {
                    R acc = 0;
                    for (int64_t q = 0; q < m; ++q) acc += AT(A2, lda2, q, ii + p) * AT(A2, lda2, q, r);
                    AT(T, ldt, p, r) = -t * acc;
                }
                for (int64_t p = 0; p < i; ++p) work[p] = AT(T, ldt, p, r);
                for (int64_t p = 0; p < i; ++p) {
                    R acc = 0;
                    for (int64_t q = p; q < i; ++q) acc += AT(T, ldt, p, ii + q) * work[q];
                    AT(T, ldt, p, r) = acc;
                }
            }
            AT(T, ldt, i, r) = t;
        }
        if (n > ii + sb)
            P(tsmqr)('L', 'C', sb, n - ii - sb, m, n - ii - sb, sb, ib, &AT(A1, lda1, ii, ii + sb), lda1,
                     &AT(A2, lda2, 0, ii + sb), lda2, &AT(A2, lda2, 0, ii), lda2, sb,
                     &AT(T, ldt, 0, ii), ldt, sb, work);
    }
}

/* ------------------------------------------------------------------------- */
/* Recursive TRSM/TRMM: src/rectrxm.jl:43-198, base kernels src/trsm.jl:5-150,
 * src/trmm.jl:43-389, GEMM_ADD!/GEMM_SUB! src/matmul.jl:5-81.
 * op(A) is represented by an accessor: element (i,j) of op(A) = A[i,j] or A[j,i]. */

typedef struct { const R *a; int64_t ld; int tr; } P(mat_t);
static inline R P(ga)(P(mat_t) A, int64_t i, int64_t j) {
    return A.tr ? AT(A.a, A.ld, j, i) : AT(A.a, A.ld, i, j);
}
static inline P(mat_t) P(sub)(P(mat_t) A, int64_t i, int64_t j) {
    P(mat_t) s = A;
    s.a = A.tr ? &AT(A.a, A.ld, j, i) : &AT(A.a, A.ld, i, j);
    return s;
}

/* C[N x M] += alpha * X[N x Rk] * Y[Rk x M] (X, Y through accessors, C plain). */
static void P(gemm_acc)(int64_t N, int64_t Rk, int64_t M, R alpha, P(mat_t) X, P(mat_t) Y, R *C, int64_t ldc) {
    if (N <= 0 || M <= 0 || Rk <= 0) return;
    #pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < M; ++j) {
        for (int64_t r0 = 0; r0 < Rk; r0 += 256) {
            int64_t r1 = r0 + 256 < Rk ? r0 + 256 : Rk;
            for (int64_t r = r0; r < r1; ++r) {
                R y = alpha * P(ga)(Y, r, j);
                if (y == 0) continue;
                if (!X.tr) {
                    const R *xc = &AT(X.a, X.ld, 0, r);
                    R *cc = &AT(C, ldc, 0, j);
                    for (int64_t i = 0; i < N; ++i) cc[i] += xc[i] * y;
                } else {
                    for (int64_t i = 0; i < N; ++i) AT(C, ldc, i, j) += P(ga)(X, i, r) * y;
                }
            }
        }
    }
}

/* Base TRSM kernels: src/trsm.jl:5-126.  The left kernels scale by the diagonal
 * first (x = b/d - sum (a/d) x), as the reference does. */
static void P(trsm_base)(char side, char uplo, int64_t n, int64_t m, P(mat_t) A, R *B, int64_t ldb) {
    if (side == 'L') {
        #pragma omp parallel for schedule(static)
        for (int64_t c = 0; c < m; ++c) {
            R *b = &AT(B, ldb, 0, c);
            for (int64_t r = 0; r < n; ++r) b[r] /= P(ga)(A, r, r);
            if (uplo == 'L') {
                for (int64_t i = 0; i < n; ++i)
                    for (int64_t r = i + 1; r < n; ++r) b[r] -= (P(ga)(A, r, i) / P(ga)(A, r, r)) * b[i];
            } else {
                for (int64_t i = n - 1; i >= 0; --i)
                    for (int64_t r = 0; r < i; ++r) b[r] -= (P(ga)(A, r, i) / P(ga)(A, r, r)) * b[i];
            }
        }
    } else {
        /* X * A = B: B is m x n, A is n x n */
        #pragma omp parallel for schedule(static)
        for (int64_t r = 0; r < m; ++r) {
            for (int64_t c = 0; c < n; ++c) AT(B, ldb, r, c) /= P(ga)(A, c, c);
            if (uplo == 'L') {
                for (int64_t i = n - 1; i >= 0; --i)
                    for (int64_t c = 0; c < i; ++c)
                        AT(B, ldb, r, c) -= AT(B, ldb, r, i) * (P(ga)(A, i, c) / P(ga)(A, c, c));
            } else {
                for (int64_t i = 0; i < n; ++i)
                    for (int64_t c = i + 1; c < n; ++c)
                        AT(B, ldb, r, c) -= AT(B, ldb, r, i) * (P(ga)(A, i, c) / P(ga)(A, c, c));
            }
        }
    }
}

/* Base TRMM kernels: src/trmm.jl:43-312 (B := A*B or B := B*A, triangle of A only). */
static void P(trmm_base)(char side, char uplo, int64_t n, int64_t m, P(mat_t) A, R *B, int64_t ldb) {
    if (side == 'L') {
        #pragma omp parallel for schedule(static)
        for (int64_t c = 0; c < m; ++c) {
            R tmp[1024];
            R *b = &AT(B, ldb, 0, c);
            for (int64_t r = 0; r < n; ++r) {
                R acc = 0;
                if (uplo == 'L') for (int64_t q = 0; q <= r; ++q) acc += P(ga)(A, r, q) * b[q];
                else             for (int64_t q = r; q < n; ++q) acc += P(ga)(A, r, q) * b[q];
                tmp[r] = acc;
            }
            for (int64_t r = 0; r < n; ++r) b[r] = tmp[r];
        }
    } else {
        #pragma omp parallel for schedule(static)
        for (int64_t r = 0; r < m; ++r) {
            R tmp[1024];
            for (int64_t c = 0; c < n; ++c) {
                R acc = 0;
                if (uplo == 'L') for (int64_t q = c; q < n; ++q) acc += AT(B, ldb, r, q) * P(ga)(A, q, c);
                else             for (int64_t q = 0; q <= c; ++q) acc += AT(B, ldb, r, q) * P(ga)(A, q, c);
                tmp[c] = acc;
            }
            for (int64_t c = 0; c < n; ++c) AT(B, ldb, r, c) = tmp[c];
        }
    }
}

/* unified_rec: src/rectrxm.jl:101-198.  B is n x m (side L) or m x n (side R). */
static void P(rec)(char func, char side, char uplo, P(mat_t) A, int64_t n, int64_t m, R *B, int64_t ldb,
                   int64_t threshold) {
    if (n <= threshold) {
        if (func == 'S') P(trsm_base)(side, uplo, n, m, A, B, ldb);
        else             P(trmm_base)(side, uplo, n, m, A, B, ldb);
        return;
    }
    int64_t mid;
    if ((n & (n - 1)) == 0) mid = n / 2;
    else { mid = 1; while (mid * 2 <= n) mid *= 2; }
    int64_t rem = n - mid;
    P(mat_t) A11 = A, A22 = P(sub)(A, mid, mid), A21 = P(sub)(A, mid, 0), A12 = P(sub)(A, 0, mid);
    R *B1 = B, *B2 = (side == 'L') ? &AT(B, ldb, mid, 0) : &AT(B, ldb, 0, mid);
    P(mat_t) mB1 = {B1, ldb, 0}, mB2 = {B2, ldb, 0};
    int forward = (side == 'L' && uplo == 'L' && func == 'S') || (side == 'R' && uplo == 'U' && func == 'S') ||
                  (side == 'L' && uplo == 'U' && func == 'M') || (side == 'R' && uplo == 'L' && func == 'M');
    if (forward) {
        P(rec)(func, side, uplo, A11, mid, m, B1, ldb, threshold);
        if (side == 'L') {
            if (func == 'S') P(gemm_acc)(rem, mid, m, (R)-1, A21, mB1, B2, ldb);   /* B2 -= A21 B1 */
            else             P(gemm_acc)(mid, rem, m, (R)1, A12, mB2, B1, ldb);    /* B1 += A12 B2 */
        } else {
            if (func == 'S') P(gemm_acc)(m, mid, rem, (R)-1, mB1, A12, B2, ldb);   /* B2 -= B1 A12 */
            else             P(gemm_acc)(m, rem, mid, (R)1, mB2, A21, B1, ldb);    /* B1 += B2 A21 */
        }
        P(rec)(func, side, uplo, A22, rem, m, B2, ldb, threshold);
    } else {
        P(rec)(func, side, uplo, A22, rem, m, B2, ldb, threshold);
        if (side == 'L') {
            if (func == 'S') P(gemm_acc)(mid, rem, m, (R)-1, A12, mB2, B1, ldb);   /* B1 -= A12 B2 */
            else             P(gemm_acc)(rem, mid, m, (R)1, A21, mB1, B2, ldb);    /* B2 += A21 B1 */
        } else {
            if (func == 'S') P(gemm_acc)(m, rem, mid, (R)-1, mB2, A21, B1, ldb);   /* B1 -= B2 A21 */
            else             P(gemm_acc)(m, mid, rem, (R)1, mB1, A12, B2, ldb);    /* B2 += B1 A12 */
        }
        P(rec)(func, side, uplo, A11, mid, m, B1, ldb, threshold);
    }
}

/* unified_rectrxm!: src/rectrxm.jl:43-76.  n = order of A; m = the other dimension of B.
 * threshold <= 0 selects the reference defaults (256 for 'S', 16 for 'M'). */
void P(rectrxm)(char side, char uplo, char trans, R alpha, char func, int64_t n, int64_t m,
                const R *A, int64_t lda, R *B, int64_t ldb, int64_t threshold) {
    if (n <= 0 || m <= 0) return;
    P(mat_t) mA = {A, lda, 0};
    if (trans == 'T' || trans == 'C') { mA.tr = 1; uplo = (uplo == 'L') ? 'U' : 'L'; }
    if (threshold <= 0) threshold = (func == 'S') ? 256 : 16;
    int64_t rows = side == 'L' ? n : m, cols = side == 'L' ? m : n;
    if (func == 'S' && alpha != 1)
        for (int64_t j = 0; j < cols; ++j) for (int64_t i = 0; i < rows; ++i) AT(B, ldb, i, j) *= alpha;
    P(rec)(func, side, uplo, mA, n, m, B, ldb, threshold);
    if (func == 'M' && alpha != 1)
        for (int64_t j = 0; j < cols; ++j) for (int64_t i = 0; i < rows; ++i) AT(B, ldb, i, j) *= alpha;
}

/* ------------------------------------------------------------------------- */
/* Tile QR driver (no reference function; the composition the reference's tests and
 * call stacks imply): flat-tree PLASMA order over mt x nt tiles of size nb:
 *   geqrt(A_kk); unmqr('L','C') on A_kj; tsqrt(A_kk, A_ik); tsmqr('L','C') on (A_kj, A_ij).
 * T holds one ib x nb block per tile: T tile (i,k) at T[i*ib, k*nb], ldt >= mt*ib.
 * Requires m and n to be multiples of nb. */
void P(tile_geqrf)(int64_t m, int64_t n, int64_t nb, int64_t ib, R *A, int64_t lda, R *T, int64_t ldt) {
    int64_t mt = m / nb, nt = n / nb, kt = mt < nt ? mt : nt;
    for (int64_t k = 0; k < kt; ++k) {
        R *Akk = &AT(A, lda, k * nb, k * nb);
        R *Tkk = &AT(T, ldt, k * ib, k * nb);
        {
            R tau[4096], work[4096 * 8];
            P(geqrt)(nb, nb, ib, Akk, lda, Tkk, ldt, tau, work);
        }
        #pragma omp parallel for schedule(dynamic)
        for (int64_t j = k + 1; j < nt; ++j) {
            R *work = (R *)malloc(sizeof(R) * nb * ib);
            P(unmqr)('L', 'C', nb, nb, nb, ib, Akk, lda, nb, Tkk, ldt, &AT(A, lda, k * nb, j * nb), lda, work);
            free(work);
        }
        for (int64_t i = k + 1; i < mt; ++i) {
            R *Aik = &AT(A, lda, i * nb, k * nb);
            R *Tik = &AT(T, ldt, i * ib, k * nb);
            {
                R tau[4096];
                R *work = (R *)malloc(sizeof(R) * nb * ib);
                P(tsqrt)(nb, nb, ib, Akk, lda, nb, Aik, lda, Tik, ldt, tau, work);
                free(work);
            }
            #pragma omp parallel for schedule(dynamic)
            for (int64_t j = k + 1; j < nt; ++j) {
                R *work = (R *)malloc(sizeof(R) * nb * ib);
                P(tsmqr)('L', 'C', nb, nb, nb, nb, nb, ib, &AT(A, lda, k * nb, j * nb), lda,
                         &AT(A, lda, i * nb, j * nb), lda, Aik, lda, nb, Tik, ldt, nb, work);
                free(work);
            }
        }
    }
}

/* Apply Q (trans='N') or Q^T (trans='T'/'C') of tile_geqrf from the left to C (m x nrhs). */
void P(tile_ormqr)(char trans, int64_t m, int64_t n, int64_t nrhs, int64_t nb, int64_t ib, const R *A, int64_t lda,
                   const R *T, int64_t ldt, R *C, int64_t ldc) {
    int64_t mt = m / nb, nt = n / nb, kt = mt < nt ? mt : nt;
    int notrans = (trans == 'N');
    for (int64_t kk = 0; kk < kt; ++kk) {
        int64_t k = notrans ? kt - 1 - kk : kk;
        const R *Akk = &AT(A, lda, k * nb, k * nb);
        const R *Tkk = &AT(T, ldt, k * ib, k * nb);
        if (!notrans) {
            R *work = (R *)malloc(sizeof(R) * (nrhs > 0 ? nrhs : 1) * ib);
            P(unmqr)('L', 'C', nb, nrhs, nb, ib, Akk, lda, nb, Tkk, ldt, &AT(C, ldc, k * nb, 0), ldc, work);
            free(work);
        }
        for (int64_t ii = 0; ii < mt - k - 1; ++ii) {
            int64_t i = notrans ? mt - 1 - ii : k + 1 + ii;
            R *work = (R *)malloc(sizeof(R) * (nrhs > 0 ? nrhs : 1) * ib);
            P(tsmqr)('L', notrans ? 'N' : 'C', nb, nrhs, nb, nrhs, nb, ib, &AT(C, ldc, k * nb, 0), ldc,
                     &AT(C, ldc, i * nb, 0), ldc, &AT(A, lda, i * nb, k * nb), lda, nb,
                     &AT(T, ldt, i * ib, k * nb), ldt, nb, work);
            free(work);
        }
        if (notrans) {
            R *work = (R *)malloc(sizeof(R) * (nrhs > 0 ? nrhs : 1) * ib);
            P(unmqr)('L', 'N', nb, nrhs, nb, ib, Akk, lda, nb, Tkk, ldt, &AT(C, ldc, k * nb, 0), ldc, work);
            free(work);
        }
    }
}

#undef AT
