/*
 * Scalar-generic CPU restatement of NextLA.jl's QR-family routines, for real AND complex
 * element types: larfg!, geqr2!, larft! (forward), larfb!, geqrt!, unmqr!, tsqrt!, tsmqr!,
 * parfb! (every side / trans / direct / storev / l), pamm!, pemv!, ttqrt!, ttmqr!.
 * TEST INFRASTRUCTURE ONLY: the checker used by tests/ and smoke(); nothing in the shipped
 * library calls it.
 *
 * Type template: nla_oracle.c includes it once per element type with
 *   R       the element type (double, float, double _Complex, float _Complex)
 *   RR      its real type
 *   P(x)    the name prefix (nlaq_d##x, nlaq_s##x, nlaq_z##x, nlaq_c##x)
 *   CONJ(x), RE(x), IM(x), MKC(re, im), RSQRT(x)  scalar helpers (identity / 0 for reals)
 * Matrices are column-major with explicit leading dimensions; indices are 0-based where
 * the reference is 1-based.  Conjugations follow the reference's complex branches, which
 * are LAPACK's: H = I - tau v v^H, the QR routines apply H^H from the left.
 */

#define QAT(A, ld, i, j) (A)[(size_t)(i) + (size_t)(j) * (size_t)(ld)]
#define ABS2(x) (RE(x) * RE(x) + IM(x) * IM(x))

static RR P(q_abs)(R x) { return RSQRT(ABS2(x)); }

/* safe sqrt(a^2 + b^2 + c^2) */
static RR P(q_hyp3)(RR a, RR b, RR c) {
    RR fa = a < 0 ? -a : a, fb = b < 0 ? -b : b, fc = c < 0 ? -c : c;
    RR mx = fa > fb ? fa : fb;
    mx = mx > fc ? mx : fc;
    if (mx == 0) return 0;
    fa /= mx; fb /= mx; fc /= mx;
    return mx * RSQRT(fa * fa + fb * fb + fc * fc);
}

static RR P(q_nrm2)(int64_t n, const R *x, int64_t incx) {
    RR scale = 0, ssq = 1;
    for (int64_t i = 0; i < n; ++i) {
        RR parts[2] = {RE(x[i * incx]), IM(x[i * incx])};
        for (int q = 0; q < 2; ++q) {
            RR a = parts[q] < 0 ? -parts[q] : parts[q];
            if (a != 0) {
                if (scale < a) { ssq = 1 + ssq * (scale / a) * (scale / a); scale = a; }
                else ssq += (a / scale) * (a / scale);
            }
        }
    }
    return scale * RSQRT(ssq);
}

static RR P(q_safmin)(void) {
    RR eps = (sizeof(RR) == 8) ? (RR)DBL_EPSILON * (RR)0.5 : (RR)FLT_EPSILON * (RR)0.5;
    RR sfmin = (sizeof(RR) == 8) ? (RR)DBL_MIN : (RR)FLT_MIN;
    return sfmin / eps;
}

/* larfg!: src/larfg.jl:43-106 with its complex branch (tau = (beta - alphr)/beta -
 * i alphi/beta, zero iff ||x|| = 0 and alphi = 0); beta via a safe 3-term norm and the
 * rescale loop bounded at 20 passes (the reference's unbounded loop at src/larfg.jl:70-77
 * never ends when beta underflows to 0). */
void P(larfg)(int64_t n, R *alpha, R *x, int64_t incx, R *tau) {
    if (n <= 1) { *tau = 0; return; }
    RR xnorm = P(q_nrm2)(n - 1, x, incx);
    RR alphr = RE(*alpha), alphi = IM(*alpha);
    if (xnorm == 0 && alphi == 0) { *tau = 0; return; }
    RR beta = -(RR)copysign((double)P(q_hyp3)(alphr, alphi, xnorm), (double)alphr);
    RR safmin = P(q_safmin)(), rsafmn = (RR)1 / safmin;
    int knt = 0;
    R a = *alpha;
    if ((beta < 0 ? -beta : beta) < safmin) {
        while ((beta < 0 ? -beta : beta) < safmin && knt < 20) {
            knt++;
            for (int64_t i = 0; i < n - 1; ++i) x[i * incx] *= rsafmn;
            beta *= rsafmn; alphr *= rsafmn; alphi *= rsafmn;
        }
        xnorm = P(q_nrm2)(n - 1, x, incx);
        a = MKC(alphr, alphi);
        beta = -(RR)copysign((double)P(q_hyp3)(alphr, alphi, xnorm), (double)alphr);
    }
    *tau = MKC((beta - alphr) / beta, -alphi / beta);
    R s = (R)1 / (a - (R)beta);
    for (int64_t i = 0; i < n - 1; ++i) x[i * incx] *= s;
    for (int j = 0; j < knt; ++j) beta *= safmin;
    *alpha = beta;
}

/* geqr2!: kernel src/geqr2.jl:23-105 (plain sum-of-squares norm; tau = 0 for a column of
 * length 1; H(i)^H applied to the trailing columns: A -= conj(tau) v (v^H A)). */
void P(geqr2)(int64_t m, int64_t n, R *A, int64_t lda, R *tau, R *work) {
    if (m <= 0 || n <= 0) return;
    int64_t k = m < n ? m : n;
    for (int64_t i = 0; i < k; ++i) {
        int64_t len = m - i;
        if (len <= 1) { tau[i] = 0; continue; }
        R alpha = QAT(A, lda, i, i);
        RR alphr = RE(alpha), alphi = IM(alpha), xs = 0;
        for (int64_t p = 1; p < len; ++p) xs += ABS2(QAT(A, lda, i + p, i));
        if (xs == 0 && alphi == 0) { tau[i] = 0; continue; }
        RR beta = -(RR)copysign(sqrt((double)(alphr * alphr + alphi * alphi + xs)), (double)alphr);
        R t = MKC((beta - alphr) / beta, -alphi / beta);
        R scale = (R)1 / (alpha - (R)beta);
        for (int64_t p = 1; p < len; ++p) QAT(A, lda, i + p, i) *= scale;
        QAT(A, lda, i, i) = beta;
        tau[i] = t;
        if (i < n - 1) {
            QAT(A, lda, i, i) = 1;
            int64_t nc = n - i - 1;
            for (int64_t j = 0; j < nc; ++j) {
                R w = 0;
                for (int64_t p = 0; p < len; ++p) w += CONJ(QAT(A, lda, i + p, i)) * QAT(A, lda, i + p, i + 1 + j);
                work[j] = w;   /* v^H a_j */
            }
            for (int64_t j = 0; j < nc; ++j)
                for (int64_t p = 0; p < len; ++p)
                    QAT(A, lda, i + p, i + 1 + j) -= CONJ(t) * QAT(A, lda, i + p, i) * work[j];
            QAT(A, lda, i, i) = beta;
        }
    }
}

/* larft! direct='F': kernel src/larft.jl:48-126, columnwise (storev 'C', V n x k,
 * T(0:i,i) = -tau_i V^H v_i) and rowwise (storev 'R', V k x n, T(0:i,i) = -tau_i V v_i^H),
 * with LAPACK's lastv trimming; then T(0:i,i) = T(0:i,0:i) T(0:i,i). */
void P(larft)(char storev, int64_t n, int64_t k, const R *V, int64_t ldv, const R *tau, R *T, int64_t ldt) {
    if (n <= 0) return;
    R work[4096];
    int64_t prevlastv = n - 1;
    for (int64_t i = 0; i < k; ++i) {
        if (prevlastv < i) prevlastv = i;
        R ti = tau[i];
        if (RE(ti) == 0 && IM(ti) == 0) {
            for (int64_t j = 0; j <= i; ++j) QAT(T, ldt, j, i) = 0;
            continue;
        }
        int64_t lastv = n - 1;
        if (storev == 'C') {
            while (lastv >= i + 1 && RE(QAT(V, ldv, lastv, i)) == 0 && IM(QAT(V, ldv, lastv, i)) == 0) lastv--;
            int64_t j = lastv < prevlastv ? lastv : prevlastv;
            for (int64_t p = 0; p < i; ++p) QAT(T, ldt, p, i) = -ti * CONJ(QAT(V, ldv, i, p));
            for (int64_t p = 0; p < i; ++p) {
                R acc = 0;
                for (int64_t r = i + 1; r <= j; ++r) acc += CONJ(QAT(V, ldv, r, p)) * QAT(V, ldv, r, i);
                QAT(T, ldt, p, i) += -ti * acc;
            }
        } else {
            while (lastv >= i + 1 && RE(QAT(V, ldv, i, lastv)) == 0 && IM(QAT(V, ldv, i, lastv)) == 0) lastv--;
            int64_t j = lastv < prevlastv ? lastv : prevlastv;
            for (int64_t p = 0; p < i; ++p) QAT(T, ldt, p, i) = -ti * QAT(V, ldv, p, i);
            for (int64_t p = 0; p < i; ++p) {
                R acc = 0;
                for (int64_t r = i + 1; r <= j; ++r) acc += QAT(V, ldv, p, r) * CONJ(QAT(V, ldv, i, r));
                QAT(T, ldt, p, i) += -ti * acc;
            }
        }
        prevlastv = (i > 0) ? (prevlastv > lastv ? prevlastv : lastv) : lastv;
        for (int64_t p = 0; p < i; ++p) work[p] = QAT(T, ldt, p, i);
        for (int64_t p = 0; p < i; ++p) {
            R acc = 0;
            for (int64_t q = p; q < i; ++q) acc += QAT(T, ldt, p, q) * work[q];
            QAT(T, ldt, p, i) = acc;
        }
        QAT(T, ldt, i, i) = ti;
    }
}

/* larfb! direct='F', storev='C': src/larfb.jl:37-142 (LAPACK zlarfb's sequence).
 * Side L: W = C^H V (n x k), W = W T^H ('N') or W T ('C'), C -= V W^H.
 * Side R: W = C V (m x k), W = W T ('N') or W T^H ('C'), C -= W V^H.  V = [V1 unit lower; V2].
 * work holds n x k (side L) or m x k (side R), leading dimension ldw. */
void P(larfb)(char side, char trans, int64_t m, int64_t n, int64_t k, const R *V, int64_t ldv,
              const R *T, int64_t ldt, R *C, int64_t ldc, R *W, int64_t ldw) {
    if (m <= 0 || n <= 0 || k <= 0) return;
    int notrans = (trans == 'N');
    R row[4096];
    if (side == 'L') {
        for (int64_t j = 0; j < k; ++j)
            for (int64_t c = 0; c < n; ++c) {
                R acc = CONJ(QAT(C, ldc, j, c));
                for (int64_t r = j + 1; r < m; ++r) acc += CONJ(QAT(C, ldc, r, c)) * QAT(V, ldv, r, j);
                QAT(W, ldw, c, j) = acc;
            }
        for (int64_t c = 0; c < n; ++c) {
            for (int64_t j = 0; j < k; ++j) row[j] = QAT(W, ldw, c, j);
            for (int64_t j = 0; j < k; ++j) {
                R acc = 0;
                if (notrans) for (int64_t q = j; q < k; ++q) acc += row[q] * CONJ(QAT(T, ldt, j, q));
                else         for (int64_t q = 0; q <= j; ++q) acc += row[q] * QAT(T, ldt, q, j);
                QAT(W, ldw, c, j) = acc;
            }
        }
        for (int64_t c = 0; c < n; ++c)
            for (int64_t r = 0; r < m; ++r) {
                R acc = 0;
                for (int64_t j = 0; j < k && j <= r; ++j) {
                    R v = (r == j) ? (R)1 : QAT(V, ldv, r, j);
                    acc += v * CONJ(QAT(W, ldw, c, j));
                }
                QAT(C, ldc, r, c) -= acc;
            }
    } else {
        for (int64_t j = 0; j < k; ++j)
            for (int64_t r = 0; r < m; ++r) {
                R acc = QAT(C, ldc, r, j);
                for (int64_t q = j + 1; q < n; ++q) acc += QAT(C, ldc, r, q) * QAT(V, ldv, q, j);
                QAT(W, ldw, r, j) = acc;
            }
        for (int64_t r = 0; r < m; ++r) {
            for (int64_t j = 0; j < k; ++j) row[j] = QAT(W, ldw, r, j);
            for (int64_t j = 0; j < k; ++j) {
                R acc = 0;
                if (notrans) for (int64_t q = 0; q <= j; ++q) acc += row[q] * QAT(T, ldt, q, j);
                else         for (int64_t q = j; q < k; ++q) acc += row[q] * CONJ(QAT(T, ldt, j, q));
                QAT(W, ldw, r, j) = acc;
            }
        }
        for (int64_t q = 0; q < n; ++q)
            for (int64_t r = 0; r < m; ++r) {
                R acc = 0;
                for (int64_t j = 0; j < k && j <= q; ++j) {
                    R v = (q == j) ? (R)1 : QAT(V, ldv, q, j);
                    acc += QAT(W, ldw, r, j) * CONJ(v);
                }
                QAT(C, ldc, r, q) -= acc;
            }
    }
}

/* geqrt!: src/geqrt.jl:22-68.  T is ib x n (block j's sb x sb triangle at T[0:sb, i:i+sb]). */
void P(geqrt)(int64_t m, int64_t n, int64_t ib, R *A, int64_t lda, R *T, int64_t ldt, R *tau, R *work) {
    if (m <= 0 || n <= 0 || ib <= 0) return;
    int64_t k = m < n ? m : n;
    for (int64_t i = 0; i < k; i += ib) {
        int64_t sb = (ib < k - i) ? ib : k - i;
        P(geqr2)(m - i, sb, &QAT(A, lda, i, i), lda, tau + i, work);
        P(larft)('C', m - i, sb, &QAT(A, lda, i, i), lda, tau + i, &QAT(T, ldt, 0, i), ldt);
        if (n > i + sb)
            P(larfb)('L', 'C', m - i, n - i - sb, sb, &QAT(A, lda, i, i), lda, &QAT(T, ldt, 0, i), ldt,
                     &QAT(A, lda, i, i + sb), lda, work, n - i - sb);
    }
}

/* unmqr!: src/unmqr.jl:52-149.  a_cols = size(A,2) for the k clamp (src/unmqr.jl:83).
 * work holds n x ib (side L) or m x ib (side R). */
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
            P(larfb)('L', trans, m - i, n, kb, &QAT(A, lda, i, i), lda, &QAT(T, ldt, 0, i), ldt,
                     &QAT(C, ldc, i, 0), ldc, work, n);
        else
            P(larfb)('R', trans, m, n - i, kb, &QAT(A, lda, i, i), lda, &QAT(T, ldt, 0, i), ldt,
                     &QAT(C, ldc, 0, i), ldc, work, m);
    }
}

/* ------------------------------------------------------------------------------------------
 * Pentagonal block reflectors: parfb! (src/parfb.jl:44-191) over pamm! (src/pamm.jl:49-473).
 * Every (direct, storev) branch of the reference is the same product with a different view
 * of V: in "column form" Vc (rows x k), Vc(r, j) = V(r, j) (storev 'C') or conj(V(j, r))
 * (storev 'R'), and H = I - Vc T Vc^H.  The pentagonal structure (order l) zeroes part of Vc:
 *   direct 'F': the last l rows are the first l rows of a k x k upper triangle:
 *               Vc(r, j) = 0 for r >= rows - l and r - (rows - l) > j;
 *   direct 'B': the first l rows are the last l rows of a k x k lower triangle:
 *               Vc(r, j) = 0 for r < l and j > k - l + r.
 * T is upper triangular for 'F', lower for 'B'; entries outside the structure are never
 * read (the reference's triangular products and block bounds skip them).
 *   side 'L' (A1 k x n, A2 rows x n):  W = A1 + Vc^H A2;  W = op(T) W;  A1 -= W;  A2 -= Vc W
 *   side 'R' (A1 m x k, A2 m x rows):  W = A1 + A2 Vc;    W = W op(T);  A1 -= W;  A2 -= W Vc^H
 * with op(T) = T ('N') or T^H ('C'/'T'). */
static int P(q_vmask)(char direct, int64_t rows, int64_t k, int64_t l, int64_t r, int64_t j) {
    if (direct == 'F') return !(r >= rows - l && r - (rows - l) > j);
    return !(r < l && j > k - l + r);
}
static R P(q_vc)(char storev, const R *V, int64_t ldv, int64_t r, int64_t j) {
    return storev == 'C' ? QAT(V, ldv, r, j) : CONJ(QAT(V, ldv, j, r));
}
/* element (i, j) of op(T) restricted to T's triangle */
static R P(q_opt)(char direct, int notrans, const R *T, int64_t ldt, int64_t i, int64_t j) {
    int64_t a = notrans ? i : j, b = notrans ? j : i;   /* T(a, b) */
    int inside = direct == 'F' ? (a <= b) : (a >= b);
    if (!inside) return 0;
    return notrans ? QAT(T, ldt, a, b) : CONJ(QAT(T, ldt, a, b));
}

/* pamm!: src/pamm.jl:49-473.  op 'W': W = A1 + Vc^H A2 (side L, W k x n) or A1 + A2 Vc
 * (side R, W m x k); op 'A': A2 -= Vc W (side L) or W Vc^H (side R).  m x n = A2's size. */
void P(pamm)(char op, char side, char storev, char direct, int64_t m, int64_t n, int64_t k, int64_t l,
             const R *A1, int64_t lda1, R *A2, int64_t lda2, const R *V, int64_t ldv, R *W, int64_t ldw) {
    if (m <= 0 || n <= 0 || k <= 0) return;
    if (side == 'L') {
        if (op == 'W') {
            for (int64_t c = 0; c < n; ++c)
                for (int64_t j = 0; j < k; ++j) {
                    R acc = QAT(A1, lda1, j, c);
                    for (int64_t r = 0; r < m; ++r)
                        if (P(q_vmask)(direct, m, k, l, r, j))
                            acc += CONJ(P(q_vc)(storev, V, ldv, r, j)) * QAT(A2, lda2, r, c);
                    QAT(W, ldw, j, c) = acc;
                }
        } else {
            for (int64_t c = 0; c < n; ++c)
                for (int64_t r = 0; r < m; ++r) {
                    R acc = 0;
                    for (int64_t j = 0; j < k; ++j)
                        if (P(q_vmask)(direct, m, k, l, r, j)) acc += P(q_vc)(storev, V, ldv, r, j) * QAT(W, ldw, j, c);
                    QAT(A2, lda2, r, c) -= acc;
                }
        }
    } else {
        if (op == 'W') {
            for (int64_t j = 0; j < k; ++j)
                for (int64_t r = 0; r < m; ++r) {
                    R acc = QAT(A1, lda1, r, j);
                    for (int64_t q = 0; q < n; ++q)
                        if (P(q_vmask)(direct, n, k, l, q, j)) acc += QAT(A2, lda2, r, q) * P(q_vc)(storev, V, ldv, q, j);
                    QAT(W, ldw, r, j) = acc;
                }
        } else {
            for (int64_t q = 0; q < n; ++q)
                for (int64_t r = 0; r < m; ++r) {
                    R acc = 0;
                    for (int64_t j = 0; j < k; ++j)
                        if (P(q_vmask)(direct, n, k, l, q, j))
                            acc += QAT(W, ldw, r, j) * CONJ(P(q_vc)(storev, V, ldv, q, j));
                    QAT(A2, lda2, r, q) -= acc;
                }
        }
    }
}

/* parfb!: src/parfb.jl:44-191.  work: k x n2 (side L) or m2 x k (side R), leading dimension ldw. */
void P(parfb)(char side, char trans, char direct, char storev, int64_t m1, int64_t n1, int64_t m2, int64_t n2,
              int64_t k, int64_t l, R *A1, int64_t lda1, R *A2, int64_t lda2, const R *V, int64_t ldv,
              const R *T, int64_t ldt, R *W, int64_t ldw) {
    if (m1 <= 0 || n1 <= 0 || n2 <= 0 || k <= 0) return;
    int notrans = (trans == 'N');
    R tmp[4096];
    P(pamm)('W', side, storev, direct, m2, n2, k, l, A1, lda1, A2, lda2, V, ldv, W, ldw);
    if (side == 'L') {
        for (int64_t c = 0; c < n2; ++c) {
            for (int64_t j = 0; j < k; ++j) tmp[j] = QAT(W, ldw, j, c);
            for (int64_t i = 0; i < k; ++i) {
                R acc = 0;
                for (int64_t j = 0; j < k; ++j) acc += P(q_opt)(direct, notrans, T, ldt, i, j) * tmp[j];
                QAT(W, ldw, i, c) = acc;
                QAT(A1, lda1, i, c) -= acc;
            }
        }
    } else {
        for (int64_t r = 0; r < m2; ++r) {
            for (int64_t j = 0; j < k; ++j) tmp[j] = QAT(W, ldw, r, j);
            for (int64_t j = 0; j < k; ++j) {
                R acc = 0;
                for (int64_t i = 0; i < k; ++i) acc += tmp[i] * P(q_opt)(direct, notrans, T, ldt, i, j);
                QAT(W, ldw, r, j) = acc;
                QAT(A1, lda1, r, j) -= acc;
            }
        }
    }
    P(pamm)('A', side, storev, direct, m2, n2, k, l, A1, lda1, A2, lda2, V, ldv, W, ldw);
}

/* tsmqr! (src/tsmqr.jl:47-146, l = 0) and ttmqr! (src/ttmqr.jl:1-95, l > 0): the ib-block
 * loop of parfb!('F','C').  For ttmqr! block i (0-based start, kb reflectors) uses the first
 * min(i + kb, m2) rows (side L; columns for side R) of A2 and V with
 * l = min(kb, max(0, m2 - i - 1)) -- the reference's src/ttmqr.jl:70-74 / 83-87.
 * work holds ib*n1 (side L) or m1*ib (side R) elements. */
static void P(q_tsmqr)(int tt, char side, char trans, int64_t m1, int64_t n1, int64_t m2, int64_t n2, int64_t k,
                       int64_t ib, R *A1, int64_t lda1, R *A2, int64_t lda2, const R *V, int64_t ldv, const R *T,
                       int64_t ldt, R *work) {
    if (m1 <= 0 || n1 <= 0 || m2 <= 0 || n2 <= 0 || k <= 0 || ib <= 0) return;
    int fwd = (side == 'L' && trans != 'N') || (side == 'R' && trans == 'N');
    int64_t nblk = (k + ib - 1) / ib;
    for (int64_t b = 0; b < nblk; ++b) {
        int64_t i = fwd ? b * ib : (nblk - 1 - b) * ib;
        int64_t kb = (ib < k - i) ? ib : k - i;
        if (side == 'L') {
            int64_t mi2 = tt ? ((i + kb < m2) ? i + kb : m2) : m2;
            int64_t l = tt ? (m2 - i - 1 > 0 ? (kb < m2 - i - 1 ? kb : m2 - i - 1) : 0) : 0;
            P(parfb)('L', trans, 'F', 'C', kb, n1, mi2, n2, kb, l, &QAT(A1, lda1, i, 0), lda1, A2, lda2,
                     &QAT(V, ldv, 0, i), ldv, &QAT(T, ldt, 0, i), ldt, work, kb);
        } else {
            int64_t ni2 = tt ? ((i + kb < n2) ? i + kb : n2) : n2;
            int64_t l = tt ? (n2 - i - 1 > 0 ? (kb < n2 - i - 1 ? kb : n2 - i - 1) : 0) : 0;
            P(parfb)('R', trans, 'F', 'C', m1, kb, m2, ni2, kb, l, &QAT(A1, lda1, 0, i), lda1, A2, lda2,
                     &QAT(V, ldv, 0, i), ldv, &QAT(T, ldt, 0, i), ldt, work, m1);
        }
    }
}

/* tsmqr!: src/tsmqr.jl:47-146.  v_cols / t_cols = size(V,2) / size(T,2) for the k clamp. */
void P(tsmqr)(char side, char trans, int64_t m1, int64_t n1, int64_t m2, int64_t n2, int64_t k, int64_t ib,
              R *A1, int64_t lda1, R *A2, int64_t lda2, const R *V, int64_t ldv, int64_t v_cols,
              const R *T, int64_t ldt, int64_t t_cols, R *work) {
    int64_t kmax = side == 'L' ? m1 : n1;
    int64_t stored = v_cols < t_cols ? v_cols : t_cols;
    if (k > kmax) k = kmax;
    if (k > stored) k = stored;
    P(q_tsmqr)(0, side, trans, m1, n1, m2, n2, k, ib, A1, lda1, A2, lda2, V, ldv, T, ldt, work);
}

/* ttmqr!: src/ttmqr.jl:1-95 (k clamped to side L ? m1 : n1, src/ttmqr.jl:38-40). */
void P(ttmqr)(char side, char trans, int64_t m1, int64_t n1, int64_t m2, int64_t n2, int64_t k, int64_t ib,
              R *A1, int64_t lda1, R *A2, int64_t lda2, const R *V, int64_t ldv, const R *T, int64_t ldt, R *work) {
    int64_t kmax = side == 'L' ? m1 : n1;
    if (k > kmax) k = kmax;
    P(q_tsmqr)(1, side, trans, m1, n1, m2, n2, k, ib, A1, lda1, A2, lda2, V, ldv, T, ldt, work);
}

/* One column of the TS / TT block QR (tsqrt_block_kernel!, src/tsqrt.jl:46-126, and
 * ttqrt_block_kernel!, src/ttqrt.jl:2-79): column r = ii + i of [A1 (upper); A2], whose
 * reflector has rows 0..mi-1 of A2 (mi = m for TS, min(r + 1, m) for TT); the in-block
 * columns r+1..ii+sb-1 get H^H; T(0:i, r) = -tau T(0:i, ii:ii+i) (A2(0:mj, ii:ii+i)^H A2(0:mj, r))
 * with mj = m (TS) or min(r, m) (TT). */
static void P(q_tcol)(int tt, int64_t m, int64_t n, int64_t ii, int64_t sb, int64_t i, R *A1, int64_t lda1, R *A2,
                      int64_t lda2, R *T, int64_t ldt, R *tau, R *work) {
    int64_t r = ii + i;
    int64_t mi = tt ? (r + 1 < m ? r + 1 : m) : m;
    R alpha = QAT(A1, lda1, r, r);
    RR alphr = RE(alpha), alphi = IM(alpha), xs = 0;
    for (int64_t p = 0; p < mi; ++p) xs += ABS2(QAT(A2, lda2, p, r));
    if (xs == 0 && alphi == 0) {
        /* the reference leaves T(:, r) unwritten; LAPACK's tpqrt2 zeroes it */
        tau[r] = 0;
        for (int64_t p = 0; p <= i; ++p) QAT(T, ldt, p, r) = 0;
        return;
    }
    RR beta = -(RR)copysign(sqrt((double)(alphr * alphr + alphi * alphi + xs)), (double)alphr);
    R t = MKC((beta - alphr) / beta, -alphi / beta);
    R scale = (R)1 / (alpha - (R)beta);
    for (int64_t p = 0; p < mi; ++p) QAT(A2, lda2, p, r) *= scale;
    QAT(A1, lda1, r, r) = beta;
    tau[r] = t;
    if (sb - i - 1 > 0 && ii + i + 1 < n) {
        for (int64_t j = 0; j < sb - i - 1; ++j) {
            R w = QAT(A1, lda1, r, r + 1 + j);    /* v^H a_j with v = [1; x] */
            for (int64_t p = 0; p < mi; ++p) w += CONJ(QAT(A2, lda2, p, r)) * QAT(A2, lda2, p, r + 1 + j);
            work[j] = w;
        }
        for (int64_t j = 0; j < sb - i - 1; ++j) QAT(A1, lda1, r, r + 1 + j) -= CONJ(t) * work[j];
        for (int64_t j = 0; j < sb - i - 1; ++j)
            for (int64_t p = 0; p < mi; ++p) QAT(A2, lda2, p, r + 1 + j) -= CONJ(t) * QAT(A2, lda2, p, r) * work[j];
    }
    if (i > 0) {
        int64_t mj = tt ? (r < m ? r : m) : m;
        for (int64_t p = 0; p < i; ++p) {
            R acc = 0;
            for (int64_t q = 0; q < mj; ++q) acc += CONJ(QAT(A2, lda2, q, ii + p)) * QAT(A2, lda2, q, r);
            QAT(T, ldt, p, r) = -t * acc;
        }
        for (int64_t p = 0; p < i; ++p) work[p] = QAT(T, ldt, p, r);
        for (int64_t p = 0; p < i; ++p) {
            R acc = 0;
            for (int64_t q = p; q < i; ++q) acc += QAT(T, ldt, p, ii + q) * work[q];
            QAT(T, ldt, p, r) = acc;
        }
    }
    QAT(T, ldt, i, r) = t;
}

/* tsqrt!: src/tsqrt.jl:128-156.  QR of [A1 (m1 x n, upper); A2 (m x n)], k = min(m1, n). */
void P(tsqrt)(int64_t m, int64_t n, int64_t ib, R *A1, int64_t lda1, int64_t m1, R *A2, int64_t lda2,
              R *T, int64_t ldt, R *tau, R *work) {
    int64_t k = m1 < n ? m1 : n;
    if (m <= 0 || n <= 0 || ib <= 0 || k <= 0) return;
    for (int64_t ii = 0; ii < k; ii += ib) {
        int64_t sb = (ib < k - ii) ? ib : k - ii;
        for (int64_t i = 0; i < sb; ++i) P(q_tcol)(0, m, n, ii, sb, i, A1, lda1, A2, lda2, T, ldt, tau, work);
        if (n > ii + sb)
            P(q_tsmqr)(0, 'L', 'C', sb, n - ii - sb, m, n - ii - sb, sb, ib, &QAT(A1, lda1, ii, ii + sb), lda1,
                       &QAT(A2, lda2, 0, ii + sb), lda2, &QAT(A2, lda2, 0, ii), lda2, &QAT(T, ldt, 0, ii), ldt, work);
    }
}

/* ttqrt!: src/ttqrt.jl:81-118.  QR of [A1 (upper, m x n); A2 (upper, m x n)], k = min(m, n);
 * the trailing update of block ii is parfb!('L','C','F','C') on the first
 * mi = min(ii + sb, m) rows of A2 with l = min(sb, mi - ii) (src/ttqrt.jl:104-116). */
void P(ttqrt)(int64_t m, int64_t n, int64_t ib, R *A1, int64_t lda1, R *A2, int64_t lda2, R *T, int64_t ldt,
              R *tau, R *work) {
    if (m <= 0 || n <= 0 || ib <= 0) return;
    int64_t k = m < n ? m : n;
    for (int64_t ii = 0; ii < k; ii += ib) {
        int64_t sb = (ib < k - ii) ? ib : k - ii;
        for (int64_t i = 0; i < sb; ++i) P(q_tcol)(1, m, n, ii, sb, i, A1, lda1, A2, lda2, T, ldt, tau, work);
        if (n > ii + sb) {
            int64_t mi = (ii + sb < m) ? ii + sb : m;
            int64_t ni = n - ii - sb;
            int64_t l = mi - ii > 0 ? (sb < mi - ii ? sb : mi - ii) : 0;
            P(parfb)('L', 'C', 'F', 'C', sb, ni, mi, ni, sb, l, &QAT(A1, lda1, ii, ii + sb), lda1,
                     &QAT(A2, lda2, 0, ii + sb), lda2, &QAT(A2, lda2, 0, ii), lda2, &QAT(T, ldt, 0, ii), ldt,
                     work, sb);
        }
    }
}

/* pemv!: src/pemv.jl:39-181.  storev 'C' with trans 'T'/'C': y (n) = alpha op(A) x + beta y
 * over A m x n whose bottom-left l x l block is upper triangular; storev 'R' with trans 'N':
 * y (m) = alpha A x + beta y over A whose top-right l x l block is lower triangular; l = 1
 * counts as l = 0.  alpha = beta = 0 leaves y unchanged; beta = 0 does not read y. */
void P(pemv)(char trans, char storev, int64_t m, int64_t n, int64_t l, const R *palpha, const R *A, int64_t lda,
             const R *x, const R *pbeta, R *y) {
    R alpha = *palpha, beta = *pbeta;
    if (m <= 0 || n <= 0) return;
    if (RE(alpha) == 0 && IM(alpha) == 0 && RE(beta) == 0 && IM(beta) == 0) return;
    if (l == 1) l = 0;
    int bz = RE(beta) == 0 && IM(beta) == 0;
    if (storev == 'C') {
        for (int64_t j = 0; j < n; ++j) {
            R acc = 0;
            for (int64_t r = 0; r < m; ++r) {
                if (j < l && r >= m - l && r - (m - l) > j) continue;
                R a = QAT(A, lda, r, j);
                acc += (trans == 'C' ? CONJ(a) : a) * x[r];
            }
            y[j] = alpha * acc + (bz ? (R)0 : beta * y[j]);
        }
    } else {
        for (int64_t r = 0; r < m; ++r) {
            R acc = 0;
            for (int64_t c = 0; c < n; ++c) {
                if (r < l && c >= n - l && c - (n - l) > r) continue;
                acc += QAT(A, lda, r, c) * x[c];
            }
            y[r] = alpha * acc + (bz ? (R)0 : beta * y[r]);
        }
    }
}

/* ------------------------------------------------------------------------------------------
 * Recursive LU with partial pivoting (src/lu.jl:185-300, LAPACK xgetrf2) and the row
 * interchanges it applies (laswp, src/lu.jl:470-522, LAPACK xlaswp).  ipiv is 1-based as in
 * the reference; info = the first i with U(i,i) exactly zero (1-based), 0 otherwise. */

/* laswp(A, k1, k2, ipiv, incx): rows k1..k2 (1-based) interchanged with ipiv(k) in order
 * (incx > 0) or in reverse (incx < 0); incx = 0 does nothing. */
void P(laswp)(int64_t n, R *A, int64_t lda, int64_t k1, int64_t k2, const int64_t *ipiv, int64_t incx) {
    if (incx == 0) return;
    for (int64_t c = 0; c < n; ++c) {
        if (incx > 0) {
            for (int64_t i = k1; i <= k2; ++i) {
                int64_t ip = ipiv[i - 1];
                if (ip != i) { R t = QAT(A, lda, i - 1, c); QAT(A, lda, i - 1, c) = QAT(A, lda, ip - 1, c); QAT(A, lda, ip - 1, c) = t; }
            }
        } else {
            for (int64_t i = k2; i >= k1; --i) {
                int64_t ip = ipiv[i - 1];
                if (ip != i) { R t = QAT(A, lda, i - 1, c); QAT(A, lda, i - 1, c) = QAT(A, lda, ip - 1, c); QAT(A, lda, ip - 1, c) = t; }
            }
        }
    }
}

static RR P(q_cabs1)(R x) { RR a = RE(x), b = IM(x); return (a < 0 ? -a : a) + (b < 0 ? -b : b); }

/* getrf2!: src/lu.jl:185-300.  Returns info. */
int64_t P(getrf2)(int64_t m, int64_t n, R *A, int64_t lda, int64_t *ipiv) {
    if (m <= 0 || n <= 0) return 0;
    if (m == 1) {
        ipiv[0] = 1;
        return (RE(QAT(A, lda, 0, 0)) == 0 && IM(QAT(A, lda, 0, 0)) == 0) ? 1 : 0;
    }
    if (n == 1) {
        RR sfmin = (sizeof(RR) == 8) ? (RR)DBL_MIN : (RR)FLT_MIN;
        int64_t p = 0;
        RR dmax = P(q_cabs1)(QAT(A, lda, 0, 0));
        for (int64_t i = 1; i < m; ++i)
            if (P(q_cabs1)(QAT(A, lda, i, 0)) > dmax) { p = i; dmax = P(q_cabs1)(QAT(A, lda, i, 0)); }
        ipiv[0] = p + 1;
        if (RE(QAT(A, lda, p, 0)) == 0 && IM(QAT(A, lda, p, 0)) == 0) return 1;
        if (p != 0) { R t = QAT(A, lda, 0, 0); QAT(A, lda, 0, 0) = QAT(A, lda, p, 0); QAT(A, lda, p, 0) = t; }
        R d = QAT(A, lda, 0, 0);
        if (P(q_abs)(d) >= sfmin) {
            R r = (R)1 / d;
            for (int64_t i = 1; i < m; ++i) QAT(A, lda, i, 0) *= r;
        } else {
            for (int64_t i = 1; i < m; ++i) QAT(A, lda, i, 0) /= d;
        }
        return 0;
    }
    int64_t mn = m < n ? m : n;
    int64_t n1 = mn / 2, n2 = n - n1;
    int64_t info = P(getrf2)(m, n1, A, lda, ipiv);
    /* interchanges on [A12; A22] */
    P(laswp)(n2, &QAT(A, lda, 0, n1), lda, 1, n1, ipiv, 1);
    /* A12 = L11^-1 A12 (L11 unit lower) */
    for (int64_t c = 0; c < n2; ++c)
        for (int64_t i = 0; i < n1; ++i) {
            R acc = QAT(A, lda, i, n1 + c);
            for (int64_t q = 0; q < i; ++q) acc -= QAT(A, lda, i, q) * QAT(A, lda, q, n1 + c);
            QAT(A, lda, i, n1 + c) = acc;
        }
    /* A22 -= A21 A12 */
    for (int64_t c = 0; c < n2; ++c)
        for (int64_t i = n1; i < m; ++i) {
            R acc = 0;
            for (int64_t q = 0; q < n1; ++q) acc += QAT(A, lda, i, q) * QAT(A, lda, q, n1 + c);
            QAT(A, lda, i, n1 + c) -= acc;
        }
    int64_t iinfo = P(getrf2)(m - n1, n2, &QAT(A, lda, n1, n1), lda, ipiv + n1);
    if (info == 0 && iinfo > 0) info = iinfo + n1;
    for (int64_t i = n1; i < mn; ++i) ipiv[i] += n1;
    P(laswp)(n1, A, lda, n1 + 1, mn, ipiv, 1);
    return info;
}

#undef QAT
#undef ABS2
