"""Pins of the scalar-generic QR-family oracle (oracle/nla_oracle_qr_t.h) for real and complex
element types, before it is trusted as the GPU checker:

* LAPACK through scipy (the routines the reference's tests call, test/lapack_helpers.jl):
  xlarfg, xgeqrt (V, R and T), xgemqrt, xtpqrt with l = 0 (tsqrt!) and l = n (ttqrt!),
  xtpmqrt with l = 0 (tsmqr!) and l > 0 (ttmqr!, and a one-block parfb! with l > 0);
* closed forms where LAPACK has no routine: every parfb! variant against the explicitly
  formed H = I - Vc T Vc^H of the pentagonal V (test/parfb.jl), pamm! against its two
  products, pemv! against the pentagonal matrix it reads (test/pemv.jl's make_pentagonal),
  larft! rowwise against the product of its reflectors (test/larft.jl:43-61);
* the real instances against the real-only oracle already pinned in test_oracle.py.
"""
import numpy as np
import pytest
from scipy.linalg import lapack

import oracle as O
import oracle_qr as Q

DTYPES = [np.float64, np.float32, np.complex128, np.complex64]
PRE = {np.float64: "d", np.float32: "s", np.complex128: "z", np.complex64: "c"}
TOL = {np.float64: 1e-12, np.float32: 1e-5, np.complex128: 1e-12, np.complex64: 1e-5}


def F(a):
    return np.asfortranarray(a)


def rnd(rng, shape, dt):
    a = rng.standard_normal(shape)
    if np.issubdtype(dt, np.complexfloating):
        a = a + 1j * rng.standard_normal(shape)
    return F(a.astype(dt))


def rel(x, y):
    return np.linalg.norm(np.asarray(x, dtype=np.complex128) - y) / max(float(np.linalg.norm(y)), 1e-300)


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("n,scale", [(2, 1.0), (7, 1.0), (33, 1.0), (6, 0.0), (12, "tiny")])
def test_larfg_vs_lapack(dt, n, scale):
    rng = np.random.default_rng(n)
    x = rnd(rng, n - 1, dt)
    alpha = rnd(rng, 1, dt)[0]
    if scale == 0.0:
        x[:] = 0
    if scale == "tiny":
        s = 1e-300 if dt in (np.float64, np.complex128) else 1e-40
        x = (x * s).astype(dt)
        alpha = dt(alpha * s)
    beta, v, tau = getattr(lapack, PRE[dt] + "larfg")(n, alpha, x.copy())
    xo = x.copy()
    b2, t2 = Q.larfg(alpha, xo)
    assert abs(b2 - beta) <= TOL[dt] * max(abs(beta), 1e-300)
    assert abs(t2 - tau) <= TOL[dt] * max(abs(tau), 1)
    assert rel(xo, np.asarray(v)[: n - 1]) < TOL[dt] or np.linalg.norm(v) == 0


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("m,n,ib", [(40, 24, 8), (64, 32, 16), (33, 20, 5), (20, 20, 10)])
def test_geqrt_vs_lapack(dt, m, n, ib):
    rng = np.random.default_rng(m * n)
    A = rnd(rng, (m, n), dt)
    qr, t, info = getattr(lapack, PRE[dt] + "geqrt")(ib, A.copy())
    Ao = A.copy(order="F")
    T, tau = Q.geqrt(Ao, ib)
    if m > n or not np.issubdtype(dt, np.complexfloating):
        assert rel(Ao, qr) < TOL[dt] and rel(T, t) < TOL[dt]
    else:
        # square complex: the reference gives a column of length 1 tau = 0 (src/geqr2.jl:32-35)
        # where LAPACK makes its diagonal real; everything else agrees
        assert rel(Ao[:, :-1], qr[:, :-1]) < TOL[dt] and rel(Ao[:-1, -1], qr[:-1, -1]) < TOL[dt]
        assert tau[-1] == 0 and abs(abs(Ao[-1, -1]) - abs(qr[-1, -1])) < TOL[dt] * abs(qr[-1, -1])
    # geqr2 (ib = n) is the same factorization as geqrt's blocks
    A2 = A.copy(order="F")
    tau2 = Q.geqr2(A2)
    assert rel(A2, Ao) < 50 * TOL[dt] and rel(tau2, tau) < 50 * TOL[dt]


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("side", "LR")
@pytest.mark.parametrize("trans", "NC")
def test_unmqr_vs_gemqrt(dt, side, trans):
    rng = np.random.default_rng(3)
    nq, k, ib = 30, 18, 6
    A = rnd(rng, (nq, k), dt)
    T, _ = Q.geqrt(A, ib)
    C = rnd(rng, (nq, 11) if side == "L" else (11, nq), dt)
    tr = trans if np.issubdtype(dt, np.complexfloating) else ("T" if trans == "C" else "N")
    want, info = getattr(lapack, PRE[dt] + "gemqrt")(A, T, C.copy(), side=side, trans=tr)
    Co = C.copy(order="F")
    Q.unmqr(side, trans, A, T, Co)
    assert rel(Co, want) < TOL[dt]


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("m,n,ib", [(32, 24, 8), (48, 32, 16), (20, 20, 10)])
def test_tsqrt_vs_tpqrt(dt, m, n, ib):
    rng = np.random.default_rng(m + n)
    A1 = F(np.triu(rnd(rng, (n, n), dt)))
    A2 = rnd(rng, (m, n), dt)
    a, b, t, info = getattr(lapack, PRE[dt] + "tpqrt")(0, ib, A1.copy(), A2.copy())
    A1o, A2o = A1.copy(order="F"), A2.copy(order="F")
    T = np.zeros((ib, n), dtype=dt, order="F")
    Q.tsqrt(A1o, A2o, T)
    assert rel(np.triu(A1o), np.triu(a)) < TOL[dt] and rel(A2o, b) < TOL[dt] and rel(T, t) < TOL[dt]


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("n,ib", [(16, 4), (24, 8), (32, 8), (48, 16)])
def test_ttqrt_vs_tpqrt(dt, n, ib):
    # test/ttqrt.jl:6-39: triangular A1, A2 with 2 added to the diagonal, against xtpqrt(l = n)
    rng = np.random.default_rng(n * ib)
    A1 = F(np.triu(rnd(rng, (n, n), dt)) + 2 * np.eye(n, dtype=dt))
    A2 = F(np.triu(rnd(rng, (n, n), dt)) + 2 * np.eye(n, dtype=dt))
    a, b, t, info = getattr(lapack, PRE[dt] + "tpqrt")(n, ib, A1.copy(), A2.copy())
    A1o, A2o = A1.copy(order="F"), A2.copy(order="F")
    T = np.zeros((ib, n), dtype=dt, order="F")
    Q.ttqrt(A1o, A2o, T)
    assert rel(np.triu(A1o), np.triu(a)) < TOL[dt]
    assert rel(np.triu(A2o), np.triu(b)) < TOL[dt]
    assert rel(T, t) < TOL[dt]


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("side", "LR")
@pytest.mark.parametrize("trans", "NC")
@pytest.mark.parametrize("tt", [False, True])
def test_tsmqr_ttmqr_vs_tpmqrt(dt, side, trans, tt):
    # test/tsmqr.jl:6-40, test/ttmqr.jl:6-39 (V upper triangular for ttmqr!)
    rng = np.random.default_rng(11 + tt)
    n, k, ib = 24, 16, 8
    A1 = rnd(rng, (n, n) if not tt else (n, n), dt)
    A2 = rnd(rng, (n, n), dt)
    V = rnd(rng, (n, k), dt)
    if tt:
        V = F(np.triu(V))
    Tm = F(np.triu(rnd(rng, (ib, k), dt)))
    tr = trans if np.issubdtype(dt, np.complexfloating) else ("T" if trans == "C" else "N")
    if side == "L":
        a1 = A1[:k, :].copy(order="F") if not tt else A1[:k, :].copy(order="F")
    else:
        a1 = A1[:, :k].copy(order="F")
    l = k if tt else 0
    if tt:   # l > 0 needs the trapezoid within V's rows: use the k x k top block as V
        V = F(V[:k, :])
        A2 = F(A2[:k, :]) if side == "L" else F(A2[:, :k])
    a, b, info = getattr(lapack, PRE[dt] + "tpmqrt")(l, V, Tm, a1.copy(), A2.copy(), side=side, trans=tr)
    A1o, A2o = a1.copy(order="F"), A2.copy(order="F")
    (Q.ttmqr if tt else Q.tsmqr)(side, trans, A1o, A2o, V, Tm)
    assert rel(A1o, a) < TOL[dt] and rel(A2o, b) < TOL[dt]


def _vc(V, storev, direct, rows, k, l):
    """Column form of the pentagonal V with its structural zeros (the reference's block views)."""
    Vc = V.copy() if storev == "C" else V.conj().T.copy()
    for r in range(rows):
        for j in range(k):
            if direct == "F" and r >= rows - l and r - (rows - l) > j:
                Vc[r, j] = 0
            if direct == "B" and r < l and j > k - l + r:
                Vc[r, j] = 0
    return Vc


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("side", "LR")
@pytest.mark.parametrize("trans", "NC")
@pytest.mark.parametrize("direct", "FB")
@pytest.mark.parametrize("storev", "CR")
def test_parfb_all_variants_closed_form(dt, side, trans, direct, storev):
    # test/parfb.jl:6-52: m2, n2, k, l = 20, 24, 8, 4 with random (unstructured) V and T
    rng = np.random.default_rng(7)
    m2, n2, k, l = 20, 24, 8, 4
    m1, n1 = (k, n2) if side == "L" else (m2, k)
    A1, A2 = rnd(rng, (m1, n1), dt), rnd(rng, (m2, n2), dt)
    rows = m2 if side == "L" else n2
    V = rnd(rng, (rows, k) if storev == "C" else (k, rows), dt)
    T = rnd(rng, (k, k), dt)
    Vc = _vc(V, storev, direct, rows, k, l)
    Tt = np.triu(T) if direct == "F" else np.tril(T)
    W = np.vstack([np.eye(k), Vc])                       # [I; Vc]
    H = np.eye(k + rows) - W @ Tt @ W.conj().T
    opH = H if trans == "N" else H.conj().T
    if side == "L":
        C = np.vstack([A1, A2])
        want = opH @ C
        w1, w2 = want[:k], want[k:]
    else:
        C = np.hstack([A1, A2])
        want = C @ opH
        w1, w2 = want[:, :k], want[:, k:]
    A1o, A2o = A1.copy(order="F"), A2.copy(order="F")
    Q.parfb(side, trans, direct, storev, m1, n1, m2, n2, k, l, A1o, A2o, V, T)
    assert rel(A1o, w1) < 10 * TOL[dt] and rel(A2o, w2) < 10 * TOL[dt]


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("side", "LR")
@pytest.mark.parametrize("trans", "NC")
def test_parfb_pentagonal_vs_tpmqrt(dt, side, trans):
    # one block of xtpmqrt (nb = k) is xtprfb with the same l
    rng = np.random.default_rng(9)
    k, l, other = 8, 5, 13
    rows = 20
    V = F(rnd(rng, (rows, k), dt))
    for r in range(rows - l, rows):                      # the pentagon's zeros, as xtpmqrt reads them
        V[r, : r - (rows - l)] = 0
    T = F(np.triu(rnd(rng, (k, k), dt)))
    A1 = rnd(rng, (k, other) if side == "L" else (other, k), dt)
    A2 = rnd(rng, (rows, other) if side == "L" else (other, rows), dt)
    tr = trans if np.issubdtype(dt, np.complexfloating) else ("T" if trans == "C" else "N")
    a, b, info = getattr(lapack, PRE[dt] + "tpmqrt")(l, V, T, A1.copy(), A2.copy(), side=side, trans=tr)
    A1o, A2o = A1.copy(order="F"), A2.copy(order="F")
    m1, n1 = A1.shape
    m2, n2 = A2.shape
    Q.parfb(side, trans, "F", "C", m1, n1, m2, n2, k, l, A1o, A2o, V, T)
    assert rel(A1o, a) < TOL[dt] and rel(A2o, b) < TOL[dt]


@pytest.mark.parametrize("dt", [np.float64, np.complex128])
@pytest.mark.parametrize("side", "LR")
@pytest.mark.parametrize("direct", "FB")
@pytest.mark.parametrize("storev", "CR")
def test_pamm_closed_form(dt, side, direct, storev):
    rng = np.random.default_rng(5)
    m, n, k, l = 40, 30, 20, 10
    rows = m if side == "L" else n
    A1 = rnd(rng, (k, n) if side == "L" else (m, k), dt)
    A2 = rnd(rng, (m, n), dt)
    V = rnd(rng, (rows, k) if storev == "C" else (k, rows), dt)
    Vc = _vc(V, storev, direct, rows, k, l)
    W = np.zeros_like(A1)
    Q.pamm("W", side, storev, direct, m, n, k, l, A1, A2.copy(order="F"), V, W)
    want = A1 + Vc.conj().T @ A2 if side == "L" else A1 + A2 @ Vc
    assert rel(W, want) < 1e-13
    A2o = A2.copy(order="F")
    Q.pamm("A", side, storev, direct, m, n, k, l, A1, A2o, V, W)
    want2 = A2 - (Vc @ W if side == "L" else W @ Vc.conj().T)
    assert rel(A2o, want2) < 1e-13


def _pentagonal(A, m, n, l, storev):
    Ap = A.copy()
    if l <= 1:
        return Ap
    if storev == "C":
        for j in range(l):
            Ap[m - l + j + 1:, j] = 0
    else:
        for j in range(l):
            Ap[:j, n - l + j] = 0
    return Ap


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("trans,storev,m,n,l", [("C", "C", 40, 30, 10), ("T", "C", 40, 30, 10),
                                                ("N", "R", 30, 40, 10), ("C", "C", 20, 15, 0),
                                                ("C", "C", 50, 40, 25), ("C", "C", 12, 8, 1)])
def test_pemv_closed_form(dt, trans, storev, m, n, l):
    if trans == "T" and np.issubdtype(dt, np.complexfloating):
        pytest.skip("test/pemv.jl uses trans 'T' with real types")
    rng = np.random.default_rng(m + n + l)
    A = rnd(rng, (m, n), dt)
    x = rnd(rng, m if storev == "C" else n, dt)
    y0 = rnd(rng, n if storev == "C" else m, dt)
    alpha, beta = dt(2.5), dt(1.25)
    if np.issubdtype(dt, np.complexfloating):
        alpha, beta = dt(2.5 + 1.5j), dt(1.2 - 0.8j)
    y = y0.copy()
    Q.pemv(trans, storev, l, alpha, A, x, beta, y)
    Ap = _pentagonal(A, m, n, l, storev)
    opA = Ap.conj().T if trans == "C" else (Ap.T if trans == "T" else Ap)
    assert rel(y, alpha * opA @ x + beta * y0) < TOL[dt]
    # beta = 0 does not read y; alpha = beta = 0 leaves y unchanged
    y = np.full_like(y0, np.nan)
    Q.pemv(trans, storev, l, alpha, A, x, dt(0), y)
    assert rel(y, alpha * opA @ x) < TOL[dt]
    y = y0.copy()
    Q.pemv(trans, storev, l, dt(0), A, x, dt(0), y)
    assert np.array_equal(y, y0)


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("storev", "CR")
@pytest.mark.parametrize("n,k", [(1, 1), (5, 3), (10, 4), (20, 8)])
def test_larft_product_of_reflectors(dt, storev, n, k):
    # test/larft.jl:6-61: I - V T V^H (columnwise) / I - V^H T V (rowwise) = H_1 ... H_k
    rng = np.random.default_rng(n * 10 + k)
    V = rnd(rng, (n, k) if storev == "C" else (k, n), dt)
    for j in range(k):
        if storev == "C":
            V[:j, j] = 0
            V[j, j] = 1
        else:
            V[j, :j] = 0
            V[j, j] = 1
    tau = rnd(rng, k, dt)
    T = np.zeros((k, k), dtype=dt, order="F")
    Q.larft(storev, V, tau, T)
    assert np.allclose(T, np.triu(T)) and np.allclose(np.diag(T), tau)
    Hc = np.eye(n) - (V @ T @ V.conj().T if storev == "C" else V.conj().T @ T @ V)
    Ht = np.eye(n, dtype=np.complex128)
    for j in range(k):
        v = V[:, j] if storev == "C" else V[j, :].conj()
        Ht = Ht @ (np.eye(n) - tau[j] * np.outer(v, v.conj()))
    assert np.linalg.norm(Hc - Ht) / max(1, np.linalg.norm(Ht)) < TOL[dt] * n


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_real_instances_match_real_oracle(dt):
    """The generic oracle's real instances agree with the real-only restatement."""
    rng = np.random.default_rng(12)
    A = rnd(rng, (48, 32), dt)
    A1, A2 = A.copy(order="F"), A.copy(order="F")
    T1, _ = Q.geqrt(A1, 8)
    T2 = np.zeros((8, 32), dtype=dt, order="F")
    O.geqrt(A2, T2)
    assert rel(A1, A2) < TOL[dt] and rel(T1, T2) < TOL[dt]
    B1 = F(np.triu(rnd(rng, (16, 16), dt)))
    C1 = rnd(rng, (24, 16), dt)
    Ta = np.zeros((4, 16), dtype=dt, order="F")
    Tb = np.zeros((4, 16), dtype=dt, order="F")
    b1a, c1a, b1b, c1b = B1.copy(order="F"), C1.copy(order="F"), B1.copy(order="F"), C1.copy(order="F")
    Q.tsqrt(b1a, c1a, Ta)
    O.tsqrt(b1b, c1b, Tb)
    assert rel(b1a, b1b) < TOL[dt] and rel(c1a, c1b) < TOL[dt] and rel(Ta, Tb) < TOL[dt]


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("m,n", [(1, 1), (1, 5), (7, 1), (10, 10), (64, 40), (40, 64), (129, 129)])
def test_getrf2_vs_lapack(dt, m, n):
    """getrf2! (src/lu.jl:185-300) against LAPACK xgetrf: same pivots, same factors."""
    rng = np.random.default_rng(m * 100 + n)
    A = rnd(rng, (m, n), dt)
    lu, piv, info = getattr(lapack, PRE[dt] + "getrf")(A.copy())
    Ao = A.copy(order="F")
    ipiv, info2 = Q.getrf2(Ao)
    assert np.array_equal(ipiv, piv + 1) and info2 == info == 0
    assert rel(Ao, lu) < 100 * TOL[dt]


def test_getrf2_singular_info_and_laswp():
    A = np.asfortranarray(np.array([[1.0, 2.0, 3.0], [2.0, 4.0, 6.0], [1.0, 1.0, 1.0]]))
    lu, piv, info = lapack.dgetrf(A.copy())
    Ao = A.copy(order="F")
    ipiv, info2 = Q.getrf2(Ao)
    assert info == info2 == 3 and np.array_equal(ipiv, piv + 1)
    # laswp forward then backward is the identity
    rng = np.random.default_rng(2)
    B = F(rng.standard_normal((6, 4)))
    ip = np.array([3, 5, 3, 6, 5, 6], dtype=np.int64)
    Bo = B.copy(order="F")
    Q.laswp(Bo, 1, 6, ip, 1)
    want = B.copy()
    for i in range(6):
        want[[i, ip[i] - 1]] = want[[ip[i] - 1, i]]
    assert np.array_equal(Bo, want)
    Q.laswp(Bo, 1, 6, ip, -1)
    assert np.array_equal(Bo, B)
