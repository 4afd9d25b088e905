"""Pins the CPU oracle (oracle/nla_oracle_t.h) against LAPACK golden vectors,
closed forms and brute force.  These are the checks the reference's own tests
make (test/larfg.jl, test/geqr2.jl, test/geqrt.jl, test/larft.jl, test/larfb.jl,
test/unmqr.jl, test/tsqrt.jl, test/tsmqr.jl, test/parfb.jl, test/unified_rectrxm.jl),
applied to the restatement instead of the Julia package."""
import numpy as np
import pytest

import oracle as O
from conftest import isapprox, relerr, rtol_for
from golden.gen_golden import GEQRT_SHAPES, TPQRT_SHAPES, TRXM_CASES, TSMQR_CASES

DTYPES = [np.float64, np.float32]
P = {np.float64: "d", np.float32: "s"}


def F(a):
    return np.asfortranarray(a)


def test_lamch():
    # src/NextLA.jl:30-53: eps is half the spacing at 1, sfmin = floatmin (1/floatmax is smaller)
    assert O.lamch(np.float64, "E") == np.finfo(np.float64).eps / 2
    assert O.lamch(np.float64, "P") == np.finfo(np.float64).eps
    assert O.lamch(np.float64, "S") == np.finfo(np.float64).tiny
    assert O.lamch(np.float32, "S") == np.finfo(np.float32).tiny


@pytest.mark.parametrize("dt", DTYPES)
def test_larfg_vs_lapack(golden, dt):
    g = golden("larfg")
    p = P[dt]
    keys = sorted({k.rsplit("_", 1)[0] for k in g if k.startswith(p + "_")})
    assert len(keys) >= 8
    for key in keys:
        x = g[key + "_x"].copy()
        alpha = g[key + "_alpha"][0]
        beta, tau = O.larfg(alpha, x)
        tol = 10 * np.finfo(dt).eps
        # test/larfg.jl:56-79 compares in absolute value
        assert abs(abs(beta) - abs(g[key + "_beta"][0])) <= tol * max(abs(beta), np.finfo(dt).tiny)
        assert abs(tau - g[key + "_tau"][0]) <= tol * max(1.0, abs(tau))
        if x.size:
            assert isapprox(np.abs(x), np.abs(g[key + "_v"]), 100 * np.finfo(dt).eps)
        # closed form: H [alpha; x] = [beta; 0] with H = I - tau v v^T, v = [1; x_out]
        v = np.concatenate([[1.0], x.astype(np.float64)])
        y = np.concatenate([[alpha], g[key + "_x"].astype(np.float64)])
        hy = y - tau * v * (v @ y)
        scale = np.abs(y).max() * np.sqrt(y.size) + np.finfo(dt).tiny   # overflow/underflow-safe
        assert abs(hy[0] - beta) <= 50 * np.finfo(dt).eps * scale
        assert np.abs(hy[1:]).max(initial=0.0) <= 50 * np.finfo(dt).eps * scale


@pytest.mark.parametrize("dt", DTYPES)
def test_larfg_special_cases(dt):
    # test/larfg.jl:42-54: n <= 1 or x = 0 gives tau = 0
    beta, tau = O.larfg(dt(3.0), np.zeros(0, dtype=dt))
    assert tau == 0 and beta == 3.0
    beta, tau = O.larfg(dt(-2.0), np.zeros(5, dtype=dt))
    assert tau == 0 and beta == -2.0


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("shape", GEQRT_SHAPES)
def test_geqrt_vs_dgeqrt(golden, dt, shape):
    g = golden("geqrt")
    idx = GEQRT_SHAPES.index(shape)
    key = f"{P[dt]}_{idx}"
    m, n, ib = shape
    A = g[key + "_A"].copy(order="F")
    T = np.zeros((ib, min(m, n)), dtype=dt, order="F")
    O.geqrt(A, T)
    rt = rtol_for(dt) * (10 if dt == np.float64 else 10)
    assert isapprox(A, g[key + "_QR"], rt)
    assert isapprox(np.triu(T[:, :]), np.triu(g[key + "_T"]) if ib <= min(m, n) else g[key + "_T"], rt) or \
        isapprox(T, g[key + "_T"], rt)
    # test/geqrt.jl:24-31 scaled residual and orthogonality
    Q = np.eye(m, dtype=dt, order="F")
    O.unmqr("L", "N", A, T, Q)
    R = np.triu(A)
    A0 = g[key + "_A"].astype(np.float64)
    eps = np.finfo(dt).eps
    res = np.linalg.norm(A0 - Q.astype(np.float64) @ R.astype(np.float64), 1) / (np.linalg.norm(A0, 1) * n * eps)
    orth = np.linalg.norm(Q.T.astype(np.float64) @ Q - np.eye(m), 1) / (m * eps)
    assert res < 10 and orth < 10


@pytest.mark.parametrize("dt", DTYPES)
def test_geqr2_equals_geqrt_ib1(dt):
    # test/geqrt.jl:54-71
    rng = np.random.default_rng(5)
    A = F(rng.standard_normal((40, 24)).astype(dt))
    A1 = A.copy(order="F")
    tau = np.zeros(24, dtype=dt)
    O.geqr2(A1, tau)
    A2 = A.copy(order="F")
    T = np.zeros((1, 24), dtype=dt, order="F")
    tau2 = O.geqrt(A2, T)
    assert isapprox(A1, A2, rtol_for(dt))
    assert isapprox(tau, tau2, rtol_for(dt))
    assert isapprox(T[0], tau, rtol_for(dt))


@pytest.mark.parametrize("dt", DTYPES)
def test_larft_invariant(dt):
    # test/larft.jl:35-60: T upper, diag = tau, I - V T V^T = H_1 ... H_k
    rng = np.random.default_rng(7)
    m, k = 30, 9
    A = F(rng.standard_normal((m, k)).astype(dt))
    tau = np.zeros(k, dtype=dt)
    O.geqr2(A, tau)
    T = np.zeros((k, k), dtype=dt, order="F")
    O.larft(A, tau, T)
    V = np.tril(A.astype(np.float64), -1) + np.eye(m, k)
    H = np.eye(m)
    for j in range(k):
        v = V[:, j]
        H = H @ (np.eye(m) - tau[j] * np.outer(v, v))
    assert np.allclose(np.tril(T, -1), 0)
    assert np.allclose(np.diag(T), tau)
    assert isapprox(np.eye(m) - V @ T.astype(np.float64) @ V.T, H, 100 * np.finfo(dt).eps)


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("side", "LR")
@pytest.mark.parametrize("trans", "NC")
def test_larfb_explicit(dt, side, trans):
    # test/larfb.jl:35-42: equals the explicitly formed H
    rng = np.random.default_rng(11)
    m, n, k = 21, 17, 6
    q = m if side == "L" else n
    V = F(rng.standard_normal((q, k)).astype(dt))
    T = F(np.triu(rng.standard_normal((k, k))).astype(dt))
    C = F(rng.standard_normal((m, n)).astype(dt))
    Vu = np.tril(V.astype(np.float64), -1) + np.eye(q, k)
    H = np.eye(q) - Vu @ np.triu(T.astype(np.float64)) @ Vu.T
    Hop = H if trans == "N" else H.T
    want = Hop @ C if side == "L" else C @ Hop
    O.larfb(side, trans, V, T, C)
    assert isapprox(C, want, 100 * np.finfo(dt).eps)


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("shape", GEQRT_SHAPES)
@pytest.mark.parametrize("side", "LR")
@pytest.mark.parametrize("trans", "NT")
def test_unmqr_vs_dgemqrt(golden, dt, shape, side, trans):
    g = golden("geqrt")
    key = f"{P[dt]}_{GEQRT_SHAPES.index(shape)}"
    QR = g[key + "_QR"]
    T = g[key + "_T"]
    m, n, ib = shape
    k = min(m, n)
    C = g[f"{key}_C{side}{trans}"].copy(order="F")
    O.unmqr(side, "C" if trans == "T" else "N", F(QR[:, :k]), F(T), C)
    assert isapprox(C, g[f"{key}_Q{side}{trans}"], 10 * rtol_for(dt))


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("shape", TPQRT_SHAPES)
def test_tsqrt_vs_dtpqrt(golden, dt, shape):
    g = golden("tsqrt")
    key = f"{P[dt]}_{TPQRT_SHAPES.index(shape)}"
    m, n, ib = shape
    A1 = g[key + "_A1"].copy(order="F")
    A2 = g[key + "_A2"].copy(order="F")
    T = np.zeros((ib, n), dtype=dt, order="F")
    O.tsqrt(A1, A2, T)
    rt = 10 * rtol_for(dt)
    assert isapprox(A1, g[key + "_R"], rt)
    assert isapprox(A2, g[key + "_V"], rt)
    assert isapprox(T, g[key + "_T"], rt)


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("case", TSMQR_CASES)
@pytest.mark.parametrize("trans", "NT")
def test_tsmqr_vs_dtpmqrt(golden, dt, case, trans):
    g = golden("tsmqr")
    key = f"{P[dt]}_{TSMQR_CASES.index(case)}"
    side = case[0]
    A1 = g[key + "_A1"].copy(order="F")
    A2 = g[key + "_A2"].copy(order="F")
    O.tsmqr(side, "C" if trans == "T" else "N", A1, A2, g[key + "_V"], g[key + "_T"])
    # random (V,T) is far from orthogonal: compare at the reference's rtol on O(1)-scaled data
    rt = rtol_for(dt) * (1e3 if dt == np.float64 else 1e2)
    assert isapprox(A1, g[f"{key}_{trans}_A1"], rt)
    assert isapprox(A2, g[f"{key}_{trans}_A2"], rt)


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("side", "LR")
@pytest.mark.parametrize("trans", "NC")
def test_parfb_explicit(dt, side, trans):
    # test/parfb.jl: W = op(T)(A1 + V^T A2); A1 -= W; A2 -= V W  (l = 0, direct F, storev C)
    rng = np.random.default_rng(3)
    k = 5
    if side == "L":
        m1, n1, m2, n2 = 9, 7, 12, 7
        V = F(rng.standard_normal((m2, k)).astype(dt))
    else:
        m1, n1, m2, n2 = 8, 9, 8, 11
        V = F(rng.standard_normal((n2, k)).astype(dt))
    A1 = F(rng.standard_normal((m1, n1)).astype(dt))
    A2 = F(rng.standard_normal((m2, n2)).astype(dt))
    T = F(np.triu(rng.standard_normal((k, k))).astype(dt))
    a1, a2, v, t = (x.astype(np.float64) for x in (A1, A2, V, T))
    opT = t if trans == "N" else t.T
    if side == "L":
        W = opT @ (a1[:k] + v.T @ a2)
        a1[:k] -= W
        a2 -= v @ W
    else:
        W = (a1[:, :k] + a2 @ v) @ opT
        a1[:, :k] -= W
        a2 -= W @ v.T
    O.parfb(side, trans, A1, A2, V, T)
    assert isapprox(A1, a1, 100 * np.finfo(dt).eps)
    assert isapprox(A2, a2, 100 * np.finfo(dt).eps)


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("case", TRXM_CASES)
def test_rectrxm_vs_blas(golden, dt, case):
    g = golden("trxm")
    key = f"{P[dt]}_{TRXM_CASES.index(case)}"
    func, side, uplo, trans, n, m, alpha, cut = case
    B = g[key + "_B"].copy(order="F")
    O.rectrxm(side, uplo, trans, alpha, func, g[key + "_A"], B, threshold=cut)
    # test/unified_rectrxm.jl:42-43 uses 1e-14 relative error for F64 at n <= 256
    tol = 1e-12 if dt == np.float64 else 1e-4
    assert relerr(B, g[key + "_X"]) < tol


@pytest.mark.parametrize("dt", DTYPES)
def test_rectrxm_cutoff_independent(dt):
    # the recursion cut-off changes the order of operations, not the result
    rng = np.random.default_rng(9)
    n, m = 384, 16
    A = np.tril(rng.standard_normal((n, n)) / np.sqrt(n), -1)
    A[np.diag_indices(n)] = rng.uniform(1, 2, n)
    A = F(A.astype(dt))
    B0 = F(rng.standard_normal((n, m)).astype(dt))
    outs = []
    for cut in (32, 64, 128, 512):
        B = B0.copy(order="F")
        O.rectrxm("L", "L", "N", 1.0, "S", A, B, threshold=cut)
        outs.append(B)
    for B in outs[1:]:
        assert relerr(B, outs[0]) < (1e-12 if dt == np.float64 else 1e-4)
    # backward error ||A X - B|| / (||A|| ||X||)
    X = outs[0].astype(np.float64)
    be = np.linalg.norm(np.tril(A.astype(np.float64)) @ X - B0) / (np.linalg.norm(np.tril(A)) * np.linalg.norm(X))
    assert be < 10 * np.finfo(dt).eps


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("idx,nb,ib", [(0, 64, 16), (0, 128, 32), (1, 64, 32), (1, 128, 32)])
def test_tile_geqrf_vs_dgeqrf(golden, dt, idx, nb, ib):
    g = golden("tileqr")
    A0 = g[f"{P[dt]}_{idx}_A"]
    m, n = A0.shape
    A = A0.copy(order="F")
    T = O.tile_geqrf(A, nb, ib)
    R = np.triu(A[:n, :n])
    eps = np.finfo(dt).eps
    # |R| is unique: matches LAPACK dgeqrf up to row signs
    assert isapprox(np.abs(R), np.abs(g[f"{P[dt]}_{idx}_R"]), 1e3 * eps)
    # ||A - QR|| / ||A|| and ||Q^T Q - I||
    QR = np.zeros((m, n), dtype=dt, order="F")
    QR[:n, :n] = R
    O.tile_ormqr("N", A, T, QR, nb, ib)
    assert np.linalg.norm(QR.astype(np.float64) - A0) / np.linalg.norm(A0) < 50 * eps
    Q = np.eye(m, dtype=dt, order="F")
    O.tile_ormqr("N", A, T, Q, nb, ib)
    assert np.linalg.norm(Q.T.astype(np.float64) @ Q - np.eye(m)) < 50 * eps * np.sqrt(m)
    Qt = Q.copy(order="F")
    O.tile_ormqr("T", A, T, Qt, nb, ib)
    assert np.linalg.norm(Qt.astype(np.float64) - np.eye(m)) < 50 * eps * np.sqrt(m)
