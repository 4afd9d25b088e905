"""GPU parity of the tile-QR kernels (HIP, through the C ABI) against the CPU oracle and
the LAPACK golden vectors, on the shapes of the reference's own tests (test/larfg.jl,
test/geqr2.jl, test/geqrt.jl, test/larft.jl, test/larfb.jl, test/unmqr.jl, test/tsqrt.jl,
test/tsmqr.jl, test/parfb.jl) plus multi-tile factorizations."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import nextla_hip as nl  # noqa: E402
import oracle as O  # noqa: E402
from conftest import isapprox, relerr, rtol_for  # noqa: E402
from golden.gen_golden import GEQRT_SHAPES, TPQRT_SHAPES, TSMQR_CASES  # noqa: E402

pytestmark = pytest.mark.gpu
DTYPES = [np.float64, np.float32]
P = {np.float64: "d", np.float32: "s"}


def F(a):
    return np.asfortranarray(a)


def dev(a):
    return nl.to_device(a)


def host(t):
    torch.cuda.synchronize()
    return nl.to_host(t)


def tol(dt, f64=1e-12, f32=1e-5):
    return f64 if dt == np.float64 else f32


@pytest.mark.parametrize("dt", DTYPES)
def test_larfg_golden(golden, dt):
    g = golden("larfg")
    p = P[dt]
    keys = sorted({k.rsplit("_", 1)[0] for k in g if k.startswith(p + "_")})
    for key in keys:
        x = g[key + "_x"]
        if x.size == 0:
            continue
        dx = dev(x.copy())
        beta, tau = nl.larfg(float(g[key + "_alpha"][0]), dx)
        xo = x.copy()
        bo, to = O.larfg(g[key + "_alpha"][0], xo)
        e = 20 * np.finfo(dt).eps
        assert abs(beta - bo) <= e * abs(bo) and abs(tau - to) <= e * max(1, abs(to))
        assert isapprox(host(dx), xo, 100 * np.finfo(dt).eps)
        assert abs(abs(beta) - abs(g[key + "_beta"][0])) <= e * abs(beta)


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("m,n", [(40, 24), (128, 32), (257, 64), (16, 30), (1, 5)])
def test_geqr2_vs_oracle(dt, m, n):
    rng = np.random.default_rng(m * 7 + n)
    A0 = F(rng.standard_normal((m, n)).astype(dt))
    dA = dev(A0)
    dtau = torch.zeros(min(m, n), dtype=dA.dtype, device="cuda")
    nl.geqr2(dA, dtau)
    A = A0.copy(order="F")
    tau = np.zeros(min(m, n), dtype=dt)
    O.geqr2(A, tau)
    assert relerr(host(dA), A) < tol(dt)
    assert relerr(host(dtau), tau) < tol(dt)


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("n,k", [(30, 9), (128, 32), (200, 64)])
def test_larft_vs_oracle(dt, n, k):
    rng = np.random.default_rng(n + k)
    A = F(rng.standard_normal((n, k)).astype(dt))
    tau = np.zeros(k, dtype=dt)
    O.geqr2(A, tau)
    To = np.zeros((k, k), dtype=dt, order="F")
    O.larft(A, tau, To)
    dT = dev(np.zeros((k, k), dtype=dt))
    nl.larft("F", "C", dev(A), dev(tau), dT)
    assert relerr(np.triu(host(dT)), np.triu(To)) < tol(dt)


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("side", "LR")
@pytest.mark.parametrize("trans", "NCT")
@pytest.mark.parametrize("m,n,k", [(21, 17, 6), (130, 70, 32), (64, 300, 64), (300, 64, 48)])
def test_larfb_vs_oracle(dt, side, trans, m, n, k):
    rng = np.random.default_rng(m + 3 * n + k)
    q = m if side == "L" else n
    if k > q:
        pytest.skip("k > order of H")
    V = F(rng.standard_normal((q, k)).astype(dt))
    T = F(np.triu(rng.standard_normal((k, k))).astype(dt))
    C0 = F(rng.standard_normal((m, n)).astype(dt))
    dC = dev(C0)
    nl.larfb(side, trans, "F", "C", dev(V), dev(T), dC)
    C = C0.copy(order="F")
    O.larfb(side, trans, V, T, C)
    assert relerr(host(dC), C) < 10 * tol(dt)


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("shape", GEQRT_SHAPES)
def test_geqrt_golden_and_oracle(golden, dt, shape):
    g = golden("geqrt")
    key = f"{P[dt]}_{GEQRT_SHAPES.index(shape)}"
    m, n, ib = shape
    k = min(m, n)
    dA = dev(g[key + "_A"])
    dT = dev(np.zeros((ib, k), dtype=dt))
    _, _, dtau = nl.geqrt(dA, dT)
    A = g[key + "_A"].copy(order="F")
    T = np.zeros((ib, k), dtype=dt, order="F")
    tau = O.geqrt(A, T)
    rt = 10 * rtol_for(dt)
    Ah, Th = host(dA), host(dT)
    assert relerr(Ah, A) < tol(dt) and relerr(Th, T) < 10 * tol(dt)
    assert isapprox(Ah, g[key + "_QR"], rt)
    # T blocks (upper triangles) equal LAPACK dgeqrt's
    for i in range(0, k, ib):
        sb = min(ib, k - i)
        assert isapprox(np.triu(Th[:sb, i:i + sb]), np.triu(g[key + "_T"][:sb, i:i + sb]), rt)
    assert relerr(host(dtau), tau[:k]) < tol(dt)


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("shape", GEQRT_SHAPES)
@pytest.mark.parametrize("side", "LR")
@pytest.mark.parametrize("trans", "NT")
def test_unmqr_golden(golden, dt, shape, side, trans):
    g = golden("geqrt")
    key = f"{P[dt]}_{GEQRT_SHAPES.index(shape)}"
    m, n, ib = shape
    k = min(m, n)
    dC = dev(g[f"{key}_C{side}{trans}"])
    nl.unmqr(side, "C" if trans == "T" else "N", dev(F(g[key + "_QR"][:, :k])), dev(g[key + "_T"]), dC)
    assert isapprox(host(dC), g[f"{key}_Q{side}{trans}"], 10 * rtol_for(dt))


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("shape", TPQRT_SHAPES)
def test_tsqrt_golden_and_oracle(golden, dt, shape):
    g = golden("tsqrt")
    key = f"{P[dt]}_{TPQRT_SHAPES.index(shape)}"
    m, n, ib = shape
    d1, d2 = dev(g[key + "_A1"]), dev(g[key + "_A2"])
    dT = dev(np.zeros((ib, n), dtype=dt))
    nl.tsqrt(d1, d2, dT)
    A1, A2 = g[key + "_A1"].copy(order="F"), g[key + "_A2"].copy(order="F")
    T = np.zeros((ib, n), dtype=dt, order="F")
    O.tsqrt(A1, A2, T)
    h1, h2, hT = host(d1), host(d2), host(dT)
    assert relerr(h1, A1) < tol(dt) and relerr(h2, A2) < tol(dt) and relerr(hT, T) < 10 * tol(dt)
    rt = 10 * rtol_for(dt)
    assert isapprox(h1, g[key + "_R"], rt) and isapprox(h2, g[key + "_V"], rt) and isapprox(hT, g[key + "_T"], rt)


def test_tsqrt_rectangular_a1():
    # test/tsqrt.jl:75-156: A1 taller than wide and wider than tall (k = min(m1, n))
    rng = np.random.default_rng(4)
    for m1, n, m, ib in [(40, 24, 30, 8), (16, 24, 20, 8)]:
        A1 = F(np.triu(rng.standard_normal((m1, n))))
        A2 = F(rng.standard_normal((m, n)))
        d1, d2 = dev(A1), dev(A2)
        dT = dev(np.zeros((ib, n)))
        nl.tsqrt(d1, d2, dT)
        T = np.zeros((ib, n), order="F")
        O.tsqrt(A1, A2, T)
        assert relerr(host(d1), A1) < 1e-12 and relerr(host(d2), A2) < 1e-12


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("case", TSMQR_CASES)
@pytest.mark.parametrize("trans", "NT")
def test_tsmqr_golden_and_oracle(golden, dt, case, trans):
    g = golden("tsmqr")
    key = f"{P[dt]}_{TSMQR_CASES.index(case)}"
    side = case[0]
    d1, d2 = dev(g[key + "_A1"]), dev(g[key + "_A2"])
    tr = "C" if trans == "T" else "N"
    nl.tsmqr(side, tr, d1, d2, dev(g[key + "_V"]), dev(g[key + "_T"]))
    A1, A2 = g[key + "_A1"].copy(order="F"), g[key + "_A2"].copy(order="F")
    O.tsmqr(side, tr, A1, A2, g[key + "_V"], g[key + "_T"])
    h1, h2 = host(d1), host(d2)
    assert relerr(h1, A1) < 10 * tol(dt) and relerr(h2, A2) < 10 * tol(dt)
    rt = rtol_for(dt) * (1e3 if dt == np.float64 else 1e2)
    assert isapprox(h1, g[f"{key}_{trans}_A1"], rt) and isapprox(h2, g[f"{key}_{trans}_A2"], rt)


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("side", "LR")
@pytest.mark.parametrize("trans", "NC")
def test_parfb_vs_oracle(dt, side, trans):
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
    d1, d2 = dev(A1), dev(A2)
    nl.parfb_(side, trans, "F", "C", m1, n1, m2, n2, k, 0, d1, d2, dev(V), dev(T))
    O.parfb(side, trans, A1, A2, V, T)
    assert relerr(host(d1), A1) < 10 * tol(dt) and relerr(host(d2), A2) < 10 * tol(dt)


def test_geqrt_ib1_equals_geqr2():
    # test/geqrt.jl:54-71
    rng = np.random.default_rng(5)
    A = F(rng.standard_normal((40, 24)))
    d1, d2 = dev(A), dev(A)
    tau1 = torch.zeros(24, dtype=torch.float64, device="cuda")
    nl.geqr2(d1, tau1)
    dT = dev(np.zeros((1, 24)))
    _, _, tau2 = nl.geqrt(d2, dT)
    assert relerr(host(d1), host(d2)) < 1e-14
    assert relerr(host(dT)[0], host(tau1)) < 1e-14


def test_zero_column_gives_tau_zero():
    # test/larfg.jl:42-54 special case through the panel kernels
    A = F(np.random.default_rng(2).standard_normal((20, 6)))
    A[1:, 2] = 0.0
    A[2, 2] = 0.0
    d = dev(A)
    tau = torch.zeros(6, dtype=torch.float64, device="cuda")
    nl.geqr2(d, tau)
    Ao = A.copy(order="F")
    to = np.zeros(6)
    O.geqr2(Ao, to)
    assert relerr(host(d), Ao) < 1e-12 and relerr(host(tau), to) < 1e-12


def test_deterministic():
    # test/geqrt.jl:83-105: same input, same output (bit for bit)
    rng = np.random.default_rng(8)
    A = F(rng.standard_normal((128, 128)))
    outs = []
    for _ in range(2):
        d = dev(A)
        dT = dev(np.zeros((32, 128)))
        nl.geqrt(d, dT)
        outs.append((host(d), host(dT)))
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("idx,nb,ib", [(0, 64, 16), (0, 128, 32), (1, 64, 32), (1, 128, 32)])
def test_tile_geqrf_golden_and_oracle(golden, dt, idx, nb, ib):
    g = golden("tileqr")
    A0 = g[f"{P[dt]}_{idx}_A"]
    m, n = A0.shape
    dA = dev(A0)
    dT = nl.tile_geqrf(dA, nb, ib)
    A = A0.copy(order="F")
    T = O.tile_geqrf(A, nb, ib)
    Ah, Th = host(dA), host(dT)
    assert relerr(Ah, A) < 10 * tol(dt) and relerr(Th, T) < 100 * tol(dt)
    eps = np.finfo(dt).eps
    assert isapprox(np.abs(np.triu(Ah[:n, :n])), np.abs(g[f"{P[dt]}_{idx}_R"]), 1e3 * eps)


@pytest.mark.parametrize("dt", DTYPES)
def test_tile_qr_config1_backward_error(dt):
    # the CPU tile-QR configuration: 1024^2, nb = 128, ib = 32; ||A-QR||/||A||, ||Q^TQ-I||
    rng = np.random.Generator(np.random.PCG64(1024))
    n, nb, ib = 1024, 128, 32
    A0 = F(rng.standard_normal((n, n)).astype(dt))
    dA = dev(A0)
    dT = nl.tile_geqrf(dA, nb, ib)
    R = nl.colmajor(torch.triu(dA))
    nl.tile_ormqr("N", dA, dT, R, nb, ib)
    QR = host(R)
    eps = np.finfo(dt).eps
    assert np.linalg.norm(QR.astype(np.float64) - A0) / np.linalg.norm(A0) < 20 * eps * np.sqrt(n)
    Q = nl.to_device(np.eye(n, dtype=dt))
    nl.tile_ormqr("N", dA, dT, Q, nb, ib)
    Qh = host(Q).astype(np.float64)
    assert np.linalg.norm(Qh.T @ Qh - np.eye(n)) < 20 * eps * np.sqrt(n)


def test_python_errors_match_reference():
    A = dev(np.zeros((4, 4)))
    with pytest.raises(nl.ArgumentError, match="illegal value of ib: 0"):
        nl.geqrt_(4, 4, 0, A, dev(np.zeros((1, 4))), torch.zeros(4, dtype=torch.float64, device="cuda"))
    with pytest.raises(nl.ArgumentError, match="For side='L', A1 and A2 must have same number of columns"):
        nl.tsmqr("L", "N", A, dev(np.zeros((4, 3))), A, dev(np.zeros((2, 4))))


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("m,n,nb", [(128, 128, 64), (256, 512, 64), (512, 256, 64), (640, 640, 128), (1024, 384, 128)])
def test_tile_geqrf_dataflow_shapes_vs_oracle(dt, m, n, nb):
    # square, wide and tall tile grids through the persistent dataflow kernel (ib = 32)
    rng = np.random.default_rng(m + 3 * n + nb)
    A0 = F(rng.standard_normal((m, n)).astype(dt))
    dA = dev(A0)
    dT = nl.tile_geqrf(dA, nb, 32)
    A = A0.copy(order="F")
    T = O.tile_geqrf(A, nb, 32)
    assert relerr(host(dA), A) < 10 * tol(dt)
    assert relerr(host(dT), T) < 100 * tol(dt)


def test_tile_geqrf_repeatable():
    # the dataflow schedule varies run to run; the arithmetic of every tile does not
    rng = np.random.default_rng(77)
    A0 = F(rng.standard_normal((512, 512)))
    outs = []
    for _ in range(3):
        dA = dev(A0)
        dT = nl.tile_geqrf(dA, 128, 32)
        outs.append((host(dA), host(dT)))
    for a, t in outs[1:]:
        assert np.array_equal(a, outs[0][0]) and np.array_equal(t, outs[0][1])
