"""GPU parity of the QR-family entry points for every element type (Float64, Float32,
ComplexF64, ComplexF32) against the scalar-generic oracle (oracle/nla_oracle_qr_t.h, pinned
by LAPACK and closed forms in test_oracle_qr.py), through the C ABI:
larfg!, geqr2!, larft! ('C' and 'R'), larfb!, geqrt!, unmqr!, tsqrt!, tsmqr!, the
triangle-on-triangle ttqrt! / ttmqr!, parfb! in every side / trans / direct / storev with
pentagon orders 0 and l > 0, pamm! and pemv!.  Shapes follow the reference's own tests
(test/ttqrt.jl, test/ttmqr.jl, test/parfb.jl, test/pamm.jl, test/pemv.jl, test/larft.jl)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import nextla_hip as nl  # noqa: E402
import oracle_qr as Q  # noqa: E402

pytestmark = pytest.mark.gpu
DTYPES = [np.float64, np.float32, np.complex128, np.complex64]
TOL = {np.float64: 1e-12, np.float32: 1e-5, np.complex128: 1e-12, np.complex64: 1e-5}
CPLX = (np.complex128, np.complex64)


def F(a):
    return np.asfortranarray(a)


def rnd(rng, shape, dt):
    a = rng.standard_normal(shape)
    if dt in CPLX:
        a = a + 1j * rng.standard_normal(shape)
    return F(a.astype(dt))


def rel(x, y):
    return np.linalg.norm(np.asarray(x, dtype=np.complex128) - y) / max(float(np.linalg.norm(y)), 1e-300)


def dev(a):
    return nl.to_device(a)


def host(t):
    torch.cuda.synchronize()
    return nl.to_host(t) if t.dim() == 2 else t.cpu().numpy()


def vdev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("n,scale", [(2, 1.0), (33, 1.0), (300, 1.0), (6, 0.0), (12, "tiny")])
def test_larfg(dt, n, scale):
    rng = np.random.default_rng(n)
    x = rnd(rng, n - 1, dt)
    alpha = rnd(rng, 1, dt)[0]
    if scale == 0.0:
        x[:] = 0
    if scale == "tiny":
        s = 1e-300 if dt in (np.float64, np.complex128) else 1e-40
        x = (x * s).astype(dt)
        alpha = dt(alpha * s)
    xo = x.copy()
    b, t = Q.larfg(alpha, xo)
    dx, da, dtau = vdev(x), vdev(np.array([alpha], dtype=dt)), vdev(np.zeros(1, dtype=dt))
    nl.larfg_(n, da, dx, 1, dtau)
    assert abs(host(da)[0] - b) <= TOL[dt] * max(abs(b), 1e-300)
    assert abs(host(dtau)[0] - t) <= TOL[dt] * max(abs(t), 1)
    assert rel(host(dx), xo) < TOL[dt] or np.linalg.norm(xo) == 0


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("m,n,ib", [(40, 24, 8), (64, 64, 16), (33, 20, 5), (20, 20, 10), (257, 64, 32)])
def test_geqrt_geqr2(dt, m, n, ib):
    rng = np.random.default_rng(m * n + ib)
    A = rnd(rng, (m, n), dt)
    Ao = A.copy(order="F")
    T, tau = Q.geqrt(Ao, ib)
    dA, dT, dtau = dev(A), dev(np.zeros((ib, n), dtype=dt)), vdev(np.zeros(min(m, n), dtype=dt))
    nl.geqrt_(m, n, ib, dA, dT, dtau)
    assert rel(host(dA), Ao) < TOL[dt] and rel(host(dT), T) < TOL[dt] and rel(host(dtau), tau) < TOL[dt]
    A2 = A.copy(order="F")
    tau2 = Q.geqr2(A2)
    dA2, dtau2 = dev(A), vdev(np.zeros(min(m, n), dtype=dt))
    nl.geqr2_(m, n, dA2, dtau2)
    assert rel(host(dA2), A2) < TOL[dt] and rel(host(dtau2), tau2) < TOL[dt]


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("storev", "CR")
@pytest.mark.parametrize("n,k", [(1, 1), (5, 3), (10, 4), (20, 8), (200, 40)])
def test_larft(dt, storev, n, k):
    rng = np.random.default_rng(n + k)
    V = rnd(rng, (n, k) if storev == "C" else (k, n), dt)
    for j in range(k):
        if storev == "C":
            V[:j, j] = 0
            V[j, j] = 1
        else:
            V[j, :j] = 0
            V[j, j] = 1
    tau = rnd(rng, k, dt)
    if n > 20:
        # random tau with long vectors makes T grow like prod(|tau| |v|^2) and overflow Float32:
        # use Householder-sized tau = 2 / |v|^2 (times a random phase) beyond the reference's sizes
        norms = np.linalg.norm(V, axis=0 if storev == "C" else 1) ** 2
        tau = (2 / norms * tau / np.abs(tau)).astype(dt)
    T = np.zeros((k, k), dtype=dt, order="F")
    Q.larft(storev, V, tau, T)
    dT = dev(np.zeros((k, k), dtype=dt))
    nl.larft_("F", storev, n, k, dev(V), vdev(tau), dT)
    assert rel(host(dT), T) < TOL[dt]


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("side", "LR")
@pytest.mark.parametrize("trans", "NC")
def test_larfb_unmqr(dt, side, trans):
    rng = np.random.default_rng(3)
    nq, k, ib, other = 70, 40, 16, 23
    A = rnd(rng, (nq, k), dt)
    T, _ = Q.geqrt(A, ib)
    C = rnd(rng, (nq, other) if side == "L" else (other, nq), dt)
    Co = C.copy(order="F")
    Q.unmqr(side, trans, A, T, Co)
    dC = dev(C)
    nl.unmqr_(side, trans, C.shape[0], C.shape[1], k, ib, dev(A), nq, dev(T), dC)
    assert rel(host(dC), Co) < TOL[dt]
    # larfb! with one block of ib reflectors (V, T from the first block)
    Vb, Tb = F(A[:, :ib]), F(T[:, :ib])
    C2 = C.copy(order="F")
    Q.larfb(side, trans, Vb, Tb, C2)
    dC2 = dev(C)
    nl.larfb_(side, trans, "F", "C", C.shape[0], C.shape[1], ib, dev(Vb), nq, dev(Tb), dC2)
    assert rel(host(dC2), C2) < TOL[dt]


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("m,n,ib", [(32, 24, 8), (48, 32, 16), (20, 20, 10), (128, 128, 32)])
def test_tsqrt_tsmqr(dt, m, n, ib):
    rng = np.random.default_rng(m + n)
    A1 = F(np.triu(rnd(rng, (n, n), dt)))
    A2 = rnd(rng, (m, n), dt)
    a1, a2 = A1.copy(order="F"), A2.copy(order="F")
    T = np.zeros((ib, n), dtype=dt, order="F")
    tau = Q.tsqrt(a1, a2, T)
    d1, d2, dT, dtau = dev(A1), dev(A2), dev(np.zeros((ib, n), dtype=dt)), vdev(np.zeros(n, dtype=dt))
    nl.tsqrt_(m, n, ib, d1, d2, dT, dtau)
    assert rel(np.triu(host(d1)), np.triu(a1)) < TOL[dt] and rel(host(d2), a2) < TOL[dt]
    assert rel(host(dT), T) < TOL[dt]
    for side in "LR":
        for trans in "NC":
            B1 = rnd(rng, (n, 19) if side == "L" else (19, n), dt)
            B2 = rnd(rng, (m, 19) if side == "L" else (19, m), dt)
            b1, b2 = B1.copy(order="F"), B2.copy(order="F")
            Q.tsmqr(side, trans, b1, b2, a2, T)
            e1, e2 = dev(B1), dev(B2)
            nl.tsmqr(side, trans, e1, e2, dev(a2), dev(T))
            assert rel(host(e1), b1) < TOL[dt] and rel(host(e2), b2) < TOL[dt], (side, trans)


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("m,n,ib", [(16, 16, 4), (24, 24, 8), (32, 32, 8), (48, 48, 16), (32, 16, 4), (48, 24, 8),
                                    (64, 32, 16), (16, 32, 4), (24, 48, 8), (128, 128, 32), (256, 256, 96)])
def test_ttqrt(dt, m, n, ib):
    # test/ttqrt.jl:6-39, 72-159: triangular tiles with 2 added to the diagonal
    rng = np.random.default_rng(m * 7 + n + ib)
    k = min(m, n)
    A1 = F(np.triu(rnd(rng, (m, n), dt)))
    A2 = F(np.triu(rnd(rng, (m, n), dt)))
    for i in range(k):
        A1[i, i] += 2
        A2[i, i] += 2
    a1, a2 = A1.copy(order="F"), A2.copy(order="F")
    T = np.zeros((ib, n), dtype=dt, order="F")
    tau = Q.ttqrt(a1, a2, T)
    d1, d2, dT, dtau = dev(A1), dev(A2), dev(np.zeros((ib, n), dtype=dt)), vdev(np.zeros(k, dtype=dt))
    nl.ttqrt_(m, n, ib, d1, d2, dT, dtau)
    assert rel(host(d1), a1) < TOL[dt] and rel(host(d2), a2) < TOL[dt]
    assert rel(host(dT), T) < TOL[dt] and rel(host(dtau), tau[:k]) < TOL[dt]


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("side", "LR")
@pytest.mark.parametrize("trans", "NC")
@pytest.mark.parametrize("n,k,ib", [(24, 16, 8), (200, 160, 40), (300, 200, 100)])
def test_ttmqr(dt, side, trans, n, k, ib):
    # test/ttmqr.jl:6-39 (n = 24, k = 16, ib = 8, V = triu, T = triu(rand)); larger shapes cover
    # several 64-wide sub-blocks (ib 40) and the general block kernel (ib 100 > 64)
    rng = np.random.default_rng(n + k + ib)
    A1, A2 = rnd(rng, (n, n), dt), rnd(rng, (n, n), dt)
    V = F(np.triu(rnd(rng, (n, k), dt)))
    Tm = F((np.triu(rnd(rng, (ib, k), dt)) / np.sqrt(k)).astype(dt))
    a1, a2 = A1.copy(order="F"), A2.copy(order="F")
    Q.ttmqr(side, trans, a1, a2, V, Tm)
    d1, d2 = dev(A1), dev(A2)
    nl.ttmqr(side, trans, d1, d2, dev(V), dev(Tm))
    assert rel(host(d1), a1) < 10 * TOL[dt] and rel(host(d2), a2) < 10 * TOL[dt]


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("side", "LR")
@pytest.mark.parametrize("trans", "NC")
@pytest.mark.parametrize("direct", "FB")
@pytest.mark.parametrize("storev", "CR")
@pytest.mark.parametrize("m2,n2,k,l", [(20, 24, 8, 4), (20, 24, 8, 0), (150, 130, 100, 37)])
def test_parfb_variants(dt, side, trans, direct, storev, m2, n2, k, l):
    # test/parfb.jl:6-52 (random V and T, every variant); k = 100 takes the general kernel
    if k > min(m2, n2) and l > 0:
        pytest.skip("pentagon larger than the block")
    rng = np.random.default_rng(m2 + k + l + "FB".index(direct) + 2 * "CR".index(storev))
    m1, n1 = (k, n2) if side == "L" else (m2, k)
    A1, A2 = rnd(rng, (m1, n1), dt), rnd(rng, (m2, n2), dt)
    rows = m2 if side == "L" else n2
    V = rnd(rng, (rows, k) if storev == "C" else (k, rows), dt)
    T = F((rnd(rng, (k, k), dt) / np.sqrt(k)).astype(dt))
    a1, a2 = A1.copy(order="F"), A2.copy(order="F")
    Q.parfb(side, trans, direct, storev, m1, n1, m2, n2, k, l, a1, a2, V, T)
    d1, d2 = dev(A1), dev(A2)
    nl.parfb_(side, trans, direct, storev, m1, n1, m2, n2, k, l, d1, d2, dev(V), dev(T))
    assert rel(host(d1), a1) < 10 * TOL[dt] and rel(host(d2), a2) < 10 * TOL[dt]


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("op", "WA")
@pytest.mark.parametrize("side", "LR")
@pytest.mark.parametrize("direct", "FB")
@pytest.mark.parametrize("storev", "CR")
def test_pamm(dt, op, side, direct, storev):
    # test/pamm.jl: m, n, k, l = 40, 30, 20, 10 (left) / 30, 40, 20, 10 (right)
    rng = np.random.default_rng(17)
    m, n, k, l = (40, 30, 20, 10) if side == "L" else (30, 40, 20, 10)
    rows = m if side == "L" else n
    A1 = rnd(rng, (k, n) if side == "L" else (m, k), dt)
    A2 = rnd(rng, (m, n), dt)
    V = rnd(rng, (rows, k) if storev == "C" else (k, rows), dt)
    W = rnd(rng, (k, n) if side == "L" else (m, k), dt)
    a2, w = A2.copy(order="F"), W.copy(order="F")
    Q.pamm(op, side, storev, direct, m, n, k, l, A1, a2, V, w)
    d2, dw = dev(A2), dev(W)
    nl.pamm_(op, side, storev, direct, m, n, k, l, dev(A1), d2, dev(V), dw)
    assert rel(host(d2), a2) < TOL[dt] and rel(host(dw), w) < TOL[dt]


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("trans,storev,m,n,l", [("C", "C", 40, 30, 10), ("T", "C", 40, 30, 10),
                                                ("N", "R", 30, 40, 10), ("C", "C", 20, 15, 0),
                                                ("C", "C", 50, 40, 25), ("C", "C", 12, 8, 1),
                                                ("N", "R", 700, 900, 300), ("T", "C", 1000, 600, 600)])
def test_pemv(dt, trans, storev, m, n, l):
    if trans == "T" and dt in CPLX:
        pytest.skip("trans 'T' is the real-type form (test/pemv.jl)")
    rng = np.random.default_rng(m + n + l)
    A = rnd(rng, (m, n), dt)
    x = rnd(rng, m if storev == "C" else n, dt)
    y0 = rnd(rng, n if storev == "C" else m, dt)
    alpha, beta = (2.5 + 1.5j, 1.2 - 0.8j) if dt in CPLX else (2.5, 1.25)
    for a, b, nan_y in ((alpha, beta, False), (alpha, 0.0, True), (0.0, 0.0, False), (0.0, beta, False)):
        y = y0.copy()
        if nan_y:
            y[:] = np.nan
        yo = y.copy()
        Q.pemv(trans, storev, l, dt(a), A, x, dt(b), yo)
        dy = vdev(y)
        nl.pemv_(trans, storev, m, n, l, a, dev(A), vdev(x), b, dy)
        got = host(dy)
        if a == 0 and b == 0:
            assert np.array_equal(got, y0)   # alpha = beta = 0 leaves y as it was
        else:
            assert rel(got, yo) < TOL[dt], (a, b)


def test_generic_argument_errors():
    z = lambda m, n: nl.zeros_cm(m, n, dtype=torch.complex128)
    with pytest.raises(nl.ArgumentError, match="illegal value of m"):
        nl.ttqrt_(-1, 8, 4, z(8, 8), z(8, 8), z(4, 8), torch.zeros(8, dtype=torch.complex128, device="cuda"))
    with pytest.raises(nl.ArgumentError, match="illegal value of side"):
        nl.ttmqr_("X", "N", 8, 8, 8, 8, 4, 2, z(8, 8), z(8, 8), z(8, 4), z(2, 4))
    with pytest.raises(nl.ArgumentError, match="not valid for complex"):
        nl.tsmqr("L", "T", z(8, 8), z(8, 8), z(8, 4), z(2, 4))
    with pytest.raises(nl.ArgumentError, match="illegal values of trans/storev"):
        nl.pemv_("N", "C", 4, 4, 0, 1.0, z(4, 4), torch.zeros(4, dtype=torch.complex128, device="cuda"), 0.0,
                 torch.zeros(4, dtype=torch.complex128, device="cuda"))
    with pytest.raises(nl.ArgumentError, match="op must be"):
        nl.pamm_("X", "L", "C", "F", 4, 4, 2, 1, z(2, 4), z(4, 4), z(4, 2), z(2, 4))
    with pytest.raises(nl.ArgumentError, match="AbstractFloat"):
        nl.unified_rectrxm_("L", "L", "N", 1.0, "S", z(4, 4), z(4, 4))
