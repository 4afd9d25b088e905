"""GPU parity of the recursive LU getrf2! (src/lu.jl:185-300) and laswp (src/lu.jl:470-522)
against the oracle's restatement (pinned to LAPACK xgetrf in test_oracle_qr.py): identical
pivots and factors within the working precision; at 4096^2 / 8192 x 2048 the residual
||P A - L U|| / ||A|| and LAPACK-level growth."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import nextla_hip as nl  # noqa: E402
import oracle_qr as Q  # noqa: E402

pytestmark = pytest.mark.gpu
TOL = {np.float64: 1e-12, np.float32: 1e-4}


def rel(x, y):
    return np.linalg.norm(x - y) / max(float(np.linalg.norm(y)), 1e-300)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("m,n", [(1, 1), (1, 40), (7, 1), (10, 10), (64, 40), (40, 64), (129, 129), (300, 77),
                                 (77, 300), (513, 512), (3000, 24), (5000, 9), (9000, 16)])
def test_getrf2_vs_oracle(dt, m, n):
    """Pivots identical to the oracle's; the tall shapes take the multi-workgroup leaves (more
    than 2048 rows in Float64, 4096 in Float32, in the first leaves)."""
    rng = np.random.default_rng(m * 1000 + n)
    A = np.asfortranarray(rng.standard_normal((m, n)).astype(dt))
    Ao = A.copy(order="F")
    ipiv_o, info_o = Q.getrf2(Ao)
    dA = nl.to_device(A)
    ipiv = torch.zeros(min(m, n), dtype=torch.int64, device="cuda")
    _, _, info = nl.getrf2_(dA, ipiv)
    assert info == info_o == 0
    assert np.array_equal(ipiv.cpu().numpy(), ipiv_o)
    assert rel(nl.to_host(dA), Ao) < TOL[dt] * (10 if m > 200 else 1)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_getrf2_tall_zero_column_info(dt):
    """A zero column in a tall panel (the multi-workgroup leaf): U(3,3) = 0 exactly, info = 3,
    no swap and no scaling for that column, as xgetrf."""
    rng = np.random.default_rng(17)
    A = np.asfortranarray(rng.standard_normal((6000, 12)).astype(dt))
    A[:, 2] = 0
    Ao = A.copy(order="F")
    ipiv_o, info_o = Q.getrf2(Ao)
    dA = nl.to_device(A)
    _, ipiv, info = nl.lu_(dA)
    assert info == info_o == 3
    assert np.array_equal(ipiv.cpu().numpy(), ipiv_o)
    assert rel(nl.to_host(dA), Ao) < TOL[dt]


def test_getrf2_zero_pivot_info():
    # a rank-deficient matrix: U(3,3) = 0 exactly, as xgetrf reports (info = 3)
    A = np.asfortranarray(np.array([[1.0, 2.0, 3.0], [2.0, 4.0, 6.0], [1.0, 1.0, 1.0]]))
    Ao = A.copy(order="F")
    ipiv_o, info_o = Q.getrf2(Ao)
    dA = nl.to_device(A)
    _, ipiv, info = nl.lu_(dA)
    assert info == info_o == 3 and np.array_equal(ipiv.cpu().numpy(), ipiv_o)
    assert rel(nl.to_host(dA), Ao) < 1e-15


@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
@pytest.mark.parametrize("m,n", [(4096, 4096), (8192, 2048), (2048, 6000)])
def test_getrf2_large_residual(dt, m, n):
    gen = torch.Generator(device="cuda")
    gen.manual_seed(m + n)
    A0 = torch.randn((n, m), generator=gen, device="cuda", dtype=torch.float64).t()
    A = nl.colmajor(A0.to(dt).clone())
    _, ipiv, info = nl.lu_(A)
    assert info == 0
    k = min(m, n)
    F_ = A.double()
    L = torch.tril(F_[:, :k], -1) + torch.eye(m, k, device="cuda", dtype=torch.float64)
    U = torch.triu(F_[:k, :])
    # P A: apply the interchanges in order to the rows of A
    perm = torch.arange(m, device="cuda")
    piv = ipiv.cpu().numpy() - 1
    p = perm.cpu().numpy()
    for i in range(k):
        p[[i, piv[i]]] = p[[piv[i], i]]
    PA = A0.to(dt).double()[torch.from_numpy(p).cuda()]
    res = (torch.linalg.norm(PA - L @ U) / torch.linalg.norm(PA)).item()
    u = 2.2e-16 if dt == torch.float64 else 1.2e-7
    assert res < 50 * u * np.sqrt(k), res
    # partial pivoting bounds |L| by 1
    assert torch.max(torch.abs(torch.tril(F_[:, :k], -1))).item() <= 1.0


def test_laswp_round_trip():
    rng = np.random.default_rng(3)
    B = np.asfortranarray(rng.standard_normal((40, 33)))
    ip = rng.integers(1, 41, 40).astype(np.int64)
    ip = np.maximum(ip, np.arange(1, 41))        # LAPACK pivots: ipiv(k) >= k
    dB = nl.to_device(B)
    dip = torch.from_numpy(ip).cuda()
    nl.laswp_(dB, 1, 40, dip, 1)
    Bo = B.copy(order="F")
    Q.laswp(Bo, 1, 40, ip, 1)
    assert np.array_equal(nl.to_host(dB), Bo)
    nl.laswp_(dB, 1, 40, dip, -1)
    assert np.array_equal(nl.to_host(dB), B)
