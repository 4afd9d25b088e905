"""The reference's tile-QR call sequence (one geqrt!/unmqr!/tsqrt!/tsmqr! call per tile or
tile pair, flat tree, PLASMA order) through the per-call entry points: nla_?tile_geqrf_calls
against the oracle's tile QR and against nla_?tile_geqrf (the persistent launch of the same
factorization), and the per-call tsmqr!/unmqr! on their own against the oracle.

Pins: the oracle's tile_geqrf (restatement of geqrt!/unmqr!/tsqrt!/tsmqr!, pinned to LAPACK
dgeqrt/dgemqrt/dtpqrt/dtpmqrt) and the golden tile-QR fixtures (LAPACK dgeqrf R factor)."""
import numpy as np
import pytest
import torch

import nextla_hip as nl
import oracle as O
import oracle_qr as Q
from conftest import relerr

pytestmark = pytest.mark.gpu

TOL = {np.float64: 1e-12, np.float32: 2e-5, np.complex128: 1e-12, np.complex64: 2e-5}


def F(a):
    return np.asfortranarray(a)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("m,n,nb,ib", [(256, 256, 64, 16), (512, 384, 128, 32), (384, 512, 128, 32),
                                       (256, 256, 128, 48), (192, 192, 64, 64)])
def test_tile_geqrf_calls_vs_oracle(dt, m, n, nb, ib):
    rng = np.random.default_rng(m * 7 + n + nb + ib)
    A0 = F(rng.standard_normal((m, n)).astype(dt))
    dA = nl.to_device(A0)
    dT = nl.tile_geqrf_calls(dA, nb, ib)
    torch.cuda.synchronize()
    A = A0.copy(order="F")
    T = O.tile_geqrf(A, nb, ib)
    assert relerr(nl.to_host(dA), A) < 10 * TOL[dt]
    assert relerr(nl.to_host(dT), T) < 100 * TOL[dt]


def oracle_calls(A, nb, ib):
    """The same call sequence on the oracle's per-call routines (any element type)."""
    m, n = A.shape
    mt, nt = m // nb, n // nb
    T = np.zeros((mt * ib, n), dtype=A.dtype, order="F")

    def get(i, j):
        return F(A[i * nb:(i + 1) * nb, j * nb:(j + 1) * nb])

    def put(X, i, j):
        A[i * nb:(i + 1) * nb, j * nb:(j + 1) * nb] = X

    for k in range(min(mt, nt)):
        Akk = get(k, k)
        Tkk, _ = Q.geqrt(Akk, ib)
        put(Akk, k, k)
        T[k * ib:(k + 1) * ib, k * nb:(k + 1) * nb] = Tkk
        for j in range(k + 1, nt):
            C = get(k, j)
            Q.unmqr("L", "C", Akk, Tkk, C)
            put(C, k, j)
        for i in range(k + 1, mt):
            Aik = get(i, k)
            Tik = np.zeros((ib, nb), dtype=A.dtype, order="F")
            Q.tsqrt(Akk, Aik, Tik, ib)
            put(Akk, k, k)
            put(Aik, i, k)
            T[i * ib:(i + 1) * ib, k * nb:(k + 1) * nb] = Tik
            for j in range(k + 1, nt):
                A1, A2 = get(k, j), get(i, j)
                Q.tsmqr("L", "C", A1, A2, Aik, Tik)
                put(A1, k, j)
                put(A2, i, j)
    return T


@pytest.mark.parametrize("dt", [np.complex128, np.complex64])
def test_tile_geqrf_calls_complex_vs_oracle(dt):
    rng = np.random.default_rng(5)
    A0 = F((rng.standard_normal((256, 192)) + 1j * rng.standard_normal((256, 192))).astype(dt))
    dA = nl.to_device(A0)
    dT = nl.tile_geqrf_calls(dA, 64, 16)
    torch.cuda.synchronize()
    A = A0.copy(order="F")
    T = oracle_calls(A, 64, 16)
    assert relerr(nl.to_host(dA), A) < 10 * TOL[dt]
    assert relerr(nl.to_host(dT), T) < 100 * TOL[dt]


def test_oracle_call_sequence_is_the_oracle_tile_qr():
    """The oracle's per-call composition above equals the oracle's tile QR (Float64)."""
    rng = np.random.default_rng(9)
    A0 = F(rng.standard_normal((256, 192)))
    A1, A2 = A0.copy(order="F"), A0.copy(order="F")
    T1 = oracle_calls(A1, 64, 16)
    T2 = O.tile_geqrf(A2, 64, 16)
    assert relerr(A1, A2) < 1e-14 and relerr(T1, T2) < 1e-13


def test_tile_geqrf_calls_matches_persistent_launch():
    """Same factorization, two schedules: the per-call sequence and the persistent launch agree
    to rounding (Float64, 1024^2, nb = 128, ib = 32)."""
    rng = np.random.default_rng(1024)
    A0 = F(rng.standard_normal((1024, 1024)))
    dA1, dA2 = nl.to_device(A0), nl.to_device(A0)
    T1 = nl.tile_geqrf_calls(dA1, 128, 32)
    T2 = nl.tile_geqrf(dA2, 128, 32)
    torch.cuda.synchronize()
    assert relerr(nl.to_host(dA1), nl.to_host(dA2)) < 1e-12
    assert relerr(nl.to_host(T1), nl.to_host(T2)) < 1e-11
