"""Q^T application by the dataflow walkers, mixed-precision tile QR and least squares with
iterative refinement (the mixed-precision least-squares workload, at reduced sizes).

Pins: the oracle's tile_ormqr (restatement of the reference's unmqr!/tsmqr! composition),
LAPACK's FP64 least-squares solution (numpy.linalg.lstsq), and the closed-form properties
of a least-squares solution (A^T r = 0, ||A - Q R|| at the working precision)."""
import numpy as np
import pytest
import torch

from conftest import relerr

pytestmark = pytest.mark.gpu


def _nl():
    import nextla_hip as nl

    nl.load()
    return nl


def ls_problem(rng, m, n, kappa, noise=1e-8):
    """A = U diag(s) V^T with s geometric from 1 to 1/kappa; b = A x0 + noise."""
    U, _ = np.linalg.qr(rng.standard_normal((m, n)))
    V, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = kappa ** (-np.arange(n) / (n - 1))
    A = (U * s) @ V.T
    x0 = rng.standard_normal(n)
    b = A @ x0 + noise * rng.standard_normal(m)
    return np.asfortranarray(A), b


@pytest.mark.parametrize("trans", ["T", "N"])
@pytest.mark.parametrize("nrhs", [5, 64, 130])
@pytest.mark.parametrize("m,n,nb", [(512, 256, 64), (768, 512, 128), (512, 512, 128)])
def test_ormqr_walkers_vs_oracle(m, n, nb, nrhs, trans):
    nl = _nl()
    import oracle as O

    rng = np.random.default_rng(3)
    A0 = np.asfortranarray(rng.standard_normal((m, n)))
    C0 = np.asfortranarray(rng.standard_normal((m, nrhs)))
    dA = nl.to_device(A0)
    dT = nl.tile_geqrf(dA, nb, 32)
    dC = nl.to_device(C0)
    nl.tile_ormqr(trans, dA, dT, dC, nb, 32)
    torch.cuda.synchronize()
    Ah, Th = nl.to_host(dA), nl.to_host(dT)
    C = C0.copy(order="F")
    O.tile_ormqr(trans, Ah, Th, C, nb, 32)
    assert relerr(nl.to_host(dC), C) < 1e-13


@pytest.mark.parametrize("dt,tol", [(np.float64, 1e-13), (np.float32, 2e-6)])
@pytest.mark.parametrize("nb,ib", [(64, 16), (64, 32), (128, 16), (128, 32)])
@pytest.mark.parametrize("m,n", [(512, 512), (1024, 256), (768, 384), (2048, 1024)])
def test_ormqr_single_vector_vs_oracle(m, n, nb, ib, dt, tol):
    """Q^T x for one column (the single-vector wavefront) against the oracle's tile_ormqr, and
    against the multi-column walkers on the same factors with x as one of several columns."""
    nl = _nl()
    import oracle as O

    rng = np.random.default_rng(m + 7 * n + ib)
    A0 = np.asfortranarray(rng.standard_normal((m, n)).astype(dt))
    x0 = np.asfortranarray(rng.standard_normal((m, 1)).astype(dt))
    dA = nl.to_device(A0)
    dT = nl.tile_geqrf(dA, nb, ib)
    dx = nl.to_device(x0)
    nl.tile_ormqr("T", dA, dT, dx, nb, ib)
    dC = nl.to_device(np.asfortranarray(np.repeat(x0, 3, axis=1)))
    nl.tile_ormqr("T", dA, dT, dC, nb, ib)
    torch.cuda.synchronize()
    x = nl.to_host(dx)
    xo = x0.copy(order="F")
    O.tile_ormqr("T", nl.to_host(dA), nl.to_host(dT), xo, nb, ib)
    assert relerr(x, xo) < tol
    assert relerr(x, nl.to_host(dC)[:, :1]) < tol
    # Q is orthogonal: ||Q^T x|| = ||x||
    assert abs(np.linalg.norm(x.astype(np.float64)) / np.linalg.norm(x0.astype(np.float64)) - 1) < 10 * tol


# backward error bounds of the update precisions: Float32 / FP16 / BF16 / split-BF16 unit
# roundoff (6e-8 / 4.9e-4 / 3.9e-3 / 2^-17 = 7.6e-6) times a modest growth factor, and well above
# the Float64 path
MIXED_BERR = {"S": (1e-9, 2e-5), "H": (1e-6, 2e-2), "B": (1e-5, 1e-1), "X": (1e-7, 2e-3)}


@pytest.mark.parametrize("prec", ["S", "H", "B", "X"])
@pytest.mark.parametrize("m,n,nb", [(1024, 512, 128), (512, 512, 64), (2048, 256, 128), (1152, 640, 128), (640, 1152, 128),
                                    (576, 448, 64)])
def test_tile_geqrf_mixed_backward_error(m, n, nb, prec):
    nl = _nl()
    rng = np.random.default_rng(9)
    A0 = np.asfortranarray(rng.standard_normal((m, n)))
    dA = nl.to_device(A0)
    dT = nl.tile_geqrf_mixed(dA, nb, 32, prec)
    # Q R from the factors: R (upper part, zero below) then Q applied from (V, T)
    QR = nl.zeros_cm(m, n)
    QR[:n, :] = torch.triu(dA[:n, :])
    nl.tile_ormqr("N", dA, dT, QR, nb, 32)
    torch.cuda.synchronize()
    berr = relerr(nl.to_host(QR), A0)
    lo, hi = MIXED_BERR[prec]
    assert lo < berr < hi, berr
    # and the factorization is not the plain Float64 one (the updates really are reduced).
    # R is compared with its rows' signs matched: a Householder step takes the sign of the
    # entry it eliminates into, so rounding at the reduced precision may flip a whole row
    # of R (and its reflector) where that entry is small, an equally valid factorization.
    dA2 = nl.to_device(A0)
    nl.tile_geqrf(dA2, nb, 32)
    R1, R2 = np.triu(nl.to_host(dA)[:n, :]), np.triu(nl.to_host(dA2)[:n, :])
    D = np.sign(np.diag(R1)) * np.sign(np.diag(R2))
    d = relerr(D[:, None] * R1, R2)
    assert lo < d < 10 * hi, d


@pytest.mark.parametrize("prec", ["S", "H", "B", "X"])
def test_tile_geqrf_mixed_large_backward_error(prec):
    """The benchmark size (16384^2, 128 tile rows and columns): every strip goes through the
    Float32 shadow for up to 127 steps, with the launch's full concurrency.  A transport
    defect of the shadow once showed only from 12288^2 up (NaN entries in R and T)."""
    nl = _nl()
    n = 16384
    torch.manual_seed(5)
    A0 = nl.zeros_cm(n, n)
    A0.copy_(torch.randn(n, n, dtype=torch.float64, device="cuda").t())
    dA = nl.zeros_cm(n, n)
    dA.copy_(A0)
    dT = nl.tile_geqrf_mixed(dA, 128, 32, prec)
    QR = nl.zeros_cm(n, n)
    QR.copy_(torch.triu(dA))
    nl.tile_ormqr("N", dA, dT, QR, 128, 32)
    torch.cuda.synchronize()
    assert not torch.isnan(dA).any() and not torch.isnan(dT).any()
    berr = (torch.linalg.norm(QR - A0) / torch.linalg.norm(A0)).item()
    lo, hi = MIXED_BERR[prec]
    assert lo < berr < hi, berr


def test_tile_geqrf_mixed_precisions_order():
    """Float32 updates are more accurate than split BF16, split BF16 than FP16, FP16 than BF16
    (unit roundoffs 6e-8 < 7.6e-6 < 4.9e-4 < 3.9e-3), on the same matrix."""
    nl = _nl()
    rng = np.random.default_rng(10)
    A0 = np.asfortranarray(rng.standard_normal((1024, 1024)))
    errs = {}
    for prec in ("S", "X", "H", "B"):
        dA = nl.to_device(A0)
        dT = nl.tile_geqrf_mixed(dA, 128, 32, prec)
        QR = nl.zeros_cm(1024, 1024)
        QR[:, :] = torch.triu(dA)
        nl.tile_ormqr("N", dA, dT, QR, 128, 32)
        torch.cuda.synchronize()
        errs[prec] = relerr(nl.to_host(QR), A0)
    assert errs["S"] < errs["X"] / 4 < errs["H"] / 100 < errs["B"] / 100, errs


@pytest.mark.parametrize("prec,kappa", [("D", 1e2), ("D", 1e6), ("S", 1e2), ("S", 1e6), ("H", 1e2), ("B", 1e2),
                                        ("H", 1e6), ("B", 1e6), ("X", 1e2), ("X", 1e6)])
def test_gels_refinement_vs_lapack(prec, kappa):
    nl = _nl()
    rng = np.random.default_rng(int(kappa) % 1000 + "DSHBX".index(prec))
    m, n, nb = 2048, 512, 128
    A, b = ls_problem(rng, m, n, kappa)
    xref = np.linalg.lstsq(A, b, rcond=None)[0]
    dA, db = nl.to_device(A), torch.tensor(b, device="cuda")
    x, iters, last = nl.tile_gels(dA, db, nb, 32, prec, maxit=200, tol=1e-15)
    xh = x.cpu().numpy()
    r = b - A @ xh
    # error of a backward-stable LS solver: ~ u (kappa + kappa^2 ||r|| / (||A|| ||x||))
    assert relerr(xh, xref) < 50 * 2.2e-16 * kappa * (1 + kappa * 1e-8), (relerr(xh, xref), iters, last)
    # normwise LS backward-error indicator ||A^T r|| / (||A|| ||r||), against LAPACK's own
    rr = b - A @ xref
    eta = np.linalg.norm(A.T @ r) / (np.linalg.norm(A, 2) * np.linalg.norm(r))
    eta_ref = np.linalg.norm(A.T @ rr) / (np.linalg.norm(A, 2) * np.linalg.norm(rr))
    assert eta < 10 * max(eta_ref, 1e-14), (eta, eta_ref)
    if prec != "D":
        assert iters >= 1


def test_gels_mixed_without_refinement_is_float32_accurate():
    nl = _nl()
    rng = np.random.default_rng(4)
    m, n = 1024, 256
    A, b = ls_problem(rng, m, n, 1e2)
    xref = np.linalg.lstsq(A, b, rcond=None)[0]
    x, iters, _ = nl.tile_gels(nl.to_device(A), torch.tensor(b, device="cuda"), 128, 32, "S", maxit=0)
    e = relerr(x.cpu().numpy(), xref)
    assert iters == 0 and 1e-12 < e < 1e-3, e


def test_gels_argument_errors():
    nl = _nl()
    A = nl.zeros_cm(256, 128)
    b = torch.zeros(256, dtype=torch.float64, device="cuda")
    with pytest.raises(nl.ArgumentError, match="prec must be"):
        nl.tile_gels(A, b, 128, 32, "Q")
    with pytest.raises(nl.ArgumentError, match="needs Float64"):
        nl.tile_gels(nl.zeros_cm(256, 128, dtype=torch.float32), b.float(), 128, 32, "H")
    with pytest.raises(nl.ArgumentError, match="Float64"):
        nl.tile_geqrf_mixed(nl.zeros_cm(256, 128, dtype=torch.float32), 128, 32)


E0_RANGE = {"S": (1e-9, 1e-5), "H": (1e-6, 1e-2), "B": (1e-5, 1e-1)}   # before refinement


@pytest.mark.parametrize("prec", ["S", "H", "B"])
@pytest.mark.parametrize("side,uplo,trans", [("L", "L", "N"), ("L", "U", "T"), ("R", "L", "N"), ("R", "U", "N")])
def test_rectrsm_mixed_refinement(side, uplo, trans, prec):
    """Float32-storage solve (GEMM updates with Float32 / FP16 / BF16 operands) + Float64
    refinement reaches the Float64 solve's backward error."""
    nl = _nl()
    rng = np.random.default_rng(21)
    n, m = 1024, 96
    A = rng.standard_normal((n, n)) / np.sqrt(n)
    A = np.tril(A, -1) if uplo == "L" else np.triu(A, 1)
    A[np.diag_indices(n)] = rng.uniform(1, 2, n)
    B = rng.standard_normal((n, m) if side == "L" else (m, n))
    dA = nl.to_device(np.asfortranarray(A))
    X64 = nl.to_device(B)
    nl.unified_rectrxm_(side, uplo, trans, 0.5, "S", dA, X64)
    X0 = nl.to_device(B)
    _, it0, _ = nl.rectrsm_mixed_(side, uplo, trans, 0.5, dA, X0, maxit=0, prec=prec)
    Xr = nl.to_device(B)
    _, it, last = nl.rectrsm_mixed_(side, uplo, trans, 0.5, dA, Xr, maxit=12, tol=1e-15, prec=prec)
    torch.cuda.synchronize()
    opA = A if trans == "N" else A.T

    def berr(X):
        R = (opA @ X if side == "L" else X @ opA) - 0.5 * B
        return np.linalg.norm(R) / (np.linalg.norm(A) * np.linalg.norm(X))

    e64, e0, er = berr(nl.to_host(X64)), berr(nl.to_host(X0)), berr(nl.to_host(Xr))
    lo, hi = E0_RANGE[prec]
    assert it0 == 0 and lo < e0 < hi, e0               # the update precision's level before refinement
    assert it >= 1 and er < max(4 * e64, 1e-16), (er, e64, it, last)
