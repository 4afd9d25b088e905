"""Adjoint kernels (TRSM / TRMM / thin QR) on the GPU.

The reference has no differentiation code, so these are pinned by
  * torch's own reverse-mode rules in Float64 on the CPU (torch.linalg.solve_triangular,
    triangular matmul and torch.linalg.qr, the counterparts of ChainRules' FP64
    LinearAlgebra rrules, which need Julia), and
  * central finite differences in Float64 (h = 1e-6, 16 random directions) at the
    differentiation workload's sizes: TRSM A 4096 x 4096, B 4096 x 512; thin QR of
    8192 x 1024.
"""
import numpy as np
import pytest
import torch

from conftest import relerr

pytestmark = pytest.mark.gpu

TOL = {np.float64: 1e-10, np.float32: 2e-4}


def _mods():
    import nextla_hip as nl
    from nextla_hip import autodiff as ad

    nl.load()
    return nl, ad


def tri(rng, n, uplo):
    A = rng.standard_normal((n, n)) / np.sqrt(n)
    A = np.tril(A, -1) if uplo == "L" else np.triu(A, 1)
    A[np.diag_indices(n)] = rng.uniform(1, 2, n)
    return A


def _torch_trx(func, side, uplo, trans, alpha, A, B):
    """Float64 CPU forward with torch (A's other triangle masked out)."""
    Am = torch.tril(A) if uplo == "L" else torch.triu(A)
    op = Am if trans == "N" else Am.T
    upper = (uplo == "U") != (trans != "N")
    if func == "S":
        return torch.linalg.solve_triangular(op, alpha * B, upper=upper, left=(side == "L"))
    return alpha * (op @ B) if side == "L" else alpha * (B @ op)


CASES = [(s, u, t) for s in "LR" for u in "LU" for t in "NT"]


@pytest.mark.parametrize("side,uplo,trans", CASES)
@pytest.mark.parametrize("func", ["S", "M"])
@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_trxm_rrule_vs_torch(side, uplo, trans, func, dt):
    nl, ad = _mods()
    rng = np.random.default_rng(11)
    n, m, alpha = 300, 48, 0.75
    A = tri(rng, n, uplo)
    shape = (n, m) if side == "L" else (m, n)
    B = rng.standard_normal(shape)
    W = rng.standard_normal(shape)
    At = torch.tensor(A, requires_grad=True)
    Bt = torch.tensor(B, requires_grad=True)
    (_torch_trx(func, side, uplo, trans, alpha, At, Bt) * torch.tensor(W)).sum().backward()
    Ag = torch.tensor(A.astype(dt), device="cuda", requires_grad=True)
    Bg = torch.tensor(B.astype(dt), device="cuda", requires_grad=True)
    f = ad.trsolve if func == "S" else ad.trmul
    Y = f(Ag, Bg, side, uplo, trans, alpha, cutoff=64)
    (Y * torch.tensor(W.astype(dt), device="cuda")).sum().backward()
    torch.cuda.synchronize()
    tol = TOL[dt]
    assert relerr(Ag.grad.cpu().numpy(), At.grad.numpy()) < tol
    assert relerr(Bg.grad.cpu().numpy(), Bt.grad.numpy()) < tol


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("m,n,nb", [(256, 128, 64), (512, 256, 128), (384, 384, 128)])
def test_qr_rrule_vs_torch(m, n, nb, dt):
    nl, ad = _mods()
    rng = np.random.default_rng(5)
    A = rng.standard_normal((m, n))
    GQ, GR = rng.standard_normal((m, n)), rng.standard_normal((n, n))
    Ag = torch.tensor(A.astype(dt), device="cuda", requires_grad=True)
    Qg, Rg = ad.thin_qr(Ag, nb, 32)
    # The tile QR's R rows may differ in sign from LAPACK's single-column Householder QR
    # (each tsqrt re-reflects the diagonal): Q_tile = Q D, R_tile = D R with D = diag(+-1).
    At = torch.tensor(A, requires_grad=True)
    Q, R = torch.linalg.qr(At, mode="reduced")
    D = torch.sign(torch.diagonal(Rg.detach().double().cpu())) * torch.sign(torch.diagonal(R.detach()))
    Qd, Rd = Q * D, D[:, None] * R
    ((Qd * torch.tensor(GQ)).sum() + (Rd * torch.tensor(GR)).sum()).backward()
    assert relerr(Qg.detach().cpu().numpy(), Qd.detach().numpy()) < (1e-11 if dt == np.float64 else 1e-4)
    assert relerr(Rg.detach().cpu().numpy(), Rd.detach().numpy()) < (1e-11 if dt == np.float64 else 1e-4)
    ((Qg * torch.tensor(GQ.astype(dt), device="cuda")).sum() + (Rg * torch.tensor(GR.astype(dt), device="cuda")).sum()).backward()
    torch.cuda.synchronize()
    assert relerr(Ag.grad.cpu().numpy(), At.grad.numpy()) < (1e-9 if dt == np.float64 else 2e-3)


def _fd_check(loss, grads, params, rng, ndir=16, h=1e-6):
    """Central differences of loss along ndir random directions vs <grad, dir>."""
    worst = 0.0
    for _ in range(ndir):
        dirs = [torch.tensor(rng.standard_normal(tuple(p.shape)), dtype=torch.float64, device="cuda") for p in params]
        dirs = [d * msk for d, msk in zip(dirs, [g[1] for g in grads])]
        plus = loss([p + h * d for p, d in zip(params, dirs)])
        minus = loss([p - h * d for p, d in zip(params, dirs)])
        fd = (plus - minus) / (2 * h)
        an = sum(float((g[0].double() * d).sum()) for g, d in zip(grads, dirs))
        worst = max(worst, abs(fd - an) / max(abs(an), 1e-30))
    return worst


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_trsm_adjoint_fd_full(dt):
    """TRSM X = A \\ B, A 4096 (lower, as the TRSM workload), B 4096 x 512, loss sum(W .* X)."""
    nl, ad = _mods()
    rng = np.random.default_rng(2024)
    n, m = 4096, 512
    A = torch.tensor(tri(rng, n, "L"), device="cuda")
    B = torch.tensor(rng.standard_normal((n, m)), device="cuda")
    W = torch.tensor(rng.standard_normal((n, m)), device="cuda")
    Ag = A.to(torch.float64 if dt == np.float64 else torch.float32).requires_grad_(True)
    Bg = B.to(Ag.dtype).requires_grad_(True)
    (ad.trsolve(Ag, Bg) * W.to(Ag.dtype)).sum().backward()
    mask_a = torch.tril(torch.ones(n, n, dtype=torch.float64, device="cuda"))

    def loss(ps):
        Ap, Bp = ps
        X = nl.colmajor(Bp.clone())
        nl.unified_rectrxm_("L", "L", "N", 1.0, "S", nl.colmajor(Ap.contiguous()), X)
        return float((X * W).sum())

    worst = _fd_check(loss, [(Ag.grad, mask_a), (Bg.grad, 1.0)], [A, B], rng)
    assert worst < (1e-6 if dt == np.float64 else 1e-3), worst


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_qr_adjoint_fd_full(dt):
    """Thin QR of A 8192 x 1024 (iid N(0,1)), loss sum(GQ .* Q) + sum(GR .* R)."""
    nl, ad = _mods()
    rng = np.random.default_rng(77)
    m, n = 8192, 1024
    A = torch.tensor(rng.standard_normal((m, n)), device="cuda")
    GQ = torch.tensor(rng.standard_normal((m, n)), device="cuda")
    GR = torch.tensor(rng.standard_normal((n, n)), device="cuda")
    Ag = A.to(torch.float64 if dt == np.float64 else torch.float32).requires_grad_(True)
    Q, R = ad.thin_qr(Ag, 128, 32)
    ((Q * GQ.to(Ag.dtype)).sum() + (R * GR.to(Ag.dtype)).sum()).backward()

    def loss(ps):
        Qp, Rp = ad.thin_qr(ps[0], 128, 32)
        return float((Qp * GQ).sum() + (Rp * GR).sum())

    worst = _fd_check(loss, [(Ag.grad, 1.0)], [A], rng)
    assert worst < (1e-6 if dt == np.float64 else 1e-3), worst


def test_rrule_argument_errors():
    nl, ad = _mods()
    A = nl.zeros_cm(8, 8)
    X = nl.zeros_cm(8, 4)
    with pytest.raises(nl.ArgumentError, match="side must be"):
        ad.trsm_rrule_("X", "L", "N", 1.0, A, X, X.clone())
