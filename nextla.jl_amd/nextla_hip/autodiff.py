"""Reverse-mode rules for the triangular solve / multiply and the thin tile QR.

NextLA.jl has no differentiation code; these are the rules a ChainRulesCore.rrule for
unified_rectrxm! (src/rectrxm.jl:43-76) and for a thin QR built from geqrt!/tsqrt! would
provide, exposed two ways:

* `trsm_rrule_`, `trmm_rrule_`, `qr_rrule_`: the raw C-ABI adjoint kernels (in place, like
  the forward routines);
* `trsolve`, `trmul`, `thin_qr`: torch.autograd.Function wrappers (the torch counterpart of
  an rrule) whose backward calls those kernels.  The QR backward applies Q from the stored
  compact-WY factors (V, T) and never forms it.

Matrices are column-major GPU tensors (see api.colmajor); nothing falls back to the CPU.
"""
import torch

from . import api
from ._lib import call


def _flip(t):
    return "N" if t in ("T", "C") else "T"


def trsm_rrule_(side, uplo, trans, alpha, A, X, Xbar, Abar=None, cutoff=0):
    """Adjoint of X = op(A)^-1 (alpha B) (side 'L') / (alpha B) op(A)^-1 ('R').
    Xbar is overwritten with Bbar; returns (Bbar, Abar) with Abar on A's uplo triangle."""
    p = api._p(A, X, Xbar)
    n = A.shape[0]
    m = X.shape[1] if side == "L" else X.shape[0]
    if Abar is None:
        Abar = api.empty_cm(n, n, dtype=A.dtype, device=A.device)
    pa, lda = api._m(A)
    px, ldx = api._m(X)
    pxb, ldxb = api._m(Xbar)
    pab, ldab = api._m(Abar)
    call(p, "trsm_rrule", api._c(side), api._c(uplo), api._c(trans), float(alpha), n, m, pa, lda, px, ldx, pxb, ldxb,
         pab, ldab, int(cutoff), api._stream())
    return Xbar, Abar


def trmm_rrule_(side, uplo, trans, alpha, A, B, Ybar, Abar=None, cutoff=0):
    """Adjoint of Y = alpha op(A) B ('L') / alpha B op(A) ('R') given the input B.
    Ybar is overwritten with Bbar; returns (Bbar, Abar)."""
    p = api._p(A, B, Ybar)
    n = A.shape[0]
    m = B.shape[1] if side == "L" else B.shape[0]
    if Abar is None:
        Abar = api.empty_cm(n, n, dtype=A.dtype, device=A.device)
    pa, lda = api._m(A)
    pb, ldb = api._m(B)
    pyb, ldyb = api._m(Ybar)
    pab, ldab = api._m(Abar)
    call(p, "trmm_rrule", api._c(side), api._c(uplo), api._c(trans), float(alpha), n, m, pa, lda, pb, ldb, pyb, ldyb,
         pab, ldab, int(cutoff), api._stream())
    return Ybar, Abar


def qr_rrule_(AF, T, Qbar, Rbar, nb, ib):
    """Adjoint of the thin QR A = Q R from tile_geqrf's (AF, T).  Qbar (m x n) is
    overwritten with Abar, which is returned."""
    p = api._p(AF, T, Qbar, Rbar)
    m, n = AF.shape
    pa, lda = api._m(AF)
    pt, ldt = api._m(T)
    pq, ldq = api._m(Qbar)
    pr, ldr = api._m(Rbar)
    call(p, "qr_rrule", m, n, nb, ib, pa, lda, pt, ldt, pq, ldq, pr, ldr, api._stream())
    return Qbar


class _TriSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, A, B, side, uplo, trans, alpha, cutoff):
        A = api.colmajor(A.detach())
        X = api.colmajor(B.detach().clone())
        api.unified_rectrxm_(side, uplo, trans, alpha, "S", A, X, cutoff=cutoff)
        ctx.save_for_backward(A, X)
        ctx.args = (side, uplo, trans, alpha, cutoff)
        return X

    @staticmethod
    def backward(ctx, Xbar):
        A, X = ctx.saved_tensors
        side, uplo, trans, alpha, cutoff = ctx.args
        Bbar, Abar = trsm_rrule_(side, uplo, trans, alpha, A, X, api.colmajor(Xbar.detach().clone()), cutoff=cutoff)
        return Abar, Bbar, None, None, None, None, None


class _TriMul(torch.autograd.Function):
    @staticmethod
    def forward(ctx, A, B, side, uplo, trans, alpha, cutoff):
        A = api.colmajor(A.detach())
        Bc = api.colmajor(B.detach())
        Y = api.colmajor(Bc.clone())
        api.unified_rectrxm_(side, uplo, trans, alpha, "M", A, Y, cutoff=cutoff)
        ctx.save_for_backward(A, Bc)
        ctx.args = (side, uplo, trans, alpha, cutoff)
        return Y

    @staticmethod
    def backward(ctx, Ybar):
        A, B = ctx.saved_tensors
        side, uplo, trans, alpha, cutoff = ctx.args
        Bbar, Abar = trmm_rrule_(side, uplo, trans, alpha, A, B, api.colmajor(Ybar.detach().clone()), cutoff=cutoff)
        return Abar, Bbar, None, None, None, None, None


class _ThinQR(torch.autograd.Function):
    @staticmethod
    def forward(ctx, A, nb, ib):
        m, n = A.shape
        AF = api.colmajor(A.detach().clone())
        T = api.tile_geqrf(AF, nb, ib)
        Q = api.zeros_cm(m, n, dtype=A.dtype, device=A.device)
        Q.diagonal().fill_(1)
        api.tile_ormqr("N", AF, T, Q, nb, ib)
        R = api.colmajor(torch.triu(AF[:n, :n]))
        ctx.save_for_backward(AF, T)
        ctx.blk = (nb, ib)
        return Q, R

    @staticmethod
    def backward(ctx, Qbar, Rbar):
        AF, T = ctx.saved_tensors
        nb, ib = ctx.blk
        m, n = AF.shape
        Qb = api.colmajor(Qbar.detach().clone()) if Qbar is not None else api.zeros_cm(m, n, dtype=AF.dtype,
                                                                                       device=AF.device)
        Rb = api.colmajor(Rbar.detach().clone()) if Rbar is not None else api.zeros_cm(n, n, dtype=AF.dtype,
                                                                                       device=AF.device)
        return qr_rrule_(AF, T, Qb, Rb, nb, ib), None, None


def trsolve(A, B, side="L", uplo="L", trans="N", alpha=1.0, cutoff=0):
    """X = op(A)^-1 (alpha B) (side 'L') or (alpha B) op(A)^-1 ('R'), differentiable in A and B."""
    return _TriSolve.apply(A, B, side, uplo, trans, alpha, cutoff)


def trmul(A, B, side="L", uplo="L", trans="N", alpha=1.0, cutoff=0):
    """Y = alpha op(A) B ('L') or alpha B op(A) ('R') with A triangular, differentiable."""
    return _TriMul.apply(A, B, side, uplo, trans, alpha, cutoff)


def thin_qr(A, nb=128, ib=32):
    """(Q, R) of the thin QR of A (m x n, n <= m, multiples of nb) through the tile QR,
    differentiable in A."""
    return _ThinQR.apply(A, nb, ib)
