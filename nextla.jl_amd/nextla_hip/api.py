"""NextLA.jl's public interface for the tile-QR / recursive TRSM-TRMM path, on MI355X.

Names follow the reference with `!` spelled `_` (geqrt! -> geqrt_).  Arguments keep the
reference's order and meaning; matrices are torch tensors on the GPU in column-major
layout (stride(0) == 1, leading dimension = stride(1)), which is what a Julia view with
unit row stride hands to the C ABI.  Convenience forms allocate like the reference's
`similar` helpers.  Illegal arguments raise ArgumentError with the reference's message.
Calls are ordered on torch's current stream; nothing here falls back to the CPU.
"""
import numpy as np
import torch

from ._lib import ArgumentError, call, load

__all__ = [
    "ArgumentError", "colmajor", "empty_cm", "zeros_cm", "to_device", "to_host",
    "larfg_", "larfg", "geqr2_", "geqr2", "larft_", "larft", "larfb_", "larfb", "geqrt_", "geqrt",
    "unmqr_", "unmqr", "tsqrt_", "tsqrt", "tsmqr_", "tsmqr", "parfb_", "parfb", "ttqrt_", "ttqrt", "ttmqr_",
    "ttmqr", "pamm_", "pemv_", "pemv",
    "unified_rectrxm_", "LeftLowerTRSM_", "LeftUpperTRSM_", "RightLowerTRSM_", "RightUpperTRSM_",
    "LeftLowerTRMM_", "LeftUpperTRMM_", "RightLowerTRMM_", "RightUpperTRMM_", "trsm", "trmm",
    "GEMM_ADD_", "GEMM_SUB_", "tile_geqrf", "tile_geqrf_calls", "tile_ormqr", "tile_geqrf_mixed", "tile_gels",
    "rectrsm_mixed_",
    "getrf2_", "lu_", "laswp_",
    "NextLAMatrix",
    "synchronize",
]


# ----------------------------------------------------------------------------------- helpers
def colmajor(t):
    """Column-major copy (or the tensor itself if it already is one)."""
    if t.dim() == 2 and (t.stride(0) == 1 or t.size(0) <= 1) and t.stride(1) >= max(1, t.size(0)):
        return t
    return t.t().contiguous().t() if t.dim() == 2 else t.contiguous()


def empty_cm(m, n, dtype=torch.float64, device="cuda"):
    return torch.empty((n, m), dtype=dtype, device=device).t()


def zeros_cm(m, n, dtype=torch.float64, device="cuda"):
    return torch.zeros((n, m), dtype=dtype, device=device).t()


def to_device(a, device="cuda"):
    """numpy (any order) -> column-major device tensor."""
    a = np.asarray(a)
    if a.ndim == 1:
        return torch.from_numpy(a.copy()).to(device)
    return torch.from_numpy(np.ascontiguousarray(a.T)).to(device).t()


def to_host(t):
    """device tensor -> Fortran-ordered numpy array."""
    return np.asfortranarray(t.detach().cpu().numpy())


def synchronize():
    torch.cuda.current_stream().synchronize()


def _stream():
    return torch.cuda.current_stream().cuda_stream


def _p(*ts):
    dt = ts[0].dtype
    for t in ts:
        if t is None:
            continue
        if t.dtype != dt:
            raise ArgumentError(f"all arrays must have the same element type, got {dt} and {t.dtype}")
        if not t.is_cuda:
            raise ArgumentError("arrays must live on the GPU")
    if dt == torch.float64:
        return "d"
    if dt == torch.float32:
        return "s"
    if dt == torch.complex128:
        return "z"
    if dt == torch.complex64:
        return "c"
    raise ArgumentError(f"unsupported element type {dt} (Float64, Float32, ComplexF64, ComplexF32)")


def _real_only(p, what):
    if p in "zc":
        raise ArgumentError(f"{what} is implemented for Float64 and Float32 (the reference's unified_rec "
                            f"requires T <: AbstractFloat, src/rectrxm.jl:101)")


def _scalar(p, v):
    """A scalar argument passed by value for element type prefix p."""
    from ._lib import CComplex, ZComplex

    if p == "z":
        v = complex(v)
        return ZComplex(v.real, v.imag)
    if p == "c":
        v = complex(v)
        return CComplex(v.real, v.imag)
    return float(v)


def _m(t):
    """(pointer, ld) of a column-major matrix view."""
    if t.dim() == 1:
        return t.data_ptr(), max(1, t.numel())
    if not (t.stride(0) == 1 or t.size(0) <= 1):
        raise ArgumentError("matrices must have unit row stride (column-major); use colmajor()")
    return t.data_ptr(), max(1, t.stride(1), t.size(0) if t.size(1) <= 1 else 1)


def _c(x):
    return x.encode() if isinstance(x, str) else x


def _ptr(t):
    return None if t is None else t.data_ptr()


# ------------------------------------------------------------------------ Householder / panel
def larfg_(n, alpha, x, incx, tau):
    """larfg!(n, alpha, x, incx, tau) (src/larfg.jl:43-106); alpha and tau are 1-element
    device tensors, overwritten with beta and tau."""
    p = _p(x, alpha, tau)
    call(p, "larfg", n, alpha.data_ptr(), x.data_ptr(), incx, tau.data_ptr(), _stream())


def larfg(alpha, x, incx=1):
    """larfg!(alpha, x, incx, tau) convenience form (src/larfg.jl:140-145): returns (beta, tau)."""
    a = torch.tensor([alpha], dtype=x.dtype, device=x.device)
    t = torch.zeros(1, dtype=x.dtype, device=x.device)
    larfg_(x.numel() // max(incx, 1) + 1 if incx > 0 else 1, a, x, incx, t)
    return a.item(), t.item()


def ttqrt_(m, n, ib, A1, A2, T, tau, work=None):
    """ttqrt!(m, n, ib, A1, A2, T, tau, work) (src/ttqrt.jl:81-118)."""
    p = _p(A1, A2, T, tau)
    p1, l1 = _m(A1)
    p2, l2 = _m(A2)
    pt, ldt = _m(T)
    call(p, "ttqrt", m, n, ib, p1, l1, p2, l2, pt, ldt, tau.data_ptr(), _ptr(work), _stream())


def ttqrt(A1, A2, T):
    """ttqrt!(A, B, T) convenience form (src/ttqrt.jl:133-144): returns (A1, A2, T, tau)."""
    m, n = A1.shape
    if A2.shape[1] != n:
        raise ArgumentError(f"A and B must have the same number of columns: got {n} and {A2.shape[1]}")
    tau = torch.zeros(min(m, n), dtype=A1.dtype, device=A1.device)
    ttqrt_(m, n, T.shape[0], A1, A2, T, tau)
    return A1, A2, T, tau


def ttmqr_(side, trans, m1, n1, m2, n2, k, ib, A1, A2, V, T, work=None):
    """ttmqr!(side, trans, m1, n1, m2, n2, k, ib, A1, A2, V, T, work) (src/ttmqr.jl:1-95)."""
    p = _p(A1, A2, V, T)
    p1, l1 = _m(A1)
    p2, l2 = _m(A2)
    pv, ldv = _m(V)
    pt, ldt = _m(T)
    call(p, "ttmqr", _c(side), _c(trans), m1, n1, m2, n2, k, ib, p1, l1, p2, l2, pv, ldv, pt, ldt, _ptr(work),
         _stream())


def ttmqr(side, trans, A1, A2, V, T):
    """ttmqr!(side, trans, A1, A2, V, T) convenience form (src/ttmqr.jl:114-134): ib = size(T,1),
    k = min(size(T,2), side == 'L' ? m1 : n1, size(V,2))."""
    m1, n1 = A1.shape
    m2, n2 = A2.shape
    ib = T.shape[0]
    k = min(T.shape[1], m1 if side == "L" else n1, V.shape[1])
    ttmqr_(side, trans, m1, n1, m2, n2, k, ib, A1, A2, V, T)
    return A1, A2


def pamm_(op, side, storev, direct, m, n, k, l, A1, A2, V, W):
    """pamm!(op, side, storev, direct, m, n, k, l, A1, A2, V, W) (src/pamm.jl:49-100)."""
    p = _p(A1, A2, V, W)
    p1, l1 = _m(A1)
    p2, l2 = _m(A2)
    pv, ldv = _m(V)
    pw, ldw = _m(W)
    call(p, "pamm", _c(op), _c(side), _c(storev), _c(direct), m, n, k, l, p1, l1, p2, l2, pv, ldv, pw, ldw,
         _stream())


def pemv_(trans, storev, m, n, l, alpha, A, x, beta, y, work=None):
    """pemv!(trans, storev, m, n, l, alpha, A, X, beta, Y, work) (src/pemv.jl:39-181)."""
    p = _p(A, x, y)
    pa, lda = _m(A)
    call(p, "pemv", _c(trans), _c(storev), m, n, l, _scalar(p, alpha), pa, lda, x.data_ptr(), _scalar(p, beta),
         y.data_ptr(), _ptr(work), _stream())


def pemv(trans, storev, alpha, A, x, beta, y):
    """pemv!(trans, storev, alpha, A, X, beta, Y) convenience form (src/pemv.jl:211-224): l = min(m, n)."""
    m, n = A.shape
    pemv_(trans, storev, m, n, min(m, n), alpha, A, x, beta, y)
    return y


def geqr2_(m, n, A, tau, work=None):
    """geqr2!(m, n, A, tau, work) (src/geqr2.jl:107-121)."""
    p = _p(A, tau)
    pa, lda = _m(A)
    call(p, "geqr2", m, n, pa, lda, tau.data_ptr(), _ptr(work), _stream())


def geqr2(A, tau):
    """geqr2!(A, tau) (src/geqr2.jl:136-140)."""
    m, n = A.shape
    geqr2_(m, n, A, tau)
    return A, tau


def larft_(direct, storev, n, k, V, tau, T):
    """larft!(direct, storev, n, k, V, tau, T) (src/larft.jl:205-213)."""
    p = _p(V, tau, T)
    pv, ldv = _m(V)
    pt, ldt = _m(T)
    call(p, "larft", _c(direct), _c(storev), n, k, pv, ldv, tau.data_ptr(), pt, ldt, _stream())


def larft(direct, storev, V, tau, T):
    """larft!(direct, storev, V, tau, T) (src/larft.jl:256-272)."""
    if storev == "C":
        n, k = V.shape
        if tau.numel() != k:
            raise ArgumentError("For columnwise storage, length(tau) must equal size(V,2)")
    else:
        k, n = V.shape
        if tau.numel() != k:
            raise ArgumentError("For rowwise storage, length(tau) must equal size(V,1)")
    larft_(direct, storev, n, k, V, tau, T)
    return T


def larfb_(side, trans, direct, storev, m, n, k, V, ldv, T, C, work=None):
    """larfb!(side, trans, direct, storev, m, n, k, V, ldv, T, C, work) (src/larfb.jl:37-142)."""
    p = _p(V, T, C)
    pv, _ = _m(V)
    pt, ldt = _m(T)
    pc, ldc = _m(C)
    call(p, "larfb", _c(side), _c(trans), _c(direct), _c(storev), m, n, k, pv, ldv, pt, ldt, pc, ldc, _ptr(work),
         0 if work is None else max(1, work.shape[0]), _stream())


def larfb(side, trans, direct, storev, V, T, C):
    """larfb!(side, trans, direct, storev, V, T, C) (src/larfb.jl:193-210): k = size(T,1)."""
    m, n = C.shape
    k = T.shape[0]
    larfb_(side, trans, direct, storev, m, n, k, V, _m(V)[1], T, C)
    return C


def geqrt_(m, n, ib, A, T, tau, work=None):
    """geqrt!(m, n, ib, A, T, tau, work) (src/geqrt.jl:22-68)."""
    p = _p(A, T, tau)
    pa, lda = _m(A)
    pt, ldt = _m(T)
    call(p, "geqrt", m, n, ib, pa, lda, pt, ldt, tau.data_ptr(), _ptr(work), _stream())


def geqrt(A, T):
    """geqrt!(A, T) (src/geqrt.jl:84-91): ib = size(T,1); returns (A, T, tau)."""
    m, n = A.shape
    ib = T.shape[0]
    tau = torch.zeros(min(m, n), dtype=A.dtype, device=A.device)
    geqrt_(m, n, ib, A, T, tau)
    return A, T, tau


# ----------------------------------------------------------------------------- apply Q
def unmqr_(side, trans, m, n, k, ib, A, lda, T, C, work=None):
    """unmqr!(side, trans, m, n, k, ib, A, lda, T, C, work) (src/unmqr.jl:52-149)."""
    p = _p(A, T, C)
    pa, _ = _m(A)
    pt, ldt = _m(T)
    pc, ldc = _m(C)
    call(p, "unmqr", _c(side), _c(trans), m, n, k, ib, pa, lda, A.shape[1], pt, ldt, pc, ldc, _ptr(work), _stream())


def unmqr(side, trans, A, T, C):
    """unmqr!(side, trans, A, T, C) (src/unmqr.jl:196-237)."""
    m, n = C.shape
    ib, k = T.shape
    nq = m if side == "L" else n
    k = min(k, nq)
    if ib <= 0:
        raise ArgumentError(f"Block size ib must be positive, got {ib}")
    if side == "L":
        if A.shape[0] != m:
            raise ArgumentError(f"For side='L', A_qr rows ({A.shape[0]}) must match C rows ({m})")
    elif A.shape[0] != n:
        raise ArgumentError(f"For side='R', A_qr rows ({A.shape[0]}) must match C columns ({n})")
    if A.shape[1] < k:
        raise ArgumentError(f"A_qr columns ({A.shape[1]}) must be ≥ k ({k})")
    unmqr_(side, trans, m, n, k, ib, A, max(1, _m(A)[1]), T, C)
    return C


def tsqrt_(m, n, ib, A1, A2, T, tau, work=None):
    """tsqrt!(m, n, ib, A1, A2, T, tau, work) (src/tsqrt.jl:128-156); m1 = size(A1,1)."""
    p = _p(A1, A2, T, tau)
    p1, l1 = _m(A1)
    p2, l2 = _m(A2)
    pt, ldt = _m(T)
    call(p, "tsqrt", m, n, ib, p1, l1, A1.shape[0], p2, l2, pt, ldt, tau.data_ptr(), _ptr(work), _stream())


def tsqrt(A1, A2, T):
    """tsqrt!(A1, A2, T) (src/tsqrt.jl:196-211): returns (A1, A2, T, tau)."""
    m1, n = A1.shape
    m = A2.shape[0]
    ib = T.shape[0]
    if ib <= 0:
        raise ArgumentError(f"Block size ib must be positive, got {ib}")
    tau = torch.zeros(min(m1, n), dtype=A1.dtype, device=A1.device)
    tsqrt_(m, n, ib, A1, A2, T, tau)
    return A1, A2, T, tau


def tsmqr_(side, trans, m1, n1, m2, n2, k, ib, A1, A2, V, T, work=None):
    """tsmqr!(side, trans, m1, n1, m2, n2, k, ib, A1, A2, V, T, work) (src/tsmqr.jl:47-146)."""
    p = _p(A1, A2, V, T)
    p1, l1 = _m(A1)
    p2, l2 = _m(A2)
    pv, ldv = _m(V)
    pt, ldt = _m(T)
    call(p, "tsmqr", _c(side), _c(trans), m1, n1, m2, n2, k, ib, p1, l1, p2, l2, pv, ldv, V.shape[1], pt, ldt,
         T.shape[1], _ptr(work), _stream())


def tsmqr(side, trans, A1, A2, V, T):
    """tsmqr!(side, trans, A1, A2, V, T) (src/tsmqr.jl:196-228): ib, k = size(T)."""
    m1, n1 = A1.shape
    m2, n2 = A2.shape
    ib, k = T.shape
    if side == "L" and n2 != n1:
        raise ArgumentError(f"For side='L', A1 and A2 must have same number of columns. Got n1={n1}, n2={n2}")
    if side == "R" and m2 != m1:
        raise ArgumentError(f"For side='R', A1 and A2 must have same number of rows. Got m1={m1}, m2={m2}")
    if ib <= 0:
        raise ArgumentError(f"Block size ib must be positive, got {ib}")
    tsmqr_(side, trans, m1, n1, m2, n2, k, ib, A1, A2, V, T)
    return A1, A2


def parfb_(side, trans, direct, storev, m1, n1, m2, n2, k, l, A1, A2, V, T, work=None):
    """parfb!(side, trans, direct, storev, m1, n1, m2, n2, k, l, A1, A2, V, T, work) (src/parfb.jl:44-191)."""
    p = _p(A1, A2, V, T)
    p1, l1 = _m(A1)
    p2, l2 = _m(A2)
    pv, ldv = _m(V)
    pt, ldt = _m(T)
    call(p, "parfb", _c(side), _c(trans), _c(direct), _c(storev), m1, n1, m2, n2, k, l, p1, l1, p2, l2, pv, ldv, pt,
         ldt, _ptr(work), 0, _stream())


def parfb(side, trans, direct, storev, A1, A2, V, T):
    """parfb!(side, trans, direct, storev, A1, A2, V, T) convenience form (src/parfb.jl:222-238):
    k = size(T,1), l = size(V,2)."""
    m1, n1 = A1.shape
    m2, n2 = A2.shape
    parfb_(side, trans, direct, storev, m1, n1, m2, n2, T.shape[0], V.shape[1], A1, A2, V, T)
    return A1, A2


# --------------------------------------------------------------------- TRSM / TRMM / GEMM
def unified_rectrxm_(side, uplo, transpose, alpha, func, A, B, cutoff=0):
    """unified_rectrxm!(side, uplo, transpose, alpha, func, A, B) (src/rectrxm.jl:43-76).
    cutoff = recursion threshold (0: the library default 256 for both functions; the reference's
    are 256 for 'S' and 16 for 'M')."""
    p = _p(A, B)
    _real_only(p, "unified_rectrxm!")
    pa, lda = _m(A)
    pb, ldb = _m(B)
    n = A.shape[0]
    m = B.shape[1] if side == "L" else B.shape[0]
    call(p, "rectrxm", _c(side), _c(uplo), _c(transpose), float(alpha), _c(func), n, m, pa, lda, pb, ldb, int(cutoff),
         _stream())
    return B


def _base(func, side, uplo, A, B):
    p = _p(A, B)
    pa, lda = _m(A)
    pb, ldb = _m(B)
    n, m = (B.shape[0], B.shape[1]) if side == "L" else (B.shape[1], B.shape[0])
    call(p, "trsm_base" if func == "S" else "trmm_base", _c(side), _c(uplo), n, m, pa, lda, pb, ldb, _stream())
    return B


def LeftLowerTRSM_(A, B):
    """LeftLowerTRSM!(A, B) (src/trsm.jl:128-132)."""
    return _base("S", "L", "L", A, B)


def LeftUpperTRSM_(A, B):
    """LeftUpperTRSM!(A, B) (src/trsm.jl:134-138)."""
    return _base("S", "L", "U", A, B)


def RightLowerTRSM_(A, B):
    """RightLowerTRSM!(A, B) (src/trsm.jl:140-144)."""
    return _base("S", "R", "L", A, B)


def RightUpperTRSM_(A, B):
    """RightUpperTRSM!(A, B) (src/trsm.jl:146-150)."""
    return _base("S", "R", "U", A, B)


def LeftLowerTRMM_(A, B):
    """LeftLowerTRMM!(A, B) (src/trmm.jl:332-338)."""
    return _base("M", "L", "L", A, B)


def LeftUpperTRMM_(A, B):
    """LeftUpperTRMM!(A, B) (src/trmm.jl:350-355)."""
    return _base("M", "L", "U", A, B)


def RightLowerTRMM_(A, B):
    """RightLowerTRMM!(A, B) (src/trmm.jl:367-372)."""
    return _base("M", "R", "L", A, B)


def RightUpperTRMM_(A, B):
    """RightUpperTRMM!(A, B) (src/trmm.jl:384-389)."""
    return _base("M", "R", "U", A, B)


def trsm(side, uplo, transa, diag, A, B, alpha=1.0):
    """trsm(side, uplo, transa, diag, A, B, alpha) (src/trsm.jl:186-205): like the reference,
    transa and diag are not used; B is scaled by alpha, then solved in place."""
    if alpha != 1.0:
        B.mul_(alpha)
    if side not in "LR" or uplo not in "LU":
        raise ArgumentError(f"Unsupported combination of side='{side}', uplo='{uplo}'")
    return _base("S", side, uplo, A, B)


def trmm(side, uplo, transa, diag, A, B, alpha=1.0):
    """trmm(side, uplo, transa, diag, A, B, alpha) (src/trmm.jl:430-448)."""
    if side not in "LR" or uplo not in "LU":
        raise ArgumentError(f"Unsupported combination of side='{side}', uplo='{uplo}'")
    _base("M", side, uplo, A, B)
    if alpha != 1.0:
        B.mul_(alpha)
    return B


def _gemm_shapes(X, Y, C, name):
    """C (N x M) += X (N x R) * Y (R x M): the operands' inner and outer sizes must agree."""
    if X.dim() != 2 or Y.dim() != 2 or C.dim() != 2:
        raise ArgumentError(f"{name}: matrices must be 2-dimensional")
    if X.shape[1] != Y.shape[0]:
        raise ArgumentError(f"{name}: inner dimensions differ ({X.shape[1]} and {Y.shape[0]})")
    if C.shape[0] != X.shape[0] or C.shape[1] != Y.shape[1]:
        raise ArgumentError(f"{name}: result is {tuple(C.shape)}, product is ({X.shape[0]}, {Y.shape[1]})")


def GEMM_ADD_(A, B, C):
    """GEMM_ADD!(A, B, C): C += A*B (src/matmul.jl:69-74)."""
    _gemm_shapes(A, B, C, "GEMM_ADD!")
    p = _p(A, B, C)
    pa, lda = _m(A)
    pb, ldb = _m(B)
    pc, ldc = _m(C)
    call(p, "gemm_acc", A.shape[0], B.shape[0], B.shape[1], 1.0, pa, lda, pb, ldb, pc, ldc, _stream())
    return C


def GEMM_SUB_(A, B, C):
    """GEMM_SUB!(A, B, C): A -= B*C (src/matmul.jl:76-81)."""
    _gemm_shapes(B, C, A, "GEMM_SUB!")
    p = _p(A, B, C)
    pa, lda = _m(A)
    pb, ldb = _m(B)
    pc, ldc = _m(C)
    call(p, "gemm_acc", A.shape[0], C.shape[0], A.shape[1], -1.0, pb, ldb, pc, ldc, pa, lda, _stream())
    return A


# --------------------------------------------------------------------------------- LU
def getrf2_(A, ipiv):
    """getrf2!(A, ipiv, info) (src/lu.jl:185-300): A overwritten with L and U, ipiv (int64
    device vector of min(m,n)) with the 1-based pivot rows.  Returns (A, ipiv, info)."""
    p = _p(A)
    _real_only(p, "getrf2!")
    m, n = A.shape
    if ipiv.dtype != torch.int64 or not ipiv.is_cuda or ipiv.numel() < min(m, n):
        raise ArgumentError("ipiv must be an int64 device vector of length min(m, n)")
    pa, lda = _m(A)
    info = torch.zeros(1, dtype=torch.int64, device=A.device)
    call(p, "getrf2", m, n, pa, lda, ipiv.data_ptr(), info.data_ptr(), _stream())
    return A, ipiv, int(info.item())


def lu_(A):
    """lu!(A, RowMaximum()) (src/lu.jl:48-58): getrf2! with a fresh ipiv; returns (A, ipiv, info)."""
    ipiv = torch.zeros(max(1, min(A.shape)), dtype=torch.int64, device=A.device)
    return getrf2_(A, ipiv)


def laswp_(A, first, last, ipiv, incx=1):
    """laswp(A, first, last, ipiv, incx) (src/lu.jl:470-522); ipiv 1-based int64 on the device."""
    p = _p(A)
    _real_only(p, "laswp")
    pa, lda = _m(A)
    call(p, "laswp", A.shape[1], pa, lda, int(first), int(last), ipiv.data_ptr(), int(incx), _stream())
    return A


# ---------------------------------------------------------------------------- tile driver
def tile_geqrf(A, nb, ib):
    """Flat-tree tile QR of A (m x n, multiples of nb): geqrt!/unmqr!/tsqrt!/tsmqr! in PLASMA
    order.  Returns T ((m/nb)*ib x n): tile (i,k)'s ib x nb block at T[i*ib:, k*nb:]."""
    p = _p(A)
    m, n = A.shape
    T = zeros_cm((m // nb) * ib, n, dtype=A.dtype, device=A.device)
    pa, lda = _m(A)
    pt, ldt = _m(T)
    call(p, "tile_geqrf", m, n, nb, ib, pa, lda, pt, ldt, _stream())
    return T


def tile_geqrf_calls(A, nb, ib):
    """The factorization of tile_geqrf issued as the reference's callers issue it: one
    geqrt!/unmqr!/tsqrt!/tsmqr! call per tile or tile pair (flat tree, PLASMA order), through
    the per-call entry points.  Returns T in tile_geqrf's layout."""
    p = _p(A)
    m, n = A.shape
    T = zeros_cm((m // nb) * ib, n, dtype=A.dtype, device=A.device)
    pa, lda = _m(A)
    pt, ldt = _m(T)
    call(p, "tile_geqrf_calls", m, n, nb, ib, pa, lda, pt, ldt, _stream())
    return T


def tile_ormqr(trans, A, T, C, nb, ib):
    """Apply Q ('N') or Q^T ('T') of tile_geqrf to C from the left, in place."""
    p = _p(A, T, C)
    m, n = A.shape
    pa, lda = _m(A)
    pt, ldt = _m(T)
    pc, ldc = _m(C)
    call(p, "tile_ormqr", _c(trans), m, n, C.shape[1], nb, ib, pa, lda, pt, ldt, pc, ldc, _stream())
    return C


def rectrsm_mixed_(side, uplo, transpose, alpha, A, B, cutoff=0, maxit=5, tol=1e-15, prec="S"):
    """unified_rectrxm!(side, uplo, transpose, alpha, 'S', A, B) for Float64 data solved in
    Float32 storage (GEMM updates with Float32 'S', FP16 'H' or BF16 'B' operands) with Float64
    iterative refinement.  Returns (B, iterations, last ||dX||/||X||)."""
    import ctypes

    if A.dtype != torch.float64 or B.dtype != torch.float64:
        raise ArgumentError("rectrsm_mixed_ needs Float64 data")
    pa, lda = _m(A)
    pb, ldb = _m(B)
    n = A.shape[0]
    m = B.shape[1] if side == "L" else B.shape[0]
    it = ctypes.c_int64(0)
    last = ctypes.c_double(0)
    call("d", "rectrsm_mixed", _c(side), _c(uplo), _c(transpose), float(alpha), n, m, pa, lda, pb, ldb, int(cutoff),
         int(maxit), float(tol), _c(prec), ctypes.cast(ctypes.pointer(it), ctypes.c_void_p),
         ctypes.cast(ctypes.pointer(last), ctypes.c_void_p), _stream())
    return B, it.value, last.value


def tile_geqrf_mixed(A, nb, ib, prec="S"):
    """tile_geqrf of a Float64 A with the trailing updates on MFMA with Float32 accumulation
    and Float32 ('S'), FP16 ('H'), BF16 ('B') or split-BF16 ('X': each operand as two BF16
    parts, three products) operands; the panels, Householder norms/scaling and T stay
    Float64.  Returns T."""
    if A.dtype != torch.float64:
        raise ArgumentError("tile_geqrf_mixed needs Float64 data")
    m, n = A.shape
    T = zeros_cm((m // nb) * ib, n, dtype=A.dtype, device=A.device)
    pa, lda = _m(A)
    pt, ldt = _m(T)
    call("d", "tile_geqrf_mixed", m, n, nb, ib, pa, lda, pt, ldt, _c(prec), _stream())
    return T


def tile_gels(A, b, nb=128, ib=32, prec="D", maxit=100, tol=1e-15, info=None):
    """Least-squares solution of min ||A x - b||: tile QR (prec 'S' / 'H' / 'B' / 'X', Float64
    only: Float32 / FP16 / BF16 / split-BF16 trailing updates); the direct solution x0 = R^-1 (Q^T b)[0:n]
    (returned for maxit = 0; prec 'D' with maxit > 0 starts from R^-1 R^-T A^T b instead),
    corrected by R-preconditioned LSQR with products by A in the working precision.  A 16-bit
    factor whose refinement stalls is replaced by a Float32-update factor (the library does
    this by itself).  Returns (x, iterations, final estimate of ||M^T r|| / (||M|| ||r||) for
    M = A R^-1); when `info` is a dict it receives "prec_used", the update precision of the
    factor that produced x."""
    import ctypes

    m, n = A.shape
    if b.numel() < m:
        raise ArgumentError(f"b must have at least m = {m} entries, got {b.numel()}")
    if m < n:
        raise ArgumentError(f"least squares needs m >= n, got m = {m}, n = {n}")
    p = _p(A, b)
    x = torch.zeros(n, dtype=A.dtype, device=A.device)
    pa, lda = _m(A)
    it = ctypes.c_int64(0)
    last = ctypes.c_double(0)
    used = ctypes.c_char(b" ")
    call(p, "tile_gels", m, n, nb, ib, pa, lda, _ptr(b.contiguous()), _ptr(x), _c(prec), int(maxit), float(tol),
         ctypes.cast(ctypes.pointer(it), ctypes.c_void_p), ctypes.cast(ctypes.pointer(last), ctypes.c_void_p),
         ctypes.cast(ctypes.pointer(used), ctypes.c_void_p), _stream())
    if info is not None:
        info["prec_used"] = used.value.decode()
    return x, it.value, last.value


# ---------------------------------------------------------------------------- NextLAMatrix
class NextLAMatrix:
    """Host-matrix wrapper (src/NextLAMatrix.jl:4-32): `data` is a numpy matrix.  size,
    length, indexing as in the reference; trsm/trmm take the full BLAS argument list
    (side, uplo, transa, diag, alpha, A, B) and run on the GPU through unified_rectrxm!
    (host -> device, compute, device -> host).  The reference's 3-Char forwards at
    src/NextLAMatrix.jl:27-28 cannot dispatch to BLAS; these are the working 4-Char form."""

    def __init__(self, data):
        self.data = np.asfortranarray(data)

    def size(self, dim=None):
        return self.data.shape if dim is None else self.data.shape[dim - 1]

    def __len__(self):
        return self.data.size

    def __getitem__(self, idx):
        return self.data[idx]

    def __setitem__(self, idx, v):
        self.data[idx] = v

    @staticmethod
    def _run(func, side, uplo, transa, diag, alpha, A, B):
        a = A.data if isinstance(A, NextLAMatrix) else np.asarray(A)
        if diag == "U":
            a = a.copy(order="F")
            np.fill_diagonal(a, 1.0)
        b = B.data if isinstance(B, NextLAMatrix) else np.asarray(B)
        dt = np.result_type(a.dtype, b.dtype)
        if dt not in (np.float64, np.float32):
            raise ArgumentError(f"unsupported element type {dt} (Float64 and Float32 are implemented)")
        a = np.asfortranarray(a, dtype=dt)
        out = np.asfortranarray(b, dtype=dt).copy(order="F")
        n = a.shape[0]
        m = out.shape[1] if side == "L" else out.shape[0]
        p = "d" if dt == np.float64 else "s"
        # host-pointer entry point: copies in, computes on the GPU, copies out, synchronises
        st = getattr(load(), f"nla_h{p}rectrxm")(_c(side), _c(uplo), _c(transa), float(alpha), _c(func), n, m,
                                                 a.ctypes.data, max(1, a.shape[0]), out.ctypes.data,
                                                 max(1, out.shape[0]), 0, _stream())
        from ._lib import check
        check(st, f"nla_h{p}rectrxm")
        if isinstance(B, NextLAMatrix):
            B.data[...] = out
            return B
        return out

    @staticmethod
    def trsm(side, uplo, transa, diag, alpha, A, B):
        return NextLAMatrix._run("S", side, uplo, transa, diag, alpha, A, B)

    @staticmethod
    def trmm(side, uplo, transa, diag, alpha, A, B):
        return NextLAMatrix._run("M", side, uplo, transa, diag, alpha, A, B)


def _loaded():
    load()
    return True
