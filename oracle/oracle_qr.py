"""CPU oracle for the QR-family routines over real AND complex element types: ctypes front
end of the scalar-generic restatement in oracle/nla_oracle_qr_t.h (nlaq_{d,s,z,c}*).

TEST INFRASTRUCTURE ONLY.  Imported by tests/ as the checker; the product path never
touches it.  Arrays are Fortran-ordered numpy arrays of float64 / float32 / complex128 /
complex64, updated in place; the reference routine each function follows is cited in
nla_oracle_qr_t.h.
"""
import ctypes

import numpy as np

from oracle import _c, _ld, _ptr, lib as _base_lib

i64 = ctypes.c_int64
ch = ctypes.c_char
vp = ctypes.c_void_p
_PRE = {np.float64: "d", np.float32: "s", np.complex128: "z", np.complex64: "c"}
_SIGS = {
    "larfg": [i64, vp, vp, i64, vp],
    "geqr2": [i64, i64, vp, i64, vp, vp],
    "larft": [ch, i64, i64, vp, i64, vp, vp, i64],
    "larfb": [ch, ch, i64, i64, i64, vp, i64, vp, i64, vp, i64, vp, i64],
    "geqrt": [i64, i64, i64, vp, i64, vp, i64, vp, vp],
    "unmqr": [ch, ch, i64, i64, i64, i64, vp, i64, i64, vp, i64, vp, i64, vp],
    "pamm": [ch, ch, ch, ch, i64, i64, i64, i64, vp, i64, vp, i64, vp, i64, vp, i64],
    "parfb": [ch, ch, ch, ch, i64, i64, i64, i64, i64, i64, vp, i64, vp, i64, vp, i64, vp, i64, vp, i64],
    "tsmqr": [ch, ch, i64, i64, i64, i64, i64, i64, vp, i64, vp, i64, vp, i64, i64, vp, i64, i64, vp],
    "ttmqr": [ch, ch, i64, i64, i64, i64, i64, i64, vp, i64, vp, i64, vp, i64, vp, i64, vp],
    "tsqrt": [i64, i64, i64, vp, i64, i64, vp, i64, vp, i64, vp, vp],
    "ttqrt": [i64, i64, i64, vp, i64, vp, i64, vp, i64, vp, vp],
    "pemv": [ch, ch, i64, i64, i64, vp, vp, i64, vp, vp, vp],
    "laswp": [i64, vp, i64, i64, i64, vp, i64],
    "getrf2": [i64, i64, vp, i64, vp],
}
_done = False


def lib():
    global _done
    L = _base_lib()
    if not _done:
        for p in "dszc":
            for name, sig in _SIGS.items():
                f = getattr(L, f"nlaq_{p}{name}")
                f.argtypes = sig
                f.restype = i64 if name == "getrf2" else None
        _done = True
    return L


def _f(a, name):
    return getattr(lib(), f"nlaq_{_PRE[a.dtype.type]}{name}")


def _work(dt, n):
    return np.zeros(max(int(n), 1), dtype=dt)


def larfg(alpha, x):
    """larfg! (src/larfg.jl:43-106): returns (beta, tau); x is overwritten with v[2:n]."""
    a = np.array([alpha], dtype=x.dtype)
    t = np.zeros(1, dtype=x.dtype)
    _f(x, "larfg")(x.size + 1, _ptr(a), _ptr(x), 1, _ptr(t))
    return a[0], t[0]


def geqr2(A):
    m, n = A.shape
    tau = np.zeros(max(min(m, n), 1), dtype=A.dtype)
    _f(A, "geqr2")(m, n, _ptr(A), _ld(A), _ptr(tau), _ptr(_work(A.dtype, n)))
    return tau[: min(m, n)]


def larft(storev, V, tau, T):
    """larft!('F', storev, n, k, V, tau, T) (src/larft.jl:48-126)."""
    n, k = V.shape if storev == "C" else V.shape[::-1]
    _f(V, "larft")(_c(storev), n, k, _ptr(V), _ld(V), _ptr(tau), _ptr(T), _ld(T))


def larfb(side, trans, V, T, C, k=None):
    m, n = C.shape
    k = T.shape[0] if k is None else k
    W = np.zeros((n if side == "L" else m, k), dtype=C.dtype, order="F")
    _f(C, "larfb")(_c(side), _c(trans), m, n, k, _ptr(V), _ld(V), _ptr(T), _ld(T), _ptr(C), _ld(C), _ptr(W), _ld(W))


def geqrt(A, ib):
    m, n = A.shape
    T = np.zeros((ib, n), dtype=A.dtype, order="F")
    tau = np.zeros(max(min(m, n), 1), dtype=A.dtype)
    _f(A, "geqrt")(m, n, ib, _ptr(A), _ld(A), _ptr(T), _ld(T), _ptr(tau), _ptr(_work(A.dtype, ib * n)))
    return T, tau[: min(m, n)]


def unmqr(side, trans, A, T, C, k=None):
    m, n = C.shape
    ib, kt = T.shape
    k = kt if k is None else k
    _f(C, "unmqr")(_c(side), _c(trans), m, n, k, ib, _ptr(A), _ld(A), A.shape[1], _ptr(T), _ld(T), _ptr(C), _ld(C),
                   _ptr(_work(C.dtype, max(m, n) * ib)))


def pamm(op, side, storev, direct, m, n, k, l, A1, A2, V, W):
    _f(A2, "pamm")(_c(op), _c(side), _c(storev), _c(direct), m, n, k, l, _ptr(A1), _ld(A1), _ptr(A2), _ld(A2),
                   _ptr(V), _ld(V), _ptr(W), _ld(W))


def parfb(side, trans, direct, storev, m1, n1, m2, n2, k, l, A1, A2, V, T):
    W = np.zeros((k, n2) if side == "L" else (m2, k), dtype=A1.dtype, order="F")
    _f(A1, "parfb")(_c(side), _c(trans), _c(direct), _c(storev), m1, n1, m2, n2, k, l, _ptr(A1), _ld(A1),
                    _ptr(A2), _ld(A2), _ptr(V), _ld(V), _ptr(T), _ld(T), _ptr(W), _ld(W))


def tsmqr(side, trans, A1, A2, V, T, k=None, ib=None):
    m1, n1 = A1.shape
    m2, n2 = A2.shape
    ibt, kt = T.shape
    ib = ibt if ib is None else ib
    k = kt if k is None else k
    _f(A1, "tsmqr")(_c(side), _c(trans), m1, n1, m2, n2, k, ib, _ptr(A1), _ld(A1), _ptr(A2), _ld(A2), _ptr(V),
                    _ld(V), V.shape[1], _ptr(T), _ld(T), T.shape[1], _ptr(_work(A1.dtype, ib * max(m1, n1))))


def ttmqr(side, trans, A1, A2, V, T, k=None, ib=None):
    m1, n1 = A1.shape
    m2, n2 = A2.shape
    ibt, kt = T.shape
    ib = ibt if ib is None else ib
    k = kt if k is None else k
    _f(A1, "ttmqr")(_c(side), _c(trans), m1, n1, m2, n2, k, ib, _ptr(A1), _ld(A1), _ptr(A2), _ld(A2), _ptr(V),
                    _ld(V), _ptr(T), _ld(T), _ptr(_work(A1.dtype, ib * max(m1, n1))))


def tsqrt(A1, A2, T, ib=None):
    m1, n = A1.shape
    m = A2.shape[0]
    ib = T.shape[0] if ib is None else ib
    tau = np.zeros(max(min(m1, n), 1), dtype=A1.dtype)
    _f(A1, "tsqrt")(m, n, ib, _ptr(A1), _ld(A1), m1, _ptr(A2), _ld(A2), _ptr(T), _ld(T), _ptr(tau),
                    _ptr(_work(A1.dtype, ib * n)))
    return tau


def ttqrt(A1, A2, T, ib=None):
    m, n = A2.shape
    ib = T.shape[0] if ib is None else ib
    tau = np.zeros(max(min(m, n), 1), dtype=A1.dtype)
    _f(A1, "ttqrt")(m, n, ib, _ptr(A1), _ld(A1), _ptr(A2), _ld(A2), _ptr(T), _ld(T), _ptr(tau),
                    _ptr(_work(A1.dtype, ib * n)))
    return tau


def pemv(trans, storev, l, alpha, A, x, beta, y):
    m, n = A.shape
    a = np.array([alpha], dtype=A.dtype)
    b = np.array([beta], dtype=A.dtype)
    _f(A, "pemv")(_c(trans), _c(storev), m, n, l, _ptr(a), _ptr(A), _ld(A), _ptr(x), _ptr(b), _ptr(y))


def laswp(A, k1, k2, ipiv, incx=1):
    """laswp(A, first, last, ipiv, incx) (src/lu.jl:470-522); ipiv 1-based int64."""
    ip = np.ascontiguousarray(ipiv, dtype=np.int64)
    _f(A, "laswp")(A.shape[1], _ptr(A), _ld(A), k1, k2, ip.ctypes.data_as(vp), incx)


def getrf2(A):
    """getrf2! (src/lu.jl:185-300): A overwritten with L and U; returns (ipiv 1-based, info)."""
    m, n = A.shape
    ipiv = np.zeros(max(min(m, n), 1), dtype=np.int64)
    info = _f(A, "getrf2")(m, n, _ptr(A), _ld(A), ipiv.ctypes.data_as(vp))
    return ipiv[: min(m, n)], int(info)
