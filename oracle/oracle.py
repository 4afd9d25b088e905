"""CPU oracle: ctypes front end of oracle/nla_oracle.c.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker; the product path never touches it.

Every routine works in place on Fortran-ordered (column-major) numpy arrays of
float64 ('d') or float32 ('s') and follows the NextLA.jl routine named in its
docstring (file:line citations are in nla_oracle_t.h).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "libnla_oracle.so")
_lib = None

i64 = ctypes.c_int64
ch = ctypes.c_char
vp = ctypes.c_void_p


def build(force=False):
    """Compile the oracle with gcc (oracle/Makefile)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []), check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        for p, real in (("d", ctypes.c_double), ("s", ctypes.c_float)):
            L = _lib
            getattr(L, f"nlao_{p}lamch").restype = real
            getattr(L, f"nlao_{p}lamch").argtypes = [ch]
            getattr(L, f"nlao_{p}larfg").argtypes = [i64, vp, vp, i64, vp]
            getattr(L, f"nlao_{p}geqr2").argtypes = [i64, i64, vp, i64, vp, vp]
            getattr(L, f"nlao_{p}larft_fc").argtypes = [i64, i64, vp, i64, vp, vp, i64]
            getattr(L, f"nlao_{p}larfb").argtypes = [ch, ch, i64, i64, i64, vp, i64, vp, i64, vp, i64, vp, i64]
            getattr(L, f"nlao_{p}geqrt").argtypes = [i64, i64, i64, vp, i64, vp, i64, vp, vp]
            getattr(L, f"nlao_{p}unmqr").argtypes = [ch, ch, i64, i64, i64, i64, vp, i64, i64, vp, i64, vp, i64, vp]
            getattr(L, f"nlao_{p}parfb").argtypes = [ch, ch, i64, i64, i64, i64, i64, vp, i64, vp, i64, vp, i64,
                                                      vp, i64, vp, i64]
            getattr(L, f"nlao_{p}tsmqr").argtypes = [ch, ch, i64, i64, i64, i64, i64, i64, vp, i64, vp, i64, vp,
                                                      i64, i64, vp, i64, i64, vp]
            getattr(L, f"nlao_{p}tsqrt").argtypes = [i64, i64, i64, vp, i64, i64, vp, i64, vp, i64, vp, vp]
            getattr(L, f"nlao_{p}rectrxm").argtypes = [ch, ch, ch, real, ch, i64, i64, vp, i64, vp, i64, i64]
            getattr(L, f"nlao_{p}tile_geqrf").argtypes = [i64, i64, i64, i64, vp, i64, vp, i64]
            getattr(L, f"nlao_{p}tile_ormqr").argtypes = [ch, i64, i64, i64, i64, i64, vp, i64, vp, i64, vp, i64]
        _lib.nlao_num_threads.restype = ctypes.c_int
    return _lib


def _p(a):
    if a.dtype == np.float64:
        return "d"
    if a.dtype == np.float32:
        return "s"
    raise TypeError(f"unsupported dtype {a.dtype}")


def _ptr(a):
    assert a.flags.f_contiguous or a.ndim == 1 or a.size == 0, "oracle arrays must be column-major"
    return a.ctypes.data_as(vp)


def _ld(a):
    return max(1, a.shape[0]) if a.ndim == 2 else max(1, a.size)


def _c(x):
    return x.encode() if isinstance(x, str) else x


def lamch(dtype, cmach):
    p = "d" if np.dtype(dtype) == np.float64 else "s"
    return getattr(lib(), f"nlao_{p}lamch")(_c(cmach))


def larfg(alpha, x):
    """larfg! (src/larfg.jl:43-106): returns (beta, tau); x is overwritten with v[2:n]."""
    p = _p(x)
    a = np.array([alpha], dtype=x.dtype)
    t = np.zeros(1, dtype=x.dtype)
    getattr(lib(), f"nlao_{p}larfg")(x.size + 1, _ptr(a), _ptr(x), 1, _ptr(t))
    return a[0], t[0]


def geqr2(A, tau):
    """geqr2! (src/geqr2.jl:107-121)."""
    m, n = A.shape
    work = np.zeros(max(n, 1), dtype=A.dtype)
    getattr(lib(), f"nlao_{_p(A)}geqr2")(m, n, _ptr(A), _ld(A), _ptr(tau), _ptr(work))


def larft(V, tau, T):
    """larft!('F','C') (src/larft.jl:205-213)."""
    n, k = V.shape
    getattr(lib(), f"nlao_{_p(V)}larft_fc")(n, k, _ptr(V), _ld(V), _ptr(tau), _ptr(T), _ld(T))


def larfb(side, trans, V, T, C):
    """larfb!(side, trans, 'F', 'C', ...) (src/larfb.jl:37-142); k = size(T,1)."""
    m, n = C.shape
    k = T.shape[0]
    W = np.zeros((n if side == "L" else m, k), dtype=C.dtype, order="F")
    getattr(lib(), f"nlao_{_p(C)}larfb")(_c(side), _c(trans), m, n, k, _ptr(V), _ld(V), _ptr(T), _ld(T),
                                         _ptr(C), _ld(C), _ptr(W), _ld(W))


def geqrt(A, T, ib=None):
    """geqrt! (src/geqrt.jl:22-68); ib defaults to size(T,1) as in src/geqrt.jl:84-91."""
    m, n = A.shape
    ib = T.shape[0] if ib is None else ib
    tau = np.zeros(max(min(m, n), 1), dtype=A.dtype)
    work = np.zeros(max(ib * n, 1), dtype=A.dtype)
    getattr(lib(), f"nlao_{_p(A)}geqrt")(m, n, ib, _ptr(A), _ld(A), _ptr(T), _ld(T), _ptr(tau), _ptr(work))
    return tau


def unmqr(side, trans, A, T, C, k=None):
    """unmqr! (src/unmqr.jl:52-149) with ib, k = size(T)."""
    m, n = C.shape
    ib, kt = T.shape
    k = kt if k is None else k
    work = np.zeros(max((n if side == "L" else m) * ib, 1), dtype=C.dtype)
    getattr(lib(), f"nlao_{_p(C)}unmqr")(_c(side), _c(trans), m, n, k, ib, _ptr(A), _ld(A), A.shape[1],
                                         _ptr(T), _ld(T), _ptr(C), _ld(C), _ptr(work))


def parfb(side, trans, A1, A2, V, T):
    """parfb!(side, trans, 'F', 'C', ..., l=0) (src/parfb.jl:44-191); k = size(T,1)."""
    m1, n1 = A1.shape
    m2, n2 = A2.shape
    k = T.shape[0]
    W = np.zeros((k, n2) if side == "L" else (m2, k), dtype=A1.dtype, order="F")
    getattr(lib(), f"nlao_{_p(A1)}parfb")(_c(side), _c(trans), m1, n1, m2, n2, k, _ptr(A1), _ld(A1),
                                          _ptr(A2), _ld(A2), _ptr(V), _ld(V), _ptr(T), _ld(T), _ptr(W), _ld(W))


def tsmqr(side, trans, A1, A2, V, T, k=None):
    """tsmqr! (src/tsmqr.jl:47-146); ib, k = size(T) as in src/tsmqr.jl:196-228."""
    m1, n1 = A1.shape
    m2, n2 = A2.shape
    ib, kt = T.shape
    k = kt if k is None else k
    work = np.zeros(max(ib * max(n1, n2), m1 * ib, 1), dtype=A1.dtype)
    getattr(lib(), f"nlao_{_p(A1)}tsmqr")(_c(side), _c(trans), m1, n1, m2, n2, k, ib, _ptr(A1), _ld(A1),
                                          _ptr(A2), _ld(A2), _ptr(V), _ld(V), V.shape[1], _ptr(T), _ld(T),
                                          T.shape[1], _ptr(work))


def tsqrt(A1, A2, T, ib=None):
    """tsqrt! (src/tsqrt.jl:128-156); ib defaults to size(T,1) (src/tsqrt.jl:196-211)."""
    m1, n = A1.shape
    m = A2.shape[0]
    ib = T.shape[0] if ib is None else ib
    tau = np.zeros(max(min(m1, n), 1), dtype=A1.dtype)
    work = np.zeros(max(ib * n, 1), dtype=A1.dtype)
    getattr(lib(), f"nlao_{_p(A1)}tsqrt")(m, n, ib, _ptr(A1), _ld(A1), m1, _ptr(A2), _ld(A2), _ptr(T), _ld(T),
                                          _ptr(tau), _ptr(work))
    return tau


def rectrxm(side, uplo, trans, alpha, func, A, B, threshold=0):
    """unified_rectrxm! (src/rectrxm.jl:43-76); threshold 0 = reference default (256 / 16)."""
    n = A.shape[0]
    m = B.shape[1] if side == "L" else B.shape[0]
    getattr(lib(), f"nlao_{_p(B)}rectrxm")(_c(side), _c(uplo), _c(trans), float(alpha), _c(func), n, m,
                                           _ptr(A), _ld(A), _ptr(B), _ld(B), int(threshold))


def tile_geqrf(A, nb, ib):
    """Flat-tree tile QR (geqrt/unmqr/tsqrt/tsmqr in PLASMA order).  Returns T of shape (mt*ib, n)."""
    m, n = A.shape
    assert m % nb == 0 and n % nb == 0
    T = np.zeros(((m // nb) * ib, n), dtype=A.dtype, order="F")
    getattr(lib(), f"nlao_{_p(A)}tile_geqrf")(m, n, nb, ib, _ptr(A), _ld(A), _ptr(T), _ld(T))
    return T


def tile_ormqr(trans, A, T, C, nb, ib):
    """Apply Q ('N') or Q^T ('T') of tile_geqrf to C from the left."""
    m, n = A.shape
    getattr(lib(), f"nlao_{_p(A)}tile_ormqr")(_c(trans), m, n, C.shape[1], nb, ib, _ptr(A), _ld(A), _ptr(T),
                                              _ld(T), _ptr(C), _ld(C))


def num_threads():
    return lib().nlao_num_threads()
