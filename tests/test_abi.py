"""CPU checks of the C ABI: the library loads, exports every symbol include/nextla_hip.h
declares, and rejects illegal arguments with the reference's ArgumentError messages.
No call here reaches the GPU (validation and quick returns happen before any HIP call)."""
import ctypes

import pytest

import nextla_hip as nl
from nextla_hip import _lib as L

NULL = None


def test_library_exports_every_header_symbol():
    lib = nl.load()
    syms = nl.header_symbols()
    assert len(syms) >= 34
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    assert lib.nla_version() >= 100


def st(name, *args):
    lib = nl.load()
    return getattr(lib, name)(*args), lib.nla_last_error().decode()


def test_geqrt_argument_errors():
    # src/geqrt.jl:24-34
    assert st("nla_dgeqrt", -1, 4, 2, NULL, 4, NULL, 2, NULL, NULL, NULL) == (-1, "illegal value of m: -1")
    assert st("nla_dgeqrt", 4, -3, 2, NULL, 4, NULL, 2, NULL, NULL, NULL) == (-2, "illegal value of n: -3")
    assert st("nla_dgeqrt", 4, 4, 0, NULL, 4, NULL, 2, NULL, NULL, NULL) == (-3, "illegal value of ib: 0")
    assert st("nla_sgeqrt", 4, 4, -1, NULL, 4, NULL, 2, NULL, NULL, NULL) == (-3, "illegal value of ib: -1")
    # quick returns: m == 0, or ib == 0 with m*n == 0
    assert st("nla_dgeqrt", 0, 4, 0, NULL, 1, NULL, 1, NULL, NULL, NULL)[0] == 0


def test_tsmqr_argument_errors():
    f = "nla_dtsmqr"
    args = lambda side, trans, m1, n1, m2, n2, k, ib: (side, trans, m1, n1, m2, n2, k, ib, NULL, 32, NULL, 32,
                                                         NULL, 32, 16, NULL, 8, 16, NULL, NULL)
    assert st(f, *args(b"X", b"N", 32, 24, 20, 24, 16, 8)) == (-1, "side must be 'L' or 'R', got 'X'")
    assert st(f, *args(b"L", b"Q", 32, 24, 20, 24, 16, 8)) == (-2, "trans must be 'N', 'C', or 'T', got 'Q'")
    assert st(f, *args(b"L", b"N", -1, 24, 20, 24, 16, 8)) == (-3, "m1 must be non-negative, got -1")
    assert st(f, *args(b"L", b"N", 32, 24, 20, 23, 16, 8)) == \
        (-6, "For side='L', n2 must equal n1. Got n1=24, n2=23")
    assert st(f, *args(b"R", b"N", 24, 32, 20, 20, 16, 8)) == \
        (-5, "For side='R', m2 must equal m1. Got m1=24, m2=20")
    assert st(f, *args(b"L", b"N", 32, 24, 20, 24, -2, 8)) == (-7, "k must be non-negative, got -2")
    assert st(f, *args(b"L", b"N", 32, 24, 20, 24, 16, -8)) == (-8, "ib must be non-negative, got -8")
    # k = 0 is a no-op (test/tsmqr.jl:75-85)
    assert st(f, *args(b"L", b"N", 32, 24, 20, 24, 0, 8))[0] == 0


def test_unmqr_and_tsqrt_errors():
    f = "nla_dunmqr"
    assert st(f, b"Z", b"N", 4, 4, 4, 2, NULL, 4, 4, NULL, 2, NULL, 4, NULL, NULL) == \
        (-1, "side must be 'L' or 'R', got 'Z'")
    assert st(f, b"L", b"N", 4, 4, -1, 2, NULL, 4, 4, NULL, 2, NULL, 4, NULL, NULL) == \
        (-5, "k must be non-negative, got -1")
    assert st(f, b"L", b"N", 8, 4, 4, 2, NULL, 4, 4, NULL, 2, NULL, 8, NULL, NULL) == \
        (-8, "lda must be >= max(1,8), got 4")
    assert st(f, b"L", b"N", 8, 4, 0, 2, NULL, 8, 4, NULL, 2, NULL, 8, NULL, NULL)[0] == 0  # k = 0 no-op
    assert st("nla_dtsqrt", -1, 4, 2, NULL, 4, 4, NULL, 4, NULL, 2, NULL, NULL, NULL) == \
        (-1, "m must be non-negative, got -1")
    assert st("nla_stsqrt", 4, 4, -2, NULL, 4, 4, NULL, 4, NULL, 2, NULL, NULL, NULL) == \
        (-3, "ib must be non-negative, got -2")


def test_larfb_parfb_rectrxm_errors():
    # larfb! quick-returns before checking flags (src/larfb.jl:38-44)
    assert st("nla_dlarfb", b"Q", b"N", b"F", b"C", 0, 4, 2, NULL, 4, NULL, 2, NULL, 4, NULL, 0, NULL)[0] == 0
    assert st("nla_dlarfb", b"L", b"N", b"B", b"C", 4, 4, 2, NULL, 4, NULL, 2, NULL, 4, NULL, 0, NULL) == \
        (-3, "larfb! currently supports only forward, columnwise reflectors")
    assert st("nla_dparfb", b"L", b"N", b"F", b"C", 4, 4, 4, 5, 2, 0, NULL, 4, NULL, 4, NULL, 4, NULL, 2, NULL, 0,
              NULL) == (-8, "illegal value of n2")
    assert st("nla_dparfb", b"L", b"N", b"F", b"C", 4, 4, 4, 4, 2, 3, NULL, 4, NULL, 4, NULL, 4, NULL, 2, NULL, 0,
              NULL) == (-10, "illegal value of l")
    r = st("nla_drectrxm", b"X", b"L", b"N", ctypes.c_double(1.0), b"S", 4, 4, NULL, 4, NULL, 4, 0, NULL)
    assert r == (-1, "side must be 'L' or 'R', got 'X'")
    r = st("nla_srectrxm", b"L", b"L", b"N", ctypes.c_float(1.0), b"S", 8, 4, NULL, 8, NULL, 8, 2048, NULL)
    assert r[0] == -12


def test_python_mirror_raises_argument_error():
    torch = pytest.importorskip("torch")
    A = torch.zeros(4, 4, dtype=torch.float64)  # host tensor: rejected before any call
    with pytest.raises(nl.ArgumentError):
        nl.tsmqr("L", "N", A, A, A, A)


def test_batched_autotune_rejects_bad_candidates():
    # a candidate outside 1..64 would never advance the block loop (ib <= 0) or overrun the
    # 64-row T blocks (ib > 64): rejected on the host before any allocation or launch
    I64 = ctypes.c_int64
    ms, ns = (I64 * 1)(512), (I64 * 1)(64)
    out = (I64 * 1)()
    for bad in (0, -4, 65):
        cand = (I64 * 2)(16, bad)
        for f in ("nla_dbatched_autotune", "nla_sbatched_autotune"):
            r = st(f, 1, ms, ns, 4, cand, 2, out, NULL, NULL)
            assert r == (-5, f"candidate ib must be in 1..64, got {bad}"), r
    cand = (I64 * 1)(16)
    r = st("nla_sbatched_autotune", 1, ms, (I64 * 1)(600), 4, cand, 1, out, NULL, NULL)
    assert r == (-2, "class 0: need 0 < n <= m")
    assert st("nla_sbatched_autotune", 1, ms, ns, 0, cand, 1, out, NULL, NULL)[0] == -4


def test_python_mirror_shape_checks():
    torch = pytest.importorskip("torch")
    z = lambda m, n: torch.zeros(m, n, dtype=torch.float64)
    with pytest.raises(nl.ArgumentError, match="inner dimensions"):
        nl.GEMM_ADD_(z(4, 3), z(2, 5), z(4, 5))
    with pytest.raises(nl.ArgumentError, match="result is"):
        nl.GEMM_ADD_(z(4, 3), z(3, 5), z(4, 6))
    with pytest.raises(nl.ArgumentError, match="inner dimensions"):
        nl.GEMM_SUB_(z(4, 5), z(4, 3), z(2, 5))
    with pytest.raises(nl.ArgumentError, match="at least m"):
        nl.tile_gels(z(256, 128), torch.zeros(100, dtype=torch.float64))


def test_noalloc_entry_rejects_oversize_on_host():
    # the size check runs before any HIP call: n = 16384 with the 256-row bases needs 36 MiB
    r = st("nla_drectrxm_noalloc", b"L", b"L", b"N", ctypes.c_double(1.0), b"S", 16384, 4, NULL, 16384, NULL,
           16384, 0, NULL)
    assert r[0] == -6 and "37748736 bytes" in r[1] and "holds 18874368" in r[1], r


def test_mixed_solve_rejects_unknown_update_precision():
    # the precision check runs before any HIP call
    r = st("nla_drectrsm_mixed", b"L", b"L", b"N", ctypes.c_double(1.0), 64, 8, NULL, 64, NULL, 64, 0, 3,
           ctypes.c_double(1e-15), b"X", NULL, NULL, NULL)
    assert r == (-14, "prec must be 'S', 'H' or 'B', got 'X'"), r
