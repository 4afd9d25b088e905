"""GPU parity of the recursive TRSM/TRMM (unified_rectrxm!), its base kernels and the
GEMM updates, against BLAS golden vectors (test/unified_rectrxm.jl, test/trsm.jl) and the
CPU oracle; large sizes are checked through backward error and cut-off independence."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import nextla_hip as nl  # noqa: E402
import oracle as O  # noqa: E402
from conftest import relerr  # noqa: E402
from golden.gen_golden import TRXM_CASES, tri_matrix  # noqa: E402

pytestmark = pytest.mark.gpu
DTYPES = [np.float64, np.float32]
P = {np.float64: "d", np.float32: "s"}


def F(a):
    return np.asfortranarray(a)


def host(t):
    torch.cuda.synchronize()
    return nl.to_host(t)


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("case", TRXM_CASES)
def test_rectrxm_golden_and_oracle(golden, dt, case):
    g = golden("trxm")
    key = f"{P[dt]}_{TRXM_CASES.index(case)}"
    func, side, uplo, trans, n, m, alpha, cut = case
    dB = nl.to_device(g[key + "_B"])
    nl.unified_rectrxm_(side, uplo, trans, alpha, func, nl.to_device(g[key + "_A"]), dB, cutoff=cut)
    B = g[key + "_B"].copy(order="F")
    O.rectrxm(side, uplo, trans, alpha, func, g[key + "_A"], B, threshold=cut)
    X = host(dB)
    # test/unified_rectrxm.jl:8,42-43 (1e-14 relative error for Float64 at n <= 256)
    tol = 1e-12 if dt == np.float64 else 1e-4
    assert relerr(X, g[key + "_X"]) < tol
    assert relerr(X, B) < tol


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("func", "SM")
@pytest.mark.parametrize("side", "LR")
@pytest.mark.parametrize("uplo", "LU")
def test_base_kernels_vs_oracle(dt, func, side, uplo):
    # LeftLowerTRSM! ... RightUpperTRMM! (src/trsm.jl:128-150, src/trmm.jl:332-389), test/trsm.jl
    rng = np.random.default_rng(hash((func, side, uplo)) % 1000)
    n, m = 200, 45
    A = tri_matrix(rng, n, uplo, dt)
    B0 = F(rng.standard_normal((n, m) if side == "L" else (m, n)).astype(dt))
    dB = nl.to_device(B0)
    fn = getattr(nl, f"{'Left' if side == 'L' else 'Right'}{'Lower' if uplo == 'L' else 'Upper'}"
                     f"{'TRSM' if func == 'S' else 'TRMM'}_")
    fn(nl.to_device(A), dB)
    B = B0.copy(order="F")
    O.rectrxm(side, uplo, "N", 1.0, func, A, B, threshold=1024)
    assert relerr(host(dB), B) < (1e-12 if dt == np.float64 else 1e-4)


@pytest.mark.parametrize("dt", DTYPES)
def test_gemm_add_sub(dt):
    rng = np.random.default_rng(1)
    A = F(rng.standard_normal((130, 70)).astype(dt))
    B = F(rng.standard_normal((70, 45)).astype(dt))
    C = F(rng.standard_normal((130, 45)).astype(dt))
    dC = nl.to_device(C)
    nl.GEMM_ADD_(nl.to_device(A), nl.to_device(B), dC)   # C += A B
    want = C.astype(np.float64) + A.astype(np.float64) @ B
    assert relerr(host(dC), want) < (1e-13 if dt == np.float64 else 1e-5)
    dA2 = nl.to_device(C)
    nl.GEMM_SUB_(dA2, nl.to_device(A), nl.to_device(B))  # A2 -= A B
    assert relerr(host(dA2), C.astype(np.float64) - A.astype(np.float64) @ B) < (1e-13 if dt == np.float64 else 1e-5)


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("cut", [64, 128, 256, 512])
def test_trsm_config_backward_error(dt, cut):
    # TRSM configuration at reduced size: A n x n lower, off-diag N(0,1)/sqrt(n), diag U[1,2];
    # property checks: backward error and agreement across cut-offs
    rng = np.random.Generator(np.random.PCG64(2))
    n, m = 2048, 256
    A = np.tril(rng.standard_normal((n, n)) / np.sqrt(n), -1)
    A[np.diag_indices(n)] = rng.uniform(1, 2, n)
    A = F(A.astype(dt))
    B0 = F(rng.standard_normal((n, m)).astype(dt))
    dA, dB = nl.to_device(A), nl.to_device(B0)
    nl.unified_rectrxm_("L", "L", "N", 1.0, "S", dA, dB, cutoff=cut)
    X = host(dB).astype(np.float64)
    A64 = np.tril(A.astype(np.float64))
    be = np.linalg.norm(A64 @ X - B0) / (np.linalg.norm(A64) * np.linalg.norm(X))
    assert be < 10 * np.finfo(dt).eps
    # TRMM round trip: A * X == B
    dX = nl.to_device(X.astype(dt))
    nl.unified_rectrxm_("L", "L", "N", 1.0, "M", dA, dX, cutoff=min(cut, 64))
    assert relerr(host(dX), B0) < (1e-12 if dt == np.float64 else 1e-4)


def test_rectrxm_mid_split_non_power_of_two():
    # src/rectrxm.jl:129-134: mid = 2^floor(log2 n) for n not a power of two
    rng = np.random.default_rng(6)
    for n in (300, 777, 1000):
        A = tri_matrix(rng, n, "U", np.float64)
        B0 = F(rng.standard_normal((n, 9)))
        dB = nl.to_device(B0)
        nl.unified_rectrxm_("L", "U", "N", 0.5, "S", nl.to_device(A), dB, cutoff=64)
        B = B0.copy(order="F")
        O.rectrxm("L", "U", "N", 0.5, "S", A, B, threshold=64)
        assert relerr(host(dB), B) < 1e-12


def test_nextlamatrix_trsm():
    rng = np.random.default_rng(10)
    n = 96
    A = tri_matrix(rng, n, "L", np.float64)
    B = nl.NextLAMatrix(rng.standard_normal((n, 7)))
    B0 = B.data.copy()
    nl.NextLAMatrix.trsm("L", "L", "N", "N", 1.0, nl.NextLAMatrix(A), B)
    assert relerr(np.tril(A) @ B.data, B0) < 1e-13
    assert B.size() == (n, 7) and len(B) == n * 7


@pytest.mark.parametrize("side", "LR")
@pytest.mark.parametrize("uplo", "LU")
@pytest.mark.parametrize("trans", "NT")
@pytest.mark.parametrize("func", "SM")
def test_rectrxm_interior_tiles_vs_oracle(side, uplo, trans, func):
    # n = 1024 with cut-off 128: every GEMM update takes the unchecked interior-tile path,
    # in all four transpose combinations of its operands
    rng = np.random.default_rng(abs(hash((side, uplo, trans, func))) % 997)
    n, m = 1024, 128
    A = tri_matrix(rng, n, uplo, np.float64)
    B0 = F(rng.standard_normal((n, m) if side == "L" else (m, n)))
    dB = nl.to_device(B0)
    nl.unified_rectrxm_(side, uplo, trans, 1.0, func, nl.to_device(A), dB, cutoff=128)
    B = B0.copy(order="F")
    O.rectrxm(side, uplo, trans, 1.0, func, A, B, threshold=128)
    assert relerr(host(dB), B) < 1e-12


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("n,m", [(512, 64), (768, 37), (1280, 50)])
def test_trsm_fused_two_block_base_vs_oracle(dt, n, m):
    """Cut-off 256, left / lower / N: every aligned 512-row block of the recursion runs as one
    launch (inverse-apply, the 256 x 256 update, inverse-apply).  n = 768 and 1280 put fused
    blocks at offsets 0, 512 and beside 256-row base blocks; m not a multiple of 16 leaves a
    partial workgroup of right-hand sides."""
    rng = np.random.default_rng(n + m)
    A = tri_matrix(rng, n, "L", dt)
    B0 = F(rng.standard_normal((n, m)).astype(dt))
    dB = nl.to_device(B0)
    nl.unified_rectrxm_("L", "L", "N", 1.5, "S", nl.to_device(A), dB, cutoff=256)
    B = B0.copy(order="F")
    O.rectrxm("L", "L", "N", 1.5, "S", A, B, threshold=256)
    tol = 1e-12 if dt == np.float64 else 1e-4
    assert relerr(host(dB), B) < tol


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("side", "LR")
@pytest.mark.parametrize("uplo", "LU")
@pytest.mark.parametrize("func", "SM")
def test_host_pointer_rectrxm(dt, side, uplo, func):
    """nla_h?rectrxm (NextLAMatrix's host-matrix path): host A and B with leading dimensions
    larger than their rows, against the oracle."""
    import ctypes

    rng = np.random.default_rng(hash((side, uplo, func)) % 991)
    n, m = 300, 37
    Abig = F(np.zeros((n + 5, n), dtype=dt))
    Abig[:n, :] = tri_matrix(rng, n, uplo, dt)
    rows, cols = (n, m) if side == "L" else (m, n)
    Bbig = F(rng.standard_normal((rows + 3, cols)).astype(dt))
    B0 = Bbig[:rows, :].copy(order="F")
    lib = nl.load()
    st = getattr(lib, f"nla_h{P[dt]}rectrxm")(side.encode(), uplo.encode(), b"T", 0.5, func.encode(), n, m,
                                             Abig.ctypes.data, n + 5, Bbig.ctypes.data, rows + 3, 64,
                                             ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert st == 0, lib.nla_last_error()
    want = B0.copy(order="F")
    O.rectrxm(side, uplo, "T", 0.5, func, F(Abig[:n, :]), want, threshold=64)
    assert relerr(Bbig[:rows, :], want) < (1e-12 if dt == np.float64 else 1e-4)
    # NextLAMatrix's 4-Char forward with a unit diagonal
    Bm = nl.NextLAMatrix(B0.copy())
    Au = F(Abig[:n, :].copy())
    getattr(nl.NextLAMatrix, "trsm" if func == "S" else "trmm")(side, uplo, "N", "U", 1.0, nl.NextLAMatrix(Au), Bm)
    np.fill_diagonal(Au, 1.0)
    want2 = B0.copy(order="F")
    O.rectrxm(side, uplo, "N", 1.0, func, Au, want2)
    assert relerr(Bm.data, want2) < (1e-12 if dt == np.float64 else 1e-4)


def _kahan_lower(k, theta):
    # transpose of Kahan's upper triangular matrix: diag sin(theta)^i, below -cos(theta) sin(theta)^j
    s, c = np.sin(theta), np.cos(theta)
    K = np.triu(-c * np.ones((k, k)), 1) + np.eye(k)
    return (np.diag(s ** np.arange(k)) @ K).T


def _ill_conditioned_lower(kind, n, rng):
    A = np.tril(rng.standard_normal((n, n)) / np.sqrt(n), -1)
    A[np.diag_indices(n)] = rng.uniform(1, 2, n)
    if kind == "graded_rows":    # row scaling 1 .. 1e-10 inside every 256-row block
        A *= (10.0 ** (-10.0 * ((np.arange(n) % 256) / 256)))[:, None]
    elif kind == "graded_cols":  # column scaling 1 .. 1e-10 inside every 256-column block
        A *= (10.0 ** (-10.0 * ((np.arange(n) % 256) / 256)))[None, :]
    else:                        # Kahan diagonal blocks (cond of a block ~1e14 / ~1e20)
        theta = 1.45 if kind == "kahan_1e14" else 1.3
        for b0 in range(0, n, 256):
            A[b0:b0 + 256, b0:b0 + 256] = _kahan_lower(256, theta)
    return A


@pytest.mark.parametrize("kind", ["graded_rows", "graded_cols", "kahan_1e14", "kahan_1e20"])
@pytest.mark.parametrize("cut", [128, 256, 512])
def test_trsm_ill_conditioned_bases(kind, cut):
    """The 256-row bases apply explicitly inverted diagonal blocks (cut-offs 256 and 512; at 512
    also the fused two-block base) where the reference and the oracle substitute; cut-off 128
    substitutes.  On ill-conditioned diagonal blocks (graded rows / columns down to 1e-10, Kahan
    blocks of condition ~1e14 and ~1e20) the solve must stay backward stable: backward error
    below 10 u (a CPU model of the inverse-apply gives at most 6e-18 on these classes, within
    10x of substitution), and the solution must agree with the oracle's substitution to 1e-11
    relative (the model: at most 5e-15)."""
    rng = np.random.Generator(np.random.PCG64(11))
    n, m = 1024, 32
    A = F(_ill_conditioned_lower(kind, n, rng))
    B0 = F(rng.standard_normal((n, m)))
    dA, dB = nl.to_device(A), nl.to_device(B0)
    nl.unified_rectrxm_("L", "L", "N", 1.0, "S", dA, dB, cutoff=cut)
    X = host(dB)
    be = np.linalg.norm(A @ X - B0) / (np.linalg.norm(A) * np.linalg.norm(X))
    assert be < 10 * np.finfo(np.float64).eps, be
    Xo = B0.copy(order="F")
    O.rectrxm("L", "L", "N", 1.0, "S", A, Xo, threshold=cut)
    assert relerr(X, Xo) < 1e-11, relerr(X, Xo)
