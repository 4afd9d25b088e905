"""The workloads at their stated sizes, and the reference's own rectrxm test grid.

* Recursive TRSM / TRMM, left / lower / N, A n = 8192 (off-diagonal N(0,1)/sqrt(n), diagonal
  U[1,2]), B 8192 x 4096 N(0,1), Float64 and Float32, recursion cut-off 64 .. 512: backward
  error ||A X - B|| / (||A|| ||X||) <= 10 u, agreement across cut-offs, the TRMM against a
  Float64 torch product, and a 32-column slice of B against the oracle at full n.
* Least squares 32768 x 8192, A = U diag(s) V^T (s geometric from 1 to 1/kappa, kappa 1e2 and
  1e6), b = A x0 + 1e-8 noise, Float64 / Float32 / FP16 / BF16 trailing updates refined to
  the Float64 factor's LS backward-error indicator ||A^T r|| / (||A|| ||r||).
* test/unified_rectrxm.jl:10-45: n in {16, 32, 128, 256} x m in {1, 8, 64} x side x uplo x
  trans in {N, T, C} x func, A = tri(rand + 1) + 10 I, B = rand + 1, against BLAS dtrsm /
  dtrmm (the reference's oracle there) at its tolerance 1e-14.
* Two solves enqueued on two streams at once (each stream has its own scratch).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import nextla_hip as nl  # noqa: E402
import oracle as O  # noqa: E402
from conftest import relerr  # noqa: E402

pytestmark = pytest.mark.gpu

N2, M2 = 8192, 4096
CUTS_S = [64, 128, 256, 512]
CUTS_M = [0, 64, 128, 256, 512]   # 0 = the reference's TRMM threshold 16
EPS = {torch.float64: 2.220446049250313e-16, torch.float32: 1.1920929e-07}


@pytest.fixture(scope="module")
def trsm_problem():
    gen = torch.Generator(device="cuda")
    gen.manual_seed(8192)
    X = torch.randn((N2, N2), generator=gen, device="cuda", dtype=torch.float64) / N2 ** 0.5
    L = torch.tril(X, -1)
    L.diagonal().copy_(torch.rand(N2, generator=gen, device="cuda", dtype=torch.float64) + 1.0)
    del X
    B0 = torch.randn((M2, N2), generator=gen, device="cuda", dtype=torch.float64).t()
    return nl.colmajor(L), nl.colmajor(B0)


def _solve(A, B0, func, cut):
    B = nl.colmajor(B0.clone())
    nl.unified_rectrxm_("L", "L", "N", 1.0, func, A, B, cutoff=cut)
    torch.cuda.synchronize()
    return B


@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
def test_config_trsm_n8192_cutoffs(trsm_problem, dt):
    L64, B64 = trsm_problem
    A, B0 = nl.colmajor(L64.to(dt)), nl.colmajor(B64.to(dt))
    Af = A.double()            # the matrix actually solved with (rounded to dt)
    nA = torch.linalg.norm(Af).item()
    Xs = {}
    for cut in CUTS_S:
        print(f"[config trsm] {dt} n={N2} m={M2} cutoff {cut}", flush=True)
        X = _solve(A, B0, "S", cut)
        Xd = X.double()
        be = (torch.linalg.norm(Af @ Xd - B0.double()) / (nA * torch.linalg.norm(Xd))).item()
        assert be <= 10 * EPS[dt], (cut, be)
        Xs[cut] = Xd
    ref = Xs[256]
    for cut, Xd in Xs.items():
        # the solves differ only in the order of their sums: forward errors of a
        # well-conditioned system (cond(A) ~ 10)
        assert relerr(Xd.cpu().numpy(), ref.cpu().numpy()) < (1e-13 if dt == torch.float64 else 1e-5), cut


@pytest.mark.parametrize("dt", [torch.float64, torch.float32])
def test_config_trmm_n8192_cutoffs(trsm_problem, dt):
    L64, B64 = trsm_problem
    A, B0 = nl.colmajor(L64.to(dt)), nl.colmajor(B64.to(dt))
    Yref = A.double() @ B0.double()
    nY = torch.linalg.norm(Yref).item()
    for cut in CUTS_M:
        print(f"[config trmm] {dt} n={N2} m={M2} cutoff {cut}", flush=True)
        Y = _solve(A, B0, "M", cut)
        err = (torch.linalg.norm(Y.double() - Yref) / nY).item()
        assert err < (1e-14 if dt == torch.float64 else 2e-6), (cut, err)


@pytest.mark.parametrize("func,cut", [("S", 256), ("M", 0)])
def test_config_slice_vs_oracle(trsm_problem, func, cut):
    """32 of the 4096 right-hand sides at full n: the GPU's columns of the full solve (each
    column's result depends only on its own data) against the oracle's recursion."""
    L64, B64 = trsm_problem
    X = _solve(L64, B64, func, cut)
    cols = slice(1000, 1032)
    A = nl.to_host(L64)
    Bs = np.asfortranarray(nl.to_host(B64)[:, cols])
    O.rectrxm("L", "L", "N", 1.0, func, A, Bs, threshold=cut)
    assert relerr(nl.to_host(X)[:, cols], Bs) < 1e-13


# ------------------------------------------------------------------ least squares 32768 x 8192
M3, N3 = 32768, 8192


@pytest.fixture(scope="module")
def ls_problems():
    cache = {}

    def get(kappa):
        if kappa not in cache:
            cache.clear()
            torch.cuda.empty_cache()
            gen = torch.Generator(device="cuda")
            gen.manual_seed(int(kappa) % 9973)
            U = torch.linalg.qr(torch.randn((M3, N3), generator=gen, device="cuda", dtype=torch.float64))[0]
            V = torch.linalg.qr(torch.randn((N3, N3), generator=gen, device="cuda", dtype=torch.float64))[0]
            sv = kappa ** (-torch.arange(N3, device="cuda", dtype=torch.float64) / (N3 - 1))
            A = nl.colmajor((U * sv) @ V.T)
            del U, V
            x0 = torch.randn(N3, generator=gen, device="cuda", dtype=torch.float64)
            b = A @ x0 + 1e-8 * torch.randn(M3, generator=gen, device="cuda", dtype=torch.float64)
            cache[kappa] = {"A": A, "b": b}
        return cache[kappa]

    return get


def _eta(A, b, x):
    r = b - A @ x
    return (torch.linalg.norm(A.T @ r) / torch.linalg.norm(r)).item()   # ||A||_2 = 1


@pytest.mark.parametrize("kappa", [1e2, 1e6])
@pytest.mark.parametrize("prec", ["D", "S", "H", "B", "X"])
def test_config_least_squares_32768x8192(ls_problems, kappa, prec):
    P = ls_problems(kappa)
    A, b = P["A"], P["b"]
    if "D" not in P:
        print(f"[config ls] kappa {kappa:g} prec D (reference level)", flush=True)
        xd, it, _ = nl.tile_gels(A, b, 128, 32, "D", maxit=200, tol=1e-15)
        torch.cuda.synchronize()
        P["D"] = (xd, _eta(A, b, xd))
    xd, eta_d = P["D"]
    print(f"[config ls] kappa {kappa:g} prec {prec}", flush=True)
    info = {}
    x, it, last = nl.tile_gels(A, b, 128, 32, prec, maxit=200, tol=1e-15, info=info)
    torch.cuda.synchronize()
    eta = _eta(A, b, x)
    assert eta <= max(3 * eta_d, 1e-13), (eta, eta_d, it, last, info)
    # forward error of a backward-stable LS solver: u kappa (1 + kappa ||r|| / (||A|| ||x||))
    fe = (torch.linalg.norm(x - xd) / torch.linalg.norm(xd)).item()
    assert fe < 50 * 2.2e-16 * kappa * (1 + kappa * 1e-8), (fe, info)
    assert info["prec_used"] in (prec, "S")
    if prec in "DS":
        assert info["prec_used"] == prec


# ------------------------------------------------------------ the reference's rectrxm grid
GRID = [(n, m, side, uplo, trans, func) for n in (16, 32, 128, 256) for m in (1, 8, 64) for side in "LR"
        for uplo in "LU" for trans in "NTC" for func in "SM"]


def _grid_inputs(n, m, side, uplo, seed):
    # test/unified_rectrxm.jl:18-24: A = tri(rand(n,n) .+ 1) + 10 I, B = rand .+ 1
    rng = np.random.default_rng(seed)
    R = rng.random((n, n)) + 1
    A = np.tril(R) if uplo == "L" else np.triu(R)
    A = np.asfortranarray(A + 10 * np.eye(n))
    B = np.asfortranarray(rng.random((n, m) if side == "L" else (m, n)) + 1)
    return A, B


@pytest.mark.parametrize("n,m,side,uplo,trans,func", GRID)
def test_reference_rectrxm_grid(n, m, side, uplo, trans, func):
    from scipy.linalg import blas

    A, B = _grid_inputs(n, m, side, uplo, GRID.index((n, m, side, uplo, trans, func)))
    dB = nl.to_device(B)
    nl.unified_rectrxm_(side, uplo, trans, 1.0, func, nl.to_device(A), dB)
    torch.cuda.synchronize()
    fn = blas.dtrsm if func == "S" else blas.dtrmm
    want = fn(1.0, A, B.copy(order="F"), side=0 if side == "L" else 1, lower=1 if uplo == "L" else 0,
              trans_a=0 if trans == "N" else 1, diag=0)
    got = nl.to_host(dB)
    assert relerr(got, want) < 1e-14           # test/unified_rectrxm.jl:8,42-43
    Bo = B.copy(order="F")
    O.rectrxm(side, uplo, trans, 1.0, func, A, Bo)
    assert relerr(got, Bo) < 1e-14


# --------------------------------------------------------- two streams, two solves at once
def test_concurrent_solves_on_two_streams():
    """Each stream keeps its own scratch (the TRSM's inverted diagonal blocks): two different
    systems solved concurrently on two streams both match the oracle, twice, the second time
    with the sizes swapped (so each stream's scratch grows while the other's is in use)."""
    rng = np.random.default_rng(77)
    probs = []
    for n, m in ((1536, 96), (1024, 160)):
        A = np.tril(rng.standard_normal((n, n)) / np.sqrt(n), -1)
        A[np.diag_indices(n)] = rng.uniform(1, 2, n)
        probs.append((np.asfortranarray(A), np.asfortranarray(rng.standard_normal((n, m)))))
    want = []
    for A, B in probs:
        Bo = B.copy(order="F")
        O.rectrxm("L", "L", "N", 1.0, "S", A, Bo, threshold=256)
        want.append(Bo)
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    for order in ((0, 1), (1, 0)):
        dA = [nl.to_device(p[0]) for p in probs]
        dB0 = [nl.to_device(p[1]) for p in probs]
        dB = [nl.colmajor(b.clone()) for b in dB0]
        torch.cuda.synchronize()
        for s, i in zip((s1, s2), order):
            with torch.cuda.stream(s):
                for _ in range(3):   # repeated solves keep both streams busy together
                    dB[i].copy_(dB0[i])
                    nl.unified_rectrxm_("L", "L", "N", 1.0, "S", dA[i], dB[i], cutoff=256)
        torch.cuda.synchronize()
        for i in range(2):
            assert relerr(nl.to_host(dB[i]), want[i]) < 1e-12, (order, i)
