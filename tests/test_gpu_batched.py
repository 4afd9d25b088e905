"""Batched uneven problems: geqrt! QRs and lower-triangular solves of many independent
problems in one persistent launch, its hipGraph replay and the ib autotuner.

Pins: each problem against the oracle's geqrt! (restatement of src/geqrt.jl, itself pinned
to LAPACK dgeqrt) and against numpy's triangular solve."""
import numpy as np
import pytest
import torch
from scipy.linalg import solve_triangular

from conftest import relerr

pytestmark = pytest.mark.gpu

TOL = {np.float64: 1e-12, np.float32: 2e-5}


def _mods():
    import nextla_hip as nl
    from nextla_hip import batched as bt

    nl.load()
    return nl, bt


def make_batch(rng, count, dt, mmax=512, nrhs=32, solve_every=1):
    probs = []
    for i in range(count):
        m = 32 * int(rng.integers(2, mmax // 32 + 1))
        n = 16 * int(rng.integers(1, min(128, m) // 16 + 1))
        A = rng.standard_normal((m, n)).astype(dt)
        L = B = None
        if i % solve_every == 0:
            L = np.tril(rng.standard_normal((n, n)) / np.sqrt(n), -1)
            L[np.diag_indices(n)] = rng.uniform(1, 2, n)
            L = L.astype(dt)
            B = rng.standard_normal((n, nrhs)).astype(dt)
        probs.append((A, L, B))
    return probs


def to_dev(nl, probs):
    As = [nl.to_device(np.asfortranarray(A)) for A, _, _ in probs]
    Ls = [None if L is None else nl.to_device(np.asfortranarray(L)) for _, L, _ in probs]
    Bs = [None if B is None else nl.to_device(np.asfortranarray(B)) for _, _, B in probs]
    return As, Ls, Bs


def check(nl, O, probs, plan, As, Bs, tol, idx=None):
    for i in idx if idx is not None else range(len(probs)):
        A0, L0, B0 = probs[i]
        ib = plan.T[i].shape[0]
        A = np.asfortranarray(A0.copy())
        T = np.zeros((ib, A.shape[1]), dtype=A.dtype, order="F")
        O.geqrt(A, T, ib)
        assert relerr(nl.to_host(As[i]), A) < tol, (i, A0.shape, ib)
        assert relerr(nl.to_host(plan.T[i]), T) < 10 * tol, (i, A0.shape, ib)
        if L0 is not None:
            X = solve_triangular(L0.astype(np.float64), B0.astype(np.float64), lower=True)
            assert relerr(nl.to_host(Bs[i]), X) < tol, i


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("ib", [8, 16, 32])
def test_batched_vs_oracle(dt, ib):
    nl, bt = _mods()
    import oracle as O

    rng = np.random.default_rng(100 + ib)
    probs = make_batch(rng, 40, dt, solve_every=2)
    As, Ls, Bs = to_dev(nl, probs)
    plan = bt.BatchedQRTRSM(As, Ls, Bs, ib=ib)
    plan.launch()
    torch.cuda.synchronize()
    check(nl, O, probs, plan, As, Bs, TOL[dt])


def test_batched_graph_replay():
    """The launch recorded in a hipGraph, replayed three times with eager work in between;
    every replay must reproduce the eager result exactly."""
    nl, bt = _mods()
    rng = np.random.default_rng(7)
    probs = make_batch(rng, 64, np.float32)
    As, Ls, Bs = to_dev(nl, probs)
    A_in = [a.clone() for a in As]
    B_in = [b.clone() for b in Bs]
    plan = bt.BatchedQRTRSM(As, Ls, Bs, ib=16)
    plan.launch()
    torch.cuda.synchronize()
    A_ref = [a.clone() for a in As]
    B_ref = [b.clone() for b in Bs]
    T_ref = [t.clone() for t in plan.T]
    for a, a0 in zip(As, A_in):
        a.copy_(a0)
    for b, b0 in zip(Bs, B_in):
        b.copy_(b0)
    plan.capture()
    for rep in range(3):
        for a, a0 in zip(As, A_in):
            a.copy_(a0)
        for b, b0 in zip(Bs, B_in):
            b.copy_(b0)
        junk = torch.randn(1024, 1024, device="cuda") @ torch.randn(1024, 1024, device="cuda")  # eager work
        plan.replay()
        torch.cuda.synchronize()
        del junk
        for x, y in zip(As + Bs + plan.T, A_ref + B_ref + T_ref):
            assert torch.equal(x, y), rep


def test_autotune_picks_a_candidate():
    nl, bt = _mods()
    res = bt.autotune(torch.float32, [(256, 64), (512, 128)], samples=8, candidates=(8, 16, 32))
    for (m, n), (ib, ms) in res.items():
        assert ib in (8, 16, 32) and ms > 0


def test_batched_full_workload_sample_checked():
    """4096 problems of the batched workload (m in 64..1024 step 32, n in 16..128 step 16,
    32 right-hand sides), one launch; 24 problems checked against the oracle."""
    nl, bt = _mods()
    import oracle as O

    rng = np.random.default_rng(4096)
    g = torch.Generator(device="cuda")
    g.manual_seed(4096)
    As, Ls, Bs, shapes = [], [], [], []
    for _ in range(4096):
        m = 32 * int(rng.integers(2, 33))
        n = 16 * int(rng.integers(1, min(128, m) // 16 + 1))
        As.append(nl.colmajor(torch.randn((n, m), generator=g, device="cuda").t()))
        L = torch.tril(torch.randn((n, n), generator=g, device="cuda") / n ** 0.5, -1)
        L.diagonal().copy_(torch.rand(n, generator=g, device="cuda") + 1)
        Ls.append(nl.colmajor(L))
        Bs.append(nl.colmajor(torch.randn((32, n), generator=g, device="cuda").t()))
    idx = rng.choice(4096, 24, replace=False)
    probs = {int(i): (nl.to_host(As[i]), nl.to_host(Ls[i]), nl.to_host(Bs[i])) for i in idx}
    plan = bt.BatchedQRTRSM(As, Ls, Bs, ib=16)
    plan.launch()
    torch.cuda.synchronize()
    check(nl, O, probs, plan, As, Bs, TOL[np.float32], idx=[int(i) for i in idx])


def test_batched_argument_errors():
    nl, bt = _mods()
    A = nl.zeros_cm(64, 128, dtype=torch.float32)  # n > m
    plan = bt.BatchedQRTRSM([A], ib=16)
    with pytest.raises(nl.ArgumentError, match="n <= m"):
        plan.launch()
