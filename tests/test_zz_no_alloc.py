"""nla_drectrxm_noalloc (the recursive TRSM/TRMM with its workspace in a library-owned device
array) against nla_drectrxm, called through the C ABI: two calls in a row, the second one
smaller, so the second reads an array that still holds the first call's contents."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import nextla_hip as nl  # noqa: E402
from golden.gen_golden import tri_matrix  # noqa: E402

pytestmark = pytest.mark.gpu


def _run(fn, side, uplo, trans, func, A, B0, cut):
    lib = nl.load()
    dA, dB = nl.to_device(A), nl.to_device(B0)
    n = A.shape[0]
    m = B0.shape[1] if side == "L" else B0.shape[0]
    pa, lda = dA.data_ptr(), dA.stride(1)
    st = getattr(lib, fn)(side.encode(), uplo.encode(), trans.encode(), ctypes.c_double(0.75), func.encode(),
                          n, m, ctypes.c_void_p(pa), lda, ctypes.c_void_p(dB.data_ptr()), dB.stride(1), cut,
                          ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert st == 0, lib.nla_last_error().decode()
    torch.cuda.synchronize()
    return nl.to_host(dB)


@pytest.mark.parametrize("case", [
    # (side, uplo, trans, func, n, m, cut) first call, then a smaller second call
    (("L", "L", "N", "S", 2048, 96, 256), ("L", "L", "N", "S", 1000, 33, 256)),
    (("R", "U", "T", "S", 1536, 70, 64), ("L", "U", "N", "S", 512, 20, 128)),
    (("L", "L", "N", "S", 8192, 40, 256), ("R", "L", "N", "M", 700, 50, 0)),
])
def test_noalloc_matches_rectrxm(case):
    rng = np.random.default_rng(len(str(case)))
    for side, uplo, trans, func, n, m, cut in case:
        print(f"[noalloc] {side}{uplo}{trans}{func} n={n} m={m} cut={cut}", flush=True)
        A = tri_matrix(rng, n, uplo, np.float64)
        B0 = np.asfortranarray(rng.standard_normal((n, m) if side == "L" else (m, n)))
        want = _run("nla_drectrxm", side, uplo, trans, func, A, B0, cut)
        got = _run("nla_drectrxm_noalloc", side, uplo, trans, func, A, B0, cut)
        # the same kernels on the same data: identical results
        assert np.array_equal(got, want), np.abs(got - want).max()


def test_noalloc_rejects_oversize():
    lib = nl.load()
    st = lib.nla_drectrxm_noalloc(b"L", b"L", b"N", ctypes.c_double(1.0), b"S", 16384, 4, None, 16384, None,
                                  16384, 256, None)
    assert st == -6 and "no-allocation entry point holds" in lib.nla_last_error().decode()
