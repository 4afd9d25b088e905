"""Generate the golden vectors for the tile-QR / TRSM-TRMM parity tests.

Inputs are drawn from numpy PCG64 with fixed seeds; expected outputs come from
LAPACK/BLAS through scipy (scipy-openblas 0.3.29, the OpenBLAS version NextLA.jl's
test suite pins in test/Manifest.toml).  The reference's own tests
(test/larfg.jl, test/geqrt.jl, test/tsqrt.jl, test/tsmqr.jl, test/unmqr.jl,
test/unified_rectrxm.jl) compare against exactly these LAPACK routines
(test/lapack_helpers.jl:18-139); the shapes below are the shapes those tests use
plus larger ones that exercise recursion and multiple inner blocks.

Run:  python tests/golden/gen_golden.py   (writes tests/golden/*.npz)
"""
import os

import numpy as np
from scipy.linalg import blas, lapack

OUT = os.path.dirname(os.path.abspath(__file__))
DT = {"d": np.float64, "s": np.float32}


def F(a):
    return np.asfortranarray(a)


def lp(p, name):
    return getattr(lapack, p + name)


def gen_larfg(rng):
    """dlarfg / slarfg (test/larfg.jl:14-79): random, x = 0, n = 1 and tiny (rescale) cases."""
    out = {}
    cases = []
    for n in (2, 5, 17, 64, 200):
        cases.append(("rand", n, 1.0))
    cases += [("zero", 6, 1.0), ("n1", 1, 1.0), ("tiny", 12, 1e-300), ("tinyf", 12, 1e-40)]
    for p in "ds":
        for idx, (kind, n, scale) in enumerate(cases):
            if p == "s" and kind == "tiny":
                continue
            if p == "d" and kind == "tinyf":
                continue
            x = rng.standard_normal(max(n - 1, 0)) * scale
            alpha = float(rng.standard_normal() * scale)
            if kind == "zero":
                x[:] = 0
            x = x.astype(DT[p])
            alpha = DT[p](alpha)
            if n >= 2:
                b, v, tau = lp(p, "larfg")(n, alpha, x.copy())
            else:
                b, v, tau = alpha, x.copy(), DT[p](0)
            key = f"{p}_{idx}"
            out[key + "_alpha"] = np.array([alpha], dtype=DT[p])
            out[key + "_x"] = x
            out[key + "_beta"] = np.array([b], dtype=DT[p])
            out[key + "_v"] = np.asarray(v, dtype=DT[p])[: max(n - 1, 0)]
            out[key + "_tau"] = np.array([tau], dtype=DT[p])
    out["ncases"] = np.array([len(cases)])
    return out


GEQRT_SHAPES = [(32, 24, 8), (48, 32, 16), (64, 64, 16), (20, 20, 10), (128, 128, 32), (96, 40, 32), (30, 50, 8)]


def gen_geqrt(rng):
    """dgeqrt with nb = ib: same V/R/T layout as NextLA geqrt! (test/geqrt.jl)."""
    out = {}
    for p in "ds":
        for idx, (m, n, ib) in enumerate(GEQRT_SHAPES):
            A = F(rng.standard_normal((m, n)).astype(DT[p]))
            qr, t, info = lp(p, "geqrt")(ib, A.copy())
            assert info == 0
            key = f"{p}_{idx}"
            out[key + "_A"] = A
            out[key + "_QR"] = F(qr)
            out[key + "_T"] = F(t)
            # unmqr parity: dgemqrt on C for each side/trans
            k = min(m, n)
            for side in "LR":
                for trans in "NT":
                    C = F(rng.standard_normal((m, 19) if side == "L" else (23, m)).astype(DT[p]))
                    c, info = lp(p, "gemqrt")(qr[:, :k], t, C.copy(), side=side, trans=trans)
                    assert info == 0
                    out[f"{key}_C{side}{trans}"] = C
                    out[f"{key}_Q{side}{trans}"] = F(c)
    return out


TPQRT_SHAPES = [(32, 24, 8), (48, 32, 16), (64, 48, 16), (20, 20, 10), (128, 128, 32), (64, 64, 64)]


def gen_tsqrt(rng):
    """dtpqrt with l = 0 (test/tsqrt.jl:6-34): A1 n x n upper, A2 m x n."""
    out = {}
    for p in "ds":
        for idx, (m, n, ib) in enumerate(TPQRT_SHAPES):
            A1 = F(np.triu(rng.standard_normal((n, n))).astype(DT[p]))
            A2 = F(rng.standard_normal((m, n)).astype(DT[p]))
            a, b, t, info = lp(p, "tpqrt")(0, ib, A1.copy(), A2.copy())
            assert info == 0
            key = f"{p}_{idx}"
            out[key + "_A1"] = A1
            out[key + "_A2"] = A2
            out[key + "_R"] = F(a)
            out[key + "_V"] = F(b)
            out[key + "_T"] = F(t)
    return out


TSMQR_CASES = [
    # side, m1, n1, m2, n2, k, ib   (first two are test/tsmqr.jl:8-13)
    ("L", 32, 24, 20, 24, 16, 8),
    ("R", 24, 32, 24, 20, 16, 8),
    ("L", 64, 64, 64, 64, 64, 32),
    ("R", 64, 64, 64, 64, 64, 32),
    ("L", 64, 200, 96, 200, 40, 16),
    ("R", 72, 64, 72, 100, 50, 32),
]


def gen_tsmqr(rng):
    """dtpmqrt with l = 0 and random (V, T) (test/tsmqr.jl:6-40)."""
    out = {}
    for p in "ds":
        for idx, (side, m1, n1, m2, n2, k, ib) in enumerate(TSMQR_CASES):
            A1 = F(rng.random((m1, n1)).astype(DT[p]))
            A2 = F(rng.random((m2, n2)).astype(DT[p]))
            V = F(rng.random((m2, k) if side == "L" else (n2, k)).astype(DT[p]))
            Tm = F(np.triu(rng.random((ib, k))).astype(DT[p]))
            key = f"{p}_{idx}"
            out[key + "_A1"] = A1
            out[key + "_A2"] = A2
            out[key + "_V"] = V
            out[key + "_T"] = Tm
            for trans in "NT":
                # LAPACK side L takes the k x n1 block of A1 that the reflectors touch
                a_in = A1[:k, :].copy(order="F") if side == "L" else A1[:, :k].copy(order="F")
                a, b, info = lp(p, "tpmqrt")(0, V, Tm, a_in, A2.copy(), side=side, trans=trans)
                assert info == 0
                A1o = A1.copy(order="F")
                if side == "L":
                    A1o[:k, :] = a
                else:
                    A1o[:, :k] = a
                out[f"{key}_{trans}_A1"] = F(A1o)
                out[f"{key}_{trans}_A2"] = F(b)
    return out


TRXM_CASES = [
    # func, side, uplo, trans, n, m, alpha, cutoff
    ("S", "L", "L", "N", 64, 33, 1.0, 0),
    ("S", "L", "U", "N", 64, 33, 2.0, 0),
    ("S", "R", "L", "N", 48, 40, 1.0, 0),
    ("S", "R", "U", "N", 48, 40, 0.5, 0),
    ("S", "L", "L", "T", 300, 20, 1.0, 0),
    ("S", "L", "U", "T", 300, 20, 1.0, 64),
    ("S", "L", "L", "N", 264, 24, 1.5, 32),
    ("S", "R", "L", "T", 200, 16, 1.0, 48),
    ("S", "R", "U", "N", 160, 12, 1.0, 32),
    ("M", "L", "L", "N", 64, 33, 1.0, 0),
    ("M", "L", "U", "N", 100, 21, 2.0, 0),
    ("M", "R", "L", "N", 77, 40, 1.0, 0),
    ("M", "R", "U", "T", 90, 12, -1.0, 0),
    ("M", "L", "L", "T", 200, 20, 1.0, 64),
]


def tri_matrix(rng, n, uplo, dtype):
    """A as in the TRSM configuration: off-diagonal N(0,1)/sqrt(n), diagonal U[1,2]."""
    A = rng.standard_normal((n, n)) / np.sqrt(n)
    A = np.tril(A, -1) if uplo == "L" else np.triu(A, 1)
    A[np.diag_indices(n)] = rng.uniform(1.0, 2.0, n)
    # garbage in the unused triangle must never be read
    junk = rng.standard_normal((n, n)) * 1e3
    A = A + (np.triu(junk, 1) if uplo == "L" else np.tril(junk, -1))
    return F(A.astype(dtype))


def gen_trxm(rng):
    """dtrsm / dtrmm (test/unified_rectrxm.jl:36-43), including n > 256 recursion."""
    out = {}
    for p in "ds":
        for idx, (func, side, uplo, trans, n, m, alpha, cut) in enumerate(TRXM_CASES):
            A = tri_matrix(rng, n, uplo, DT[p])
            B = F(rng.standard_normal((n, m) if side == "L" else (m, n)).astype(DT[p]))
            fn = getattr(blas, p + ("trsm" if func == "S" else "trmm"))
            X = fn(alpha, A, B.copy(), side=0 if side == "L" else 1, lower=1 if uplo == "L" else 0,
                   trans_a=0 if trans == "N" else 1, diag=0)
            key = f"{p}_{idx}"
            out[key + "_A"] = A
            out[key + "_B"] = B
            out[key + "_X"] = F(X)
    return out


def gen_tileqr(rng):
    """dgeqrf R factor of matrices the tile QR driver factors (|R| is unique)."""
    out = {}
    for p in "ds":
        for idx, (m, n) in enumerate([(256, 256), (384, 256)]):
            A = F(rng.standard_normal((m, n)).astype(DT[p]))
            qr, tau, work, info = lp(p, "geqrf")(A.copy())
            assert info == 0
            out[f"{p}_{idx}_A"] = A
            out[f"{p}_{idx}_R"] = F(np.triu(qr[:n, :n]))
    return out


def main():
    gens = [("larfg", gen_larfg, 11), ("geqrt", gen_geqrt, 12), ("tsqrt", gen_tsqrt, 13),
            ("tsmqr", gen_tsmqr, 14), ("trxm", gen_trxm, 15), ("tileqr", gen_tileqr, 16)]
    for name, g, seed in gens:
        rng = np.random.Generator(np.random.PCG64(seed))
        data = g(rng)
        path = os.path.join(OUT, f"{name}.npz")
        np.savez_compressed(path, **data)
        print(f"wrote {path} ({os.path.getsize(path)} bytes, {len(data)} arrays)")


if __name__ == "__main__":
    main()
