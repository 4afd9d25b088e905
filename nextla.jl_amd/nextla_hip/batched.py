"""Batched uneven tile problems: many independent geqrt! QRs and lower-triangular solves in
one persistent launch (nla_?batched_qr_trsm), with an autotuned inner block per problem
width and optional hipGraph capture of the launch for replay.

The work list is an array of the C struct nla_?batch_problem (include/nextla_hip.h) in
device memory, ordered largest problem first so that the persistent workgroups finish
together.  Matrices are column-major GPU tensors; nothing runs on the CPU."""
import ctypes

import numpy as np
import torch

from . import api
from ._lib import call, load

# field order of nla_?batch_problem: 13 eight-byte fields
_FIELDS = ("m", "n", "ib", "A", "lda", "T", "ldt", "nt", "nrhs", "L", "ldl", "B", "ldb")


def autotune(dtype, classes, samples=16, candidates=(8, 16, 32)):
    """Fastest ib for each (m, n) class, by timing the batched launch (nla_?batched_autotune)."""
    p = "d" if dtype == torch.float64 else "s"
    load()
    nc = len(classes)
    mv = (ctypes.c_int64 * nc)(*[int(c[0]) for c in classes])
    nv = (ctypes.c_int64 * nc)(*[int(c[1]) for c in classes])
    cv = (ctypes.c_int64 * len(candidates))(*candidates)
    out = (ctypes.c_int64 * nc)()
    ms = (ctypes.c_double * nc)()
    call(p, "batched_autotune", nc, ctypes.cast(mv, ctypes.c_void_p), ctypes.cast(nv, ctypes.c_void_p), int(samples),
         ctypes.cast(cv, ctypes.c_void_p), len(candidates), ctypes.cast(out, ctypes.c_void_p),
         ctypes.cast(ms, ctypes.c_void_p), api._stream())
    return {tuple(c): (int(out[i]), float(ms[i])) for i, c in enumerate(classes)}


class BatchedQRTRSM:
    """QR (geqrt!) of every A_i and solve L_i X_i = B_i (X overwrites B) in one launch.

    As, Ls, Bs: lists of column-major GPU tensors (Ls/Bs may be None or contain None for
    problems without a solve).  ib: an int, or a dict {n: ib} (e.g. from `autotune`).
    The T factors are allocated here (self.T[i], ib x n)."""

    def __init__(self, As, Ls=None, Bs=None, ib=32):
        if not As:
            raise api.ArgumentError("empty batch")
        self.dtype = As[0].dtype
        self.p = api._p(*As)
        self.As, self.Ls, self.Bs = As, Ls or [None] * len(As), Bs or [None] * len(As)
        self.T = []
        rows = []
        for A, L, B in zip(self.As, self.Ls, self.Bs):
            m, n = A.shape
            b = ib.get(n, 32) if isinstance(ib, dict) else int(ib)
            b = max(1, min(b, n, 64))
            T = api.zeros_cm(max(b, 1), n, dtype=self.dtype, device=A.device)
            self.T.append(T)
            pa, lda = api._m(A)
            pt, ldt = api._m(T)
            if L is not None and B is not None:
                pl, ldl = api._m(L)
                pb, ldb = api._m(B)
                nt, nrhs = L.shape[0], B.shape[1]
            else:
                pl = pb = 0
                ldl = ldb = 1
                nt = nrhs = 0
            flops = 2.0 * m * n * n + nt * nt * nrhs
            rows.append((flops, (m, n, b, pa, lda, pt, ldt, nt, nrhs, pl or 0, ldl, pb or 0, ldb)))
        rows.sort(key=lambda r: -r[0])  # largest first
        host = np.array([r[1] for r in rows], dtype=np.int64)
        self.work = torch.from_numpy(host.reshape(-1)).to(As[0].device)
        self.counter = torch.zeros(1, dtype=torch.int32, device=As[0].device)
        self.count = len(rows)
        self.flops = sum(r[0] for r in rows)
        self.graph = None
        self._validated = False

    def launch(self):
        """Enqueue the batched launch on the current stream."""
        call(self.p, "batched_qr_trsm", self.count, self.work.data_ptr(), self.counter.data_ptr(),
             0 if self._validated else 1, api._stream())
        self._validated = True

    def capture(self):
        """Record the launch into a hipGraph (after one validating launch)."""
        if not self._validated:
            raise api.ArgumentError("run launch() once (it validates the work list) before capture()")
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                self.launch()
        torch.cuda.current_stream().wait_stream(s)
        self.graph = g
        return g

    def replay(self):
        if self.graph is None:
            raise api.ArgumentError("capture() first")
        self.graph.replay()
