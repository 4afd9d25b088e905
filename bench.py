"""Benchmark of the NextLA.jl hot path on one MI355X.

Workload (`value`): the recursive TRSM configuration -- unified_rectrxm!('L','L','N', 1,
'S', A, B) with A 8192 x 8192 lower triangular (off-diagonal N(0,1)/sqrt(n), diagonal
U[1,2]), B 8192 x 4096 N(0,1), Float64, recursion cut-off 256 (the reference default).
One step = the solve of one of K copies of B (K = the step count, at most 32; each pass
solves every copy once in turn), all made from one pristine B before timing starts, so the
timed region holds nothing but the solves.  TFLOP/s use the LAPACK count n^2 m.

Extra fields: Float32 TRSM and Float64 TRMM of the same shape, the solve's backward error,
the mixed-precision solve (Float32 + Float64 refinement), Float64 and reduced-precision-update
tile QR (flat-tree geqrt/unmqr/tsqrt/tsmqr) at 16384^2 with backward errors, the same
factorizations compared at equal backward error as refined solves of a 16384^2 system, the
recursive LU (getrf2!) at 8192^2, the batched uneven workload (4096 problems, one
graph-replayed launch), the reference's per-tile call sequence (one geqrt!/unmqr!/tsqrt!/
tsmqr! launch per tile, 8192^2) beside the persistent tile QR, adjoint timings, the
32768 x 8192 least-squares workload; the roofline of the dominant kernel (the 128 x 128-tile
MFMA GEMM of the level-1/2 off-diagonal updates, timed live with hipEvents on its launch stream
in a second pass of the same steps, so that the event packets do not slow the pass that gives
`value`; HBM traffic from the committed counter profile) and CPU baselines on the host cores
(the oracle's TRSM on a bounded sample, vendor LAPACK legs, the 1024^2 tile QR), each the
median of 5 runs after a warm-up.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--qr-n N]
                    [--extras trsm32,trmm,trsm_mixed,qr,qr_solve,lu,batched,percall,adjoints,ls] [--no-extras]
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(REPO, "nextla.jl_amd"), os.path.join(REPO, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

PEAK = {"f64": 78.6, "f32": 157.3, "f16": 2516.6, "bf16": 2516.6}   # TFLOP/s, MI355X dense MFMA
PEAK["bf16x3"] = PEAK["bf16"] / 3   # split BF16: three BF16 products per product


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def gemm_flops_of_recursion(n, m, cut, off=0):
    """Flops of the GEMM launches unified_rec issues (src/rectrxm.jl:101-198) for the lower,
    left, untransposed solve: with 256-row base blocks (cut-off in [256, 512)) an aligned
    512-row block runs as one fused launch (base, update, base) and issues no GEMM."""
    if n <= cut:
        return 0
    if 256 <= cut < 512 and n == 512 and off % 256 == 0:
        return 0
    mid = n // 2 if (n & (n - 1)) == 0 else 1 << (n.bit_length() - 1)
    rem = n - mid
    return (2 * mid * rem * m + gemm_flops_of_recursion(mid, m, cut, off)
            + gemm_flops_of_recursion(rem, m, cut, off + mid))


def gemm128_flops_of_recursion(n, m, cut, off=0):
    """Flops of those GEMM launches of the same recursion that run on 128 x 128 tiles (the
    library picks them when the update's C has >= 512 such tiles: C is rem x m here)."""
    if n <= cut or (256 <= cut < 512 and n == 512 and off % 256 == 0):
        return 0
    mid = n // 2 if (n & (n - 1)) == 0 else 1 << (n.bit_length() - 1)
    rem = n - mid
    big = ((rem + 127) // 128) * ((m + 127) // 128) >= 512
    return ((2 * mid * rem * m if big else 0) + gemm128_flops_of_recursion(mid, m, cut, off)
            + gemm128_flops_of_recursion(rem, m, cut, off + mid))


def make_tri(torch, nl, n, dtype, gen):
    X = torch.randn((n, n), generator=gen, device="cuda", dtype=torch.float64) / (n ** 0.5)
    L = torch.tril(X, -1)
    L.diagonal().copy_(torch.rand(n, generator=gen, device="cuda", dtype=torch.float64) + 1.0)
    return nl.colmajor(L.to(dtype))


def time_steps(torch, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def host_info():
    """CPU model, logical CPUs available to this process, physical cores (lscpu)."""
    import subprocess

    model, phys = None, None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        kv = {l.split(":", 1)[0].strip(): l.split(":", 1)[1].strip() for l in out.splitlines() if ":" in l}
        model = kv.get("Model name")
        phys = int(kv.get("Core(s) per socket", "0")) * int(kv.get("Socket(s)", "0")) or None
    except Exception:
        pass
    if model is None:
        try:
            with open("/proc/cpuinfo") as f:
                model = next(l.split(":", 1)[1].strip() for l in f if l.startswith("model name"))
        except Exception:
            model = "unknown"
    return {"model": model, "logical_cpus_available": len(os.sched_getaffinity(0)), "physical_cores": phys}


def median_of(fn, runs=5):
    """1 warm-up, then the median wall time of `runs` runs (the CPU-baseline protocol)."""
    fn()
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return sorted(ts)[len(ts) // 2]


def cpu_baseline(n, cut, target_s=2.5):
    """CPU baselines on the host cores, each 1 warm-up + median of 5 runs:
    (value) the oracle's recursive TRSM (a C/OpenMP restatement of unified_rectrxm!) on n x n
    with a bounded number of the 4096 right-hand sides, chosen for ~target_s per run;
    (legs) vendor LAPACK/BLAS through scipy (OpenBLAS) dtrsm on the same sample, the tile QR
    workload 1024^2 (nb 128, ib 32) through the oracle, and LAPACK dgeqrf of the same matrix."""
    import numpy as np
    from scipy.linalg import blas, lapack

    import oracle as O

    try:
        from threadpoolctl import threadpool_info
        blas_threads = max((d.get("num_threads", 0) for d in threadpool_info() if d.get("user_api") == "blas"),
                           default=None)
    except Exception:
        blas_threads = None
    rng = np.random.default_rng(7)
    A = np.tril(rng.standard_normal((n, n)) / np.sqrt(n), -1)
    A[np.diag_indices(n)] = rng.uniform(1, 2, n)
    A = np.asfortranarray(A)
    B = np.asfortranarray(rng.standard_normal((n, 32)))
    t0 = time.perf_counter()
    O.rectrxm("L", "L", "N", 1.0, "S", A, B.copy(order="F"), threshold=cut)
    rate = n * n * 32 / (time.perf_counter() - t0)
    m2 = int(max(32, min(4096, rate * target_s / (n * n))))
    B = np.asfortranarray(rng.standard_normal((n, m2)))
    dt = median_of(lambda: O.rectrxm("L", "L", "N", 1.0, "S", A, B.copy(order="F"), threshold=cut))
    dt_l = median_of(lambda: blas.dtrsm(1.0, A, B, side=0, lower=1, trans_a=0, diag=0))
    # tile QR 1024^2, nb 128, ib 32 (oracle) and LAPACK dgeqrf
    Q = np.asfortranarray(rng.standard_normal((1024, 1024)))
    qflops = 2.0 * 1024 ** 3 - 2.0 / 3.0 * 1024 ** 3

    def tqr():
        Qw = Q.copy(order="F")
        O.tile_geqrf(Qw, 128, 32)
    dt_q = median_of(tqr)
    dt_qf = median_of(lambda: lapack.dgeqrf(Q.copy(order="F")))
    info = host_info()
    threads = O.num_threads()
    return {"value": round(n * n * m2 / dt / 1e12, 6), "unit": "TFLOP/s", "cores": threads, "kind": "port",
            "sample": f"oracle unified_rectrxm TRSM L/L/N Float64 n={n} with {m2} of the 4096 right-hand sides, "
                      f"cut-off {cut}; median of 5 runs after 1 warm-up, {dt:.2f} s each; {threads} OpenMP threads",
            "host": info,
            "legs": [
                {"what": f"LAPACK/BLAS dtrsm (scipy, OpenBLAS) on the same {n} x {m2} sample", "kind": "vendor",
                 "value": round(n * n * m2 / dt_l / 1e12, 6), "unit": "TFLOP/s", "s": round(dt_l, 3),
                 "threads": blas_threads},
                {"what": "tile QR 1024^2, nb 128, ib 32 (geqrt/unmqr/tsqrt/tsmqr), oracle", "kind": "port",
                 "value": round(qflops / dt_q / 1e12, 6), "unit": "TFLOP/s", "s": round(dt_q, 4), "threads": threads},
                {"what": "LAPACK dgeqrf 1024^2 (scipy, OpenBLAS)", "kind": "vendor",
                 "value": round(qflops / dt_qf / 1e12, 6), "unit": "TFLOP/s", "s": round(dt_qf, 4),
                 "threads": blas_threads},
            ]}


def cutoff_sweep(torch, nl, n, m, funcs="SM", dtypes=("f64", "f32"), cuts=(16, 32, 64, 128, 256, 512)):
    """TFLOP/s (n^2 m) of unified_rectrxm! L/L/N, TRSM and TRMM, Float64 and Float32, over the
    recursion cut-off (src/rectrxm.jl:52,63), one JSON line."""
    gen = torch.Generator(device="cuda")
    gen.manual_seed(8192)
    A64 = make_tri(torch, nl, n, torch.float64, gen)
    B64 = nl.colmajor(torch.randn((m, n), generator=gen, device="cuda", dtype=torch.float64).t())
    out = {"n": n, "m": m, "tflops": {}}
    for dt, tag in ((torch.float64, "f64"), (torch.float32, "f32")):
        if tag not in dtypes:
            continue
        A = nl.colmajor(A64.to(dt))
        B0 = nl.colmajor(B64.to(dt))
        B = nl.colmajor(torch.empty_like(B0))
        for func in funcs:
            for cut in cuts:
                def step():
                    B.copy_(B0)
                    nl.unified_rectrxm_("L", "L", "N", 1.0, func, A, B, cutoff=cut)
                t = time_steps(torch, step, 5, 2)
                key = f"{'trsm' if func == 'S' else 'trmm'}_{tag}_cut{cut}"
                out["tflops"][key] = round(float(n) * n * m / t / 1e12, 3)
                log(f"[sweep] {key}: {out['tflops'][key]} TFLOP/s ({t*1e3:.3f} ms incl. the copy of B)")
    print(json.dumps(out), flush=True)


def least_squares(torch, nl, gen, m=32768, n=8192, nb=128, ib=32):
    """Mixed-precision tile QR + least squares: A = U diag(s) V^T (U, V from QR of Gaussian,
    s geometric from 1 to 1/kappa), b = A x0 + 1e-8 N(0,1); tile QR with Float64 / Float32 /
    FP16 / BF16 / split-BF16 trailing updates, each refined by
    R-preconditioned LSQR (products by A in Float64) towards the Float64 solution."""
    out = {"m": m, "n": n}
    for kappa in (1e2, 1e6):
        U = torch.linalg.qr(torch.randn((m, n), generator=gen, device="cuda", dtype=torch.float64))[0]
        V = torch.linalg.qr(torch.randn((n, n), generator=gen, device="cuda", dtype=torch.float64))[0]
        sv = kappa ** (-torch.arange(n, device="cuda", dtype=torch.float64) / (n - 1))
        A = nl.colmajor((U * sv) @ V.T)
        del U, V
        x0 = torch.randn(n, generator=gen, device="cuda", dtype=torch.float64)
        b = A @ x0 + 1e-8 * torch.randn(m, generator=gen, device="cuda", dtype=torch.float64)
        row = {}
        for prec in ("D", "S", "H", "B", "X"):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            info = {}
            x, it, last = nl.tile_gels(A, b, nb, ib, prec, maxit=200, tol=1e-15, info=info)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            r = b - A @ x
            # ||A^T r|| / (||A||_2 ||r||) with ||A||_2 = s_1 = 1 by construction
            eta = (torch.linalg.norm(A.T @ r) / torch.linalg.norm(r)).item()
            row[prec] = {"s": round(dt, 3), "iterations": it, "eta": eta, "factor_used": info["prec_used"], "x": x}
            log(f"[bench] LS kappa={kappa:g} prec {prec}: {dt:.3f} s, {it} steps, eta {eta}, "
                f"factor {info['prec_used']}")
        xd = row["D"]["x"]
        for p in ("S", "H", "B", "X"):
            row[p]["rel_diff_vs_f64_solution"] = (torch.linalg.norm(row[p]["x"] - xd) / torch.linalg.norm(xd)).item()
            row[p]["vs_f64_time"] = round(row["D"]["s"] / row[p]["s"], 3)
        for v in row.values():
            v.pop("x")
        out[f"kappa_{kappa:g}"] = {"f64": row["D"], "mixed_f32": row["S"], "mixed_f16": row["H"], "mixed_bf16": row["B"],
                                   "mixed_bf16x3": row["X"]}
        del A, b
    return out


def qr_solve_equal_error(torch, nl, gen, n=16384, kappa=1e2, nb=128, ib=32):
    """Tile QR at 16384^2 compared at equal backward error: solve A x = b (A = U diag(s) V^T, s
    geometric from 1 to 1/kappa, b = A x0) with the Float64 factorization and with the Float32 /
    FP16 / BF16 / split-BF16-update factorizations, each refined by R-preconditioned LSQR (products by A in
    Float64) until LSQR's estimate reaches 1e-15; the time covers factorization + refinement
    and every row states the normwise backward error ||b - A x|| / (||A||_2 ||x|| + ||b||)."""
    U = torch.linalg.qr(torch.randn((n, n), generator=gen, device="cuda", dtype=torch.float64))[0]
    V = torch.linalg.qr(torch.randn((n, n), generator=gen, device="cuda", dtype=torch.float64))[0]
    sv = kappa ** (-torch.arange(n, device="cuda", dtype=torch.float64) / (n - 1))
    A = nl.colmajor((U * sv) @ V.T)
    del U, V
    x0 = torch.randn(n, generator=gen, device="cuda", dtype=torch.float64)
    b = A @ x0
    out = {"n": n, "kappa": kappa, "nb": nb, "ib": ib,
           "time": "factorization + direct start + LSQR refinement, one call, wall clock"}
    for prec, tag in (("D", "f64"), ("S", "mixed_f32"), ("H", "mixed_f16"), ("B", "mixed_bf16"), ("X", "mixed_bf16x3")):
        info = {}
        nl.tile_gels(A, b, nb, ib, prec, maxit=100, tol=1e-15, info=info)   # warm-up
        times = []
        for _ in range(3):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            x, it, last = nl.tile_gels(A, b, nb, ib, prec, maxit=100, tol=1e-15, info=info)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
        t = sorted(times)[1]
        be = (torch.linalg.norm(b - A @ x) / (torch.linalg.norm(x) + torch.linalg.norm(b))).item()  # ||A||_2 = 1
        out[tag] = {"ms": round(t * 1e3, 2), "iterations": it, "factor_used": info["prec_used"],
                    "backward_error": be, "forward_error": (torch.linalg.norm(x - x0) / torch.linalg.norm(x0)).item()}
        log(f"[bench] QR solve {n}^2 kappa {kappa:g} prec {prec}: {t*1e3:.1f} ms, {it} its, berr {be:.2e}, "
            f"factor {info['prec_used']}")
    for tag in ("mixed_f32", "mixed_f16", "mixed_bf16", "mixed_bf16x3"):
        out[tag]["vs_f64"] = round(out["f64"]["ms"] / out[tag]["ms"], 3)
    del A, b
    return out


def lu_factor(torch, nl, gen, n=8192):
    """Recursive LU with partial pivoting (getrf2!, src/lu.jl:185-300) of an n x n N(0,1)
    matrix, Float64: TFLOP/s with the LAPACK count 2n^3/3 and the residual ||P A - L U|| / ||A||
    estimated on 4 random vectors."""
    A0 = nl.colmajor(torch.randn((n, n), generator=gen, device="cuda", dtype=torch.float64))
    A = nl.colmajor(torch.empty_like(A0))
    ipiv = torch.zeros(n, dtype=torch.int64, device="cuda")
    holder = {}

    def run():
        A.copy_(A0)
        holder["r"] = nl.getrf2_(A, ipiv)
    t = time_steps(torch, run, 3, 1)
    # residual on random vectors: ||P A x - L (U x)|| / (||A|| ||x||)
    X = torch.randn((n, 4), generator=gen, device="cuda", dtype=torch.float64)
    U = torch.triu(A)
    L = torch.tril(A, -1) + torch.eye(n, device="cuda", dtype=torch.float64)
    piv = ipiv.cpu().numpy() - 1
    import numpy as np
    p = np.arange(n)
    for i in range(n):
        p[[i, piv[i]]] = p[[piv[i], i]]
    PAx = A0[torch.from_numpy(p).cuda()] @ X
    res = (torch.linalg.norm(PAx - L @ (U @ X)) / (torch.linalg.norm(A0) * torch.linalg.norm(X))).item()
    flops = 2.0 / 3.0 * n ** 3
    log(f"[bench] LU getrf2 {n}^2 f64: {t*1e3:.2f} ms, {flops / t / 1e12:.2f} TFLOP/s, residual {res:.2e}")
    return {"n": n, "ms": round(t * 1e3, 3), "tflops": round(flops / t / 1e12, 3),
            "frac_of_f64_peak": round(flops / t / 1e12 / PEAK["f64"], 4), "residual": res,
            "flops": "2n^3/3 (LAPACK count)", "info": holder["r"][2]}


def batched(torch, nl, gen, count=4096, nrhs=32, fixed_ib=0):
    """Batched uneven workload: `count` Float32 problems, m uniform over multiples of 32 in
    [64, 1024], n uniform over multiples of 16 in [16, min(128, m)]; geqrt! of each A and a
    lower solve with 32 right-hand sides, all in one persistent launch, replayed from a hipGraph.
    ib per problem width from the autotuner."""
    import numpy as np

    from nextla_hip import batched as bt

    rng = np.random.default_rng(4096)
    shapes = []
    for _ in range(count):
        m = 32 * int(rng.integers(2, 33))
        n = 16 * int(rng.integers(1, min(128, m) // 16 + 1))
        shapes.append((m, n))
    As, Ls, Bs = [], [], []
    for m, n in shapes:
        As.append(nl.colmajor(torch.randn((n, m), generator=gen, device="cuda").t()))
        L = torch.tril(torch.randn((n, n), generator=gen, device="cuda") / n ** 0.5, -1)
        L.diagonal().copy_(torch.rand(n, generator=gen, device="cuda") + 1)
        Ls.append(nl.colmajor(L))
        Bs.append(nl.colmajor(torch.randn((nrhs, n), generator=gen, device="cuda").t()))
    A0 = [a.clone() for a in As]
    B0 = [b.clone() for b in Bs]
    t0 = time.perf_counter()
    if fixed_ib > 0:
        ibs = {n: fixed_ib for n in range(16, 129, 16)}
    else:
        tune = bt.autotune(torch.float32, [(512, n) for n in range(16, 129, 16)], samples=64, candidates=(8, 16, 32))
        ibs = {n: ib for (m, n), (ib, ms) in tune.items()}
    t_tune = time.perf_counter() - t0
    plan = bt.BatchedQRTRSM(As, Ls, Bs, ib=ibs)
    qflops = sum(2.0 * m * n * n - 2.0 / 3.0 * n ** 3 for m, n in shapes)
    sflops = sum(float(n) * n * nrhs for m, n in shapes)
    nbytes = 4.0 * sum(2 * m * n + n * n + 2 * n * nrhs for m, n in shapes)

    def restore():
        for a, a0 in zip(As, A0):
            a.copy_(a0)
        for b, b0 in zip(Bs, B0):
            b.copy_(b0)

    plan.launch()  # validates the work list
    torch.cuda.synchronize()
    restore()
    plan.capture()
    times = []
    for _ in range(5):
        restore()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        plan.replay()
        e1.record()
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1) * 1e-3)
    t = sorted(times)[len(times) // 2]
    log(f"[bench] batched {count} problems: {t*1e3:.2f} ms per graph replay, autotune {t_tune:.1f} s, ib {ibs}")
    return {"problems": count, "dtype": "f32", "ms": round(t * 1e3, 3),
            "tflops": round((qflops + sflops) / t / 1e12, 3), "qr_gflop": round(qflops / 1e9, 2),
            "trsm_gflop": round(sflops / 1e9, 3), "effective_gbps": round(nbytes / t / 1e9, 1),
            "frac_of_f32_peak": round((qflops + sflops) / t / 1e12 / PEAK["f32"], 4), "ib_by_n": ibs,
            "launch": "one persistent kernel replayed from a hipGraph"}


def percall(torch, nl, gen, n=8192, nb=128, ib=32):
    """The reference's tile-QR call sequence, one geqrt!/unmqr!/tsqrt!/tsmqr! launch per tile or
    tile pair in flat-tree order (nla_dtile_geqrf_calls: the loop a NextLA caller writes), beside
    the persistent nla_dtile_geqrf, Float64 n x n, nb = 128, ib = 32."""
    nt = n // nb
    calls = {"geqrt": nt, "unmqr": nt * (nt - 1) // 2, "tsqrt": nt * (nt - 1) // 2,
             "tsmqr": sum(j * j for j in range(nt))}
    A0 = nl.colmajor(torch.randn((n, n), generator=gen, device="cuda", dtype=torch.float64))
    warm = nl.colmajor(A0[:1024, :1024].clone())
    nl.tile_geqrf_calls(warm, nb, ib)
    torch.cuda.synchronize()
    res = {}
    for name, fn in (("persistent", nl.tile_geqrf), ("per_call", nl.tile_geqrf_calls)):
        A = nl.colmajor(A0.clone())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn(A, nb, ib)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        res[name] = {"ms": round(t * 1e3, 2), "tflops": round(4.0 / 3.0 * n ** 3 / t / 1e12, 3)}
        del A
    res["per_call"]["launches"] = calls
    res["per_call"]["us_per_call"] = round(res["per_call"]["ms"] * 1e3 / sum(calls.values()), 2)
    res["persistent_vs_per_call"] = round(res["per_call"]["ms"] / res["persistent"]["ms"], 1)
    log(f"[bench] tile QR call sequence n={n}: per-call {res['per_call']['ms']} ms "
        f"({res['per_call']['us_per_call']} us per call), persistent {res['persistent']['ms']} ms")
    return {"n": n, "nb": nb, "ib": ib, "dtype": "f64", **res}


def gemm_rates(torch, nl, sizes=(4096, 8192)):
    """Square GEMM rates that bound an integer-sliced (Ozaki-type) Float64 emulation of the
    TRSM's off-diagonal updates: int8 x int8 -> int32 and FP16 through torch's library GEMMs,
    Float64 through torch and through this library's MFMA GEMM (GEMM_SUB!)."""
    def bench_fn(fn, reps=10):
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) * 1e-3 / reps
    out = {}
    for n in sizes:
        fl = 2.0 * n ** 3
        a8 = torch.randint(-127, 128, (n, n), dtype=torch.int8, device="cuda")
        b8 = torch.randint(-127, 128, (n, n), dtype=torch.int8, device="cuda")
        x = nl.colmajor(torch.randn((n, n), device="cuda", dtype=torch.float64))
        y = nl.colmajor(torch.randn((n, n), device="cuda", dtype=torch.float64))
        c = nl.colmajor(torch.zeros((n, n), device="cuda", dtype=torch.float64))
        r = {"int8_tops": fl / bench_fn(lambda: torch._int_mm(a8, b8)) / 1e12,
             "f16_tflops": fl / bench_fn(lambda: x.half() @ y.half()) / 1e12,
             "f64_torch_tflops": fl / bench_fn(lambda: x @ y) / 1e12,
             "f64_mfma_gemm_tflops": fl / bench_fn(lambda: nl.GEMM_SUB_(x, y, c)) / 1e12}
        out[str(n)] = {k: round(v, 1) for k, v in r.items()}
        log(f"[bench] GEMM rates n={n}: {out[str(n)]}")
        del a8, b8, x, y, c
    return out


def adjoints(torch, nl, gen):
    """Adjoint kernels at the differentiation workload's sizes: TRSM A 4096, B 4096 x 512;
    thin QR of 8192 x 1024 (Float64); time of the backward pass alone."""
    from nextla_hip import autodiff as ad

    out = {}
    n, m = 4096, 512
    A = make_tri(torch, nl, n, torch.float64, gen)
    B = torch.randn((n, m), generator=gen, device="cuda", dtype=torch.float64)
    W = torch.randn((n, m), generator=gen, device="cuda", dtype=torch.float64)
    X = nl.colmajor(B.clone())
    nl.unified_rectrxm_("L", "L", "N", 1.0, "S", A, X)
    Xb = nl.colmajor(W.clone())

    def trsm_bw():
        Xb.copy_(W)
        ad.trsm_rrule_("L", "L", "N", 1.0, A, X, Xb)
    t = time_steps(torch, trsm_bw, 5, 2)
    out["trsm_rrule_ms"] = round(t * 1e3, 3)
    out["trsm_rrule_tflops"] = round(2.0 * n * n * m / t / 1e12, 3)  # solve n^2 m + masked product n^2 m
    mq, nq = 8192, 1024
    Aq = torch.randn((mq, nq), generator=gen, device="cuda", dtype=torch.float64, requires_grad=True)
    GQ = torch.randn((mq, nq), generator=gen, device="cuda", dtype=torch.float64)
    GR = torch.randn((nq, nq), generator=gen, device="cuda", dtype=torch.float64)
    for _ in range(2):
        Q, R = ad.thin_qr(Aq, 128, 32)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ((Q * GQ).sum() + (R * GR).sum()).backward()
        torch.cuda.synchronize()
        tb = time.perf_counter() - t0
        Aq.grad = None
    out["qr_rrule_ms"] = round(tb * 1e3, 3)
    log(f"[bench] adjoints: TRSM rrule {t*1e3:.2f} ms, thin-QR rrule {tb*1e3:.1f} ms")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--m", type=int, default=4096)
    ap.add_argument("--cutoff", type=int, default=256)
    ap.add_argument("--qr-n", type=int, default=16384)
    ap.add_argument("--ls", action="store_true", help="also run the 32768 x 8192 least-squares workload")
    ap.add_argument("--extras", default="trsm32,trmm,trsm_mixed,qr,qr_solve,lu,batched,percall,adjoints,ls",
                    help="comma-separated extra workloads: trsm32, trmm, trsm_mixed, qr, qr_solve, lu, batched, "
                         "percall, adjoints, ls")
    ap.add_argument("--qr-no-check", action="store_true", help="skip the tile-QR backward-error check")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--only", default="",
                    help="run only these extra workloads (comma-separated, as for --extras) and print their JSON; "
                         "no headline solve, no CPU baseline (for profiling one workload)")
    ap.add_argument("--batched-ib", type=int, default=0, help="fixed ib for the batched workload (0: autotune)")
    ap.add_argument("--cutoff-sweep", action="store_true",
                    help="only time TRSM and TRMM (Float64, Float32) at n x m over cut-offs 16..512 and print JSON")
    ap.add_argument("--sweep-funcs", default="SM", help="functions of the cut-off sweep: S and/or M")
    ap.add_argument("--sweep-dtypes", default="f64,f32", help="element types of the cut-off sweep")
    ap.add_argument("--sweep-cuts", default="16,32,64,128,256,512", help="cut-offs of the sweep")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on exactly one GPU: the library has no multi-GPU path "
              "(the workloads are measured on one GPU's share); rerun with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    import nextla_hip as nl

    nl.load()
    n, m, cut = args.n, args.m, args.cutoff
    if args.only:
        gen = torch.Generator(device="cuda")
        gen.manual_seed(8192)
        only = set(x for x in args.only.split(",") if x)
        runners = {"batched": lambda: batched(torch, nl, gen, fixed_ib=args.batched_ib),
                   "percall": lambda: percall(torch, nl, gen),
                   "gemm_rates": lambda: gemm_rates(torch, nl),
                   "lu": lambda: lu_factor(torch, nl, gen), "adjoints": lambda: adjoints(torch, nl, gen),
                   "ls": lambda: least_squares(torch, nl, gen),
                   "qr_solve": lambda: qr_solve_equal_error(torch, nl, gen, n=args.qr_n)}
        unknown = only - set(runners)
        if unknown:
            print(f"--only supports {sorted(runners)}, not {sorted(unknown)}", file=sys.stderr)
            sys.exit(2)
        print(json.dumps({k: runners[k]() for k in sorted(only)}))
        return
    if args.cutoff_sweep:
        cutoff_sweep(torch, nl, n, m, args.sweep_funcs, args.sweep_dtypes.split(","),
                     [int(c) for c in args.sweep_cuts.split(",")])
        return
    gen = torch.Generator(device="cuda")
    gen.manual_seed(8192)
    A = make_tri(torch, nl, n, torch.float64, gen)
    B0 = nl.colmajor(torch.randn((m, n), generator=gen, device="cuda", dtype=torch.float64).t())
    B = nl.colmajor(torch.empty_like(B0))
    flops = float(n) * n * m
    # one right-hand side matrix per timed step (268 MB each), refilled from B0 between passes
    Bk = [nl.colmajor(torch.empty_like(B0)) for _ in range(max(1, min(args.steps, 32)))]

    def refill():
        for b in Bk:
            b.copy_(B0)

    def step(k):
        nl.unified_rectrxm_("L", "L", "N", 1.0, "S", A, Bk[k % len(Bk)], cutoff=cut)

    for _ in range(args.warmup):
        Bk[0].copy_(B0)
        step(0)
    refill()
    # Two passes of the same K steps, each bracketed by synchronisations: the first gives
    # `value` (nothing but the solves on the stream); the second runs with the library's launch
    # profiler, which records a HIP event pair on the launch stream around every GEMM and base
    # launch (the per-kernel durations of the roofline) and whose event packets between the
    # ~66 launches of a solve would otherwise slow the timed solves by about 10 %.
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(k)
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / args.steps
    refill()
    nl.prof_reset()
    nl.prof_enable(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(k)
    torch.cuda.synchronize()
    sec_prof = (time.perf_counter() - t0) / args.steps
    nl.prof_enable(False)
    gemm_ms, gemm_n = nl.prof_read("gemm")
    g128_ms, g128_n = nl.prof_read("gemm128")
    base_ms, base_n = nl.prof_read("trsm_base")
    nl.prof_reset()
    value = flops / sec / 1e12
    gemm_flops = gemm_flops_of_recursion(n, m, cut)
    g128_flops = gemm128_flops_of_recursion(n, m, cut)
    all_gemm_tf = gemm_flops * args.steps / (gemm_ms * 1e-3) / 1e12 if gemm_ms > 0 else None
    # the roofline's kernel: gemm_kernel<double,128,128,...>, the dominant launches of the solve
    achieved = g128_flops * args.steps / (g128_ms * 1e-3) / 1e12 if g128_ms > 0 else None
    log(f"[bench] TRSM f64 n={n} m={m} cut={cut}: {sec*1e3:.3f} ms/step, {value:.2f} TFLOP/s "
        f"({sec_prof*1e3:.3f} ms/step with the launch profiler); "
        f"gemm {gemm_ms/args.steps:.3f} ms/step over {gemm_n//args.steps} launches ({all_gemm_tf} TFLOP/s); "
        f"128x128 gemm {g128_ms/args.steps:.3f} ms/step over {g128_n//args.steps} launches ({achieved} TFLOP/s); "
        f"base {base_ms/args.steps:.3f} ms/step")

    # backward error of the same solve of the pristine B: ||A X - B|| / (||A|| ||X||)
    del Bk
    B.copy_(B0)
    nl.unified_rectrxm_("L", "L", "N", 1.0, "S", A, B, cutoff=cut)
    Lf = torch.tril(A)
    R = Lf @ B - B0
    berr = (torch.linalg.norm(R) / (torch.linalg.norm(Lf) * torch.linalg.norm(B))).item()
    del R, Lf

    extras = {}
    chosen = set() if args.no_extras else set(x for x in args.extras.split(",") if x)
    if args.ls:
        chosen.add("ls")
    if "trsm32" in chosen:
        # Float32 TRSM, same shape
        A32 = A.to(torch.float32)
        A32 = nl.colmajor(A32)
        B032 = nl.colmajor(B0.to(torch.float32))
        B32 = nl.colmajor(torch.empty_like(B032))

        def s32():
            B32.copy_(B032)
            nl.unified_rectrxm_("L", "L", "N", 1.0, "S", A32, B32, cutoff=cut)
        t32 = time_steps(torch, s32, max(3, args.steps // 2), 1)
        extras["trsm_f32_tflops"] = round(flops / t32 / 1e12, 3)
        del A32, B032, B32

    if "trmm" in chosen:
        for cut_m, key in ((0, "trmm_f64_tflops"), (16, "trmm_f64_tflops_cutoff16")):
            def trmm():
                B.copy_(B0)
                nl.unified_rectrxm_("L", "L", "N", 1.0, "M", A, B, cutoff=cut_m)
            tm = time_steps(torch, trmm, max(3, args.steps // 2), 1)
            extras[key] = round(flops / tm / 1e12, 3)
        log(f"[bench] TRMM f64: {extras['trmm_f64_tflops']} TFLOP/s (cut-off 256), "
            f"{extras['trmm_f64_tflops_cutoff16']} (the reference's 16)")

    if "trsm_mixed" in chosen:
        # mixed-precision TRSM (Float32-storage recursive solve whose GEMM updates take Float32,
        # FP16 or BF16 operands, + Float64 residual refinement), same problem
        Bm = nl.colmajor(torch.empty_like(B0))
        res = {}
        for prec in ("S", "H", "B"):
            info = {}

            def smix():
                Bm.copy_(B0)
                _, info["it"], info["last"] = nl.rectrsm_mixed_("L", "L", "N", 1.0, A, Bm, cutoff=cut, maxit=12,
                                                                tol=1e-15, prec=prec)
            tmx = time_steps(torch, smix, 3, 1)
            Rm = torch.tril(A) @ Bm - B0
            berr_m = (torch.linalg.norm(Rm) / (torch.linalg.norm(torch.tril(A)) * torch.linalg.norm(Bm))).item()
            del Rm
            res[{"S": "f32", "H": "f16", "B": "bf16"}[prec]] = {
                "ms": round(tmx * 1e3, 3), "tflops": round(flops / tmx / 1e12, 3), "vs_f64": round(sec / tmx, 3),
                "iterations": info["it"], "backward_error": berr_m}
            log(f"[bench] TRSM mixed ({prec} updates): {tmx*1e3:.2f} ms, {info['it']} refinement steps, "
                f"berr {berr_m:.2e}")
        del Bm
        extras["trsm_mixed"] = {"scheme": "Float32-storage recursive solve (GEMM updates with Float32 / FP16 / "
                                          "BF16 operands, Float32 accumulation) + Float64 residual refinement",
                                "backward_error_f64": berr, **res}

    if "qr" in chosen:
        # tile QR at the north-star size: Float64 and Float32-update mixed, nb = 128, ib = 32
        qn = args.qr_n
        nb, ib = 128, 32
        Q0 = nl.colmajor(torch.randn((qn, qn), generator=gen, device="cuda", dtype=torch.float64))
        Qw = nl.colmajor(torch.empty_like(Q0))
        qflops = 2.0 * qn ** 3 - 2.0 / 3.0 * qn ** 3
        res = {}
        # the panels (geqrt/tsqrt, Float64 in every variant) are about nb/n of the flops: n^2 nb
        panel_share = nb / qn
        variants = (("f64", "f64", lambda: nl.tile_geqrf(Qw, nb, ib)),
                    ("mixed_f32", "f32", lambda: nl.tile_geqrf_mixed(Qw, nb, ib, "S")),
                    ("mixed_f16", "f16", lambda: nl.tile_geqrf_mixed(Qw, nb, ib, "H")),
                    ("mixed_bf16", "bf16", lambda: nl.tile_geqrf_mixed(Qw, nb, ib, "B")),
                    ("mixed_bf16x3", "bf16x3", lambda: nl.tile_geqrf_mixed(Qw, nb, ib, "X")))
        for tag, upd, fn in variants:
            holder = {}

            def qr():
                Qw.copy_(Q0)
                holder["T"] = fn()
            tq = time_steps(torch, qr, 2, 1)
            be = None
            if not args.qr_no_check:
                # backward error ||A - Q R|| / ||A|| with Q applied from the factors
                QR = nl.zeros_cm(qn, qn)
                QR.copy_(torch.triu(Qw))
                nl.tile_ormqr("N", Qw, holder["T"], QR, nb, ib)
                be = (torch.linalg.norm(QR - Q0) / torch.linalg.norm(Q0)).item()
                del QR
            tf = qflops / tq / 1e12
            # flop-weighted peak: trailing updates at the update dtype's peak, panels at FP64's
            blend = 1.0 / ((1 - panel_share) / PEAK[upd] + panel_share / PEAK["f64"])
            res[tag] = {"ms": round(tq * 1e3, 2), "tflops": round(tf, 3), "update_dtype": upd,
                        "frac_of_f64_peak": round(tf / PEAK["f64"], 4),
                        "frac_of_blended_peak": round(tf / blend, 4), "blended_peak_tflops": round(blend, 1),
                        "backward_error": be}
            if tag != "f64":
                res[tag]["vs_f64"] = round(res["f64"]["ms"] / res[tag]["ms"], 3)
            log(f"[bench] tile QR {tag} n={qn}: {tq*1e3:.1f} ms, {tf:.2f} TFLOP/s, berr {be}")
        extras["tile_qr"] = {"n": qn, "nb": nb, "ib": ib, "flops": "2n^3 - 2n^3/3 (LAPACK count)", **res}
        del Q0, Qw

    if "qr_solve" in chosen:
        extras["tile_qr_solve_equal_error"] = qr_solve_equal_error(torch, nl, gen, n=args.qr_n)
    if "lu" in chosen:
        extras["lu"] = lu_factor(torch, nl, gen)
    if "batched" in chosen:
        extras["batched"] = batched(torch, nl, gen, fixed_ib=args.batched_ib)
    if "percall" in chosen:
        extras["tile_qr_call_sequence"] = percall(torch, nl, gen)
    if "adjoints" in chosen:
        extras["adjoints"] = adjoints(torch, nl, gen)
    if "ls" in chosen:
        extras["least_squares"] = least_squares(torch, nl, gen)

    traffic = None
    tpath = os.path.join(REPO, "profiles", "trsm_f64_gemm_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch_128")

    out = {
        "metric": "tile-QR & TRSM TFLOP/s (% of dtype peak), mixed vs FP64 at equal backward err",
        "value": round(value, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sec * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"rectrxm TRSM left/lower/N Float64 n={n}, B {n}x{m}, cut-off {cut}",
                   "n": n, "m": m, "cutoff": cut, "frac_of_peak": round(value / PEAK["f64"], 4),
                   "backward_error": berr},
        "roofline": {"bound": "mfma", "achieved": None if achieved is None else round(achieved, 3),
                     "peak": PEAK["f64"], "unit": "TFLOP/s",
                     "frac": None if achieved is None else round(achieved / PEAK["f64"], 4),
                     "traffic": traffic,
                     "kernel": "nla::gemm_kernel<double, 128, 128, 16, false, false, false, 1, true> (the level-1/2 "
                               "off-diagonal GEMM_SUB! updates of the recursion)",
                     "launches_per_step": g128_n // max(1, args.steps),
                     "flop_per_step": g128_flops,
                     "avg_launch_ms": round(g128_ms / max(1, g128_n), 4),
                     "all_gemm_launches": {"launches_per_step": gemm_n // max(1, args.steps),
                                           "tflops": None if all_gemm_tf is None else round(all_gemm_tf, 3),
                                           "avg_launch_ms": round(gemm_ms / max(1, gemm_n), 4)},
                     "measured": "HIP events on the launch stream around every 128x128-tile GEMM launch, over a "
                                 "second pass of the same steps (ms_per_step there: "
                                 f"{round(sec_prof * 1e3, 4)})"},
        "extras": extras,
    }
    if not args.no_cpu:
        try:
            out["cpu_baseline"] = cpu_baseline(n, cut)
        except Exception as e:  # the baseline must not hide the GPU result
            out["cpu_baseline"] = {"value": None, "unit": "TFLOP/s", "cores": os.cpu_count(), "kind": "port",
                                   "sample": f"failed: {e}"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
