"""Benchmark: converged eigenpairs/s of the lowest-eigenpair solve on one MI355X.

Workload (one "step" = one whole solve, operator resident in HBM):
  3-D 7-point Laplacian on a 128^3 periodic grid (n = 2,097,152, CSR int32)
  plus an Anderson diagonal potential U[-2,2] from seed 1; lowest 64
  eigenpairs, subspace increment 32, buffer 32, tolerance 1e-10 with the
  residual-only stop; FP64 throughout (filter, orthonormalisation,
  Rayleigh-Ritz); start vectors from seed 0.  Each step runs the Lanczos
  bounds, all filter passes and returns eigenvalues + eigenvectors on device.

value = converged eigenpairs (||Av - lambda v|| / ||A|| <= 1e-10, checked on
the device after the timed region) per second of timed wall time, FP64
filter, with the final FP64 Rayleigh-Ritz refinement over the found pairs.
The same solve with the mixed-precision filter (FP32 Chebyshev steps on each
Ritz column's correction, FP64 for everything else and the same refinement)
is timed the same way
and reported under "mixed", with its speed-up and its largest eigenvalue
difference from the FP64 solve.

roofline: the fused Chebyshev step kernel (csr_block_kernel, mode kCheb),
timed with HIP events on the library's stream around every filter
application inside the timed solves; its algorithmic bytes per launch are
nnz*(8+4) + 4*(n+1) + 3*n*b*8 (CSR values+columns+row pointers, one read of
X_{j-1}, one read of X_{j-2}, one write of X_j).

cpu_baseline: the CPU oracle (a restatement of the reference's
RayleighChebyshevLM, OpenMP) timed on a bounded sample of the same solve (the
Lanczos bounds and one inner iteration per outer pass), extrapolated to the
whole solve with the GPU solve's own counts (filter degrees, orthogonalisations,
Rayleigh-Ritz steps).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "converged eigenpairs/s @ resid≤1e-10, mixed vs FP64; p50/p99 solve latency"
K, INC, BUF, TOL = 64, 32, 32, 1e-10
GRID = 128
# workloads: the default is the 128^3 Anderson solve; "hofstadter1024" is the
# complex-Hermitian 1024^2 lattice (k=128, inc=64, buf=32, complex64 mixed)
WORKLOADS = {
    "anderson128": dict(k=64, inc=32, buf=32, tol=1e-10, complex=False,
                        label="3-D 7-point periodic Laplacian 128^3 + Anderson U[-2,2] (seed 1)",
                        profile="128cube"),
    "hofstadter1024": dict(k=128, inc=64, buf=32, tol=1e-10, complex=True,
                           label=("2-D periodic Hofstadter lattice 1024^2, Peierls phase "
                                  "exp(2 pi i x/64), onsite U[-1,1] (seed 2), complex128 CSR"),
                           profile="hofstadter1024"),
    "dense16384": dict(k=256, inc=128, buf=128, tol=1e-10, complex=False, dense=True,
                       label=("dense symmetric A = Q diag(linspace(0,1,n)^2) Q^T, Q the Haar QR "
                              "factor of a Gaussian matrix (seed 3), n=16384"),
                       profile="dense16384"),
}
WL = WORKLOADS["anderson128"]


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def dense_operator_host(n, seed=3):
    """A = Q diag(linspace(0,1,n)^2) Q^T (the dense workload), built on the GPU by
    torch in a child process started before this process touches the GPU, read
    back from a .npy file; returns (A column-major, the exact eigenvalues)."""
    import subprocess
    import tempfile
    path = os.path.join(tempfile.gettempdir(), f"rcg_dense_{n}_{seed}.npy")
    if not os.path.exists(path):
        subprocess.run([sys.executable, "-m", "rayleighchebyshevgpgpu_amd.dense_gen", str(n),
                        str(seed), path], check=True, cwd=ROOT)
    A = np.load(path, mmap_mode="r")
    return A, np.linspace(0.0, 1.0, n) ** 2


def build_operator(grid=GRID):
    from rayleighchebyshevgpgpu_amd import operators as ops
    if WL["complex"]:
        return ops.hofstadter_2d(1024, phi=1.0 / 64, W=2.0, seed=2)
    n = grid ** 3
    pot = ops.anderson_potential(n, 4.0, 1)
    return ops.laplacian_3d(grid, bc="periodic", potential=pot)


def traffic_from_profiles(precision="fp64"):
    """HBM bytes per Chebyshev-step launch from the committed rocprofv3 --pmc summary."""
    name = f"cheb_step_{WL['profile']}_" + ("f32src" if precision == "mixed" else "f64")
    path = os.path.join(ROOT, "profiles", name + "_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d.get("hbm_bytes_per_launch")


def cpu_baseline(A, gpu_counts, max_inner=1):
    """Oracle on a bounded sample of the same solve, extrapolated with the GPU counts."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t0 = time.time()
    r = orc.solve(A, K, TOL, INC, BUF, seed=0, stop="RESIDUAL_ONLY", max_inner=max_inner)
    wall = time.time() - t0
    t = r["timing"]
    per_deg = t["filter"] / max(t["degree_sum"], 1)
    per_ortho = t["ortho"] / max(t["n_ortho"], 1)
    per_defl = t["deflate"] / max(t["n_deflate"], 1)
    per_rr = t["rr"] / max(t["n_rr"], 1)
    est = (t["lanczos"] + per_deg * gpu_counts["degrees"] + per_ortho * gpu_counts["ortho"]
           + (per_defl + per_rr) * gpu_counts["rr"])
    return {
        "value": K / est,
        "unit": "eigenpairs/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"CPU oracle (restated RayleighChebyshevLM, OpenMP, {threads} threads) on the "
                   f"same operator ({WL['label']}): Lanczos bounds + {max_inner} inner iteration(s) per outer "
                   f"pass, {wall:.1f} s wall; per-degree filter {per_deg:.3f} s, per-MGS "
                   f"{per_ortho:.3f} s, per-deflation {per_defl:.3f} s, per-Rayleigh-Ritz "
                   f"{per_rr:.3f} s, scaled by the GPU solve's {gpu_counts['degrees']} filter "
                   f"degrees, {gpu_counts['ortho']} orthogonalisations and {gpu_counts['rr']} "
                   f"Rayleigh-Ritz steps to an estimated {est:.1f} s solve"),
    }


def step_ceiling(nx=GRID, iters=10):
    """The FP64 step's memory pattern without the CSR operator (bench_ceiling.hip):
    the same row reads and writes of the 64-column block alone ("stream") and
    with the 7-point periodic gathers computed from the row index ("stencil"),
    HIP-event timed.  Rates are the block streams' 3 n b 8 bytes per launch."""
    import ctypes as C
    path = os.path.join(ROOT, "bench_ceiling.so")
    if not os.path.exists(path):
        return None
    L = C.CDLL(path)
    L.ceiling_ms.restype = C.c_double
    L.ceiling_ms.argtypes = [C.c_int, C.c_int64, C.c_int]
    n = nx ** 3
    out = {"block_bytes_per_launch": 3 * n * 64 * 8}
    for kind, name in ((0, "stream"), (1, "stencil")):
        ms = L.ceiling_ms(kind, nx, iters)
        if ms <= 0:
            return None
        out[name] = {"us_per_launch": round(ms * 1e3, 1),
                     "GBps": round(3 * n * 64 * 8 / (ms * 1e-3) / 1e9, 1)}
    return out


def cpu_model_check(grid=32):
    """The cpu_baseline model against a whole timed oracle solve on a smaller
    operator of the same family ({grid}^3 periodic Anderson, k 64, inc 32,
    buf 32): the model's per-phase unit costs from a bounded sample (Lanczos +
    one inner iteration per outer pass) times the whole solve's own counts,
    next to the solve's measured wall time."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc
    A = build_operator(grid)
    t0 = time.time()
    full = orc.solve(A, K, TOL, INC, BUF, seed=0, stop="RESIDUAL_ONLY")
    wall_full = time.time() - t0
    f = full["timing"]
    t1 = time.time()
    smp = orc.solve(A, K, TOL, INC, BUF, seed=0, stop="RESIDUAL_ONLY", max_inner=1)
    wall_sample = time.time() - t1
    t = smp["timing"]
    est = (t["lanczos"] + t["filter"] / max(t["degree_sum"], 1) * f["degree_sum"]
           + t["ortho"] / max(t["n_ortho"], 1) * f["n_ortho"]
           + (t["deflate"] / max(t["n_deflate"], 1) + t["rr"] / max(t["n_rr"], 1)) * f["n_rr"])
    return {"operator": f"{grid}^3 periodic Laplacian + Anderson U[-2,2] (seed 1), k={K} "
                        f"inc={INC} buf={BUF}",
            "measured_solve_s": round(wall_full, 2), "model_estimate_s": round(est, 2),
            "model_over_measured": round(est / wall_full, 3), "sample_s": round(wall_sample, 2),
            "found": int(full["found"]), "filter_degrees": int(f["degree_sum"])}


def cpu_baseline_dense(A, gpu_counts, b):
    """The dense workload on the host cores: each filter degree is one n x n by
    n x b product, timed with numpy's BLAS on the full block (two products),
    the block orthonormalisation with the oracle's MGS; both scaled by the GPU
    solve's counts (degrees, products of the Rayleigh-Ritz steps, Lanczos
    matrix-vector products, orthonormalisations)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    n = A.shape[0]
    A = np.ascontiguousarray(A)  # symmetric: row- and column-major agree
    rng = np.random.default_rng(0)
    X = rng.uniform(-1.0, 1.0, (n, b))
    t0 = time.time()
    for _ in range(2):
        X = A @ X
        X /= np.linalg.norm(X, axis=0)
    per_deg = (time.time() - t0) / 2
    t1 = time.time()
    orc.mgs(X)
    per_ortho = time.time() - t1
    products = gpu_counts["degrees"] + 2 * gpu_counts["rr"]
    est = (per_deg * products + per_ortho * gpu_counts["ortho"]
           + per_deg / b * gpu_counts["lanczos"])
    return {
        "value": K / est,
        "unit": "eigenpairs/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"host cores ({threads} BLAS threads): two n x n by n x {b} products with "
                   f"numpy ({per_deg:.2f} s each) and the oracle's MGS of the {n} x {b} block "
                   f"({per_ortho:.2f} s), scaled by the GPU solve's {products} block products, "
                   f"{gpu_counts['ortho']} orthonormalisations and {gpu_counts['lanczos']} "
                   f"Lanczos steps to an estimated {est:.0f} s solve"),
    }


def kernel_only_dense(args):
    """The dense workload's FP64 (or, with --kernel-mixed, FP32 correction)
    MFMA filter steps alone, on its 256-column block, for rocprofv3 counter
    runs.  The matrix comes from the cache file dense_operator_host writes, so
    run the dense workload once (without a profiler) first."""
    from rayleighchebyshevgpgpu_amd import capi, solver
    Ad, _ = dense_operator_host(args.dense_n)
    ctx = solver.Context(0)
    op = solver.DenseOperator(np.asarray(Ad), ctx)
    b = INC + BUF
    X = solver.Block(op.n, b, ctx=ctx)
    T = solver.Block(op.n, b, ctx=ctx)
    X.randomize(0, b, 0, 0)
    m = args.kernel_only + (args.kernel_only % 2)
    theta = np.full(b, 0.01)
    ritz = np.ones(b, dtype=np.int32)

    def run():
        if args.kernel_mixed:
            capi.check(capi.lib().rcg_cheb_filter_mixed(op.h, X.h, T.h, b, m, 1, 1.001, -0.001,
                                                        theta.ctypes.data, ritz.ctypes.data))
        else:
            capi.check(capi.lib().rcg_cheb_filter(op.h, X.h, T.h, b, m, 1, 1.001, -0.001))
        ctx.synchronize()
    run()  # warm-up (FP32 copy of A for the mixed steps)
    t0 = time.perf_counter()
    run()
    dt = time.perf_counter() - t0
    nn = float(op.n)
    print(json.dumps({"mode": "kernel-only", "workload": "dense16384",
                      "precision": "mixed" if args.kernel_mixed else "fp64", "cheb_steps": m,
                      "n": op.n, "b": b, "flops_per_step": 2.0 * nn * nn * b,
                      "ms_per_step": round(dt / m * 1e3, 4),
                      "tflops": round(2.0 * nn * nn * b * m / dt / 1e12, 2)}), flush=True)


def kernel_only(args):
    """Filter steps alone (for rocprofv3 counter runs) plus a device-to-device
    copy of the 64-column block (a plain copy kernel of known size) whose
    FETCH_SIZE/WRITE_SIZE calibrate the counters."""
    from rayleighchebyshevgpgpu_amd import capi, solver
    if WL.get("dense"):
        return kernel_only_dense(args)
    A = build_operator(args.grid)
    ctx = solver.Context(0)
    op = solver.CsrOperator(A, ctx)
    X = solver.Block(A.n, K, ctx=ctx)
    T = solver.Block(A.n, K, ctx=ctx)
    X.randomize(0, K, 0, 0)
    m = args.kernel_only + (args.kernel_only % 2)
    if args.kernel_mixed:
        # all columns as Ritz columns: the FP32 correction step with its source block
        theta = np.full(K, -0.3)
        ritz = np.ones(K, dtype=np.int32)
        capi.check(capi.lib().rcg_cheb_filter_mixed(op.h, X.h, T.h, K, m, 1, 14.2, 2.0,
                                                    theta.ctypes.data, ritz.ctypes.data))
    else:
        capi.check(capi.lib().rcg_cheb_filter(op.h, X.h, T.h, K, m, 1, 14.2, 2.0))
    capi.check(capi.lib().rcg_block_copy(T.h, X.h))
    # column norms + normalisation: col_sumsq reads the block once (n*ld*8
    # bytes), normalize_cols reads and writes it once -- known-size kernels of
    # the same access type as the filter step's own-row streams
    capi.check(capi.lib().rcg_deflate(X.h, K, T.h, 0))
    ctx.synchronize()
    csr = A.nnz * 12 + 4 * (A.n + 1)
    blk = A.n * X.ld * 8
    if args.kernel_mixed:
        step = A.nnz * 8 + 4 * (A.n + 1) + 4 * A.n * K * 4
    else:
        step = csr + 3 * A.n * K * 8
    print(json.dumps({"mode": "kernel-only", "precision": "mixed" if args.kernel_mixed else "fp64",
                      "cheb_steps": m, "n": A.n, "nnz": A.nnz, "b": K,
                      "alg_bytes_per_step": step,
                      "alg_bytes_first_step": csr + 2 * A.n * K * 8,
                      "copy_bytes_read": blk, "copy_bytes_written": blk}), flush=True)


def serving_cpu_baseline(reqs, inc_buf, budget_s=15.0):
    """The oracle (FP64, OpenMP) on the stream's requests in order until
    budget_s of CPU solving has passed."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    t0 = time.time()
    found = done = 0
    for r in reqs:
        inc, buf = inc_buf(r.k)
        o = orc.solve(r.operator(), r.k, r.tol, inc, buf, seed=0, stop="RESIDUAL_ONLY")
        found += int(o["found"])
        done += 1
        if time.time() - t0 > budget_s:
            break
    wall = time.time() - t0
    return {"value": round(found / wall, 3), "unit": "eigenpairs/s", "cores": threads,
            "kind": "port",
            "sample": (f"CPU oracle (restated RayleighChebyshevLM, OpenMP, {threads} threads), the "
                       f"first {done} requests of the same stream one by one, FP64: {found} "
                       f"eigenpairs in {wall:.1f} s (operator construction included)")}


def serving_run(args, cpu=True, cpu_budget_s=15.0):
    """Batched serving: the request stream through rcg_server (mixed precision,
    continuous batching) against the plain FP64 path solving requests one by
    one.  Operators are uploaded before the timed region, one sparsity pattern
    per grid shape with each request's own values.  Returns the measurements."""
    from rayleighchebyshevgpgpu_amd import serving, solver
    reqs = serving.serving_requests(args.requests, rate=args.rate, seed=4)
    ctx = solver.Context(0)
    t0 = time.time()
    ops_ = []
    cache = serving.OperatorCache(ctx)
    for i, r in enumerate(reqs):
        ops_.append(cache.operator(r))
        if i % 1024 == 1023:
            log(f"operators built: {i + 1}/{len(reqs)} ({time.time() - t0:.0f}s)")
    log(f"{len(reqs)} operators on the device in {time.time() - t0:.1f}s "
        f"(total n {sum(r.n for r in reqs)})")

    def inc_buf(k):
        return k, max(4, k // 2)

    def serve(saturate, count=None, scale=1.0):
        """One pass of the stream through a fresh server: the first `count`
        requests (all by default), arrival times multiplied by `scale` (a
        Poisson stream at rate/scale), or all queued at once when saturate."""
        srv = serving.Server(ctx, hbm_budget=args.hbm_budget << 30, max_active=args.slots,
                             window_us=args.window_us)
        import resource
        sel = reqs[:count] if count else reqs
        tickets = []
        ru0 = resource.getrusage(resource.RUSAGE_SELF)
        start = time.perf_counter()
        for r, op in zip(sel, ops_):
            if not saturate:
                delay = scale * r.arrival - (time.perf_counter() - start)
                if delay > 0:
                    time.sleep(delay)
            inc, buf = inc_buf(r.k)
            tickets.append(srv.submit(op, r.k, r.tol, inc, buf, precision="mixed"))
        outs = [srv.wait(t) for t in tickets]
        makespan = time.perf_counter() - start
        ru1 = resource.getrusage(resource.RUSAGE_SELF)
        cpu_s = (ru1.ru_utime - ru0.ru_utime) + (ru1.ru_stime - ru0.ru_stime)
        st = srv.stats()
        fp = srv.filter_profile()
        # latency = completion - arrival, on the library's steady clock
        t_first = min(o["t_submit"] for o in outs)
        arrival0 = t_first - (0.0 if saturate else scale * sel[0].arrival)
        lat = np.array([o["t_done"] - (arrival0 + (0.0 if saturate else scale * r.arrival))
                        for o, r in zip(outs, sel)]) * 1e3
        bad = [i for i, o in enumerate(outs) if o["status"] != 0 or o["found"] < sel[i].k]
        found = sum(o["found"] for o in outs)
        # per-request host wall time by phase (admitted to done), averaged
        res = [o["result"] for o in outs]
        phases = {k: round(1e3 * float(np.mean([getattr(r, k) for r in res])), 2)
                  for k in ("t_total", "t_lanczos", "t_filter", "t_ortho", "t_rr")}
        phases.update({k: round(float(np.mean([getattr(r, k) for r in res])), 1)
                       for k in ("lanczos_iters", "count_eig", "filter_degree_sum")})
        for t in tickets:
            srv.release(t)
        srv.close()
        return {"requests": len(sel), "eigenpairs": found, "makespan_s": round(makespan, 3),
                "value": round(found / makespan, 2), "p50_ms": round(float(np.percentile(lat, 50)), 2),
                "p99_ms": round(float(np.percentile(lat, 99)), 2), "failed": len(bad),
                "filter_batches": st["filter_batches"],
                "requests_per_batch": round(st["batched_requests"] / max(st["filter_batches"], 1), 2),
                "hbm_peak_gb": round(st["hbm_peak"] / 2**30, 2), "per_request_mean": phases,
                "host_cpu_s": round(cpu_s, 2),
                "shared_filter": {"device_s": round(fp["device_s"], 3), "launches": fp["launches"],
                                  "alg_GBps": round(fp["alg_bytes"] / max(fp["device_s"], 1e-9) / 1e9, 1),
                                  "busy_fraction": round(fp["device_s"] / makespan, 3)}}

    offered = saturated = light = None
    if not args.plain_only:
        log("serving (Poisson arrivals) ...")
        offered = serve(False)
        log(f"offered-load run: {offered}")
        saturated = serve(True) if args.saturate else None
        if saturated:
            log(f"saturated run: {saturated}")
        if args.light_requests > 0 and args.light_rate > 0:
            # the same stream's first requests at a rate the server keeps up
            # with: latency is then service time rather than queueing
            light = serve(False, args.light_requests, args.rate / args.light_rate)
            light["rate"] = args.light_rate
            log(f"light-load run: {light}")
    # plain FP64 path: the first requests one by one through rcg_solve_lowest
    m = min(args.plain_requests, len(reqs))
    tp = time.perf_counter()
    found_plain = 0
    for r, op in zip(reqs[:m], ops_[:m]):
        rc = solver.RayleighChebyshevLM()
        rc.setStopCondition("RESIDUAL_ONLY")
        inc, buf = inc_buf(r.k)
        ev = []
        found_plain += rc.getMinEigenSystem(r.k, r.tol, inc, buf, None, op, solver.RandomizeOp(0),
                                            ev, None)
    ctx.synchronize()
    plain_s = time.perf_counter() - tp
    plain = found_plain / plain_s
    plain_rec = {"requests": m, "eigenpairs": found_plain, "seconds": round(plain_s, 3),
                 "value": round(plain, 2)}
    log(f"plain FP64: {m} requests in {plain_s:.3f} s, {plain:.1f} eigenpairs/s")
    if args.plain_only:
        return {"plain_fp64": plain_rec}
    best = saturated or offered
    sf = best["shared_filter"]
    return {
        "value": best["value"], "unit": "eigenpairs/s",
        "ms_per_step": round(best["makespan_s"] * 1e3, 1),
        "workload": (f"batched serving: {len(reqs)} requests, Poisson arrivals at {args.rate:g}/s "
                     "(seed 4), 2-D/3-D Dirichlet Laplacians n log-uniform in [1e3, 2.56e5] + "
                     "U[-W/2,W/2], W in [0,8], k in {8,16,32}, tol 1e-10, mixed precision"),
        "slots": args.slots, "window_us": args.window_us, "hbm_budget_gb": args.hbm_budget,
        "p50_ms": offered["p50_ms"], "p99_ms": offered["p99_ms"],
        "offered_load": offered, "saturated": saturated, "light_load": light,
        "plain_fp64": plain_rec,
        "speedup_vs_plain_fp64": round(best["value"] / plain, 2) if plain > 0 else None,
        # the shared filter launches: algorithmic bytes over their device time
        "roofline": {"bound": "hbm", "achieved": sf["alg_GBps"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(sf["alg_GBps"] / HBM_PEAK_GBS, 4),
                     "traffic": None,
                     "kernel": "csr_batch_kernel<float,...,SRC> (batched FP32 correction steps)",
                     "launches": sf["launches"], "device_s": sf["device_s"]},
        "cpu_baseline": (serving_cpu_baseline(reqs, inc_buf, cpu_budget_s) if cpu else None),
    }


def batch_filter_bench(args):
    """--workload batchfilter: the shared filter launches alone.  The first
    --requests requests of the served stream (their operators, k + max(4, k/2)
    FP64 columns) filtered together by rcg_cheb_filter_batch, degree 24 each,
    HIP-synchronised wall time per call; the roofline is the requests' summed
    algorithmic step bytes (CSR + 3 block passes per step) over that time."""
    import ctypes as C
    from rayleighchebyshevgpgpu_amd import capi, serving, solver
    reqs = serving.serving_requests(args.requests, rate=args.rate, seed=4)
    ctx = solver.Context(0)
    cache = serving.OperatorCache(ctx)
    m = 24
    held = []
    for r in reqs:
        op = cache.operator(r)
        cols = r.k + max(4, r.k // 2)
        X, T = solver.Block(r.n, cols, ctx=ctx), solver.Block(r.n, cols, ctx=ctx)
        X.randomize(0, cols, 1, 0)
        held.append((r, op, X, T, cols))
    n = len(held)
    arr = lambda t, v: (t * n)(*v)  # noqa: E731
    argv = (n, arr(C.c_void_p, [h[1].h.value for h in held]), arr(C.c_void_p, [h[2].h.value for h in held]),
            arr(C.c_void_p, [h[3].h.value for h in held]), arr(C.c_int64, [h[4] for h in held]),
            arr(C.c_int64, [m] * n), arr(C.c_double, [4.0 * h[0].dim + 4.1 for h in held]),
            arr(C.c_double, [0.0] * n))
    L = capi.lib()
    for _ in range(args.warmup):
        capi.check(L.rcg_cheb_filter_batch(*argv))
    t0 = time.perf_counter()
    for _ in range(args.steps):
        capi.check(L.rcg_cheb_filter_batch(*argv))
    dt = (time.perf_counter() - t0) / args.steps
    alg = sum(m * (h[1].nnz * 12 + 4 * (h[0].n + 1) + 3 * h[0].n * h[4] * 8) for h in held)
    rows = sum(h[0].n for h in held)
    line = {"workload": f"shared filter launches: {n} served requests ({rows} rows), FP64, degree {m}",
            "ms_per_call": round(dt * 1e3, 3), "steps_per_call": m,
            "alg_GBps": round(alg / dt / 1e9, 1), "frac_of_8TBps": round(alg / dt / 8e12, 4)}
    print(json.dumps(line), flush=True)


def serving_bench(args):
    """--workload serving: the serving measurements as the bench line."""
    r = serving_run(args, cpu=not args.no_cpu)
    if args.plain_only:
        print(json.dumps(r), flush=True)
        return
    line = {
        "metric": METRIC,
        "value": r["value"],
        "unit": "eigenpairs/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 0,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 filter steps (correction form), f64 elsewhere",
        "data": "synthetic",
        "config": {"workload": r["workload"], "slots": r["slots"], "window_us": r["window_us"],
                   "hbm_budget_gb": r["hbm_budget_gb"]},
    }
    line.update({k: r[k] for k in ("p50_ms", "p99_ms", "offered_load", "saturated", "light_load",
                                   "plain_fp64", "speedup_vs_plain_fp64", "roofline",
                                   "cpu_baseline")})
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, default=GRID)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=1)
    ap.add_argument("--workload", default="anderson128", choices=sorted(WORKLOADS) + ["serving", "batchfilter"])
    ap.add_argument("--dense-n", type=int, default=16384, help="dense16384: matrix size")
    ap.add_argument("--requests", type=int, default=4096, help="serving: stream length")
    ap.add_argument("--rate", type=float, default=200.0, help="serving: arrivals per second")
    ap.add_argument("--slots", type=int, default=64, help="serving: concurrent solves")
    ap.add_argument("--window-us", type=int, default=300, help="serving: filter batching window")
    ap.add_argument("--hbm-budget", type=int, default=200, help="serving: GiB for admitted solves")
    ap.add_argument("--saturate", action="store_true",
                    help="serving: also run with every request queued at once (capacity)")
    ap.add_argument("--light-requests", type=int, default=512,
                    help="serving: requests in the light-load latency run (0 skips it)")
    ap.add_argument("--light-rate", type=float, default=40.0,
                    help="serving: arrivals per second of the light-load run")
    ap.add_argument("--plain-only", action="store_true",
                    help="serving: only the plain FP64 path (one request at a time)")
    ap.add_argument("--plain-requests", type=int, default=256,
                    help="serving: requests solved one by one on the plain FP64 path")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision solves")
    ap.add_argument("--no-serving", action="store_true",
                    help="anderson128: skip the serving workload reported under 'serving'")
    ap.add_argument("--kernel-only", type=int, default=0, metavar="M",
                    help="profiling mode: run M Chebyshev steps of the filter on the 64-column "
                         "block and one 2 GiB device copy (the HBM calibration kernel), no solve")
    ap.add_argument("--kernel-mixed", action="store_true",
                    help="with --kernel-only: the FP32 correction steps of the mixed filter")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the solve runs "
              "on a single MI355X; multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    global K, INC, BUF, TOL, WL
    if args.workload == "serving":
        return serving_bench(args)
    if args.workload == "batchfilter":
        return batch_filter_bench(args)
    WL = WORKLOADS[args.workload]
    K, INC, BUF, TOL = WL["k"], WL["inc"], WL["buf"], WL["tol"]
    from rayleighchebyshevgpgpu_amd import capi, solver

    if args.kernel_only:
        if WL["complex"]:
            print("--kernel-only profiles the real workload only", file=sys.stderr)
            sys.exit(2)
        return kernel_only(args)

    t0 = time.time()
    if WL.get("dense"):
        Ad, lam_exact = dense_operator_host(args.dense_n)  # before this process uses the GPU
        ctx = solver.Context(0)
        op = solver.DenseOperator(np.asarray(Ad), ctx)
        A = op
        log(f"dense operator n={args.dense_n} built in {time.time() - t0:.1f}s")
    else:
        ctx = solver.Context(0)
        A = build_operator(args.grid)
        lam_exact = None
        log(f"operator n={A.n} nnz={A.nnz} built in {time.time() - t0:.1f}s")
        op = solver.CsrOperator(A, ctx)
    bdt = np.complex128 if WL["complex"] else np.float64
    V = solver.Block(A.n, max(K, INC + BUF), dtype=bdt, ctx=ctx)

    def run(precision, steps, warmup):
        """warmup + timed solves at one filter precision; returns the measurements."""
        def one_solve():
            rc = solver.RayleighChebyshevLM()
            rc.setStopCondition("RESIDUAL_ONLY")
            rc.filterPrecision = np.float32 if precision == "mixed" else np.float64
            rc.refine = True  # the final FP64 Rayleigh-Ritz sweep over the found pairs
            ev = []
            t = time.perf_counter()
            f = rc.getMinEigenSystem(K, TOL, INC, BUF, None, op, solver.RandomizeOp(0), ev, V)
            ctx.synchronize()
            return time.perf_counter() - t, f, np.array(ev), rc

        for w in range(warmup):
            dt, f, _, _ = one_solve()
            log(f"{precision} warmup {w}: {dt:.3f}s found={f}")
        ctx.synchronize()
        lat, results = [], []
        t_start = time.perf_counter()
        for s in range(steps):
            dt, f, ev, rc = one_solve()
            lat.append(dt)
            results.append((f, ev, rc.getResult()))
            log(f"{precision} step {s}: {dt:.3f}s found={f} iters={rc.getResult().count_eig} "
                f"degrees={rc.getResult().filter_degree_sum}")
        ctx.synchronize()
        total = time.perf_counter() - t_start

        # residual check of the last solve's pairs on the device (outside the timed region)
        f, ev, res = results[-1]
        r = np.zeros(f)
        capi.check(capi.lib().rcg_operator_apply(op.h, V.h, Wb.h, 1.0, 0.0, f))
        capi.check(capi.lib().rcg_residuals(op.h, V.h, Wb.h, f, ev.ctypes.data_as(C.c_void_p),
                                            r.ctypes.data_as(C.c_void_p)))
        normA = max(abs(res.min_est), abs(res.max_est))
        rel = r / normA
        converged_per_solve = int(np.sum(rel <= 1e-10))
        converged_total = sum(int(x[0]) for x in results) * converged_per_solve / max(f, 1)

        # dominant kernel: the Chebyshev step, HIP-event timed inside the solves
        launches = sum(x[2].filter_launches for x in results)
        t_dev = sum(x[2].t_filter_device for x in results)
        alg_bytes = sum(x[2].filter_bytes for x in results)
        achieved = alg_bytes / t_dev / 1e9 if t_dev > 0 else 0.0
        if WL["complex"]:
            kern = ("csr_block_kernel<float,...,kChebSrc,complex> (complex64 correction step)"
                    if precision == "mixed" else
                    "csr_block_kernel<double,...,kCheb,complex> (fused complex128 Chebyshev step)")
        else:
            kern = ("csr_block_kernel<float,8,2,kChebSrc> (FP32 correction step)"
                    if precision == "mixed" else
                    "csr_block_kernel<double,16,2,kCheb> (fused FP64 Chebyshev step)")
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": traffic_from_profiles(precision), "kernel": kern,
                    "avg_launch_us": round(t_dev / max(launches, 1) * 1e6, 2),
                    "alg_bytes_per_launch": int(alg_bytes / max(launches, 1)),
                    "launches": int(launches)}
        if WL.get("dense"):
            # dense steps are MFMA-bound: 2 n^2 b flops per launch, b recovered from
            # the step's byte count (FP64: 8n^2 + 24nb; FP32 correction: 4n^2 + 16nb)
            nn = float(A.n)
            bpl = alg_bytes / max(launches, 1)
            b_eff = (bpl - 4 * nn * nn) / (16 * nn) if precision == "mixed" else \
                (bpl - 8 * nn * nn) / (24 * nn)
            flops = 2.0 * nn * nn * b_eff * launches
            tf = flops / t_dev / 1e12 if t_dev > 0 else 0.0
            peak = 157.3 if precision == "mixed" else 78.6
            roofline = {"bound": "mfma", "achieved": round(tf, 2), "peak": peak,
                        "unit": "TFLOP/s", "frac": round(tf / peak, 4), "traffic": None,
                        "kernel": ("dense_step_kernel<float,kChebSrc> (FP32 MFMA 16x16x4)"
                                   if precision == "mixed" else
                                   "dense_step_kernel<double,kCheb> (FP64 MFMA 16x16x4)"),
                        "avg_launch_us": round(t_dev / max(launches, 1) * 1e6, 2),
                        "alg_flops_per_launch": int(2.0 * nn * nn * b_eff),
                        "launches": int(launches)}
        last = results[-1][2]
        lat_ms = np.array(lat) * 1e3
        return {
            "value": converged_total / total, "total": total, "lat_ms": lat_ms, "f": f, "ev": ev,
            "last": last, "roofline": roofline,
            "solve": {"found": int(f), "converged": converged_per_solve,
                      "max_rel_residual": float(rel.max()) if f else None,
                      "inner_iterations": int(last.count_eig),
                      "filter_degrees": int(last.filter_degree_sum),
                      "lanczos_iters": int(last.lanczos_iters),
                      "t_lanczos_s": round(last.t_lanczos, 4),
                      "t_filter_steps_device_s": round(last.t_filter_device, 4),
                      "t_ortho_s": round(last.t_ortho, 4), "t_rr_s": round(last.t_rr, 4),
                      "t_rr_host_eig_s": round(last.t_eig, 4),
                      "lambda_min": float(ev[0]) if f else None,
                      "lambda_k": float(ev[-1]) if f else None},
            "p50_ms": round(float(np.percentile(lat_ms, 50)), 2),
            "p99_ms": round(float(np.percentile(lat_ms, 99)), 2),
        }

    import ctypes as C
    Wb = solver.Block(A.n, max(K, INC + BUF), dtype=bdt, ctx=ctx)
    fp64 = run("fp64", args.steps, args.warmup)
    mixed = None if args.no_mixed else run("mixed", args.steps, args.warmup)
    if mixed is not None:
        ev64 = fp64["ev"]
        floor = np.maximum(np.abs(ev64), 1e-11 / TOL)
        n_cmp = min(len(ev64), len(mixed["ev"]))
        mixed["max_rel_eig_diff_vs_fp64"] = float(np.max(
            np.abs(mixed["ev"][:n_cmp] - ev64[:n_cmp]) / floor[:n_cmp])) if n_cmp else None

    if lam_exact is not None:  # the dense operator's spectrum is known exactly
        for r in (fp64, mixed):
            if r is None or not len(r["ev"]):
                continue
            ref = lam_exact[: len(r["ev"])]
            r["solve"]["max_rel_err_vs_known_lambda"] = float(np.max(
                np.abs(r["ev"] - ref) / np.maximum(np.abs(ref), 1e-11 / TOL)))
    total, f, ev, last = fp64["total"], fp64["f"], fp64["ev"], fp64["last"]
    value, roofline = fp64["value"], fp64["roofline"]
    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "eigenpairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(total / args.steps * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128" if WL["complex"] else "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"lowest-eigenpair solve, {WL['label']}, k={K} inc={INC} buf={BUF} "
                         f"tol={TOL:g} residual-only, FP64 filter"),
            "n": A.n, "nnz": A.nnz, "k": K, "inc": INC, "buf": BUF, "tol": TOL,
            "precision": "fp64",
        },
        "p50_ms": fp64["p50_ms"],
        "p99_ms": fp64["p99_ms"],
        "solve": fp64["solve"],
        "roofline": roofline,
        "mixed": None if mixed is None else {
            "value": round(mixed["value"], 3), "unit": "eigenpairs/s",
            "speedup_vs_fp64": round(mixed["value"] / value, 3) if value > 0 else None,
            "ms_per_step": round(mixed["total"] / args.steps * 1e3, 2),
            "p50_ms": mixed["p50_ms"], "p99_ms": mixed["p99_ms"],
            "max_rel_eig_diff_vs_fp64": mixed["max_rel_eig_diff_vs_fp64"],
            "dtype": (("c64" if WL["complex"] else "f32") + " filter steps, f64 start/"
                      "recombination, orthonormalisation, Rayleigh-Ritz and refinement"),
            "solve": mixed["solve"], "roofline": mixed["roofline"]},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        log("cpu baseline sample ...")
        counts = {"degrees": int(last.filter_degree_sum), "ortho": int(last.count_ortho),
                  "rr": int(last.count_eig)}
        if WL.get("dense"):
            counts["lanczos"] = int(last.lanczos_iters)
            line["cpu_baseline"] = cpu_baseline_dense(Ad, counts, INC + BUF)
        else:
            line["cpu_baseline"] = cpu_baseline(A, counts, args.cpu_iters)
            if args.workload == "anderson128":
                log("cpu baseline model check on 32^3 ...")
                chk = cpu_model_check(32)
                line["cpu_baseline"]["model_check"] = chk
                line["cpu_baseline"]["sample"] += (
                    f"; the same model on a whole {chk['operator']} oracle solve estimates "
                    f"{chk['model_estimate_s']} s against {chk['measured_solve_s']} s measured "
                    f"(x{chk['model_over_measured']})")
    if args.workload == "anderson128":
        log("step-pattern ceiling kernels ...")
        line["roofline"]["ceiling"] = step_ceiling(args.grid)
    if args.workload == "anderson128" and not args.no_serving:
        # the batched serving workload beside the single-operator solve: the
        # mixed-precision served stream against the plain FP64 path
        log("serving workload ...")
        line["serving"] = serving_run(args, cpu=not args.no_cpu, cpu_budget_s=10.0)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
