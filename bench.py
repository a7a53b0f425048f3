"""Benchmark: PCG solves/s to 1e-8 relative residual with the rank-structured
preconditioner resident in HBM, mixed (FP32 dense factor + BF16 sketches)
against the FP64 factor, on one MI355X.

Default workload (poisson3d_128): 3D 7-point variable-coefficient Poisson,
128^3 cells (n = 2,097,152), log-normal face coefficients (mu 0, sigma 2,
seed 1234), Dirichlet boundary; separators with more than 255 unknowns
compressed at the reference's rank heuristic with oversampling 10; 16
uniform[-1,1] right-hand sides (seed 5678); FP64 block PCG to 1e-8.

A step = one block PCG solve of the 16 right-hand sides (each column to
recurrence relres <= 1e-8) with the factor and all vectors already in HBM.
`value` is the mixed factor's solves/s (right-hand sides per second); the
FP64 factor's is reported beside it, with both factorization times, both
end-to-end times (symbolic + numeric factor + one solve) and their ratio,
and each solution's relative error against an FP64 direct solve (the exact
Cholesky factor on the device, no compression, refined by PCG to 1e-14).

Prints ONE JSON line.  `python bench.py --gpus N` with N > 1 exits: every
configuration here runs on one GPU.
"""
import argparse
import ctypes as C
import json
import os
import statistics
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

CONFIGS = {
    # the mixed-precision + level-batched form at 128^3 (headline)
    "poisson3d_128": dict(N=128, sigma=2.0, max_dense_block=255, max_rank=0, oversampling=10,
                          nrhs=16, cpu_sample_N=48),
    # plain FP64 GPU form against the CPU path
    "poisson3d_64": dict(N=64, sigma=1.0, max_dense_block=255, max_rank=64, oversampling=5,
                         nrhs=1, cpu_sample_N=48),
    # plain method, 2D
    "poisson2d_256": dict(N2=256, max_dense_block=127, max_rank=32, oversampling=5, nrhs=1,
                          cpu_sample_N=256),
    # mixed multi-right-hand-side elasticity
    "elasticity_48": dict(NE=48, sigma=1.0, nu=0.3, max_dense_block=255, max_rank=96,
                          oversampling=5, nrhs=64, cpu_sample_N=16),
    # request serving over eight resident factors (continuous batching)
    "serving": dict(grids=(32, 48, 64, 80, 96, 96, 112, 128), sigma=1.0,
                    max_dense_block=255, max_rank=64, oversampling=5, requests=24,
                    max_rhs=256, rate=20.0, slots=64),
}

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="poisson3d_128", choices=sorted(CONFIGS))
    ap.add_argument("--eps", type=float, default=1e-8)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64-factor leg")
    ap.add_argument("--no-direct", action="store_true",
                    help="skip the error against the exact (direct) factor")
    ap.add_argument("--requests", type=int, default=0,
                    help="serving: number of requests (default from the config)")
    ap.add_argument("--max-rhs", type=int, default=0,
                    help="serving: largest right-hand-side count of a request")
    ap.add_argument("--rate", type=float, default=0.0,
                    help="serving: mean request arrivals per second")
    ap.add_argument("--traffic-run", action="store_true",
                    help="for counter passes: factor, 3 applies, then 3 plain 256 MiB "
                         "copies as the calibration kernel; no timing line")
    ap.add_argument("--fp64-factor", action="store_true",
                    help="--traffic-run with the FP64 factor (default: the mixed factor)")
    return ap.parse_args()


def serving_leg(cfg, eps):
    """Eight log-normal Poisson systems factored and kept resident; a seeded
    Poisson stream of requests (matrix uniform, 1..max_rhs uniform[-1,1]
    right-hand sides each) served (a) by the continuous-batching server and
    (b) one request at a time in arrival order with block PCG (columns in
    chunks of the slot count).  Right-hand sides are drawn on the device
    before the clock starts.  Returns throughput (right-hand sides solved per
    second) and request latency percentiles of both."""
    import threading

    import torch

    import rank_structured_cholesky_amd as R
    from rank_structured_cholesky_amd import _lib
    L = _lib.lib()
    mats, facs, tf = [], [], []
    for i, N in enumerate(cfg["grids"]):
        A, xyz = R.poisson3d(N, cfg["sigma"], 1234 + i)
        M, t = factor(A, xyz, cfg, 0)
        inf = M.info()
        print("factored %d^3: n %d, %.2f s, factor %.2f GB, workspace %.2f GB" % (
            N, A._nrow, t, inf["factor_bytes"] / 1e9, inf["workspace_bytes"] / 1e9),
            file=sys.stderr, flush=True)
        mats.append(A)
        facs.append(M)
        tf.append(t)
    rng = np.random.default_rng(91011)
    nreq = cfg["requests"]
    gaps = rng.exponential(1.0 / cfg["rate"], nreq)
    arrive = np.cumsum(gaps) - gaps[0]
    which = rng.integers(0, len(facs), nreq)
    nrhs = rng.integers(1, cfg["max_rhs"] + 1, nreq)
    gen = torch.Generator(device="cuda").manual_seed(91011)
    reqs = []
    for q in range(nreq):
        n = mats[which[q]]._nrow
        B = torch.rand((int(nrhs[q]), n), generator=gen, device="cuda",
                       dtype=torch.float64) * 2 - 1
        reqs.append((int(which[q]), int(nrhs[q]), B, torch.zeros_like(B)))
    torch.cuda.synchronize()

    def pct(v, p):
        v = sorted(v)
        return v[min(len(v) - 1, int(round(p * (len(v) - 1))))]

    # (a) continuous batching
    srv = R.RequestServer(facs, eps=eps, max_iters=10000, slots=cfg["slots"])
    slots_given = [srv.slots(f) for f in range(len(facs))]
    tickets = [None] * nreq
    t0 = time.perf_counter()

    def client():
        for q in range(nreq):
            d = t0 + arrive[q] - time.perf_counter()
            if d > 0:
                time.sleep(d)
            f, m, B, X = reqs[q]
            tickets[q] = srv.submit(f, m, B.data_ptr(), B.shape[1], X.data_ptr(), X.shape[1])

    th = threading.Thread(target=client)
    th.start()
    lat_a, iters_a, worst = [], [], 0.0
    for q in range(nreq):
        while tickets[q] is None:
            time.sleep(0.0005)
        res = srv.wait(tickets[q], reqs[q][1])
        print("served request %d/%d (%d right-hand sides)" % (q + 1, nreq, reqs[q][1]),
              file=sys.stderr, flush=True)
        lat_a.append(res[0]["seconds"])
        iters_a += [r["iterations"] for r in res]
        worst = max(worst, max(r["true_relres"] for r in res))
    th.join()
    wall_a = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    # (b) one request at a time (FIFO, no overlap between requests)
    lat_b = []
    stream = torch.cuda.Stream()
    t0 = time.perf_counter()
    for q in range(nreq):
        d = t0 + arrive[q] - time.perf_counter()
        if d > 0:
            time.sleep(d)
        f, m, B, X = reqs[q]
        X.zero_()
        torch.cuda.synchronize()
        n = B.shape[1]
        for c0 in range(0, m, cfg["slots"]):
            k = min(cfg["slots"], m - c0)
            res = (_lib.PcgResult * k)()
            L.rsc_block_pcg_solve(facs[f].handle, k, C.c_void_p(B[c0:].data_ptr()), n,
                                  C.c_void_p(X[c0:].data_ptr()), n, eps, 10000, res,
                                  C.c_void_p(stream.cuda_stream))
        torch.cuda.synchronize()
        lat_b.append(time.perf_counter() - (t0 + arrive[q]))
        print("one at a time: request %d/%d" % (q + 1, nreq), file=sys.stderr, flush=True)
    wall_b = time.perf_counter() - t0
    total = int(nrhs.sum())
    return dict(
        rhs_total=total, requests=nreq, wall_s=wall_a, solves_per_s=total / wall_a,
        p50_ms=1e3 * pct(lat_a, 0.5), p99_ms=1e3 * pct(lat_a, 0.99),
        steps=st["steps"], max_busy_slots=st["max_busy"], slots_given=slots_given,
        workspace_gb_after_release=round(sum(M.info()["workspace_bytes"] for M in facs) / 1e9, 3),
        worst_true_relres=worst,
        iterations_mean=float(np.mean(iters_a)),
        one_at_a_time=dict(wall_s=wall_b, solves_per_s=total / wall_b,
                           p50_ms=1e3 * pct(lat_b, 0.5), p99_ms=1e3 * pct(lat_b, 0.99)),
        factor_s=[round(t, 3) for t in tf],
        n=[m._nrow for m in mats])


def traffic_run(cfg, precision):
    """The program the FETCH_SIZE / WRITE_SIZE counter passes profile: the
    apply dispatches (k_apply_df) and a copy kernel of known size."""
    import torch

    from rank_structured_cholesky_amd import _lib
    L = _lib.lib()
    A, xyz, B = build_problem(cfg)
    M, _ = factor(A, xyz, cfg, precision)
    n, k = B.shape
    dB = torch.from_numpy(np.ascontiguousarray(B.T)).to("cuda")
    dY = torch.empty_like(dB)
    stream = torch.cuda.current_stream()
    for _ in range(3):
        _lib.check(L.rsc_apply(M.handle, k, C.c_void_p(dB.data_ptr()), n,
                               C.c_void_p(dY.data_ptr()), n,
                               C.c_void_p(stream.cuda_stream)))
    _lib.check(L.rsc_apply_status(M.handle, C.c_void_p(stream.cuda_stream)))
    torch.cuda.synchronize()
    src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")  # 256 MiB
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"traffic_run": True, "apply_bytes": apply_bytes(M, A, k), "nrhs": k,
                      "factor": "FP64" if precision == 0 else "mixed"}), flush=True)


def build_problem(cfg, N=None):
    """The workload's matrix, positions and right-hand sides (n x nrhs);
    N overrides the grid size (the CPU baseline's sample)."""
    import rank_structured_cholesky_amd as R
    if "N2" in cfg:
        A, xyz = R.poisson2d(N or cfg["N2"])
    elif "NE" in cfg:
        A, xyz = R.elasticity3d(N or cfg["NE"], cfg["sigma"], cfg["nu"], 1234)
    else:
        A, xyz = R.poisson3d(N or cfg["N"], cfg["sigma"], 1234)
    B = R.uniform_rhs(A._nrow * cfg["nrhs"], 5678).reshape(cfg["nrhs"], A._nrow).T
    return A, xyz, np.asfortranarray(B)


def options(cfg, precision):
    """precision 0: FP64 factor; 1: the mixed form (FP32 dense factor, BF16
    range-finder products with FP32 accumulation); 2: the exact factor (no
    compression, FP64): a direct solver."""
    import rank_structured_cholesky_amd as R
    o = dict(max_rank=cfg["max_rank"], oversampling=cfg["oversampling"])
    if precision == 1:
        o.update(factor_precision=R.RSC_FP32, sketch_precision=R.RSC_BF16)
    return o


def factor(A, xyz, cfg, precision):
    import rank_structured_cholesky_amd as R
    M = R.FactorPreconditioner()
    mdb = 2 ** 30 if precision == 2 else cfg["max_dense_block"]
    t0 = time.perf_counter()
    M.initialize(A, xyz, mdb, 400, **options(cfg, precision))
    return M, time.perf_counter() - t0


def apply_bytes(M, A, k):
    """Algorithmic bytes of one apply with k right-hand sides: every stored
    factor element once per traversal (forward and backward) in the factor
    precision, plus 56 B per row per right-hand side of vector traffic."""
    vec = 56.0 * A._nrow
    return M.info()["apply_bytes"] - vec + vec * k


def gpu_leg(A, xyz, B, cfg, precision, steps, warmup, eps, keep_x=False):
    """Factor, then W untimed and K timed block PCG solves of B, then ten
    back-to-back applies timed with HIP events on the stream they run on."""
    import torch

    from rank_structured_cholesky_amd import _lib
    L = _lib.lib()
    M, t_factor = factor(A, xyz, cfg, precision)
    info = M.info()
    n, k = B.shape
    dB = torch.from_numpy(np.ascontiguousarray(B.T)).to("cuda")  # k x n rows = columns
    dX = torch.zeros_like(dB)
    stream = torch.cuda.Stream()
    stream.wait_stream(torch.cuda.current_stream())  # dB's upload
    per_step = []

    def solve():
        res = (_lib.PcgResult * k)()
        with torch.cuda.stream(stream):
            dX.zero_()  # ordered before the solve on the solve's own stream
        st = L.rsc_block_pcg_solve(M.handle, k, C.c_void_p(dB.data_ptr()), n,
                                   C.c_void_p(dX.data_ptr()), n, eps, 10000, res,
                                   C.c_void_p(stream.cuda_stream))
        _lib.check(st)  # RSC_NOT_CONVERGED fails the step
        per_step.append(([r.iterations for r in res], [r.true_relres for r in res]))

    t0 = time.perf_counter()
    solve()
    first_solve = time.perf_counter() - t0
    for _ in range(max(0, warmup - 1)):
        solve()
    torch.cuda.synchronize()
    lat = []
    t0 = time.perf_counter()
    for _ in range(steps):
        t1 = time.perf_counter()
        solve()
        lat.append(time.perf_counter() - t1)
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    iters, relres = per_step[-1]
    for its, _ in per_step:
        if its != iters:
            raise RuntimeError("iteration counts differ between steps: %s vs %s" % (its, iters))
    # preconditioner apply timed with events on the stream it runs on
    dY = torch.empty_like(dB)
    napp = 10
    with torch.cuda.stream(stream):
        L.rsc_apply(M.handle, k, C.c_void_p(dB.data_ptr()), n, C.c_void_p(dY.data_ptr()),
                    n, C.c_void_p(stream.cuda_stream))
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(napp):
            L.rsc_apply(M.handle, k, C.c_void_p(dB.data_ptr()), n,
                        C.c_void_p(dY.data_ptr()), n, C.c_void_p(stream.cuda_stream))
        e1.record(stream)
    e1.synchronize()
    _lib.check(L.rsc_apply_status(M.handle, C.c_void_p(stream.cuda_stream)))
    apply_s = e0.elapsed_time(e1) / 1e3 / napp
    out = dict(t_factor=t_factor, info=info, total=total, lat=lat, iters=iters,
               relres=relres, apply_s=apply_s, k=k, first_solve=first_solve,
               apply_bytes=apply_bytes(M, A, k))
    if keep_x:
        out["X"] = dX
    del M
    torch.cuda.synchronize()
    return out


def direct_error(A, xyz, B, cfg, sols):
    """Relative error of each solution column against an FP64 direct solve:
    the exact supernodal Cholesky factor on the device (no compression),
    refined by PCG with itself as preconditioner to 1e-14.  sols: name ->
    device array (k x n).  Returns max/median error per name and the direct
    solve's own figures."""
    import torch

    from rank_structured_cholesky_amd import _lib
    L = _lib.lib()
    t0 = time.perf_counter()
    E, t_factor = factor(A, xyz, cfg, 2)
    n, k = B.shape
    dB = torch.from_numpy(np.ascontiguousarray(B.T)).to("cuda")
    dXd = torch.zeros_like(dB)
    res = (_lib.PcgResult * k)()
    st = L.rsc_block_pcg_solve(E.handle, k, C.c_void_p(dB.data_ptr()), n,
                               C.c_void_p(dXd.data_ptr()), n, 1e-14, 20, res, None)
    if st not in (_lib.RSC_OK, _lib.RSC_NOT_CONVERGED):
        _lib.check(st)
    torch.cuda.synchronize()
    info = E.info()
    del E
    out = {"direct": dict(factor_s=round(t_factor, 3), factor_gb=round(info["factor_bytes"] / 1e9, 2),
                          refinement_iterations=max(r.iterations for r in res),
                          true_relres=max(r.true_relres for r in res),
                          wall_s=round(time.perf_counter() - t0, 3))}
    ref = dXd.double()
    nrm = torch.linalg.norm(ref, dim=1)
    for name, X in sols.items():
        e = (torch.linalg.norm(X - ref, dim=1) / nrm).cpu().numpy()
        out[name] = dict(max=float(e.max()), median=float(np.median(e)))
    del dXd, ref
    torch.cuda.synchronize()
    return out


def cpu_leg(cfg, eps, gpu_steps=3):
    """The oracle (FP64 CPU restatement of the reference method, built for
    this host with -march=native and OpenMP) on a bounded sample of the
    workload: the same options and right-hand-side count on a smaller grid
    (cfg["cpu_sample_N"]).  Timed: symbolic + numeric factor once, then a
    warm-up and five block solves (median), on all host threads; one
    column's solve on one thread (the columns are independent, so one
    thread's block solve is k times that).  The GPU runs the same sample in
    the same process for a same-size ratio."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from helpers import oracle_for, symbolic
    from oracle import build as obuild
    from oracle import oracle as OR
    native = obuild.build_native(os.path.join(tempfile.gettempdir(), "rsc_oracle_native"))
    OR.lib(native)
    threads = OR.max_threads()
    sN = cfg["cpu_sample_N"]
    A, xyz, B = build_problem(cfg, sN)
    k = B.shape[1]
    t0 = time.perf_counter()
    sym = symbolic(A, xyz, max_dense_block=cfg["max_dense_block"], max_rank=cfg["max_rank"],
                   oversampling=cfg["oversampling"])
    O = oracle_for(A, sym, max_rank=cfg["max_rank"], oversampling=cfg["oversampling"])
    t_factor = time.perf_counter() - t0
    O.block_pcg(B, eps, 10000)
    ts = []
    for _ in range(5):
        t1 = time.perf_counter()
        _, info = O.block_pcg(B, eps, 10000)
        ts.append(time.perf_counter() - t1)
    t_solve = statistics.median(ts)
    OR.set_threads(1)
    t1 = time.perf_counter()
    O.block_pcg(B[:, :1], eps, 10000)
    t_one = time.perf_counter() - t1
    OR.set_threads(threads)
    del O
    # the GPU on the same sample (mixed and FP64 factors)
    same = {}
    for prec, name in ((1, "mixed"), (0, "fp64")):
        g = gpu_leg(A, xyz, B, cfg, prec, gpu_steps, 1, eps)
        same[name] = dict(solves_per_s=round(gpu_steps * g["k"] / g["total"], 2),
                          factor_s=round(g["t_factor"], 4),
                          end_to_end_s=round(g["t_factor"] + g["total"] / gpu_steps, 4))
    kind = "-march=native" if native else "-march=x86-64-v3 (native build failed)"
    cpu_e2e = t_factor + t_solve
    return dict(
        value=k / t_solve, cores=threads, kind="port",
        sample=("%s at %d^3 (n=%d), %d right-hand sides, same options; oracle built %s, "
                "OpenMP %d threads: factor %.2f s, block solve %.3f s (median of 5 after a "
                "warm-up, %d iterations); 1 thread: one column %.3f s" % (
                    "poisson3d" if "N" in cfg else ("elasticity3d" if "NE" in cfg
                                                    else "poisson2d"),
                    sN, A._nrow, k, kind, threads, t_factor, t_solve,
                    int(max(info["iterations"])), t_one)),
        factor_s=round(t_factor, 3), solve_s=round(t_solve, 4), end_to_end_s=round(cpu_e2e, 3),
        one_thread_solves_per_s=round(1.0 / t_one, 3),
        gpu_same_sample=same,
        speedup_same_sample=dict(
            solves=round(same["mixed"]["solves_per_s"] / (k / t_solve), 1),
            end_to_end=round(cpu_e2e / same["mixed"]["end_to_end_s"], 1)))


def warm_up_device():
    """One small factorization and solve before anything is timed, so the
    factor times do not include the process's first HIP module loads."""
    import rank_structured_cholesky_amd as R
    A, xyz = R.poisson3d(12, 1.0, 1)
    for extra in ({}, dict(factor_precision=R.RSC_FP32, sketch_precision=R.RSC_BF16)):
        M = R.FactorPreconditioner()
        M.initialize(A, xyz, 63, 400, **extra)
        x = np.zeros((A._nrow, 2))
        B = np.stack([R.uniform_rhs(A._nrow, 1)] * 2, axis=1)
        R.PCG_SOLVER(A, 1e-8, 100).solvePCG(x, B, M)
        del M


def measured_traffic(workload, precision):
    """HBM bytes per apply (all right-hand sides of the workload) from the
    committed counter summary of this workload and factor
    (profiles/apply_traffic_<workload>[_mixed].md)."""
    name = "apply_traffic_%s%s.md" % (workload, "_mixed" if precision == 1 else "")
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None, None
    for line in open(path):
        if line.startswith("traffic_bytes_per_apply:"):
            return int(line.split(":")[1]), os.path.relpath(path, ROOT)
    return None, None


def pct(v, p):
    v = sorted(v)
    return v[min(len(v) - 1, int(round(p * (len(v) - 1))))]


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "single-GPU build: run with --gpus 1 (no multi-GPU "
                                   "path exists in this build)"}))
        sys.exit(2)
    cfg = CONFIGS[a.config]
    import torch
    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU visible"}))
        sys.exit(1)
    if a.traffic_run:
        traffic_run(cfg, 0 if a.fp64_factor else 1)
        return
    if a.config == "serving":
        cfg = dict(cfg)
        if a.requests:
            cfg["requests"] = a.requests
        if a.max_rhs:
            cfg["max_rhs"] = a.max_rhs
        if a.rate:
            cfg["rate"] = a.rate
        r = serving_leg(cfg, a.eps)
        print(json.dumps({
            "metric": "served PCG solves/s to 1e-8 (right-hand sides per second), "
                      "p50/p99 request latency",
            "value": round(r["solves_per_s"], 3), "unit": "solves/s", "n_gpus": 1,
            "higher_is_better": True, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "serving", "grids": list(cfg["grids"]),
                       "requests": cfg["requests"], "max_rhs": cfg["max_rhs"],
                       "rate_per_s": cfg["rate"], "slots_per_factor": cfg["slots"],
                       "eps": a.eps}, **r}))
        return
    A, xyz, B = build_problem(cfg)
    warm_up_device()
    print("workload %s: n %d, nnz %d, %d right-hand sides" % (a.config, A._nrow, A._nzmax,
                                                             B.shape[1]),
          file=sys.stderr, flush=True)
    mixed = gpu_leg(A, xyz, B, cfg, 1, a.steps, a.warmup, a.eps, keep_x=True)
    print("mixed leg: factor %.3f s, %.4f s per solve" % (mixed["t_factor"],
                                                        mixed["total"] / a.steps),
          file=sys.stderr, flush=True)
    g64 = None
    if not a.no_fp64:
        g64 = gpu_leg(A, xyz, B, cfg, 0, a.steps, a.warmup, a.eps, keep_x=True)
        print("fp64 leg: factor %.3f s, %.4f s per solve" % (g64["t_factor"],
                                                           g64["total"] / a.steps),
              file=sys.stderr, flush=True)
    err = None
    if not a.no_direct:
        sols = {"mixed": mixed["X"]}
        if g64 is not None:
            sols["fp64"] = g64["X"]
        err = direct_error(A, xyz, B, cfg, sols)
        print("direct solve: %s" % err["direct"], file=sys.stderr, flush=True)
    mixed.pop("X", None)
    if g64 is not None:
        g64.pop("X", None)
    torch.cuda.empty_cache()
    cpu = None if a.no_cpu else cpu_leg(cfg, a.eps)
    steps, k = a.steps, mixed["k"]
    value = steps * k / mixed["total"]
    lat = mixed["lat"]
    achieved = mixed["apply_bytes"] / mixed["apply_s"] / 1e9
    traffic, traffic_src = measured_traffic(a.config, 1)
    e2e_mixed = mixed["t_factor"] + mixed["total"] / steps
    line = {
        "metric": "PCG solves/s to 1e-8 relres, mixed vs FP64 factor; factor s; "
                  "p50/p99 latency",
        "value": round(value, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": a.warmup,
        "ms_per_step": round(mixed["total"] / steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 factor, bf16 sketches (f32 accumulation), f64 PCG",
        "data": "synthetic",
        "config": {"workload": a.config, "n": A._nrow, "nnz": int(A._nzmax),
                   "sigma": cfg.get("sigma"), "max_dense_block": cfg["max_dense_block"],
                   "max_rank": cfg["max_rank"], "oversampling": cfg["oversampling"],
                   "nrhs": k, "eps": a.eps, "parallelism": "single GPU"},
        "pcg_iterations": mixed["iters"],
        "true_relres_max": max(mixed["relres"]),
        "p50_ms": round(statistics.median(lat) * 1e3, 3),
        "p99_ms": round(pct(lat, 0.99) * 1e3, 3),
        "factor_s": round(mixed["t_factor"], 4),
        "numeric_factor_s": round(mixed["info"]["numeric_seconds"], 4),
        "symbolic_s": round(mixed["info"]["symbolic_seconds"], 4),
        "factor_bytes": mixed["info"]["factor_bytes"],
        "pivot_replacements": mixed["info"]["pivot_replacements"],
        "end_to_end_s": round(e2e_mixed, 4),
        "roofline": {"bound": "hbm", "kernel": "preconditioner apply, mixed factor "
                                               "(k_apply_df<float, true>: fwd+bwd sweeps)",
                     "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "traffic_source": traffic_src,
                     "apply_ms": round(mixed["apply_s"] * 1e3, 4),
                     "apply_bytes": mixed["apply_bytes"]},
    }
    if g64 is not None:
        e2e64 = g64["t_factor"] + g64["total"] / steps
        line["fp64"] = {"solves_per_s": round(steps * k / g64["total"], 3),
                        "factor_s": round(g64["t_factor"], 4),
                        "numeric_factor_s": round(g64["info"]["numeric_seconds"], 4),
                        "symbolic_s": round(g64["info"]["symbolic_seconds"], 4),
                        "pcg_iterations": g64["iters"],
                        "true_relres_max": max(g64["relres"]),
                        "apply_ms": round(g64["apply_s"] * 1e3, 4),
                        "apply_achieved_gbs": round(g64["apply_bytes"] / g64["apply_s"] / 1e9,
                                                    2),
                        "factor_bytes": g64["info"]["factor_bytes"],
                        "end_to_end_s": round(e2e64, 4)}
        line["speedup_mixed_vs_fp64"] = {"end_to_end": round(e2e64 / e2e_mixed, 3),
                                         "solve": round(g64["total"] / mixed["total"], 3),
                                         "numeric_factor": round(
                                             g64["info"]["numeric_seconds"] /
                                             mixed["info"]["numeric_seconds"], 3)}
    if err is not None:
        line["error_vs_direct_solve"] = err
    if cpu is not None:
        line["cpu_baseline"] = {"value": round(cpu["value"], 4), "unit": "solves/s",
                                "cores": cpu["cores"], "kind": cpu["kind"],
                                "sample": cpu["sample"],
                                "one_thread_solves_per_s": cpu["one_thread_solves_per_s"],
                                "end_to_end_s": cpu["end_to_end_s"],
                                "gpu_same_sample": cpu["gpu_same_sample"],
                                "speedup_same_sample": cpu["speedup_same_sample"]}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
