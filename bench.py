"""Benchmark: PCG solves/s to 1e-8 relative residual with the rank-structured
preconditioner resident in HBM, on one MI355X.

Workload (default): 3D 7-point variable-coefficient Poisson, 64^3 cells
(n = 262,144), log-normal face coefficients (mu 0, sigma 1, seed 1234),
Dirichlet boundary; separators with more than 255 unknowns compressed,
ranks capped at 64; one uniform[-1,1] right-hand side (seed 5678).
A step = one PCG solve (recurrence relres <= 1e-8) of that right-hand side
with the factor and all vectors already in HBM.  The FP64 factor is the
headline; the mixed variant (FP32 dense factor, BF16 range-finder products,
FP64 PCG) is reported beside it.

Prints ONE JSON line.  `python bench.py --gpus N` with N > 1 exits: every
configuration here runs on one GPU.
"""
import argparse
import ctypes as C
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

CONFIGS = {
    "poisson3d_64": dict(N=64, sigma=1.0, max_dense_block=255, max_rank=64, nrhs=1),
    "poisson3d_128": dict(N=128, sigma=2.0, max_dense_block=255, max_rank=0, nrhs=16),
    "poisson2d_256": dict(N2=256, max_dense_block=127, max_rank=32, nrhs=1),
    # request serving over eight resident factors (continuous batching)
    "serving": dict(grids=(32, 48, 64, 80, 96, 96, 112, 128), sigma=1.0,
                    max_dense_block=255, max_rank=64, requests=24, max_rhs=256,
                    rate=20.0, slots=32),
}

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="poisson3d_64", choices=sorted(CONFIGS))
    ap.add_argument("--eps", type=float, default=1e-8)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed FP32/BF16-factor leg")
    ap.add_argument("--requests", type=int, default=0,
                    help="serving: number of requests (default from the config)")
    ap.add_argument("--max-rhs", type=int, default=0,
                    help="serving: largest right-hand-side count of a request")
    ap.add_argument("--rate", type=float, default=0.0,
                    help="serving: mean request arrivals per second")
    ap.add_argument("--traffic-run", action="store_true",
                    help="for counter passes: factor, 3 applies, then 3 plain 256 MiB "
                         "copies as the calibration kernel; no timing line")
    ap.add_argument("--mixed-factor", action="store_true",
                    help="--traffic-run with the mixed factor (FP32 + BF16 sketches)")
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
        steps=st["steps"], max_busy_slots=st["max_busy"], worst_true_relres=worst,
        iterations_mean=float(np.mean(iters_a)),
        one_at_a_time=dict(wall_s=wall_b, solves_per_s=total / wall_b,
                           p50_ms=1e3 * pct(lat_b, 0.5), p99_ms=1e3 * pct(lat_b, 0.99)),
        factor_s=[round(t, 3) for t in tf],
        n=[m._nrow for m in mats])


def traffic_run(cfg, precision=0):
    """The program the FETCH_SIZE / WRITE_SIZE counter passes profile: the
    apply dispatches (k_apply_df) and a copy kernel of known size."""
    import torch

    import rank_structured_cholesky_amd as R  # noqa: F401
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
    torch.cuda.synchronize()
    src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda")  # 256 MiB
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    vec = 56.0 * n
    print(json.dumps({"traffic_run": True,
                      "apply_bytes": M.info()["apply_bytes"] - vec + vec * k,
                      "nrhs": k}), flush=True)


def build_problem(cfg):
    import rank_structured_cholesky_amd as R
    if "N2" in cfg:
        A, xyz = R.poisson2d(cfg["N2"])
    else:
        A, xyz = R.poisson3d(cfg["N"], cfg["sigma"], 1234)
    B = R.uniform_rhs(A._nrow * cfg["nrhs"], 5678).reshape(cfg["nrhs"], A._nrow).T
    return A, xyz, np.asfortranarray(B)


def factor(A, xyz, cfg, precision):
    """precision 0: FP64 factor; 1: the mixed form (FP32 dense factor, BF16
    range-finder products with FP32 accumulation)."""
    import rank_structured_cholesky_amd as R
    M = R.FactorPreconditioner()
    extra = dict(sketch_precision=R.RSC_BF16) if precision == 1 else {}
    t0 = time.perf_counter()
    M.initialize(A, xyz, cfg["max_dense_block"], 400, max_rank=cfg["max_rank"],
                 factor_precision=precision, **extra)
    return M, time.perf_counter() - t0


def gpu_leg(A, xyz, B, cfg, precision, steps, warmup, eps):
    import torch

    import rank_structured_cholesky_amd as R
    from rank_structured_cholesky_amd import _lib
    L = _lib.lib()
    M, t_factor = factor(A, xyz, cfg, precision)
    info = M.info()
    n, k = B.shape
    dB = torch.from_numpy(np.ascontiguousarray(B.T)).to("cuda")  # k x n rows = columns
    dX = torch.zeros_like(dB)
    res = (_lib.PcgResult * k)()
    stream = torch.cuda.Stream()

    def solve():
        dX.zero_()
        st = L.rsc_block_pcg_solve(M.handle, k, C.c_void_p(dB.data_ptr()), n,
                                   C.c_void_p(dX.data_ptr()), n, eps, 10000, res,
                                   C.c_void_p(stream.cuda_stream))
        if st not in (_lib.RSC_OK, _lib.RSC_NOT_CONVERGED):
            _lib.check(st)

    for _ in range(warmup):
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
    iters = [r.iterations for r in res]
    relres = [r.true_relres for r in res]
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
    apply_s = e0.elapsed_time(e1) / 1e3 / napp
    x = dX.cpu().numpy().T
    err = float(np.linalg.norm(A.multiply(x[:, 0]) - B[:, 0]) / np.linalg.norm(B[:, 0]))
    out = dict(t_factor=t_factor, info=info, total=total, lat=lat, iters=iters,
               relres=relres, apply_s=apply_s, check_relres=err, k=k)
    del M
    return out


def cpu_leg(A, xyz, B, cfg, eps, full):
    """Oracle (FP64 CPU restatement of the reference, one thread) on a bounded
    sample of the same workload family, scaled to the full workload by the
    ratio of per-solve algorithmic bytes (the solve is memory-bound on the
    CPU as on the GPU)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from helpers import oracle_for, symbolic
    import rank_structured_cholesky_amd as R
    N = cfg.get("N", 0)
    if N:
        sN = min(N, 40)
        As, xs = R.poisson3d(sN, cfg["sigma"], 1234)
        desc = f"poisson3d {sN}^3 (n={As._nrow}), same options"
    else:
        sN = min(cfg["N2"], 128)
        As, xs = R.poisson2d(sN)
        desc = f"poisson2d {sN}^2 (n={As._nrow}), same options"
    bs = R.uniform_rhs(As._nrow, 5678)
    t0 = time.perf_counter()
    sym = symbolic(As, xs, max_dense_block=cfg["max_dense_block"], max_rank=cfg["max_rank"])
    O = oracle_for(As, sym, max_rank=cfg["max_rank"])
    tf = time.perf_counter() - t0
    reps = 3
    t1 = time.perf_counter()
    for _ in range(reps):
        _, info = O.pcg(bs, eps, 10000)
    ts = (time.perf_counter() - t1) / reps

    def spmv_bytes(M):
        return 12.0 * M._nzmax + 4.0 * (M._nrow + 1) + 16.0 * M._nrow

    def solve_bytes(iters, factor_bytes, Mx):
        return iters * (2.0 * factor_bytes + spmv_bytes(Mx) + 80.0 * Mx._nrow) + spmv_bytes(Mx)

    w_sample = solve_bytes(info["iterations"], 8.0 * sym["factor_elems"], As)
    w_full = solve_bytes(max(full["iters"]), float(full["info"]["factor_bytes"]), A)
    value = (1.0 / ts) * (w_sample / w_full)
    return dict(value=value, cores=1,
                sample=desc + "; factor %.1f s, one PCG solve %.3f s (%d it, mean of %d); "
                "scaled by per-solve algorithmic bytes %.3g / %.3g" % (
                    tf, ts, info["iterations"], reps, w_sample, w_full))


def warm_up_device():
    """One small factorization and solve before anything is timed, so the
    factor times do not include the process's first HIP module loads."""
    import rank_structured_cholesky_amd as R
    A, xyz = R.poisson3d(12, 1.0, 1)
    M = R.FactorPreconditioner()
    M.initialize(A, xyz, 63, 400)
    x = np.zeros(A._nrow)
    R.PCG_SOLVER(A, 1e-8, 100).solvePCG(x, R.uniform_rhs(A._nrow, 1), M)
    del M


def measured_traffic(workload, precision):
    """HBM bytes per apply (all right-hand sides of the workload) from the
    committed counter summary of this workload
    (profiles/apply_traffic_<workload>.md), FP64 factor only."""
    path = os.path.join(ROOT, "profiles", "apply_traffic_%s.md" % workload)
    if precision != 0 or not os.path.exists(path):
        return None, None
    for line in open(path):
        if line.startswith("traffic_bytes_per_apply:"):
            return int(line.split(":")[1]), os.path.relpath(path, ROOT)
    return None, None


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
        traffic_run(cfg, 1 if a.mixed_factor else 0)
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
    g64 = gpu_leg(A, xyz, B, cfg, 0, a.steps, a.warmup, a.eps)
    mixed = None
    if not a.no_mixed:
        mixed = gpu_leg(A, xyz, B, cfg, 1, a.steps, a.warmup, a.eps)
    cpu = None if a.no_cpu else cpu_leg(A, xyz, B, cfg, a.eps, g64)
    steps = a.steps
    ms = g64["total"] / steps * 1e3
    value = steps * g64["k"] / g64["total"]
    lat = sorted(g64["lat"])
    p50 = statistics.median(lat) * 1e3
    p99 = lat[min(len(lat) - 1, int(round(0.99 * (len(lat) - 1))))] * 1e3
    # algorithmic bytes of one apply with k right-hand sides: the stored
    # factor once, the vector terms (56 B per row: gather, scatter, front
    # and rank vectors) k times
    vec_bytes = 56.0 * A._nrow
    apply_bytes = (g64["info"]["apply_bytes"] - vec_bytes) + vec_bytes * g64["k"]
    traffic, traffic_src = measured_traffic(a.config, 0)
    achieved = apply_bytes / g64["apply_s"] / 1e9
    line = {
        "metric": "PCG solves/s to 1e-8 relres (FP64 factor; mixed FP32/BF16 factor beside)",
        "value": round(value, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": a.config, "n": A._nrow, "nnz": int(A._nzmax),
                   "max_dense_block": cfg["max_dense_block"], "max_rank": cfg["max_rank"],
                   "nrhs": g64["k"], "eps": a.eps, "parallelism": "single GPU"},
        "pcg_iterations": g64["iters"],
        "true_relres": g64["relres"],
        "p50_ms": round(p50, 3),
        "p99_ms": round(p99, 3),
        "factor_s": round(g64["t_factor"], 4),
        "numeric_factor_s": round(g64["info"]["numeric_seconds"], 4),
        "symbolic_s": round(g64["info"]["symbolic_seconds"], 4),
        "factor_bytes": g64["info"]["factor_bytes"],
        "roofline": {"bound": "hbm", "kernel": "preconditioner apply (fwd+bwd sweeps)",
                     "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "traffic_source": traffic_src,
                     "apply_ms": round(g64["apply_s"] * 1e3, 4),
                     "apply_bytes": apply_bytes},
    }
    if mixed is not None:
        line["mixed"] = {"factor": "FP32 dense blocks, BF16 range-finder products "
                                   "(FP32 accumulation), FP64 PCG",
                         "solves_per_s": round(steps * mixed["k"] / mixed["total"], 3),
                         "factor_s": round(mixed["t_factor"], 4),
                         "numeric_factor_s": round(mixed["info"]["numeric_seconds"], 4),
                         "pcg_iterations": mixed["iters"], "true_relres": mixed["relres"],
                         "apply_ms": round(mixed["apply_s"] * 1e3, 4),
                         "pivot_replacements": mixed["info"]["pivot_replacements"],
                         "end_to_end_s": round(mixed["t_factor"] + mixed["total"] / steps, 4)}
        line["end_to_end_s_fp64"] = round(g64["t_factor"] + g64["total"] / steps, 4)
    if cpu is not None:
        line["cpu_baseline"] = {"value": round(cpu["value"], 4), "unit": "solves/s",
                                "cores": cpu["cores"], "kind": "port",
                                "sample": cpu["sample"]}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
