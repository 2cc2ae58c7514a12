"""The non-headline configurations of bench.py (`--workload ...`).

coulomb80k   3D Coulomb, N = 80,000 uniform points in a cube of side N^(1/3),
             reltol 1e-6, leaf 300, AOT (stored B and D), one FP64 RHS
             U(-0.5, 0.5); GPU matvec vs the oracle's AOT matvec on this host.
rpy_matmul   RPY tensor kernel (a = 1, eta = 1), N = 300,000 particles
             uniform at 20% volume fraction, reltol 1e-7, leaf 200, H2P_matmul
             with 32 FP64 columns ~ N(0, 1), column-major; JIT and AOT.
precision2m  3D Gaussian (l = 1) and 3D Coulomb at N = 2,000,000, density 10,
             reltol 1e-6, leaf 300, JIT: FP64 vs FP32-mixed matvec.
periodic     periodic RPY-Ewald (H2P_matvec_periodic), radius 1 at 20% volume
             fraction, --periodic-n particles, FP64 JIT; error against rows of
             the dense Ewald matrix.
serving      Poisson arrivals of Matern-3/2 GP requests (lengthscale U(0.5, 2),
             nugget 1e-2, N log-uniform in [1e4, 4e5], density 10, reltol
             1e-6, leaf 300): half 8-RHS products, half CG solves to 1e-6;
             continuous batching on the FP32-mixed path against serving each
             request alone on the FP64 path (--requests of them).
Every figure is on one GPU with vectors resident in HBM.
"""
from __future__ import annotations

import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _time(fn, K, W, sync):
    for _ in range(W):
        fn()
    sync()
    t = time.perf_counter()
    for _ in range(K):
        fn()
    sync()
    return (time.perf_counter() - t) / K


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _host_threads():
    """Host threads the builds may use: the GPU box's share is 16 CPUs,
    whatever the machine shows; fewer where fewer exist (BLAS given more
    threads than CPUs has crashed in the workers)."""
    return max(1, min(16, os.cpu_count() or 1))


def _build_worker(spec, kw, threads):
    os.environ["H2P_BUILD_THREADS"] = str(threads)
    from threadpoolctl import threadpool_limits
    from h2pack_amd.serve import build_request
    with threadpool_limits(limits=threads):
        t = time.perf_counter()
        r = build_request(spec, **kw)
        return r, time.perf_counter() - t


def start_request_builds(n, seed, rate, cg_max_iter, workers):
    """Build the serving requests' H2 matrices in worker processes (which
    never touch the GPU); returns (executor, futures).  Call this before
    the GPU is first used: the workers are started by this call."""
    import multiprocessing as mp
    from concurrent.futures import ProcessPoolExecutor
    from h2pack_amd.serve import request_specs
    specs = request_specs(n, seed, 10_000, 400_000, rate)
    kw = {"max_iter": cg_max_iter}
    workers = max(1, workers)
    threads = max(1, _host_threads() // workers)
    # the workers inherit these at start: bounded BLAS / OpenMP threads that
    # sleep instead of spinning when idle, so the processes do not starve
    # each other (restored in this process once they have started)
    keys = {"OMP_WAIT_POLICY": "PASSIVE", "OPENBLAS_NUM_THREADS": str(threads), "H2P_BUILD_THREADS": str(threads)}
    saved = {k: os.environ.get(k) for k in keys}
    os.environ.update(keys)
    ex = ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("spawn"))
    futs = [(s["rid"], ex.submit(_build_worker, s, kw, threads)) for s in specs]
    for k, v in saved.items():
        if v is None:
            os.environ.pop(k, None)
        else:
            os.environ[k] = v
    return ex, futs


def start_stream(n, seed, rate, cg_max_iter, workers, window, first=0, count=None):
    """Requests built while they are served: placeholders in arrival order,
    each given a build future by a feeder thread that keeps at most
    ``window`` requests built or building but not yet admitted by the server
    (so host memory stays bounded); requests ``first`` .. ``first + count``
    of the ``n``-request sequence.  Call before the GPU is first used."""
    import math as _m
    import multiprocessing as mp
    import threading
    from concurrent.futures import ProcessPoolExecutor
    from h2pack_amd.serve import Request, request_specs
    specs = request_specs(n, seed, 10_000, 400_000, rate)
    specs = specs[first:] if count is None else specs[first:first + count]
    kw = {"max_iter": cg_max_iter}
    workers = max(1, workers)
    threads = max(1, _host_threads() // workers)
    # the workers start at the feeder's first submit and inherit these then
    # (bounded BLAS / OpenMP threads that sleep when idle), so they stay set
    # in this process, which builds nothing itself
    os.environ.update({"OMP_WAIT_POLICY": "PASSIVE", "OPENBLAS_NUM_THREADS": str(threads),
                       "H2P_BUILD_THREADS": str(threads)})
    ex = ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("spawn"))
    reqs = [Request(rid=s["rid"], h2=None, kind=s["kind"], rhs=None, arrival=s["arrival"]) for s in specs]

    def feed():
        k = 0
        while k < len(specs):
            waiting = sum(1 for r in reqs[:k] if _m.isnan(r.admitted) and not r.rejected)
            if waiting >= window:
                time.sleep(0.02)
                continue
            reqs[k].future = ex.submit(_build_worker, specs[k], kw, threads)
            k += 1

    th = threading.Thread(target=feed, daemon=True)
    th.start()
    return reqs, ex, th


def collect_requests(ex, futs):
    reqs = [None] * len(futs)
    for rid, f in futs:
        r, tb = f.result()
        reqs[rid] = r
        log(f"[bench] request {r.rid}: {r.kind} N={r.N} built in {tb:.1f}s")
    ex.shutdown()
    return reqs


def serve_one(reqs, name, prec, slots, hbm_frac=0.8, drop_host=False, cg_active=None):
    """One serving leg over ``reqs``; returns its summary dict."""
    import torch
    from h2pack_amd.serve import Server
    free, total = torch.cuda.mem_get_info()
    budget = int(hbm_frac * free)
    srv = Server(hbm_budget=budget, max_active=slots, prec=prec, cg_chunk=8, drop_host=drop_host,
                 cg_active=cg_active)
    log(f"[bench] serving {len(reqs)} requests: {name} ({prec}, up to {slots} at once)")
    st = srv.run(reqs, realtime=True, log=log)
    srv.close()
    cg = [r for r in reqs if r.kind == "cg"]
    d = st.as_dict()
    d["cg_converged"] = sum(1 for r in cg if r.relres <= r.tol)
    d["cg_requests"] = len(cg)
    d["cg_mean_iters"] = float(np.mean([r.iters for r in cg])) if cg else 0.0
    late = [r.loaded - r.arrival for r in reqs if r.loaded == r.loaded]
    d["build_late_max_s"] = float(max(0.0, max(late))) if late else 0.0
    d["hbm_budget_bytes"] = budget
    d["mean_N"] = float(np.mean([r.N for r in reqs]))
    log(f"[bench] {name}: {st.n_done} done in {st.makespan:.1f}s, {st.throughput:.3f} req/s, "
        f"p50 {st.p50:.2f}s p99 {st.p99:.2f}s, max active {st.max_active}, "
        f"CG {d['cg_converged']}/{d['cg_requests']} converged ({d['cg_mean_iters']:.0f} iterations mean, "
        f"{st.cg_replacements} FP64 residual replacements), builds late by at most {d['build_late_max_s']:.1f}s")
    return d


def serve_compare(reqs, slots, hbm_frac=0.8, cg_active=16):
    """The requests through the continuous-batching loop (FP32-mixed JIT,
    co-scheduled launches) and, on the same arrivals, one at a time on the
    plain FP64 path; returns the summary dict."""
    import copy
    import torch
    from h2pack_amd.serve import Server
    free, total = torch.cuda.mem_get_info()
    budget = int(hbm_frac * free)

    def fresh():
        out = []
        for r in reqs:
            c = copy.copy(r)
            c.result = None
            out.append(c)
        return out

    res = {}
    for name, prec, sl in (("batched_fp32mix", "fp32mix", slots), ("alone_fp64", "fp64", 1)):
        rs = fresh()
        srv = Server(hbm_budget=budget, max_active=sl, prec=prec, cg_chunk=8, cg_active=cg_active)
        log(f"[bench] serving {len(rs)} requests: {name} ({prec}, up to {sl} at once)")
        st = srv.run(rs, realtime=True, log=log)
        srv.close()
        cg = [r for r in rs if r.kind == "cg"]
        d = st.as_dict()
        d["cg_converged"] = sum(1 for r in cg if r.relres <= r.tol)
        d["cg_requests"] = len(cg)
        d["cg_mean_iters"] = float(np.mean([r.iters for r in cg])) if cg else 0.0
        res[name] = (d, rs)
        log(f"[bench] {name}: {st.n_done} done in {st.makespan:.1f}s, {st.throughput:.3f} req/s, "
            f"p50 {st.p50:.2f}s p99 {st.p99:.2f}s, max active {st.max_active}")
    # the FP32-mixed results against the FP64 ones (the matmul requests)
    errs = []
    for a, b in zip(res["batched_fp32mix"][1], res["alone_fp64"][1]):
        if a.kind == "matmul":
            errs.append(_rel(a.result, b.result))
    d32, d64 = res["batched_fp32mix"][0], res["alone_fp64"][0]
    return {"batched_fp32mix": d32, "alone_fp64": d64, "hbm_budget_bytes": budget,
            "throughput_speedup": d32["throughput"] / d64["throughput"] if d64["throughput"] else None,
            "p50_speedup": d64["p50"] / d32["p50"] if d32["p50"] else None,
            "p99_speedup": d64["p99"] / d32["p99"] if d32["p99"] else None,
            "matmul_max_rel_diff_fp32mix_vs_fp64": max(errs) if errs else None,
            "matmul_min_rel_diff_fp32mix_vs_fp64": min(errs) if errs else None,
            "mean_N": float(np.mean([r.N for r in reqs]))}


def run_alone_service(args):
    """Service times of the FP64-alone server: requests ``--serve-first`` ..
    ``+ --serve-count`` of the ``--requests`` sequence served one at a time,
    back to back (no arrival clock), each timed from the start of its upload
    to its result.  The arrival times of a rate only scale the same draws,
    so these times give the FIFO latencies at every rate (serving-replay)."""
    count = args.serve_count if args.serve_count > 0 else args.requests - args.serve_first
    reqs, ex, th = start_stream(args.requests, args.seed, args.rate, args.cg_max_iter, args.build_workers,
                                args.build_window, first=args.serve_first, count=count)
    head = min(args.build_head, args.build_window, count)
    t0 = time.perf_counter()
    while sum(1 for r in reqs[:head] if r.future is not None and r.future.done()) < head:
        time.sleep(0.2)
    log(f"[bench] head start: {head} requests built in {time.perf_counter() - t0:.1f}s")
    import torch
    from h2pack_amd.serve import Server
    free, _ = torch.cuda.mem_get_info()
    srv = Server(hbm_budget=int(0.8 * free), max_active=1, prec="fp64", cg_chunk=8, drop_host=True)
    log(f"[bench] FP64 alone, back to back: requests {args.serve_first}..{args.serve_first + count - 1}")
    st = srv.run(reqs, realtime=False, log=log)
    srv.close()
    th.join()
    ex.shutdown()
    rows = [{"rid": r.rid, "kind": r.kind, "N": r.N, "service_s": r.finished - r.admit_start,
             "iters": r.iters, "relres": r.relres}
            for r in sorted(reqs, key=lambda q: q.rid)]
    cg = [x for x in rows if x["kind"] == "cg"]
    out = {"metric": "FP64-alone service times (Matern-3/2 GP requests)", "unit": "s",
           "n_gpus": 1, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"serving_matern32_{args.requests}req_N1e4-4e5_density10_reltol1e-6_leaf300"
                                  f"_8rhs_cg_tol1e-6_nugget1e-2_alone_fp64_service",
                      "requests": args.requests, "first": args.serve_first, "count": count,
                      "cg_max_iter": args.cg_max_iter, "seed": args.seed},
           "n_done": st.n_done, "sum_service_s": float(sum(x["service_s"] for x in rows)),
           "cg_converged": sum(1 for x in cg if x["relres"] <= 1e-6), "cg_requests": len(cg),
           "requests_detail": rows}
    log(f"[bench] FP64 alone: {st.n_done} requests, service {out['sum_service_s']:.1f}s in total, "
        f"CG {out['cg_converged']}/{len(cg)} converged")
    print(json.dumps(out), flush=True)


def run_serving_replay(args):
    """FP64-alone latencies at each of ``--replay-rates`` from measured
    service times (the JSON lines of run_alone_service covering every
    request once), by the FIFO recurrence; no GPU."""
    from h2pack_amd.serve import fifo_latencies, request_specs
    rows = {}
    for f in args.replay_files:
        with open(f) as fh:
            line = [ln for ln in fh.read().splitlines() if ln.startswith("{")][-1]
        for x in json.loads(line)["requests_detail"]:
            rows[x["rid"]] = x
    n = args.requests
    missing = [k for k in range(n) if k not in rows]
    if missing:
        raise SystemExit(f"service times missing for {len(missing)} requests (first {missing[:5]})")
    svc = np.array([rows[k]["service_s"] for k in range(n)])
    res = {"sum_service_s": float(svc.sum()), "capacity_req_per_s": n / float(svc.sum())}
    for rate in args.replay_rates:
        arr = np.array([s["arrival"] for s in request_specs(n, args.seed, 10_000, 400_000, rate)])
        lat = fifo_latencies(arr, svc)
        ends = arr + lat
        res[f"rate_{rate}"] = {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                               "mean": float(lat.mean()), "makespan": float(ends.max()),
                               "throughput": n / float(ends.max()), "last_arrival": float(arr[-1])}
        log(f"[bench] FP64 alone at {rate}/s (FIFO replay): {n / ends.max():.3f} req/s, "
            f"p50 {np.percentile(lat, 50):.1f}s p99 {np.percentile(lat, 99):.1f}s")
    print(json.dumps({"metric": "FP64-alone served latency by FIFO replay of measured service times",
                      "unit": "s", "n_gpus": 1, "dtype": "f64",
                      "config": {"workload": f"serving_matern32_{n}req_alone_fp64_replay", "seed": args.seed},
                      **res}), flush=True)


def run_serving(args):
    """Matern-3/2 GP requests: continuous batching on the FP32-mixed path
    against serving each alone on the FP64 path.  Host construction happens
    before the clock starts, in worker processes; with ``--serve-legs
    batched`` or ``alone`` one leg runs on requests built while they are
    served (the configuration's 512 requests do not fit in host memory at
    once)."""
    if args.serve_legs != "both":
        leg = {"batched": ("batched_fp32mix", "fp32mix", args.slots), "alone": ("alone_fp64", "fp64", 1)}[args.serve_legs]
        t0 = time.perf_counter()
        reqs, ex, th = start_stream(args.requests, args.seed, args.rate, args.cg_max_iter,
                                    args.build_workers, args.build_window)
        head = min(args.build_head, args.build_window, args.requests)
        t_log = time.perf_counter()
        while True:   # a head start: the first requests built before the clock starts
            done = sum(1 for r in reqs[:head] if r.future is not None and r.future.done())
            if done >= head:
                break
            if time.perf_counter() - t_log > 15.0:
                t_log = time.perf_counter()
                log(f"[bench] head start: {done}/{head} built after {t_log - t0:.0f}s")
            time.sleep(0.2)
        log(f"[bench] head start: {head} requests built in {time.perf_counter() - t0:.1f}s")
        d = serve_one(reqs, leg[0], leg[1], leg[2], drop_host=True,
                      cg_active=args.cg_active if args.cg_active > 0 else None)
        th.join()
        ex.shutdown()
        out = {"metric": "served GP requests/s and p50/p99 latency (Matern-3/2)", "value": d["throughput"],
               "unit": "requests/s", "higher_is_better": True, "n_gpus": 1,
               "dtype": "f32-mixed" if leg[1] == "fp32mix" else "f64", "data": "synthetic",
               "config": {"workload": f"serving_matern32_{args.requests}req_N1e4-4e5_density10_reltol1e-6_leaf300"
                                      f"_rate{args.rate}_8rhs_cg_tol1e-6_nugget1e-2_{leg[0]}",
                          "requests": args.requests, "rate_per_s": args.rate, "slots": leg[2],
                          "cg_active": args.cg_active if args.cg_active > 0 else None,
                          "cg_max_iter": args.cg_max_iter, "build_workers": args.build_workers},
               leg[0]: d, "vs_baseline": None}
        print(json.dumps(out), flush=True)
        return
    t0 = time.perf_counter()
    ex, futs = start_request_builds(args.requests, args.seed, args.rate, args.cg_max_iter, args.build_workers)
    reqs = collect_requests(ex, futs)
    t_build = time.perf_counter() - t0
    r = serve_compare(reqs, args.slots, cg_active=args.cg_active if args.cg_active > 0 else None)
    d32 = r["batched_fp32mix"]
    out = {"metric": "served GP requests/s and p50/p99 latency (Matern-3/2, FP32-mixed batched vs FP64 alone)",
           "value": d32["throughput"], "unit": "requests/s", "higher_is_better": True,
           "n_gpus": 1, "dtype": "f32-mixed", "data": "synthetic",
           "config": {"workload": f"serving_matern32_{args.requests}req_N1e4-4e5_density10_reltol1e-6_leaf300"
                                  f"_rate{args.rate}_8rhs_cg_tol1e-6_nugget1e-2",
                      "requests": args.requests, "rate_per_s": args.rate, "slots": args.slots,
                      "hbm_budget_bytes": r.pop("hbm_budget_bytes"), "cg_max_iter": args.cg_max_iter},
           **r, "host_build_s": t_build, "vs_baseline": None}
    print(json.dumps(out), flush=True)


def run(args):
    if args.workload == "serving-replay":
        return run_serving_replay(args)
    if args.workload == "serving" and args.serve_legs == "alone-service":
        return run_alone_service(args)
    if args.workload == "serving":
        return run_serving(args)
    import torch
    from h2pack_amd import abi
    from h2pack_amd import construct as C
    from h2pack_amd.h2mat import H2MatHIP, direct_nbody
    dev = torch.device("cuda:0")
    sptr = torch.cuda.current_stream().cuda_stream
    sync = torch.cuda.synchronize
    K, W = max(1, args.steps), max(1, args.warmup)
    out = {"workload": args.workload, "n_gpus": 1, "steps": K, "warmup": W, "data": "synthetic"}

    if args.workload == "coulomb80k":
        N = 80_000
        coord = C.uniform_cube_points(N, 1.0, args.seed)
        t0 = time.perf_counter()
        h2 = C.build_h2(coord, C.KRNL_COULOMB, [], 1e-6, 300, BD_JIT=0, pp_kind="numerical", host_blocks=True)
        tb = time.perf_counter() - t0
        x = np.random.default_rng(args.seed + 1).random(N) - 0.5
        xd = torch.from_numpy(x).to(dev); yd = torch.empty_like(xd)
        m = H2MatHIP(h2, mode="aot", host_blocks=True)
        tune = m.autotune(0)
        mv = lambda: m.matvec_dev(xd.data_ptr(), yd.data_ptr(), sptr)   # noqa: E731
        dt = _time(mv, K, W, sync)
        # the same steps again with HIP events around every phase
        m.set_profiling(True)
        _time(mv, K, 0, sync)
        info = m.info()
        m.set_profiling(False)
        blk_bytes = 8.0 * (info["B_elems"] + info["D_elems"])
        y = yd.cpu().numpy()
        rows = np.sort(np.random.default_rng(args.seed + 2).choice(N, 1000, replace=False))
        yref = direct_nbody(C.KRNL_COULOMB, [], coord, x, rows)
        out.update({"metric": "H2 matvecs/s (Coulomb N=80k AOT, FP64)", "value": 1.0 / dt, "unit": "matvecs/s",
                    "ms_per_step": dt * 1e3, "dtype": "f64",
                    "hbm_GBps_model": info["hbm_bytes_model"] / dt / 1e9,
                    "hbm_frac_model": info["hbm_bytes_model"] / dt / 1e9 / 8000.0,
                    "block_variant": tune["variants"][0],
                    "roofline": {"bound": "hbm", "kernel": "k_blocks_aot", "kernel_ms": info["t_blocks"],
                                 "achieved": blk_bytes / (info["t_blocks"] * 1e-3) / 1e9, "peak": 8000.0,
                                 "unit": "GB/s",
                                 "frac": blk_bytes / (info["t_blocks"] * 1e-3) / 1e9 / 8000.0},
                    "phases_ms": {k: info[k] for k in ("t_fwd", "t_blocks", "t_reduce", "t_bwd", "t_final",
                                                       "t_total")},
                    "stored_GB": blk_bytes / 1e9,
                    "rel_err_vs_direct_nbody_1000rows": _rel(y[rows], yref), "build_s": tb})
        log(f"[bench] coulomb80k AOT: {dt * 1e3:.3f} ms per matvec, model {info['hbm_bytes_model'] / dt / 1e9:.0f} GB/s; "
            f"block launch {info['t_blocks']:.3f} ms for {blk_bytes / 1e9:.2f} GB; phases {out['phases_ms']}")
        if not args.no_cpu:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import oracle as O
            o = O.OracleH2(h2, BD_JIT=0)
            nt = os.cpu_count() or 1
            o.matvec(x, n_thread=nt)
            t = time.perf_counter(); calls = 0
            while calls < 10 and time.perf_counter() - t < 10:
                yo = o.matvec(x, n_thread=nt); calls += 1
            el = (time.perf_counter() - t) / calls
            out["cpu_baseline"] = {"value": 1.0 / el, "unit": "matvecs/s", "cores": nt, "kind": "port",
                                   "sample": f"{calls} AOT FP64 matvecs of the same H2 matrix (oracle, OpenMP)"}
            out["rel_err_vs_oracle"] = _rel(y, yo)

    elif args.workload == "rpy_matmul":
        N, nv = 300_000, 32
        L = (N * (4.0 / 3.0) * math.pi / 0.2) ** (1.0 / 3.0)
        rng = np.random.default_rng(args.seed)
        coord = np.concatenate([rng.random((3, N)) * L, np.ones((1, N))], axis=0)
        t0 = time.perf_counter()
        h2 = C.build_h2(coord, C.KRNL_RPY, [1.0], 1e-7, 200, BD_JIT=1, pp_kind="surface", max_leaf_size=4.0)
        tb = time.perf_counter() - t0
        log(f"[bench] RPY N={N}: built in {tb:.1f}s, levels {h2.max_level + 1}, adm {h2.n_r_adm_pair}")
        Nk = h2.krnl_mat_size
        leaf_npt = [h2.pt_cluster[2 * n + 1] - h2.pt_cluster[2 * n] + 1 for n in h2.leaf_nodes]
        diag_pairs = float(sum(k * k for k in leaf_npt))
        X = np.asfortranarray(rng.standard_normal((Nk, nv)))
        Xd = torch.from_numpy(X.T.copy()).to(dev)          # column-major: column v at v * Nk
        Yd = torch.empty_like(Xd)
        res = {}
        for mode in ("jit", "aot"):
            m = H2MatHIP(h2, mode=mode, host_blocks=False)
            mm = lambda: m.matmul_dev(abi.COL_MAJOR, nv, Xd.data_ptr(), Nk, Yd.data_ptr(), Nk, sptr)   # noqa: E731
            dt_default = _time(mm, K, W, sync)
            Y0 = Yd.cpu().numpy().T
            t_t = time.perf_counter()
            tune = m.autotune(nv)
            tune["tune_s"] = time.perf_counter() - t_t
            dt = _time(mm, K, W, sync)
            Y = Yd.cpu().numpy().T
            log(f"[bench] rpy {mode}: tile variants {tune['variants']} (phases {tune['default_ms']:.1f} -> "
                f"{tune['tuned_ms']:.1f} ms), matmul {dt_default * 1e3:.1f} -> {dt * 1e3:.1f} ms, "
                f"identical: {bool(np.array_equal(Y0, Y))}")
            xv = torch.empty(Nk, dtype=torch.float64, device=dev); yv = torch.empty_like(xv)
            def cols():
                for v in range(nv):
                    m.matvec_dev(Xd[v].data_ptr(), yv.data_ptr(), sptr)
            dtv = _time(cols, 1, 1, sync)
            # matmul == matvec per column
            errs = []
            for v in (0, nv - 1):
                m.matvec_dev(Xd[v].data_ptr(), yv.data_ptr(), sptr); sync()
                errs.append(_rel(Y[:, v], yv.cpu().numpy()))
            info = m.info()
            # 2 n_vec (2 sum U + 2 sum B + sum D_diag + 2 sum D_offdiag) entries of
            # the 3x3 kernel: the diagonal D blocks act once, every other block in
            # both directions
            d_diag = 9.0 * diag_pairs
            flops = 2.0 * nv * (2 * info["U_elems"] + 2 * 9 * info["B_pairs"] + d_diag
                                + 2 * (9 * info["D_pairs"] - d_diag))
            res[mode] = {"matmuls_per_s": 1.0 / dt, "ms_per_matmul": dt * 1e3,
                         "ms_per_matmul_default_tiles": dt_default * 1e3, "autotune": tune,
                         "tuned_equals_default_bitwise": bool(np.array_equal(Y0, Y)),
                         "ms_per_32_matvecs": dtv * 1e3, "speedup_vs_matvec_loop": dtv / dt,
                         "model_TFLOPs": flops / dt / 1e12, "rel_err_vs_matvec_cols": max(errs)}
            log(f"[bench] rpy {mode}: {dt * 1e3:.2f} ms per 32-column matmul ({flops / dt / 1e12:.2f} TFLOP/s model), "
                f"32 matvecs {dtv * 1e3:.2f} ms")
            m.close()
        rows = np.sort(rng.choice(N, 1000, replace=False))
        yref = direct_nbody(C.KRNL_RPY, [1.0], coord, np.ascontiguousarray(X[:, 0]), rows)
        Yr = Y[:, 0].reshape(N, 3)[rows].reshape(-1)
        # single-RHS matvecs, JIT: FP64 against FP32-mixed (the tensor in FP32)
        xv = torch.empty(Nk, dtype=torch.float64, device=dev)
        xv.copy_(Xd[0])
        yv = torch.empty_like(xv)
        mv_res = {}
        for prec in ("fp64", "fp32mix"):
            m = H2MatHIP(h2, mode="jit", prec=prec)
            dtm = _time(lambda: m.matvec_dev(xv.data_ptr(), yv.data_ptr(), sptr), K, W, sync)
            mv_res[prec] = (dtm, yv.cpu().numpy())
            m.close()
        rpy_mv = {"fp64_ms": mv_res["fp64"][0] * 1e3, "fp32mix_ms": mv_res["fp32mix"][0] * 1e3,
                  "fp32mix_speedup": mv_res["fp64"][0] / mv_res["fp32mix"][0],
                  "fp32mix_rel_err_vs_fp64_gpu": _rel(mv_res["fp32mix"][1], mv_res["fp64"][1])}
        log(f"[bench] rpy jit matvec: FP64 {rpy_mv['fp64_ms']:.2f} ms, FP32-mixed {rpy_mv['fp32mix_ms']:.2f} ms "
            f"({rpy_mv['fp32mix_speedup']:.2f}x), rel diff {rpy_mv['fp32mix_rel_err_vs_fp64_gpu']:.2e}")
        out.update({"metric": "H2 matmuls/s (RPY N=300k, 32 RHS, FP64)", "value": res["jit"]["matmuls_per_s"],
                    "unit": "matmuls/s", "dtype": "f64", "jit": res["jit"], "aot": res["aot"],
                    "jit_matvec": rpy_mv,
                    "rel_err_vs_direct_nbody_1000rows_col0": _rel(Yr, yref), "build_s": tb})

    elif args.workload == "periodic":
        # periodic RPY (H2P_matvec_periodic): radius 1, 20 % volume fraction,
        # eta = 1/(6 pi), Ewald xi = 0.015 with 6 layers each way (converged
        # to machine precision for this cell), reltol 1e-6, leaf 200
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        N = args.periodic_n
        L = (N * (4.0 / 3.0) * math.pi / 0.2) ** (1.0 / 3.0)
        rng = np.random.default_rng(args.seed)
        coord = np.concatenate([rng.random((3, N)) * L, np.ones((1, N))], axis=0)
        ewald = (L, 0.015, 6, 6)
        t0 = time.perf_counter()
        h2 = C.build_periodic(coord, [0.0, 0.0, 0.0, L, L, L], 1.0 / (6.0 * math.pi), ewald, 1e-6, 200)
        tb = time.perf_counter() - t0
        log(f"[bench] periodic RPY N={N}: built in {tb:.1f}s, levels {h2.max_level + 1}, adm {h2.n_r_adm_pair}, "
            f"inadm {h2.n_r_inadm_pair}, root rank {h2.U_ncol[h2.root_idx]}")
        Nk = h2.krnl_mat_size
        x = rng.standard_normal(Nk)
        xd = torch.from_numpy(x).to(dev); yd = torch.empty_like(xd)
        m = H2MatHIP(h2, mode="jit")
        dt = _time(lambda: m.matvec_dev(xd.data_ptr(), yd.data_ptr(), sptr), K, W, sync)
        y = yd.cpu().numpy()
        log(f"[bench] periodic: {1 / dt:.2f} matvecs/s ({dt * 1e3:.2f} ms)")
        # accuracy: rows of the dense Ewald matrix (host, OpenMP)
        rows = np.sort(rng.choice(N, args.periodic_rows, replace=False))
        t0 = time.perf_counter()
        Kr = C.rpy_ewald_eval(coord[:, rows], coord, *ewald)
        te = time.perf_counter() - t0
        yr = Kr @ x
        yg = y.reshape(N, 3)[rows].reshape(-1)
        err = _rel(yg, yr)
        log(f"[bench] periodic: rel err vs dense Ewald on {rows.size} rows {err:.2e} ({te:.1f}s)")
        cpu = None
        if not args.no_cpu:
            t0 = time.perf_counter()
            yo = O.periodic_matvec(h2, x)
            tc = time.perf_counter() - t0
            cpu = {"value": 1.0 / tc, "unit": "matvecs/s", "cores": os.cpu_count() and min(16, os.cpu_count()),
                   "kind": "port", "sample": "1 full matvec of the same representation by the oracle's periodic "
                                             "restatement (numpy + the oracle's C kernel, BLAS threads)"}
            log(f"[bench] periodic oracle: {tc:.1f}s per matvec; GPU vs oracle {_rel(y, yo):.2e}")
        m.close()
        out.update({"metric": "H2 matvecs/s (periodic RPY-Ewald, N=%d, FP64 JIT)" % N, "value": 1.0 / dt,
                    "unit": "matvecs/s", "ms_per_step": dt * 1e3, "dtype": "f64",
                    "rel_err_vs_dense_ewald_rows": err, "check_rows": int(rows.size), "build_s": tb,
                    "cpu_baseline": cpu, "root_rank": int(h2.U_ncol[h2.root_idx]),
                    "adm_pairs": h2.n_r_adm_pair, "inadm_pairs": h2.n_r_inadm_pair})

    else:   # precision2m
        N = 2_000_000
        res = {}
        for kid, name, param in ((C.KRNL_GAUSSIAN, "gaussian", [1.0]), (C.KRNL_COULOMB, "coulomb", [])):
            coord = C.uniform_cube_points(N, 10.0, args.seed)
            t0 = time.perf_counter()
            h2 = C.build_h2(coord, kid, param, 1e-6, 300, BD_JIT=1, pp_kind="numerical")
            tb = time.perf_counter() - t0
            x = np.random.default_rng(args.seed + 1).random(N) - 0.5
            xd = torch.from_numpy(x).to(dev); yd = torch.empty_like(xd)
            ys, ts, ph, tn = {}, {}, {}, {}
            for prec in ("fp64", "fp32mix"):
                m = H2MatHIP(h2, mode="jit", prec=prec)
                tn[prec] = m.autotune(0)
                mv = lambda: m.matvec_dev(xd.data_ptr(), yd.data_ptr(), sptr)   # noqa: E731
                ts[prec] = _time(mv, K, W, sync)
                ys[prec] = yd.cpu().numpy()
                # the same steps again with HIP events around every phase (in order on one stream)
                m.set_profiling(True)
                _time(mv, K, 0, sync)
                info = m.info()
                m.set_profiling(False)
                ph[prec] = {k: round(info[k], 3) for k in ("t_fwd", "t_blocks", "t_reduce", "t_bwd", "t_final", "t_total")}
                log(f"[bench] {name} 2M {prec}: {ts[prec] * 1e3:.2f} ms per matvec; profiled phases ms {ph[prec]}; "
                    f"tune {tn[prec]}")
                m.close()
            rows = np.sort(np.random.default_rng(args.seed + 2).choice(N, 1000, replace=False))
            yref = direct_nbody(kid, param, coord, x, rows)
            res[name] = {"fp64_matvecs_per_s": 1 / ts["fp64"], "fp32mix_matvecs_per_s": 1 / ts["fp32mix"],
                         "fp32mix_speedup": ts["fp64"] / ts["fp32mix"],
                         "fp32mix_rel_err_vs_fp64_gpu": _rel(ys["fp32mix"], ys["fp64"]),
                         "fp64_rel_err_vs_direct_nbody_1000rows": _rel(ys["fp64"][rows], yref),
                         "fp32mix_rel_err_vs_direct_nbody_1000rows": _rel(ys["fp32mix"][rows], yref),
                         "build_s": tb, "levels": h2.max_level + 1, "adm_pairs": h2.n_r_adm_pair,
                         "phases_ms": ph, "tune": {k: {"variant": v["variants"][0], "d_concurrent": v["d_concurrent"],
                                                       "level_split": v["level_split"]} for k, v in tn.items()}}
            log(f"[bench] {name} 2M: fp64 {1 / ts['fp64']:.2f}/s fp32mix {1 / ts['fp32mix']:.2f}/s "
                f"errs {res[name]['fp32mix_rel_err_vs_fp64_gpu']:.2e} {res[name]['fp64_rel_err_vs_direct_nbody_1000rows']:.2e}")
            del h2
        out.update({"metric": "H2 matvecs/s at N=2M, FP64 vs FP32-mixed (Gaussian, Coulomb)",
                    "value": res["gaussian"]["fp32mix_matvecs_per_s"], "unit": "matvecs/s",
                    "dtype": "f32-mixed", "gaussian": res["gaussian"], "coulomb": res["coulomb"]})
    print(json.dumps(out), flush=True)
