#!/usr/bin/env python
"""H2 matvec throughput on one MI355X.

Workload (the drop-in configuration): 3D Gaussian kernel exp(-|x-y|^2/l^2)
with l = 1 (H2Pack parameter 1/l^2 = 1), N = 1,000,000 points i.i.d. uniform
in a cube at density 10 points per unit volume, H2 built at reltol 1e-6 with
at most 300 points per leaf (numerical proxy points), one FP64 right-hand
side x ~ U(-0.5, 0.5), nugget 1e-2 added on the diagonal outside the H2
product (y = A_H2 x + 1e-2 x).

A step is one (A_H2 + sigma^2 I) x on vectors already resident in HBM.
`value` is FP64 JIT-mode matvecs per second.  The same run also reports the
FP32-mixed JIT leg, the AOT mode (stored blocks, HBM GB/s against the H2Pack
byte model), the relative error against an FP64 direct n-body product on
1000 sampled rows and against the oracle, the roofline of the dominant
kernel, a serving leg (Matern-3/2 GP requests with Poisson arrivals:
continuous batching on the FP32-mixed path against one request at a time on
the FP64 path, p50/p99 latency and requests/s; the requests' H2 matrices are
built in worker processes while the N = 1M legs run), and the oracle's CPU
time on this host.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N] [--no-aot]
                    [--serve-requests R] [--serve-rate Q]
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

HBM_PEAK_GBS = 8000.0     # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 (vector and matrix)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(h2, x, budget_s: float):
    """Time the oracle (C restatement of H2P_matvec, OpenMP) on this host."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    nt = max(1, min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "0") or (os.cpu_count() or 1))))
    o = O.OracleH2(h2, BD_JIT=1)
    t0 = time.perf_counter()
    y = o.matvec(x, n_thread=nt)
    calls = 1
    el = time.perf_counter() - t0
    while el < budget_s * 0.5 and calls < 10:
        o.matvec(x, n_thread=nt)
        calls += 1
        el = time.perf_counter() - t0
    return y, calls / el, nt, calls, el


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--no-aot", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--check-rows", type=int, default=1000)
    ap.add_argument("--workload", default="gaussian1m",
                    choices=["gaussian1m", "coulomb80k", "rpy_matmul", "precision2m", "serving", "periodic",
                             "serving-replay"],
                    help="gaussian1m: the headline (default); the others are the remaining "
                         "configurations, each printed as its own JSON line")
    ap.add_argument("--requests", type=int, default=32, help="serving: number of requests")
    ap.add_argument("--rate", type=float, default=2.0, help="serving: Poisson arrivals per second")
    ap.add_argument("--slots", type=int, default=32, help="serving: most requests on the GPU at once")
    ap.add_argument("--cg-max-iter", type=int, default=3000, help="serving: CG iteration cap")
    ap.add_argument("--build-workers", type=int, default=4, help="serving: host construction processes")
    ap.add_argument("--cg-active", type=int, default=16,
                    help="serving: most CG solves advanced per tick, the longest admitted first (0: all)")
    ap.add_argument("--serve-legs", default="both", choices=["both", "batched", "alone", "alone-service"],
                    help="serving: both legs on prebuilt requests, one leg on requests built while served, "
                         "or the FP64-alone service time of each request (back to back, no arrival clock)")
    ap.add_argument("--serve-first", type=int, default=0, help="serving, alone-service: first request")
    ap.add_argument("--serve-count", type=int, default=0, help="serving, alone-service: requests (0: the rest)")
    ap.add_argument("--replay-files", nargs="*", default=[],
                    help="serving-replay: JSON outputs of --serve-legs alone-service covering every request")
    ap.add_argument("--replay-rates", type=float, nargs="*", default=[0.3, 0.6],
                    help="serving-replay: arrival rates to replay")
    ap.add_argument("--build-window", type=int, default=64,
                    help="serving, one leg: most requests built but not yet admitted")
    ap.add_argument("--build-head", type=int, default=16,
                    help="serving, one leg: requests built before the clock starts")
    ap.add_argument("--serve-requests", type=int, default=24,
                    help="gaussian1m: requests of the serving leg (0: no serving leg)")
    ap.add_argument("--serve-rate", type=float, default=1.5, help="gaussian1m: serving leg arrivals per second")
    ap.add_argument("--periodic-n", type=int, default=100_000, help="periodic: particles")
    ap.add_argument("--periodic-rows", type=int, default=10, help="periodic: dense Ewald check rows")
    ap.add_argument("--copy-calib", action="store_true",
                    help="run one 2 GiB device-to-device copy first, the reference kernel "
                         "for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE readings")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU sharding is a design note, not a measured path.", file=sys.stderr)
        sys.exit(2)

    import torch
    from h2pack_amd import construct as C
    from h2pack_amd.h2mat import H2MatHIP, direct_nbody

    if args.copy_calib:
        src = torch.ones(2 << 27, dtype=torch.float64, device="cuda:0")   # 2 GiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log("[bench] copy calibration: 2 GiB read + 2 GiB written by one copy kernel")
        del src, dst

    if args.workload != "gaussian1m":
        import bench_workloads
        bench_workloads.run(args)
        return

    N, K, W = args.n, args.steps, args.warmup
    sigma2 = 1e-2
    t0 = time.perf_counter()
    coord = C.uniform_cube_points(N, 10.0, args.seed)
    h2 = C.build_h2(coord, C.KRNL_GAUSSIAN, [1.0], 1e-6, 300, BD_JIT=1, pp_kind="numerical",
                    host_blocks=False)
    t_build = time.perf_counter() - t0
    log(f"[bench] built H2 for N={N} in {t_build:.1f}s: levels={h2.max_level + 1} "
        f"leaves={h2.n_leaf_node} adm={h2.n_r_adm_pair} inadm={h2.n_r_inadm_pair}")
    # the serving leg's requests are built in worker processes while the
    # N = 1M legs run on the GPU (started before this process uses the GPU)
    serve_builds = None
    if args.serve_requests > 0:
        import bench_workloads
        t_sb = time.perf_counter()
        serve_builds = bench_workloads.start_request_builds(args.serve_requests, args.seed, args.serve_rate,
                                                            args.cg_max_iter, args.build_workers)

    rng = np.random.default_rng(args.seed + 1)
    x = rng.random(N) - 0.5
    dev = torch.device("cuda:0")
    xd = torch.from_numpy(x).to(dev)
    yd = torch.empty_like(xd)
    stream = torch.cuda.current_stream()
    sptr = stream.cuda_stream

    def tune(m, label):
        t = time.perf_counter()
        r = m.autotune(0)
        r["tune_s"] = time.perf_counter() - t
        log(f"[bench] {label}: autotuned block launch -> variant {r['variants'][0]} "
            f"(default {r['default_ms']:.3f} ms, chosen {r['tuned_ms']:.3f} ms, {r['tune_s']:.1f}s)"
            + ("" if "aot" in label else
               f"; D blocks {'beside the sweeps' if r['d_concurrent'] else 'in order'}"
               f"; level split {'on' if r['level_split'] else 'off'}"))
        return r

    def timed(m, label):
        def step():
            m.matvec_dev(xd.data_ptr(), yd.data_ptr(), sptr)
            yd.add_(xd, alpha=sigma2)
        for _ in range(W):
            step()
        torch.cuda.synchronize()
        # timed region: the production path (captured hipGraph for the interior)
        t = time.perf_counter()
        for _ in range(K):
            step()
        torch.cuda.synchronize()
        el = time.perf_counter() - t
        # the same K steps again with HIP events around every phase (eager
        # launches on the same stream) for the per-kernel durations
        m.set_profiling(True)
        t = time.perf_counter()
        for _ in range(K):
            step()
        torch.cuda.synchronize()
        el_prof = time.perf_counter() - t
        info = m.info()
        m.set_profiling(False)
        info["ms_per_step_profiled"] = el_prof / K * 1e3
        log(f"[bench] {label}: {K} steps in {el * 1e3:.2f} ms -> {K / el:.2f} matvec/s "
            f"(profiled pass {el_prof / K * 1e3:.3f} ms/step); phases ms fwd {info['t_fwd']:.3f} "
            f"blocks {info['t_blocks']:.3f} reduce {info['t_reduce']:.3f} bwd {info['t_bwd']:.3f} final {info['t_final']:.3f}")
        return el, info, (yd.cpu().numpy() - sigma2 * x)

    # ---- FP64 JIT (headline)
    t = time.perf_counter()
    mj = H2MatHIP(h2, mode="jit")
    log(f"[bench] JIT upload {time.perf_counter() - t:.1f}s")
    tune_j = tune(mj, "fp64-jit")
    el_j, info_j, y_j = timed(mj, "fp64-jit")
    # ---- accuracy: direct FP64 n-body on sampled rows (device utility)
    rows = np.sort(rng.choice(N, args.check_rows, replace=False))
    y_dir = direct_nbody(C.KRNL_GAUSSIAN, [1.0], coord, x, rows)
    err_direct = float(np.linalg.norm(y_j[rows] - y_dir) / np.linalg.norm(y_dir))
    log(f"[bench] fp64-jit rel err vs direct n-body ({args.check_rows} rows) = {err_direct:.3e}")
    mj.close()
    del mj

    # ---- FP32-mixed JIT
    t = time.perf_counter()
    mf = H2MatHIP(h2, mode="jit", prec="fp32mix")
    tune_f = tune(mf, "fp32mix-jit")
    el_f, info_f, y_f = timed(mf, "fp32mix-jit")
    mf.close()
    del mf
    err_f_direct = float(np.linalg.norm(y_f[rows] - y_dir) / np.linalg.norm(y_dir))
    err_f_fp64 = float(np.linalg.norm(y_f - y_j) / np.linalg.norm(y_j))
    fp32 = {"matvecs_per_s": K / el_f, "ms_per_step": el_f / K * 1e3,
            "speedup_vs_fp64": el_j / el_f,
            "rel_err_vs_fp64_gpu": err_f_fp64,
            "rel_err_vs_direct_nbody_1000rows": err_f_direct,
            "kernel_ms": info_f["t_blocks"], "block_variant": tune_f["variants"][0],
            "d_concurrent": tune_f["d_concurrent"], "level_split": tune_f["level_split"]}
    log(f"[bench] fp32mix-jit: {K / el_f:.2f} matvec/s ({el_j / el_f:.2f}x FP64); rel err vs FP64 GPU "
        f"{err_f_fp64:.3e}, vs direct n-body {err_f_direct:.3e}")

    # ---- AOT (stored blocks evaluated into HBM at upload)
    aot = None
    if not args.no_aot:
        try:
            t = time.perf_counter()
            ma = H2MatHIP(h2, mode="aot", host_blocks=False)
            log(f"[bench] AOT upload (device block build) {time.perf_counter() - t:.1f}s")
            tune_a = tune(ma, "fp64-aot")
            el_a, info_a, y_a = timed(ma, "fp64-aot")
            ma.close()
            del ma
            t = time.perf_counter()
            ma32 = H2MatHIP(h2, mode="aot", prec="fp32mix", host_blocks=False)
            log(f"[bench] AOT FP32 upload (device block build) {time.perf_counter() - t:.1f}s")
            tune_a32 = tune(ma32, "fp32mix-aot")
            el_a32, info_a32, y_a32 = timed(ma32, "fp32mix-aot")
            ma32.close()
            del ma32
            blk32 = 4.0 * (info_a32["B_elems"] + info_a32["D_elems"])
            gbs_model = info_a["hbm_bytes_model"] / (el_a / K) / 1e9
            blk_bytes = 8.0 * (info_a["B_elems"] + info_a["D_elems"])
            gbs_blk = blk_bytes / (info_a["t_blocks"] * 1e-3) / 1e9
            aot = {"matvecs_per_s": K / el_a, "ms_per_step": el_a / K * 1e3,
                   "hbm_GBps_model": gbs_model,
                   "roofline": {"bound": "hbm", "achieved": gbs_blk, "peak": HBM_PEAK_GBS,
                                "unit": "GB/s", "frac": gbs_blk / HBM_PEAK_GBS, "traffic": None,
                                "kernel": ("k_blocks_aot_occ<double, 6>" if tune_a["variants"][0] == 1
                                           else "k_blocks_aot<double>"), "kernel_ms": info_a["t_blocks"]},
                   "rel_err_vs_jit": float(np.linalg.norm(y_a - y_j) / np.linalg.norm(y_j)),
                   "stored_GB": blk_bytes / 1e9,
                   "fp32mix": {"matvecs_per_s": K / el_a32, "ms_per_step": el_a32 / K * 1e3,
                               "block_variant": tune_a32["variants"][0],
                               "stored_GB": blk32 / 1e9,
                               "kernel_GBps": blk32 / (info_a32["t_blocks"] * 1e-3) / 1e9,
                               "rel_err_vs_fp64_gpu": float(np.linalg.norm(y_a32 - y_j) / np.linalg.norm(y_j))}}
        except Exception as e:   # report, do not hide the headline
            aot = {"error": str(e)}
            log(f"[bench] AOT leg failed: {e}")

    # ---- serving: Matern-3/2 GP requests, continuous batching (FP32-mixed)
    # against one request at a time on the FP64 path
    serving = None
    if serve_builds is not None:
        try:
            reqs = bench_workloads.collect_requests(*serve_builds)
            t_sb = time.perf_counter() - t_sb
            cg_active = args.cg_active if args.cg_active > 0 else None
            r = bench_workloads.serve_compare(reqs, slots=args.serve_requests, cg_active=cg_active)
            d32, d64 = r["batched_fp32mix"], r["alone_fp64"]
            serving = {"workload": f"serving_matern32_{args.serve_requests}req_N1e4-4e5_density10_reltol1e-6_"
                                   f"leaf300_poisson{args.serve_rate}_8rhs_cg_tol1e-6_nugget1e-2",
                       "requests": args.serve_requests, "rate_per_s": args.serve_rate,
                       "cg_active": cg_active, "p50_s": d32["p50"], "p99_s": d32["p99"], "throughput_req_per_s": d32["throughput"],
                       "alone_fp64_p50_s": d64["p50"], "alone_fp64_p99_s": d64["p99"],
                       "alone_fp64_throughput_req_per_s": d64["throughput"],
                       "host_build_s_overlapped": t_sb, **r}
        except Exception as e:   # report, do not hide the headline
            serving = {"error": str(e)}
            log(f"[bench] serving leg failed: {e}")

    # ---- CPU baseline: the oracle on this host, also the FP64 parity check
    cpu = None
    err_oracle = None
    if not args.no_cpu:
        y_o, rate, nt, calls, el = cpu_baseline(h2, x, args.cpu_budget)
        err_oracle = float(np.linalg.norm(y_j - y_o) / np.linalg.norm(y_o))
        cpu = {"value": rate, "unit": "matvecs/s", "cores": nt, "kind": "port",
               "sample": f"{calls} full FP64 JIT matvec(s) of the same N={N} H2 matrix, "
                         f"{el:.1f}s, oracle/h2oracle.c (OpenMP, no BLAS)"}
        log(f"[bench] oracle CPU: {rate:.3f} matvec/s on {nt} threads; GPU vs oracle rel err {err_oracle:.3e}")

    # HBM bytes of the named kernel in one matvec (its launches summed where it
    # runs as two, like the JIT block kernels), from the committed rocprofv3
    # FETCH_SIZE / WRITE_SIZE summary of this workload (null when absent)
    def traffic(kernel):
        try:
            with open(os.path.join(ROOT, "profiles", "gaussian_1M_traffic.json")) as f:
                t = json.load(f)
            if N != 1_000_000:
                return None
            e = t["kernels"][kernel]
            return e.get("hbm_bytes_per_matvec", e["hbm_bytes"])
        except (OSError, KeyError, ValueError):
            return None

    if aot and "roofline" in aot:
        aot["roofline"]["traffic"] = traffic(aot["roofline"]["kernel"])
    # the dominant launch is the B-block launch (the D blocks are a second
    # launch, concurrent with the sweeps in the timed path); the flop model
    # splits by kernel pairs
    flops = info_j["jit_flops_model"]
    pairs = info_j["B_pairs"] + info_j["D_pairs"]
    flops_b = flops * info_j["B_pairs"] / pairs
    flops_d = flops * info_j["D_pairs"] / pairs
    ach = flops_b / (info_j["t_blocks_b"] * 1e-3) / 1e12
    jv = tune_j["variants"][0]
    out = {
        "metric": "H2 matvecs/s at N=1M (FP64 vs FP32-mixed, rel err); served p50/p99 latency",
        "value": K / el_j,
        "unit": "matvecs/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": el_j / K * 1e3,
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"gaussian3d_l1_N{N}_density10_reltol1e-6_leaf300_jit_fp64_nugget1e-2",
                   "N": N, "mode": "jit", "precision": "fp64", "kernel": "gaussian", "reltol": 1e-6,
                   "leaf": 300, "parallelism": "single-gpu"},
        "roofline": {"bound": "mfma",
                     "pipe": "FP64 issue: the launch is compute bound on the shared FP64 pipe (VALU kernel "
                             "evaluation plus the MFMA distance tiles), not on HBM; 'mfma' names the compute "
                             "bound, the peak is the FP64 one",
                     "achieved": ach, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": ach / FP64_PEAK_TFLOPS,
                     "kernel": ("k_blocks_jit2_occ<1, double, 4, 0>" if jv == 0
                                else "k_blocks_jit2<1, double, 0>"), "kernel_ms": info_j["t_blocks_b"],
                     "flops_per_launch": flops_b, "block_variant": jv,
                     "d_concurrent": tune_j["d_concurrent"], "level_split": tune_j["level_split"],
                     "d_launch": {"kernel": ("k_blocks_jit2_occ<1, double, 4, 1>" if jv == 0
                                             else "k_blocks_jit2<1, double, 1>"),
                                  "kernel_ms": info_j["t_blocks_d"], "flops_per_launch": flops_d,
                                  "achieved_TFLOPs": flops_d / (info_j["t_blocks_d"] * 1e-3) / 1e12},
                     "traffic_source": "profiles/gaussian_1M_hbm_traffic.md (rocprofv3 FETCH_SIZE/WRITE_SIZE)",
                     "note": "achieved = H2Pack JIT_flops model (14 flops per Gaussian pair) / kernel time; "
                             "the B blocks are two launches of this kernel per matvec (the deep levels' blocks "
                             "and, with the level split, the shallower levels'), and kernel_ms is their sum "
                             "measured with HIP events in the profiled pass, where they run in order; "
                             "FP64 VALU and FP64 MFMA share the 78.6 TFLOP/s peak (one FP64 pipe: "
                             "SQ_VALU_MFMA_COEXEC_CYCLES = 0); evaluating a pair issues ~18 pipe slots "
                             "(2^x 12 FP64/integer operations with its 128-entry LDS table, the squared "
                             "distance 4 slots of MFMA, the two products 2), and rocprofv3 counters put the "
                             "pipe ~90% busy (profiles/gaussian_1M_jit_pipe_counters.md)"},
        "cpu_baseline": cpu,
        "serving": serving,
        "rel_err_vs_direct_nbody_1000rows": err_direct,
        "rel_err_vs_oracle": err_oracle,
        "fp32mix": fp32,
        "aot": aot,
        "phases_ms": {k: info_j[k] for k in ("t_fwd", "t_blocks", "t_blocks_b", "t_blocks_d", "t_reduce", "t_bwd",
                                             "t_final", "t_total", "ms_per_step_profiled")},
        "h2": {"build_s": t_build, "levels": h2.max_level + 1, "leaves": h2.n_leaf_node,
               "adm_pairs": h2.n_r_adm_pair, "inadm_pairs": h2.n_r_inadm_pair,
               "B_pairs": info_j["B_pairs"], "D_pairs": info_j["D_pairs"]},
    }
    out["roofline"]["traffic"] = traffic(out["roofline"]["kernel"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
