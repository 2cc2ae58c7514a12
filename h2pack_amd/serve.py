"""Continuous-batching request loop over resident H2 matrices.

A request is a kernel matrix already built on the host (the reference's
construction: ``H2P_partition_points`` + proxy points + ``H2P_build``), plus
its work: either a multi-RHS product Y = (K + shift I) X (``H2P_matmul``,
src/H2Pack_matmul.c:455) or a CG solve of (K + shift I) x = b to a relative
residual (the unpreconditioned ``pcg`` of examples/PCG/pcg.c:11-152 with the
shift wrapper of examples/SPDHSS-H2/pcg_tests.c:15-21).

The loop:
  * admits arrived requests first-come-first-served while their device
    footprint (``h2p_hip_footprint`` plus the request's vectors) fits the HBM
    budget and a slot is free; an admitted request is uploaded once and its
    state (matrix, vectors, CG scalars) stays on the device until it retires;
  * runs the CG requests together: every step is one co-scheduled batch of
    ``cg_chunk`` iterations over all active solves (``h2p_hip_cg_step``), so
    the blocks of different matrices share each grouped launch, and solves
    admitted mid-flight join at the next step;
  * runs the products of newly admitted matmul requests as one co-scheduled
    ``h2p_hip_matmul_batch`` on a second stream, overlapping the CG steps;
  * retires finished requests (result copied to the host), frees their
    memory, and admits more.

Each request's result does not depend on which requests it shared launches
with: the device code per task is the same and every sum has a fixed order.
"""
from __future__ import annotations

import math
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import abi
from .h2mat import H2Batch, H2MatHIP, footprint


@dataclass
class Request:
    rid: int
    h2: object                 # host-built representation (construct.H2Pack)
    kind: str                  # "cg" or "matmul"
    rhs: np.ndarray            # b (N,) or X (N, n_vec)
    arrival: float = 0.0       # seconds after the start of the run
    shift: float = 0.0         # nugget: the operator is K + shift I
    tol: float = 1e-6
    max_iter: int = 1000
    # filled by the server
    nbytes: int = 0            # device bytes while active, for the server's mode
    admitted: float = math.nan
    finished: float = math.nan
    result: np.ndarray | None = None
    iters: int = 0
    relres: float = math.nan
    rejected: bool = False
    # a request built while the server runs: h2 and rhs are None until
    # ``future`` (resolving to a built Request) is done; ``loaded`` is when
    # the server took it over (requests whose build is late enter the queue
    # then, not at their arrival)
    future: object = None
    loaded: float = math.nan
    n_mat: int = 0
    # when its admission (upload) began: with one request at a time,
    # finished - admit_start is the request's whole service time
    admit_start: float = math.nan

    @property
    def N(self) -> int:
        return self.h2.krnl_mat_size if self.h2 is not None else self.n_mat

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def request_bytes(req: Request, mode: str, prec: str, refine: bool = False) -> int:
    """Device bytes the request holds while active (with ``refine``, a CG
    request also holds the FP64 handle of its residual replacements)."""
    n_vec = req.rhs.shape[1] if req.kind == "matmul" else 0
    b = footprint(req.h2, mode, prec, n_vec=n_vec, with_cg=req.kind == "cg")
    vec = 2 * req.rhs.size * 8            # b and x, or X and Y
    if refine and req.kind == "cg":
        b += footprint(req.h2, mode, "fp64")
    return b + vec


@dataclass
class _Active:
    req: Request
    mat: H2MatHIP
    a: object                  # device b / X
    b: object                  # device x / Y
    # FP64 handle of an FP32-mixed CG solve's residual replacements (see Server)
    mat64: H2MatHIP | None = None
    refines: int = 0           # residual replacements so far


@dataclass
class ServeStats:
    n_done: int = 0
    n_rejected: int = 0
    makespan: float = 0.0
    throughput: float = 0.0    # requests per second
    p50: float = 0.0
    p99: float = 0.0
    mean: float = 0.0
    max_active: int = 0
    peak_bytes: int = 0
    cg_iters: int = 0
    cg_replacements: int = 0   # FP64 residual replacements of FP32-mixed solves
    ticks: int = 0
    latencies: list = field(default_factory=list)

    def as_dict(self):
        d = dict(self.__dict__)
        d.pop("latencies")
        return d


class Server:
    """Continuous batching of GP requests on one GPU.

    CG requests on the FP32-mixed path get residual replacements (``refine``,
    default on for FP32-mixed): the FP32 operator's rounding is relative to
    |A| |x|, so the residual that CG's recurrence carries drifts from the
    FP64 operator's b - (A64 + shift) x by up to ~1e-7 kappa (kappa the
    condition number).  Each solve pauses on the device when its recurrence
    residual claims convergence or has dropped by ``replace_ratio`` since its
    residual was last formed; the server then forms r = b - (A64 + shift) x
    with an FP64 handle of the same representation (h2p_hip_cg_replace) and
    the solve goes on from there with the same search direction (the
    reliable updates of Sleijpen and van der Vorst, 1996), or ends once that
    residual is below tol |b|.  Unlike restarting CG on a correction
    equation, nothing of the Krylov space is lost, so the iteration count
    stays near the FP64 solve's.  ``relres`` is then the FP64 residual.
    Where a solve pauses depends only on its own iterates, so its result is
    the same whichever requests share its launches.
    """

    def __init__(self, hbm_budget: int, max_active: int = 64, mode: str = "jit",
                 prec: str = "fp32mix", cg_chunk: int = 8, device: str = "cuda:0",
                 refine: bool | None = None, replace_ratio: float = 0.1, drop_host: bool = False,
                 cg_active: int | None = None):
        import torch
        self.torch = torch
        self.budget = int(hbm_budget)
        self.max_active = int(max_active)
        self.mode, self.prec = mode, prec
        self.refine = (prec == "fp32mix") if refine is None else bool(refine)
        self.replace_ratio = float(replace_ratio)
        # at most this many CG solves (the longest admitted) advance per tick;
        # the others stay resident and wait (None: all of them)
        self.cg_active = None if cg_active is None else max(1, int(cg_active))
        self.drop_host = bool(drop_host)   # free a request's host representation once uploaded
        self.cg_chunk = int(cg_chunk)
        self.dev = torch.device(device)
        # two streams at the default priority
        self.s_cg = torch.cuda.Stream(device=self.dev)
        self.s_mm = torch.cuda.Stream(device=self.dev)
        self.bt_cg = H2Batch()
        self.bt_mm = H2Batch()

    def close(self):
        self.bt_cg.close()
        self.bt_mm.close()

    # ------------------------------------------------------------------
    def _admit(self, req: Request, stream):
        torch = self.torch
        mat = H2MatHIP(req.h2, mode=self.mode, prec=self.prec, host_blocks=False)
        with torch.cuda.stream(stream):
            a = torch.from_numpy(np.ascontiguousarray(req.rhs.T if req.kind == "matmul" else req.rhs))
            a = a.to(self.dev, non_blocking=False)
            b = torch.zeros_like(a)
        act = _Active(req, mat, a, b)
        if self.refine and req.kind == "cg":
            act.mat64 = H2MatHIP(req.h2, mode=self.mode, prec="fp64", host_blocks=False)
            mat.set_cg_replace(self.replace_ratio)
        return act

    def run(self, requests: list[Request], realtime: bool = True, log=None,
            log_every: float = 15.0) -> ServeStats:
        """Serve ``requests`` (arrival times in seconds from the start; with
        ``realtime`` False every request is available at once).  ``log``, if
        given, receives a progress line every ``log_every`` seconds."""
        torch = self.torch
        pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
        for r in pending:
            if r.h2 is not None:
                r.nbytes = request_bytes(r, self.mode, self.prec, self.refine)
                r.n_mat = r.h2.krnl_mat_size
            r.iters = 0
        queue: deque[Request] = deque()
        cg: list[_Active] = []
        mm_inflight: list[tuple[list[_Active], object]] = []
        to_free: list[_Active] = []
        used = 0
        st = ServeStats()
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0   # noqa: E731
        i = 0
        t_log = 0.0
        while i < len(pending) or queue or cg or mm_inflight:
            t = now()
            if log and t - t_log >= log_every:
                t_log = t
                log(f"[serve] t={t:.1f}s done {st.n_done}/{len(pending)} active cg {len(cg)} "
                    f"products in flight {sum(len(g) for g, _ in mm_inflight)} queued {len(queue)} "
                    f"bytes {used / 1e9:.1f} GB")
            late = False
            while i < len(pending) and (not realtime or pending[i].arrival <= t):
                r = pending[i]
                if r.h2 is None:
                    f = r.future
                    if f is None or not f.done():
                        late = True
                        break
                    built = f.result()
                    built = built[0] if isinstance(built, tuple) else built
                    r.h2, r.rhs, r.future = built.h2, built.rhs, None
                    r.shift, r.tol, r.max_iter = built.shift, built.tol, built.max_iter
                    r.n_mat = r.h2.krnl_mat_size
                    r.nbytes = request_bytes(r, self.mode, self.prec, self.refine)
                r.loaded = now()
                queue.append(r)
                i += 1
            # ---- admission, first come first served
            new_cg, new_mm = [], []
            n_active = len(cg) + sum(len(g) for g, _ in mm_inflight)
            while queue and n_active < self.max_active:
                r = queue[0]
                if r.nbytes > self.budget:
                    queue.popleft()
                    r.rejected = True
                    st.n_rejected += 1
                    continue
                if used + r.nbytes > self.budget:
                    break
                queue.popleft()
                r.admit_start = now()
                act = self._admit(r, self.s_cg if r.kind == "cg" else self.s_mm)
                r.admitted = now()
                if self.drop_host:
                    r.h2 = None
                    act.mat.h2 = None
                    if act.mat64 is not None:
                        act.mat64.h2 = None
                used += r.nbytes
                st.peak_bytes = max(st.peak_bytes, used)
                n_active += 1
                (new_cg if r.kind == "cg" else new_mm).append(act)
            st.max_active = max(st.max_active, n_active)
            # ---- products of the new matmul requests, one co-scheduled call
            if new_mm:
                with torch.cuda.stream(self.s_mm):
                    self.bt_mm.matmul([a.mat for a in new_mm], abi.COL_MAJOR,
                                      [a.req.rhs.shape[1] for a in new_mm],
                                      [a.a.data_ptr() for a in new_mm], [a.mat.N for a in new_mm],
                                      [a.b.data_ptr() for a in new_mm], [a.mat.N for a in new_mm],
                                      stream=self.s_mm.cuda_stream)
                    for a in new_mm:
                        if a.req.shift:
                            a.b.add_(a.a, alpha=a.req.shift)
                    ev = torch.cuda.Event()
                    ev.record(self.s_mm)
                mm_inflight.append((new_mm, ev))
            # ---- CG: start the new solves, then one step over every solve
            if new_cg:
                self.bt_cg.cg_start([a.mat for a in new_cg], [a.a.data_ptr() for a in new_cg],
                                    [a.b.data_ptr() for a in new_cg], [a.req.shift for a in new_cg],
                                    [a.req.tol for a in new_cg], [a.req.max_iter for a in new_cg],
                                    stream=self.s_cg.cuda_stream)
                cg.extend(new_cg)
            if cg:
                run = cg if self.cg_active is None else \
                    sorted(cg, key=lambda a: (a.req.admitted, a.req.rid))[:self.cg_active]
                self.bt_cg.cg_step([a.mat for a in run], self.cg_chunk, stream=self.s_cg.cuda_stream)
                st.ticks += 1
            # ---- retire finished products
            still = []
            for grp, ev in mm_inflight:
                if ev.query():
                    for a in grp:
                        a.req.result = a.b.T.cpu().numpy()
                        a.req.finished = now()
                        to_free.append(a)
                        st.n_done += 1
                else:
                    still.append((grp, ev))
            mm_inflight = still
            # ---- poll the solves (waits for the CG stream)
            if cg:
                done, its, rr = self.bt_cg.cg_poll([a.mat for a in cg], stream=self.s_cg.cuda_stream)
                keep, fin, replace = [], [], []
                for a, d, it, r in zip(cg, done, its, rr):
                    a.req.iters = it
                    if d == 0:
                        keep.append(a)
                    elif a.mat64 is not None:
                        replace.append(a)      # paused, or at max_iter: its FP64 residual decides
                    else:
                        a.req.relres = r
                        fin.append(a)
                # residual replacement: r = b - (A64 + shift) x, one co-scheduled FP64 product
                if replace:
                    mats = [a.mat for a in replace]
                    self.bt_cg.cg_replace(mats, [a.mat64 for a in replace], stream=self.s_cg.cuda_stream)
                    done, its, rr = self.bt_cg.cg_poll(mats, stream=self.s_cg.cuda_stream)
                    for a, d, it, r in zip(replace, done, its, rr):
                        a.refines += 1
                        a.req.iters = it
                        if d == 1:
                            a.req.relres = r
                            fin.append(a)
                        else:
                            keep.append(a)
                for a in fin:
                    st.cg_iters += a.req.iters
                    st.cg_replacements += a.refines
                    a.req.result = a.b.cpu().numpy()
                    a.req.finished = now()
                    to_free.append(a)
                    st.n_done += 1
                cg = keep
            # ---- free retired requests once no product is in flight (a free
            # waits for the device); their bytes return to the budget then
            if to_free and not mm_inflight:
                for a in to_free:
                    a.mat.close()
                    if a.mat64 is not None:
                        a.mat64.close()
                    a.a = a.b = None
                    used -= a.req.nbytes
                to_free = []
            # ---- only products in flight: wait for the oldest one
            if not cg and mm_inflight:
                mm_inflight[0][1].synchronize()
            # ---- idle: wait for the next arrival (or its build)
            elif not cg and not queue and i < len(pending) and realtime:
                time.sleep(0.005 if late else max(0.0, min(pending[i].arrival - now(), 0.05)))
        torch.cuda.synchronize()
        for a in to_free:
            a.mat.close()
            if a.mat64 is not None:
                a.mat64.close()
        st.makespan = now()
        lat = sorted(r.latency for r in requests if not r.rejected)
        st.latencies = lat
        if lat:
            st.p50 = float(np.percentile(lat, 50))
            st.p99 = float(np.percentile(lat, 99))
            st.mean = float(np.mean(lat))
        st.throughput = st.n_done / st.makespan if st.makespan > 0 else 0.0
        return st


def fifo_latencies(arrivals, services) -> np.ndarray:
    """Latencies of a single first-come-first-served server: request k starts
    at max(its arrival, the end of request k - 1) and takes services[k]
    seconds.  This is exactly the one-request-at-a-time server
    (``max_active`` = 1) when each request's service time does not depend on
    the others, so its latencies at any arrival rate follow from service
    times measured once, back to back."""
    arr = np.asarray(arrivals, dtype=np.float64)
    svc = np.asarray(services, dtype=np.float64)
    if arr.shape != svc.shape:
        raise ValueError("one service time per arrival")
    if np.any(np.diff(arr) < 0) or np.any(svc < 0):
        raise ValueError("arrivals must be sorted and service times non-negative")
    lat = np.empty_like(arr)
    end = -np.inf
    for k in range(arr.size):
        end = max(arr[k], end) + svc[k]
        lat[k] = end - arr[k]
    return lat


def request_specs(n: int, seed: int, n_min: int = 10_000, n_max: int = 400_000,
                  rate: float = 4.0) -> list[dict]:
    """The random draws of ``n`` GP requests: Poisson arrivals at ``rate`` per
    second, N log-uniform in [n_min, n_max], Matern-3/2 lengthscale U(0.5, 2),
    alternating 8-RHS products and CG solves."""
    rng = np.random.default_rng(seed)
    out, t = [], 0.0
    for k in range(n):
        t += rng.exponential(1.0 / rate)
        N = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
        out.append({"rid": k, "arrival": t, "N": N, "ell": float(rng.uniform(0.5, 2.0)),
                    "kind": "matmul" if k % 2 == 0 else "cg",
                    "seed": int(rng.integers(1 << 31))})
    return out


def build_request(spec: dict, density: float = 10.0, reltol: float = 1e-6, leaf: int = 300,
                  n_vec: int = 8, nugget: float = 1e-2, tol: float = 1e-6,
                  max_iter: int = 1000) -> Request:
    """Host construction of one request (uniform-cube points at ``density``,
    Matern-3/2 with l = 1 / lengthscale, the nugget as the shift)."""
    from . import construct as C
    N = spec["N"]
    rng = np.random.default_rng(spec["seed"])
    coord = C.uniform_cube_points(N, density, seed=spec["seed"])
    h2 = C.build_h2(coord, C.KRNL_MATERN32, [1.0 / spec["ell"]], reltol, leaf, BD_JIT=1)
    kind = spec["kind"]
    rhs = rng.standard_normal((N, n_vec)) if kind == "matmul" else rng.standard_normal(N)
    return Request(rid=spec["rid"], h2=h2, kind=kind, rhs=rhs, arrival=spec["arrival"],
                   shift=nugget, tol=tol, max_iter=max_iter)


def make_requests(n: int, seed: int, n_min: int = 10_000, n_max: int = 400_000,
                  rate: float = 4.0, log=None, **kw) -> list[Request]:
    """Build ``n`` requests in this process (see request_specs / build_request)."""
    reqs = []
    for spec in request_specs(n, seed, n_min, n_max, rate):
        t_b = time.perf_counter()
        reqs.append(build_request(spec, **kw))
        if log:
            log(f"request {spec['rid']}: N={spec['N']} lengthscale={spec['ell']:.3f} "
                f"built in {time.perf_counter() - t_b:.1f}s")
    return reqs
