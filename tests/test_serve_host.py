"""CPU: host logic of the serving loop (no GPU).

fifo_latencies (the one-request-at-a-time server's latencies from measured
service times) is checked against closed forms and against the max-plus
form of the same recurrence on random inputs; request_specs against the
rate scaling the replay relies on (arrivals at rate r are the rate-1 draws
divided by r, every other draw the same).
"""
import numpy as np
import pytest

from h2pack_amd.serve import fifo_latencies, request_specs


def test_fifo_no_queueing_when_arrivals_are_sparse():
    arr = np.arange(10) * 5.0
    svc = np.linspace(0.5, 4.5, 10)
    assert np.allclose(fifo_latencies(arr, svc), svc)


def test_fifo_all_at_once_is_the_running_sum():
    svc = np.array([3.0, 1.0, 2.0, 0.5])
    assert np.allclose(fifo_latencies(np.zeros(4), svc), np.cumsum(svc))


def test_fifo_matches_the_max_plus_form():
    # end_k = max over j <= k of (arrival_j + svc_j + ... + svc_k)
    rng = np.random.default_rng(3)
    for _ in range(20):
        n = int(rng.integers(1, 40))
        arr = np.sort(rng.exponential(1.0, n).cumsum() * rng.uniform(0.2, 3.0))
        svc = rng.exponential(1.0, n)
        ends = np.array([max(arr[j] + svc[j:k + 1].sum() for j in range(k + 1)) for k in range(n)])
        assert np.allclose(fifo_latencies(arr, svc), ends - arr)


def test_fifo_rejects_bad_input():
    with pytest.raises(ValueError):
        fifo_latencies([0.0, 1.0], [1.0])
    with pytest.raises(ValueError):
        fifo_latencies([1.0, 0.0], [1.0, 1.0])
    with pytest.raises(ValueError):
        fifo_latencies([0.0, 1.0], [1.0, -1.0])


def test_request_specs_scale_with_the_rate():
    a = request_specs(64, 2024, rate=1.0)
    b = request_specs(64, 2024, rate=0.25)
    for x, y in zip(a, b):
        assert y["arrival"] == pytest.approx(x["arrival"] / 0.25, rel=1e-12)
        assert (x["N"], x["ell"], x["kind"], x["seed"]) == (y["N"], y["ell"], y["kind"], y["seed"])


def _small_specs(monkeypatch):
    import h2pack_amd.serve as sv
    real = sv.request_specs
    monkeypatch.setattr(sv, "request_specs",
                        lambda n, seed, n_min, n_max, rate: real(n, seed, 2000, 3000, rate))


def test_request_builds_in_worker_processes(monkeypatch):
    """bench.py's host side of the serving legs: requests built in spawned
    worker processes, all at once or streamed with a window and a slice."""
    import os
    import bench_workloads as bw
    _small_specs(monkeypatch)
    env = dict(os.environ)
    try:
        ex, futs = bw.start_request_builds(2, 7, 1.0, 50, 1)
        reqs = bw.collect_requests(ex, futs)
        assert [r.rid for r in reqs] == [0, 1] and all(r.h2 is not None for r in reqs)
        assert {r.kind for r in reqs} == {"matmul", "cg"} and reqs[1].max_iter == 50
        reqs, ex, th = bw.start_stream(4, 7, 1.0, 50, 1, window=4, first=1, count=2)
        th.join()
        built = [r.future.result()[0] for r in reqs]
        ex.shutdown()
        assert [r.rid for r in reqs] == [1, 2] and [b.rid for b in built] == [1, 2]
        assert all(b.h2 is not None and 2000 <= b.h2.krnl_mat_size <= 3000 for b in built)
    finally:
        os.environ.clear()
        os.environ.update(env)
