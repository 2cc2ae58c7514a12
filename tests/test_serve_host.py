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
