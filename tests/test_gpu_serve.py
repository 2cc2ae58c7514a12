"""GPU: the continuous-batching request loop (h2pack_amd/serve.py).

Bars: every request's result is bitwise the same when it is served alone and
when it shares launches with the others; the FP32-mixed results agree with
the plain FP64 serving of the same requests (CG: true residual of the FP64
operator within 10x the requested tolerance; products: 1e-6 relative);
admission never holds more device bytes than the budget.
"""
import numpy as np
import pytest

import oracle as O
from h2pack_amd.serve import Server, make_requests, request_bytes
from tests.problems import rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def reqs():
    return make_requests(6, seed=21, n_min=3000, n_max=9000, rate=50.0, leaf=150,
                         max_iter=400, nugget=0.5)


def _clone(rs):
    import copy
    out = []
    for r in rs:
        c = copy.copy(r)
        c.result = None
        c.finished = c.admitted = float("nan")
        out.append(c)
    return out


def test_serving_same_alone_and_in_the_mix(reqs):
    mixed = _clone(reqs)
    srv = Server(hbm_budget=64 << 30, max_active=8, prec="fp32mix", cg_chunk=4)
    st = srv.run(mixed, realtime=False)
    print(f"mixed: {st.as_dict()}", flush=True)
    assert st.n_done == len(reqs) and st.max_active > 1
    for r in reqs:
        alone = _clone([r])
        Server(hbm_budget=64 << 30, max_active=1, prec="fp32mix", cg_chunk=7).run(alone, realtime=False)
        m = [x for x in mixed if x.rid == r.rid][0]
        print(f"request {r.rid} {r.kind} N={r.N}: iters alone {alone[0].iters} mixed {m.iters}", flush=True)
        assert np.array_equal(alone[0].result, m.result)
        assert alone[0].iters == m.iters


def test_serving_cap_on_stepped_solves_keeps_the_results(reqs):
    """Advancing only the longest-admitted solves each tick (cg_active)
    changes when solves run, not what they compute."""
    full = _clone(reqs)
    Server(hbm_budget=64 << 30, max_active=8, prec="fp32mix", cg_chunk=4).run(full, realtime=False)
    capped = _clone(reqs)
    st = Server(hbm_budget=64 << 30, max_active=8, prec="fp32mix", cg_chunk=4, cg_active=1).run(
        capped, realtime=False)
    assert st.n_done == len(reqs)
    for a, b in zip(full, capped):
        assert np.array_equal(a.result, b.result) and a.iters == b.iters


def test_serving_fp32mix_against_fp64(reqs):
    mix = _clone(reqs)
    Server(hbm_budget=64 << 30, max_active=8, prec="fp32mix").run(mix, realtime=False)
    ref = _clone(reqs)
    Server(hbm_budget=64 << 30, max_active=1, prec="fp64").run(ref, realtime=False)
    for a, b, r in zip(mix, ref, reqs):
        if r.kind == "matmul":
            assert rel(a.result, b.result) < 1e-6
        else:
            orc = O.OracleH2(r.h2, BD_JIT=1)
            res = rel(orc.matvec(a.result, n_thread=8) + r.shift * a.result, r.rhs)
            print(f"request {r.rid}: CG fp32mix iters {a.iters} (fp64 {b.iters}), true residual "
                  f"{res:.3e}", flush=True)
            assert a.relres <= r.tol and res <= 10 * r.tol
            res64 = rel(orc.matvec(b.result, n_thread=8) + r.shift * b.result, r.rhs)
            assert b.relres <= r.tol and res64 <= 2 * r.tol
            assert abs(a.iters - b.iters) <= 0.5 * b.iters


def test_admission_respects_the_budget(reqs):
    rs = _clone(reqs)
    sizes = sorted(request_bytes(r, "jit", "fp32mix", refine=True) for r in rs)
    budget = sizes[-1] + sizes[0]     # room for the largest plus the smallest
    st = Server(hbm_budget=budget, max_active=8, prec="fp32mix").run(rs, realtime=False)
    assert st.n_done == len(rs) and st.peak_bytes <= budget
    # a request larger than the budget is rejected, the others still run
    rs = _clone(reqs)
    st = Server(hbm_budget=sizes[-1] - 1, max_active=8, prec="fp32mix").run(rs, realtime=False)
    assert st.n_rejected >= 1 and st.n_done == len(rs) - st.n_rejected
