"""GPU: co-scheduled calls over several resident matrices (h2p_hip_*_batch,
the batched CG).  Every call goes through the C ABI.

Bars: a batched call computes exactly what the single-matrix call computes
(bitwise, the same device code per task); a matrix's result does not depend
on the other matrices in the batch (bitwise); the batched CG follows the
reference's unpreconditioned PCG (examples/PCG/pcg.c:11-152) on the shifted
system of examples/SPDHSS-H2/pcg_tests.c:15-21, checked against the oracle's
restatement of it.
"""
import numpy as np
import pytest

import oracle as O
from h2pack_amd import abi
from h2pack_amd.h2mat import H2Batch, H2MatHIP, H2PHipError, footprint
from tests.problems import problem, rel

pytestmark = pytest.mark.gpu

# (problem, mode, precision) of the matrices mixed in one batch
MIX = [("gaussian", "jit", "fp64"), ("matern32", "jit", "fp32mix"), ("coulomb", "aot", "fp64"),
       ("rpy", "jit", "fp64"), ("coulomb_diag", "aot", "fp32mix"), ("rpy", "jit", "fp32mix")]


def _torch():
    import torch
    return torch


def _mats(specs):
    out = []
    for name, mode, prec in specs:
        h2, _ = problem(name)
        out.append(H2MatHIP(h2, mode=mode, prec=prec, host_blocks=False))
    return out


def _dev(a):
    torch = _torch()
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def test_matvec_batch_equals_single_calls():
    torch = _torch()
    mats = _mats(MIX)
    rng = np.random.default_rng(3)
    xs = [_dev(rng.random(m.N) - 0.5) for m in mats]
    ys_single = [torch.empty_like(x) for x in xs]
    for m, x, y in zip(mats, xs, ys_single):
        m.matvec_dev(x.data_ptr(), y.data_ptr())
    ys = [torch.empty_like(x) for x in xs]
    b = H2Batch()
    b.matvec(mats, [x.data_ptr() for x in xs], [y.data_ptr() for y in ys])
    torch.cuda.synchronize()
    for (name, mode, prec), a, c in zip(MIX, ys, ys_single):
        print(f"[{name} {mode} {prec}] batch vs single max |diff| = "
              f"{(a - c).abs().max().item():.3e}", flush=True)
        assert torch.equal(a, c)


def test_matvec_batch_shift():
    """y = A x + shift x, the PCG wrapper of pcg_tests.c:15-21."""
    torch = _torch()
    mats = _mats(MIX[:3])
    rng = np.random.default_rng(4)
    xs = [_dev(rng.random(m.N) - 0.5) for m in mats]
    shifts = [0.25, 1e-2, 3.0]
    ys = [torch.empty_like(x) for x in xs]
    H2Batch().matvec(mats, [x.data_ptr() for x in xs], [y.data_ptr() for y in ys], shifts)
    for m, x, y, s in zip(mats, xs, ys, shifts):
        y0 = torch.empty_like(x)
        m.matvec_dev(x.data_ptr(), y0.data_ptr())
        torch.cuda.synchronize()
        ref = (y0 + s * x).cpu().numpy()
        assert rel(y.cpu().numpy(), ref) < 1e-15


def test_result_is_the_same_alone_and_in_the_mix():
    torch = _torch()
    mats = _mats(MIX)
    rng = np.random.default_rng(5)
    xs = [_dev(rng.random(m.N) - 0.5) for m in mats]
    b = H2Batch()
    alone = []
    for m, x in zip(mats, xs):
        y = torch.empty_like(x)
        b.matvec([m], [x.data_ptr()], [y.data_ptr()], [0.5])
        alone.append(y)
    mixed = [torch.empty_like(x) for x in xs]
    order = [3, 5, 0, 4, 1, 2]
    b.matvec([mats[i] for i in order], [xs[i].data_ptr() for i in order],
             [mixed[i].data_ptr() for i in order], [0.5] * len(order))
    torch.cuda.synchronize()
    for a, c in zip(alone, mixed):
        assert torch.equal(a, c)


def test_matmul_batch_equals_single_calls():
    torch = _torch()
    specs = [("gaussian", "jit", "fp64"), ("rpy", "aot", "fp64"), ("matern32", "jit", "fp32mix")]
    mats = _mats(specs)
    nvs = [8, 20, 70]   # 70 crosses the 64-column chunk
    rng = np.random.default_rng(6)
    Xs = [_dev(rng.standard_normal((nv, m.N))) for m, nv in zip(mats, nvs)]   # column-major N x nv
    Ys = [torch.empty_like(X) for X in Xs]
    b = H2Batch()
    b.matmul(mats, abi.COL_MAJOR, nvs, [X.data_ptr() for X in Xs], [m.N for m in mats],
             [Y.data_ptr() for Y in Ys], [m.N for m in mats])
    for m, nv, X, Y in zip(mats, nvs, Xs, Ys):
        Y1 = torch.empty_like(X)
        m.matmul_dev(abi.COL_MAJOR, nv, X.data_ptr(), m.N, Y1.data_ptr(), m.N)
        torch.cuda.synchronize()
        assert torch.equal(Y, Y1)
    # and against the oracle for the FP64 matrix
    h2, _ = problem("gaussian")
    ref = O.OracleH2(h2, BD_JIT=1).matmul(Xs[0].cpu().numpy().T, "col", n_thread=8)
    assert rel(Ys[0].cpu().numpy().T, ref) < 1e-12


def _shifted_oracle(h2, shift):
    orc = O.OracleH2(h2, BD_JIT=1)
    return lambda v: orc.matvec(v, n_thread=8) + shift * v


@pytest.mark.parametrize("name,shift,max_iter", [("matern32", 1.0, 400), ("gaussian", 0.5, 400),
                                                  ("matern32", 0.3, 20)])
def test_cg_matches_oracle_pcg(name, shift, max_iter):
    """Converged solves (shift 1.0 / 0.5) and one stopped by max_iter.  The
    capped case is chosen where 20 iterations of CG amplify rounding little
    (a 1e-15 relative perturbation of the operator moves the residual by
    ~1e-15 and the iterate by ~3e-11); on the nugget 1e-2 system after 60
    iterations the same perturbation moves the residual by ~1 %, so no
    tight comparison with the oracle is possible there."""
    torch = _torch()
    h2, _ = problem(name)
    m = H2MatHIP(h2, mode="jit", prec="fp64")
    rng = np.random.default_rng(7)
    bvec = rng.standard_normal(m.N)
    tol = 1e-6
    x_ref, flag, relres_ref, it_ref = O.pcg(_shifted_oracle(h2, shift), bvec, tol, max_iter)
    bd = _dev(bvec)
    xd = torch.zeros_like(bd)
    bt = H2Batch()
    bt.cg_start([m], [bd.data_ptr()], [xd.data_ptr()], [shift], [tol], [max_iter])
    done = [False]
    while not done[0]:
        bt.cg_step([m], 8)
        done, its, rr = bt.cg_poll([m])
    x = xd.cpu().numpy()
    true_res = rel(_shifted_oracle(h2, shift)(x), bvec)
    print(f"[{name}] CG iters gpu {its[0]} oracle {it_ref}; relres gpu {rr[0]:.3e} "
          f"oracle {relres_ref:.3e}; true residual {true_res:.3e}; "
          f"x rel diff {rel(x, x_ref):.3e}", flush=True)
    if max_iter == 20:
        # stopped by the cap: same count, residuals and iterates as the oracle's
        assert flag == 1 and it_ref == its[0] == max_iter
        assert abs(rr[0] - relres_ref) <= 1e-8 * relres_ref
        assert abs(true_res - relres_ref) <= 1e-6 * relres_ref
        assert rel(x, x_ref) < 1e-8
        return
    assert flag == 0 and it_ref < max_iter
    assert abs(its[0] - it_ref) <= 2
    assert rr[0] <= tol and true_res <= 2 * tol
    assert rel(x, x_ref) < 1e-6


def test_cg_same_alone_and_in_the_mix():
    torch = _torch()
    specs = [("matern32", "jit", "fp32mix"), ("gaussian", "jit", "fp32mix"), ("coulomb_diag", "aot", "fp64")]
    mats = _mats(specs)
    shifts = [1e-2, 0.1, 1.0]
    rng = np.random.default_rng(9)
    bs = [_dev(rng.standard_normal(m.N)) for m in mats]
    bt = H2Batch()

    def solve(idx, step):
        xs = {i: torch.zeros_like(bs[i]) for i in idx}
        ms = [mats[i] for i in idx]
        bt.cg_start(ms, [bs[i].data_ptr() for i in idx], [xs[i].data_ptr() for i in idx],
                    [shifts[i] for i in idx], [1e-6] * len(idx), [300] * len(idx))
        while True:
            bt.cg_step(ms, step)
            done, its, rr = bt.cg_poll(ms)
            if all(done):
                return {i: (xs[i].clone(), its[k], rr[k]) for k, i in enumerate(idx)}

    mixed = solve([0, 1, 2], 5)
    for i in range(3):
        x, it, r = solve([i], 3)[i]
        print(f"[{specs[i][0]}] alone {it} iters relres {r:.3e}; mixed {mixed[i][1]} iters", flush=True)
        assert it == mixed[i][1]
        assert torch.equal(x, mixed[i][0])


def test_cg_residual_replacement():
    """An FP32-mixed solve with a replacement ratio pauses (done = 2) instead
    of finishing on its recurrence residual; h2p_hip_cg_replace forms the
    residual with the FP64 handle, and the solve ends (done = 1) only once
    that residual is below tol.  The reported residual then matches the
    oracle's b - (A + shift) x, and the whole run is the same whatever the
    polling chunk."""
    torch = _torch()
    h2, _ = problem("matern32")
    shift, tol, max_iter = 1e-2, 1e-8, 3000
    m32 = H2MatHIP(h2, mode="jit", prec="fp32mix")
    m64 = H2MatHIP(h2, mode="jit", prec="fp64")
    bvec = np.random.default_rng(17).standard_normal(h2.krnl_mat_size)
    bd = _dev(bvec)
    bt = H2Batch()

    def solve(chunk):
        xd = torch.zeros_like(bd)
        m32.set_cg_replace(0.1)
        bt.cg_start([m32], [bd.data_ptr()], [xd.data_ptr()], [shift], [tol], [max_iter])
        replacements = 0
        while True:
            bt.cg_step([m32], chunk)
            done, its, rr = bt.cg_poll([m32])
            if done[0] == 1:
                break
            if done[0] == 2:
                bt.cg_replace([m32], [m64])
                replacements += 1
                done, its, rr = bt.cg_poll([m32])
                assert done[0] in (0, 1)
                if done[0] == 1:
                    break
        return xd.cpu().numpy(), its[0], rr[0], replacements

    x, its, relres, nrep = solve(8)
    true_res = rel(_shifted_oracle(h2, shift)(x), bvec)
    print(f"[matern32 fp32mix, replacement] {its} iterations, {nrep} replacements, reported {relres:.3e}, "
          f"oracle {true_res:.3e}", flush=True)
    assert nrep >= 2 and its < max_iter
    assert relres <= tol and abs(true_res - relres) <= 0.05 * tol
    x2, its2, relres2, nrep2 = solve(3)
    assert (its2, nrep2) == (its, nrep) and np.array_equal(x2, x) and relres2 == relres
    m32.set_cg_replace(0.0)   # the next solve never pauses again
    with pytest.raises(H2PHipError) as e:
        m32.set_cg_replace(1.0)
    assert e.value.code == abi.EINVAL
    m32.close()
    m64.close()


def test_batch_rejects_bad_calls():
    torch = _torch()
    mats = _mats(MIX[:1])
    x = torch.zeros(mats[0].N, dtype=torch.float64, device="cuda:0")
    bt = H2Batch()
    with pytest.raises(H2PHipError) as e:
        bt.matvec([mats[0], mats[0]], [x.data_ptr()] * 2, [x.data_ptr()] * 2)
    assert e.value.code == abi.EINVAL
    with pytest.raises(H2PHipError) as e:
        bt.cg_step(mats, 1)
    assert e.value.code == abi.EINVAL
    with pytest.raises(H2PHipError) as e:
        bt.cg_replace(mats, mats)       # no solve started on it
    assert e.value.code == abi.EINVAL
    other = _mats([("gaussian_leaf300", "jit", "fp64")])[0]
    bt.cg_start(mats, [x.data_ptr()], [x.data_ptr()], [1.0], [1e-6], [1])
    with pytest.raises(H2PHipError) as e:
        bt.cg_replace(mats, [other])    # another order
    assert e.value.code == abi.EINVAL
    other.close()


def test_footprint_matches_device_bytes():
    torch = _torch()
    h2, _ = problem("matern32")
    for mode, prec in (("jit", "fp32mix"), ("aot", "fp64"), ("aot", "fp32mix")):
        m = H2MatHIP(h2, mode=mode, prec=prec, host_blocks=False)
        assert footprint(h2, mode, prec) == m.info()["device_bytes"]
        if mode == "jit":
            X = torch.zeros((8, m.N), dtype=torch.float64, device="cuda:0")
            Y = torch.empty_like(X)
            m.matmul_dev(abi.COL_MAJOR, 8, X.data_ptr(), m.N, Y.data_ptr(), m.N)
            torch.cuda.synchronize()
            assert footprint(h2, mode, prec, n_vec=8) == m.info()["device_bytes"]
            bt = H2Batch()
            b = torch.ones(m.N, dtype=torch.float64, device="cuda:0")
            xx = torch.zeros_like(b)
            bt.cg_start([m], [b.data_ptr()], [xx.data_ptr()], [1e-2], [1e-6], [5])
            torch.cuda.synchronize()
            assert footprint(h2, mode, prec, n_vec=8, with_cg=True) == m.info()["device_bytes"]
        m.close()
