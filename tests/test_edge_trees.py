"""Degenerate trees: a single point, one leaf holding every point, leaves
that are all neighbours of each other (no admissible pair, so no B block and
no transfer matrix), two one-point leaves, and a one-leaf RPY matrix.

CPU: the oracle equals the direct product on every row (there is no
low-rank block, so only rounding separates them), and the library's
host-side upload validation (``h2p_hip_footprint``, a dry run that checks
every task against its buffers without touching the device) accepts each
tree in every mode.  GPU: matvec, matmul and a batched matvec through the C
ABI equal the oracle (FP64 1e-12, FP32-mixed 1e-6); on these trees the JIT
handles' B-block launch is empty and only the D-block launch runs.
"""
import functools

import numpy as np
import pytest

import oracle as O
from h2pack_amd import abi
from h2pack_amd import construct as C
from h2pack_amd.h2mat import H2Batch, H2MatHIP, footprint
from tests.problems import coords_for, problem

# name: (kernel id, params, N, density, leaf, proxy points)
CASES = {
    "one_point": (C.KRNL_GAUSSIAN, [1.0], 1, 10.0, 100, "numerical"),
    "one_leaf": (C.KRNL_GAUSSIAN, [1.0], 50, 10.0, 100, "numerical"),
    "no_admissible": (C.KRNL_COULOMB, [], 200, 10.0, 40, "surface"),
    "two_points": (C.KRNL_COULOMB, [], 2, 10.0, 1, "surface"),
    "one_leaf_rpy": (C.KRNL_RPY, [1.0], 40, 0.03, 100, "surface"),
}


@functools.lru_cache(maxsize=None)
def edge_problem(name):
    kid, param, n, dens, leaf, ppk = CASES[name]
    coord = coords_for(kid, n, dens, 77)
    h2 = C.build_h2(coord, kid, param, 1e-8, leaf, BD_JIT=0, pp_kind=ppk, host_blocks=True)
    return h2, coord


def _x(h2, seed=5):
    return np.random.default_rng(seed).random(h2.krnl_mat_size) - 0.5


def _close(y, ref, tol):
    nr = np.linalg.norm(ref)
    if nr == 0.0:
        return np.array_equal(y, np.zeros_like(y))
    return np.linalg.norm(y - ref) / nr < tol


def _modes(name):
    out = [("jit", "fp64"), ("aot", "fp64"), ("aot", "fp32mix")]
    if CASES[name][0] != C.KRNL_RPY:       # RPY JIT has no FP32-mixed path
        out.append(("jit", "fp32mix"))
    return out


@pytest.mark.parametrize("name", list(CASES))
def test_oracle_equals_direct_product(name):
    h2, coord = edge_problem(name)
    kid, param = CASES[name][0], CASES[name][1]
    assert h2.n_r_adm_pair == 0
    x = _x(h2)
    y = O.OracleH2(h2, BD_JIT=1).matvec(x, n_thread=1)
    rows = np.arange(coord.shape[1])
    ref = O.direct_nbody(kid, param, coord, x, rows)
    assert _close(y, ref, 1e-13)


@pytest.mark.parametrize("name", list(CASES))
def test_host_validation_accepts(name):
    h2, _ = edge_problem(name)
    for mode, prec in _modes(name):
        assert footprint(h2, mode, prec) > 0
        assert footprint(h2, mode, prec, n_vec=3, with_cg=True) > footprint(h2, mode, prec)


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(CASES))
def test_matvec_and_matmul_match_oracle(name):
    h2, _ = edge_problem(name)
    orc = O.OracleH2(h2, BD_JIT=1)
    x = _x(h2)
    X = np.random.default_rng(6).standard_normal((h2.krnl_mat_size, 3))
    y_ref = orc.matvec(x, n_thread=1)
    Y_ref = orc.matmul(X, "col", n_thread=1)
    for mode, prec in _modes(name):
        m = H2MatHIP(h2, mode=mode, prec=prec, host_blocks=(mode == "aot"))
        tol = 1e-12 if prec == "fp64" else 1e-6
        y = m.matvec(x)
        print(f"[{name} {mode} {prec}] matvec", flush=True)
        assert _close(y, y_ref, tol), (mode, prec)
        if m.info()["prec"] == abi.FP64 or CASES[name][0] != C.KRNL_RPY:   # FP32-mixed matmul: scalar kernels
            Y = m.matmul(X, "col")
            assert _close(Y, Y_ref, tol), (mode, prec)
        m.close()


@pytest.mark.gpu
def test_edge_trees_in_a_batch():
    """Edge matrices co-scheduled with a regular one: each result equals the
    matrix's own call, bit for bit."""
    import torch
    hs = [edge_problem("one_leaf")[0], edge_problem("no_admissible")[0], problem("gaussian")[0],
          edge_problem("one_point")[0]]
    mats = [H2MatHIP(h2, mode="jit") for h2 in hs]
    rng = np.random.default_rng(8)
    xs = [torch.from_numpy(rng.random(m.N) - 0.5).to("cuda:0") for m in mats]
    ys = [torch.empty_like(x) for x in xs]
    H2Batch().matvec(mats, [x.data_ptr() for x in xs], [y.data_ptr() for y in ys])
    for m, x, y in zip(mats, xs, ys):
        y1 = torch.empty_like(x)
        m.matvec_dev(x.data_ptr(), y1.data_ptr())
        torch.cuda.synchronize()
        assert torch.equal(y, y1)
