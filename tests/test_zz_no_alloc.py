"""GPU: h2p_hip_matmul_dev_noalloc, the matmul whose panels live in the
library's static per-GPU scratch, against h2p_hip_matmul_dev on the same
handles.

Bars: bitwise the same products (the same kernels and tile variants run in
the same order) on two calls in a row, the second on a smaller matrix with
fewer columns, so it runs on the scratch the first call left behind; both
also within 1e-12 of the oracle's H2P_matmul restatement; a request whose
panels exceed the scratch is refused with H2P_HIP_EINVAL.
"""
import numpy as np
import pytest
import torch

import oracle as O
from h2pack_amd import abi
from h2pack_amd import construct as C
from h2pack_amd.h2mat import H2MatHIP, H2PHipError
from tests.problems import problem, rel

pytestmark = pytest.mark.gpu


def _both(m, X, layout):
    """(matmul_dev, matmul_dev_noalloc) of X on the device, as numpy arrays."""
    dev = torch.device("cuda:0")
    N, n_vec = X.shape
    if layout == abi.ROW_MAJOR:
        Xd = torch.from_numpy(np.ascontiguousarray(X)).to(dev)
        ld = n_vec
    else:
        Xd = torch.from_numpy(np.ascontiguousarray(X.T)).to(dev)   # column j at j * N
        ld = N
    Ya = torch.full_like(Xd, np.nan)
    Yb = torch.full_like(Xd, np.nan)
    s = torch.cuda.current_stream().cuda_stream
    m.matmul_dev(layout, n_vec, Xd.data_ptr(), ld, Ya.data_ptr(), ld, s)
    m.matmul_dev_noalloc(layout, n_vec, Xd.data_ptr(), ld, Yb.data_ptr(), ld, s)
    torch.cuda.synchronize()
    a, b = Ya.cpu().numpy(), Yb.cpu().numpy()
    if layout == abi.COL_MAJOR:
        a, b = a.T, b.T
    return a, b


@pytest.mark.parametrize("prec", ["fp64", "fp32mix"])
def test_noalloc_matmul_equals_matmul_dev(prec):
    h2_big, _ = problem("gaussian_leaf300")
    h2_small, _ = problem("gaussian")
    rng = np.random.default_rng(11)
    # first call: 40 columns, column-major, on the larger matrix
    m1 = H2MatHIP(h2_big, mode="jit", prec=prec)
    X1 = rng.standard_normal((h2_big.krnl_mat_size, 40))
    a1, b1 = _both(m1, X1, abi.COL_MAJOR)
    print(f"[noalloc {prec}] call 1: N={h2_big.krnl_mat_size} n_vec=40 col-major", flush=True)
    assert np.array_equal(a1, b1)
    # second call: a smaller matrix and 8 columns, row-major
    m2 = H2MatHIP(h2_small, mode="jit", prec=prec)
    X2 = rng.standard_normal((h2_small.krnl_mat_size, 8))
    a2, b2 = _both(m2, X2, abi.ROW_MAJOR)
    print(f"[noalloc {prec}] call 2: N={h2_small.krnl_mat_size} n_vec=8 row-major", flush=True)
    assert np.array_equal(a2, b2)
    if prec == "fp64":
        assert rel(b1, O.OracleH2(h2_big, BD_JIT=1).matmul(X1, "col", n_thread=8)) < 1e-12
        assert rel(b2, O.OracleH2(h2_small, BD_JIT=1).matmul(X2, "row", n_thread=8)) < 1e-12
    m1.close()
    m2.close()


def test_noalloc_matmul_aot_and_rpy():
    """Stored blocks (AOT) and the 3x3 kernel take the same route."""
    for name, mode in (("coulomb", "aot"), ("rpy", "jit")):
        h2, _ = problem(name)
        m = H2MatHIP(h2, mode=mode, host_blocks=False)
        X = np.random.default_rng(12).standard_normal((h2.krnl_mat_size, 20))
        a, b = _both(m, X, abi.COL_MAJOR)
        print(f"[noalloc {name} {mode}] N={h2.krnl_mat_size} n_vec=20", flush=True)
        assert np.array_equal(a, b)
        m.close()


def test_noalloc_matmul_refuses_panels_beyond_the_scratch():
    # N = 300k: (2 N + 2 S) x 64 doubles exceed the 256 MiB scratch, 16
    # columns fit
    coord = C.uniform_cube_points(300_000, 10.0, 5)
    h2 = C.build_h2(coord, C.KRNL_GAUSSIAN, [1.0], 1e-6, 300, BD_JIT=1, pp_kind="numerical",
                    host_blocks=False)
    m = H2MatHIP(h2, mode="jit")
    N = h2.krnl_mat_size
    dev = torch.device("cuda:0")
    Xd = torch.zeros((64, N), dtype=torch.float64, device=dev)
    Yd = torch.zeros_like(Xd)
    with pytest.raises(H2PHipError) as e:
        m.matmul_dev_noalloc(abi.COL_MAJOR, 64, Xd.data_ptr(), N, Yd.data_ptr(), N)
    assert e.value.code == abi.EINVAL
    X = np.random.default_rng(13).standard_normal((N, 16))
    a, b = _both(m, X, abi.COL_MAJOR)
    print(f"[noalloc] N={N} n_vec=16 within the scratch", flush=True)
    assert np.array_equal(a, b)
    m.close()
