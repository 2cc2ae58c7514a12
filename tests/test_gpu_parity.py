"""GPU parity: libh2p_hip.so against the oracle on the same H2 representation.

Bar: FP64 results agree with the oracle's restatement of
H2P_matvec to 1e-12 relative 2-norm; FP32-mixed results to the build
tolerance.  Every call goes through the C ABI (include/h2p_hip.h).
"""
import numpy as np
import pytest

import oracle as O
from h2pack_amd import abi
from h2pack_amd.h2mat import H2MatHIP, H2PHipError, direct_nbody
from tests.problems import problem, rel

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
KERNELS = ["coulomb", "coulomb_diag", "gaussian", "matern32", "rpy", "gaussian_leaf300", "matern32_leaf300",
           "rpy_leaf300"]


def _x(h2, seed=11):
    return np.random.default_rng(seed).random(h2.krnl_mat_size) - 0.5


@pytest.mark.parametrize("name", KERNELS)
@pytest.mark.parametrize("mode", ["jit", "aot", "aot_devbuild"])
def test_matvec_fp64_matches_oracle(name, mode):
    h2, _ = problem(name)
    x = _x(h2)
    y_ref = O.OracleH2(h2, BD_JIT=1).matvec(x, n_thread=8)
    m = H2MatHIP(h2, mode="jit" if mode == "jit" else "aot", host_blocks=(mode == "aot"))
    y = m.matvec(x)
    err = rel(y, y_ref)
    print(f"[{name} {mode}] rel err vs oracle = {err:.3e}", flush=True)
    assert err < FP64_TOL


@pytest.mark.parametrize("name", ["coulomb", "rpy"])
def test_matvec_is_deterministic(name):
    h2, _ = problem(name)
    x = _x(h2)
    for mode in ("jit", "aot"):
        m = H2MatHIP(h2, mode=mode, host_blocks=False)
        a, b = m.matvec(x), m.matvec(x)
        assert np.array_equal(a, b)


@pytest.mark.parametrize("name", ["coulomb_diag", "gaussian", "rpy"])
@pytest.mark.parametrize("mode", ["jit", "aot"])
@pytest.mark.parametrize("layout", ["col", "row"])
def test_matmul_matches_oracle_and_matvec(name, mode, layout):
    """MFMA pull-form matmul: against the oracle's H2P_matmul restatement and
    against one matvec per column (extra/test_H2_matmul.h:18-116)."""
    h2, _ = problem(name)
    rng = np.random.default_rng(5)
    X = rng.standard_normal((h2.krnl_mat_size, 7))
    m = H2MatHIP(h2, mode=mode, host_blocks=False)
    Y = m.matmul(X, layout)
    Y_ref = O.OracleH2(h2, BD_JIT=1).matmul(X, layout, n_thread=8)
    err = rel(Y, Y_ref)
    print(f"[{name} {mode} {layout}] matmul rel err vs oracle = {err:.3e}", flush=True)
    assert err < FP64_TOL
    cols = np.stack([m.matvec(X[:, i]) for i in range(X.shape[1])], axis=1)
    assert rel(Y, cols) < FP64_TOL


def test_matmul_many_columns_and_leading_dimensions():
    """70 columns cross the 64-column chunk; ld larger than the panel."""
    import ctypes as ct
    h2, _ = problem("gaussian")
    N = h2.krnl_mat_size
    rng = np.random.default_rng(8)
    X = rng.standard_normal((N, 70))
    m = H2MatHIP(h2, mode="jit")
    Y_ref = O.OracleH2(h2, BD_JIT=1).matmul(X, "col", n_thread=8)
    assert rel(m.matmul(X, "col"), Y_ref) < FP64_TOL
    # column-major with ldx = N + 5, through the C ABI directly
    ld = N + 5
    Xb = np.zeros((70, ld)); Xb[:, :N] = X.T
    Yb = np.full((70, ld), 7.0)
    rc = m.lib.h2p_hip_matmul(m.handle, abi.COL_MAJOR, 70, Xb.ctypes.data, ld, Yb.ctypes.data, ld)
    assert rc == 0
    assert rel(Yb[:, :N].T, Y_ref) < FP64_TOL
    assert np.all(Yb[:, N:] == 7.0)          # padding untouched


@pytest.mark.parametrize("name", ["coulomb", "gaussian", "rpy"])
def test_direct_nbody_matches_oracle(name):
    h2, coord = problem(name)
    rng = np.random.default_rng(3)
    x = rng.random(h2.krnl_mat_size) - 0.5
    rows = np.sort(rng.choice(h2.n_point, 257, replace=False))
    y = direct_nbody(h2.krnl_id, list(h2.krnl_param), coord, x, rows)
    y_ref = O.direct_nbody(h2.krnl_id, list(h2.krnl_param), coord, x, rows, n_thread=8)
    assert rel(y, y_ref) < 1e-13


def test_matvec_dev_on_torch_tensors():
    import torch
    h2, _ = problem("gaussian")
    x = _x(h2)
    m = H2MatHIP(h2, mode="jit")
    xd = torch.from_numpy(x).cuda()
    yd = torch.empty_like(xd)
    stream = torch.cuda.current_stream().cuda_stream
    m.matvec_dev(xd.data_ptr(), yd.data_ptr(), stream)
    torch.cuda.synchronize()
    assert rel(yd.cpu().numpy(), m.matvec(x)) < 1e-15


def test_bad_calls_raise():
    h2, _ = problem("tiny_coulomb")
    m = H2MatHIP(h2, mode="jit")
    with pytest.raises(ValueError):
        m.matvec(np.zeros(3))
    with pytest.raises(H2PHipError) as ei:
        m.matmul_dev(999, 2, 0, 1, 0, 1)
    assert ei.value.code == abi.EINVAL


@pytest.mark.parametrize("name", ["coulomb", "coulomb_diag", "gaussian", "matern32"])
@pytest.mark.parametrize("mode", ["jit", "aot", "aot_devbuild"])
def test_matvec_fp32mix_within_build_tolerance(name, mode):
    """FP32 kernel values (or FP32-stored blocks) and block products, FP64 U
    sweeps and outputs: error against the FP64 oracle stays within the H2
    build tolerance."""
    h2, _ = problem(name)
    x = _x(h2)
    y_ref = O.OracleH2(h2, BD_JIT=1).matvec(x, n_thread=8)
    m = H2MatHIP(h2, mode="jit" if mode == "jit" else "aot", prec="fp32mix",
                 host_blocks=(mode == "aot"))
    y = m.matvec(x)
    err = rel(y, y_ref)
    print(f"[{name} {mode} fp32mix] rel err vs FP64 oracle = {err:.3e}", flush=True)
    assert err < h2.QR_stop_tol
    assert np.array_equal(y, m.matvec(x))   # deterministic


@pytest.mark.parametrize("mode", ["jit", "aot"])
def test_rpy_fp32mix_within_build_tolerance(mode):
    """RPY on the FP32-mixed path: FP32-stored blocks (AOT) or the tensor
    evaluated in FP32 from block-local coordinates (JIT), FP32 products
    folded into FP64.  Bar: the build tolerance where FP32 rounding allows
    it (1e-7 here for the stored blocks), the FP32-mixed 1e-6 for the
    evaluated tensor; deterministic, and not the FP64 product in disguise.
    Matmul on these handles has no FP32 tiles and is refused (ENOSYS)."""
    h2, _ = problem("rpy")
    x = _x(h2)
    y_ref = O.OracleH2(h2, BD_JIT=1).matvec(x, n_thread=8)
    m = H2MatHIP(h2, mode=mode, prec="fp32mix", host_blocks=False)
    y = m.matvec(x)
    err = rel(y, y_ref)
    print(f"[rpy {mode} fp32mix] rel err vs FP64 oracle = {err:.3e}", flush=True)
    assert err < (h2.QR_stop_tol if mode == "aot" else 1e-6)
    assert np.array_equal(y, m.matvec(x))
    y64 = H2MatHIP(h2, mode=mode, host_blocks=False).matvec(x)
    assert not np.array_equal(y, y64)
    with pytest.raises(H2PHipError) as ei:
        m.matmul(np.ones((h2.krnl_mat_size, 2)), "col")
    assert ei.value.code == abi.ENOSYS
    m.close()


@pytest.mark.parametrize("name", ["gaussian", "matern32", "coulomb"])
@pytest.mark.parametrize("mode", ["jit", "aot"])
def test_matmul_fp32mix_within_build_tolerance(name, mode):
    """FP32-mixed H2P_matmul (FP32 kernel values or FP32-stored blocks in
    the B/D products on FP32 MFMA, folded into FP64; FP64 U phases) against
    the FP64 oracle, both layouts, 8 and 32 columns; it is not the FP64
    product in disguise (the results differ from it bitwise)."""
    h2, _ = problem(name)
    orc = O.OracleH2(h2, BD_JIT=1)
    m32 = H2MatHIP(h2, mode=mode, prec="fp32mix", host_blocks=False)
    m64 = H2MatHIP(h2, mode=mode, host_blocks=False)
    rng = np.random.default_rng(21)
    for nv in (8, 32):
        X = rng.standard_normal((h2.krnl_mat_size, nv))
        Y_ref = orc.matmul(X, "col", n_thread=8)
        Yc = m32.matmul(X, "col")
        Yr = m32.matmul(X, "row")
        err_c, err_r = rel(Yc, Y_ref), rel(Yr, Y_ref)
        print(f"[{name} {mode} fp32mix matmul n_vec={nv}] rel err vs FP64 oracle col {err_c:.3e} "
              f"row {err_r:.3e}", flush=True)
        assert err_c < h2.QR_stop_tol and err_r < h2.QR_stop_tol
        assert np.array_equal(Yc, Yr)
        Y64 = m64.matmul(X, "col")
        assert not np.array_equal(Yc, Y64)
        assert rel(Y64, Y_ref) < FP64_TOL
    m32.close()
    m64.close()


def test_fp32mix_rpy_jit_matches_aot():
    """The evaluated FP32 tensor (JIT) and the FP32-stored FP64 tensor (AOT)
    are two roundings of the same blocks: their products agree to FP32
    accuracy on a tree with every pair kind (overlapping particles
    included: volume fraction 0.03 at radius 1 puts some pairs inside 2a)."""
    h2, _ = problem("tiny_rpy")
    x = _x(h2)
    yj = H2MatHIP(h2, mode="jit", prec="fp32mix").matvec(x)
    ya = H2MatHIP(h2, mode="aot", prec="fp32mix", host_blocks=False).matvec(x)
    assert rel(yj, ya) < 1e-6


@pytest.mark.parametrize("mode", ["jit", "aot"])
def test_graph_replays_match_eager_path(mode):
    """The matvec interior is a captured hipGraph replayed on every call; the
    profiled path launches the same kernels eagerly.  Several replays with
    different inputs and unrelated device work in between must each equal
    the eager result bit for bit."""
    import torch
    h2, _ = problem("gaussian")
    m = H2MatHIP(h2, mode=mode, host_blocks=False)
    rng = np.random.default_rng(21)
    junk = torch.randn(1 << 20, device="cuda:0", dtype=torch.float64)
    for it in range(4):
        x = rng.random(h2.krnl_mat_size) - 0.5
        y_graph = m.matvec(x)
        junk.mul_(1.0001).add_(0.5)          # eager work between replays
        torch.cuda.synchronize()
        m.set_profiling(True)                # profiled calls do not use the graph
        y_eager = m.matvec(x)
        m.info()
        m.set_profiling(False)
        assert np.array_equal(y_graph, y_eager), it


@pytest.mark.parametrize("name,mode", [("gaussian", "jit"), ("coulomb", "aot"), ("rpy", "jit")])
def test_whole_graph_follows_caller_pointers(name, mode):
    """The whole matvec (gather of x, interior, leaf outputs into y) is one
    graph; calls with other x / y device addresses patch its two boundary
    kernel nodes.  Replays alternating between buffers, reusing one, and
    writing into a buffer that was an input before, each equal the eager
    (profiled) result bit for bit, and leave the other buffers untouched."""
    import torch
    h2, _ = problem(name)
    m = H2MatHIP(h2, mode=mode, host_blocks=False)
    N = h2.krnl_mat_size
    rng = np.random.default_rng(31)
    xs = [torch.from_numpy(rng.random(N) - 0.5).to("cuda:0") for _ in range(3)]
    ys = [torch.full((N,), np.nan, dtype=torch.float64, device="cuda:0") for _ in range(3)]
    eager = []
    m.set_profiling(True)                       # profiled calls launch eagerly
    for x in xs:
        y = torch.empty_like(x)
        m.matvec_dev(x.data_ptr(), y.data_ptr())
        torch.cuda.synchronize()
        eager.append(y.cpu().numpy())
    m.info()
    m.set_profiling(False)
    order = [(0, 0), (1, 1), (2, 2), (0, 1), (2, 0), (1, 2), (0, 0)]
    for i, j in order:
        before = [y.clone() for y in ys]
        m.matvec_dev(xs[i].data_ptr(), ys[j].data_ptr())
        torch.cuda.synchronize()
        assert np.array_equal(ys[j].cpu().numpy(), eager[i]), (i, j)
        for k in range(3):
            if k != j:
                assert torch.equal(ys[k], before[k]) or (torch.isnan(ys[k]).all() and torch.isnan(before[k]).all())
    # y into an earlier input buffer: the graph's x and y move independently
    x_copy = xs[1].clone()
    m.matvec_dev(x_copy.data_ptr(), xs[0].data_ptr())
    torch.cuda.synchronize()
    assert np.array_equal(xs[0].cpu().numpy(), eager[1])
    m.close()


@pytest.mark.parametrize("d_late", [None, "0", "1"])
@pytest.mark.parametrize("name,prec", [("gaussian", "fp64"), ("coulomb", "fp64"), ("matern32", "fp32mix")])
def test_split_block_launches_agree(name, prec, d_late, monkeypatch):
    """JIT handles launch the B blocks and the D blocks separately, the D
    blocks on a second stream concurrent with the sweeps (a fork/join inside
    the captured interior, or eager with graphs off): part of the D blocks
    from the start, the rest (H2P_HIP_D_LATE of their pairs: the default,
    none, all) after the B blocks, and the leaf outputs in two launches.
    Both must equal the profiled path, which runs every launch in order on
    one stream, bit for bit, and the profiled D launch must carry the
    near-field work."""
    import torch
    h2, _ = problem(name)
    if d_late is not None:
        monkeypatch.setenv("H2P_HIP_D_LATE", d_late)
    g = H2MatHIP(h2, mode="jit", prec=prec)
    monkeypatch.setenv("H2P_HIP_NO_GRAPH", "1")
    e = H2MatHIP(h2, mode="jit", prec=prec)
    monkeypatch.delenv("H2P_HIP_NO_GRAPH")
    rng = np.random.default_rng(23)
    junk = torch.randn(1 << 20, device="cuda:0", dtype=torch.float64)
    for it in range(3):
        x = rng.random(h2.krnl_mat_size) - 0.5
        y_graph = g.matvec(x)
        junk.mul_(1.0001).add_(0.5)
        y_eager = e.matvec(x)
        torch.cuda.synchronize()
        g.set_profiling(True)
        y_serial = g.matvec(x)
        info = g.info()
        g.set_profiling(False)
        assert np.array_equal(y_graph, y_serial), it
        assert np.array_equal(y_eager, y_serial), it
        assert info["t_blocks_d"] > 0.0 and info["t_blocks_b"] > 0.0
        assert abs(info["t_blocks_b"] + info["t_blocks_d"] - info["t_blocks"]) < 1e-6 * max(1.0, info["t_blocks"])


@pytest.mark.parametrize("name", ["gaussian", "rpy"])
def test_matmul_graph_replays_match_eager_handle(name, monkeypatch):
    """Each matmul chunk's product phases are a captured hipGraph (one per
    panel width), re-captured when a tile variant changes; every replay must
    equal a handle uploaded with graphs switched off, bit for bit, with
    unrelated device work between replays."""
    import torch
    h2, _ = problem(name)
    g = H2MatHIP(h2, mode="jit")
    monkeypatch.setenv("H2P_HIP_NO_GRAPH", "1")
    e = H2MatHIP(h2, mode="jit")
    monkeypatch.delenv("H2P_HIP_NO_GRAPH")
    rng = np.random.default_rng(22)
    junk = torch.randn(1 << 20, device="cuda:0", dtype=torch.float64)
    for it, n_vec in enumerate([8, 8, 32, 8, 70]):
        if it == 3:
            g.mm_variant(-1, 8, 2)          # new variant: the width-16 graph is re-captured
        X = rng.standard_normal((h2.krnl_mat_size, n_vec))
        Yg = g.matmul(X, "col")
        junk.mul_(1.0001).add_(0.5)
        torch.cuda.synchronize()
        Ye = e.matmul(X, "col")
        assert np.array_equal(Yg, Ye), it


@pytest.mark.parametrize("name", ["hss_coulomb", "hss_gaussian"])
@pytest.mark.parametrize("mode", ["jit", "aot"])
def test_hss_mode_matches_oracle(name, mode):
    """HSS representations (weak admissibility, hybrid bases; the HSS pair
    lists in the ordinary fields, as H2P_matvec reads them in HSS mode) run
    through the same engine: FP64 matvec and matmul against the oracle to
    1e-12, FP32-mixed within the build tolerance."""
    h2, _ = problem(name)
    x = _x(h2)
    orc = O.OracleH2(h2, BD_JIT=1)
    y_ref = orc.matvec(x, n_thread=8)
    m = H2MatHIP(h2, mode=mode, host_blocks=False)
    err = rel(m.matvec(x), y_ref)
    X = np.random.default_rng(9).standard_normal((h2.krnl_mat_size, 5))
    err_mm = rel(m.matmul(X, "col"), orc.matmul(X, "col", n_thread=8))
    m32 = H2MatHIP(h2, mode=mode, prec="fp32mix", host_blocks=False)
    err32 = rel(m32.matvec(x), y_ref)
    print(f"[{name} {mode}] HSS rel err vs oracle: matvec {err:.3e}, matmul {err_mm:.3e}, "
          f"fp32mix {err32:.3e}", flush=True)
    assert err < FP64_TOL and err_mm < FP64_TOL
    assert err32 < 1e-6


def test_periodic_matvec_and_matmul_match_oracle():
    """Periodic RPY (H2P_matvec_periodic): one-sided shifted pairs and the
    root's periodic block on the device, against the oracle's restatement
    to 1e-12; matmul columns against the matvec."""
    from tests.problems import periodic_problem
    h2, _, _ = periodic_problem()
    x = np.random.default_rng(12).standard_normal(h2.krnl_mat_size)
    y_ref = O.periodic_matvec(h2, x)
    m = H2MatHIP(h2, mode="jit")
    err = rel(m.matvec(x), y_ref)
    X = np.random.default_rng(13).standard_normal((h2.krnl_mat_size, 6))
    Y = m.matmul(X, "col")
    err_mm = max(rel(Y[:, j], O.periodic_matvec(h2, X[:, j])) for j in range(6))
    print(f"[periodic rpy] rel err vs oracle: matvec {err:.3e}, matmul {err_mm:.3e}", flush=True)
    assert err < FP64_TOL and err_mm < FP64_TOL
    with pytest.raises(H2PHipError) as e:
        H2MatHIP(h2, mode="aot")
    assert e.value.code == abi.ENOSYS


@pytest.mark.parametrize("name,prec", [("gaussian_deep", "fp64"), ("gaussian_deep", "fp32mix"),
                                       ("coulomb_deep", "fp64")])
def test_matvec_schedules_agree(name, prec, monkeypatch):
    """A JIT handle of the scalar kernels has four matvec schedules (D blocks
    beside the sweeps or in order; the levels above the leaf level as a chain
    of their own on a third stream beside the deep levels' B blocks, or not):
    every schedule, as a captured graph and eagerly, gives bitwise the y of
    the profiled path, which runs every launch in order on one stream."""
    h2, _ = problem(name)
    g = H2MatHIP(h2, mode="jit", prec=prec)
    monkeypatch.setenv("H2P_HIP_NO_GRAPH", "1")
    e = H2MatHIP(h2, mode="jit", prec=prec)
    monkeypatch.delenv("H2P_HIP_NO_GRAPH")
    x = np.random.default_rng(29).random(h2.krnl_mat_size) - 0.5
    g.set_profiling(True)
    y_serial = g.matvec(x)
    g.set_profiling(False)
    assert g.info()["mv_level_split"] == 1      # the default where the tree allows it
    for k in range(4):
        g.mm_variant(5, 0, k)
        e.mm_variant(5, 0, k)
        assert g.mm_variant(5, 0) == k
        assert g.info()["mv_level_split"] == (0 if k & 2 else 1)
        for it in range(2):
            assert np.array_equal(g.matvec(x), y_serial), (k, it)
        assert np.array_equal(e.matvec(x), y_serial), k
    ref = O.OracleH2(h2, BD_JIT=1).matvec(x, n_thread=8)
    bar = 1e-12 if prec == "fp64" else 1e-6
    assert rel(y_serial, ref) < bar
