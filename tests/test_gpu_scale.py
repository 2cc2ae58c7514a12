"""GPU parity at larger sizes than the other parity tests.

Bars (as in the other parity tests): the FP64 product agrees with the
oracle's H2P_matvec / H2P_matmul restatement on the same representation to
1e-12 relative 2-norm, in JIT mode and with the blocks evaluated into HBM
(AOT, compared with the oracle's on-the-fly evaluation of the same blocks);
FP32-mixed within 1e-6 of the FP64 oracle (the build tolerance); the
served CG solves at the serving configuration's nugget 1e-2 reach their
tolerance on the FP64 operator.
"""
import numpy as np
import pytest

import oracle as O
from h2pack_amd import construct as C
from h2pack_amd.h2mat import H2MatHIP
from h2pack_amd.serve import Server, make_requests
from tests.problems import rel

pytestmark = pytest.mark.gpu


def _build(kid, param, n, leaf, seed):
    coord = C.uniform_cube_points(n, 10.0, seed)
    return C.build_h2(coord, kid, param, 1e-6, leaf, BD_JIT=1, pp_kind="numerical", host_blocks=False)


@pytest.mark.parametrize("kid,param,n", [(C.KRNL_GAUSSIAN, [1.0], 200_000),
                                         (C.KRNL_COULOMB, [], 100_000),
                                         (C.KRNL_MATERN32, [1.0], 100_000)])
def test_large_matvec_matches_oracle(kid, param, n):
    h2 = _build(kid, param, n, 300, 17)
    x = np.random.default_rng(4).random(h2.krnl_mat_size) - 0.5
    ref = O.OracleH2(h2, BD_JIT=1).matvec(x, n_thread=16)
    for mode, prec, bar in (("jit", "fp64", 1e-12), ("aot", "fp64", 1e-12), ("jit", "fp32mix", 1e-6)):
        m = H2MatHIP(h2, mode=mode, prec=prec, host_blocks=False)
        err = rel(m.matvec(x), ref)
        print(f"[scale kid={kid} N={n} leaf 300 {mode} {prec}] rel err vs oracle {err:.3e}", flush=True)
        assert err < bar
        m.close()


def test_rpy_at_the_matmul_configuration_shape():
    """RPY (radius 1, 20% volume fraction, reltol 1e-7, leaf 200: the shape
    of the RPY matmul configuration at N = 30k): matvec in JIT FP64, AOT
    FP64 and JIT FP32-mixed, and an 8-column matmul in JIT and AOT, against
    the oracle."""
    import math
    n = 30_000
    L = (n * (4.0 / 3.0) * math.pi / 0.2) ** (1.0 / 3.0)
    rng = np.random.default_rng(19)
    coord = np.concatenate([rng.random((3, n)) * L, np.ones((1, n))], axis=0)
    h2 = C.build_h2(coord, C.KRNL_RPY, [1.0], 1e-7, 200, BD_JIT=1, pp_kind="surface", max_leaf_size=4.0,
                    host_blocks=False)
    orc = O.OracleH2(h2, BD_JIT=1)
    x = rng.random(h2.krnl_mat_size) - 0.5
    X = rng.standard_normal((h2.krnl_mat_size, 8))
    y_ref = orc.matvec(x, n_thread=16)
    Y_ref = orc.matmul(X, "col", n_thread=16)
    for mode, prec, bar in (("jit", "fp64", 1e-12), ("aot", "fp64", 1e-12), ("jit", "fp32mix", 1e-6)):
        m = H2MatHIP(h2, mode=mode, prec=prec, host_blocks=False)
        err = rel(m.matvec(x), y_ref)
        msg = f"[rpy N={n} {mode} {prec}] matvec rel err vs oracle {err:.3e}"
        if prec == "fp64":
            err_mm = rel(m.matmul(X, "col"), Y_ref)
            msg += f", 8-column matmul {err_mm:.3e}"
            assert err_mm < bar
        print(msg, flush=True)
        assert err < bar
        m.close()


def test_large_matmul_matches_oracle():
    h2 = _build(C.KRNL_MATERN32, [1.0], 100_000, 300, 18)
    X = np.random.default_rng(5).standard_normal((h2.krnl_mat_size, 8))
    ref = O.OracleH2(h2, BD_JIT=1).matmul(X, "col", n_thread=16)
    for mode, prec, bar in (("jit", "fp64", 1e-12), ("aot", "fp64", 1e-12), ("jit", "fp32mix", 1e-6)):
        m = H2MatHIP(h2, mode=mode, prec=prec, host_blocks=False)
        err = rel(m.matmul(X, "col"), ref)
        print(f"[scale matmul N=100k 8 RHS {mode} {prec}] rel err vs oracle {err:.3e}", flush=True)
        assert err < bar
        m.close()


def test_serving_at_the_configured_nugget():
    """CG at nugget 1e-2 (condition numbers in the hundreds and more): the
    FP64 path reaches the tolerance on the FP64 operator.  On the FP32-mixed
    path the FP32 operator's rounding (relative to |A| |x|, not |A x|) alone
    would leave a residual of ~1e-7 kappa on the FP64 operator; the server's
    residual replacements (the residual formed again with the FP64 handle,
    the search direction kept) bring it to the tolerance, which the oracle
    confirms, in about as many iterations as the FP64 solve."""
    import copy
    base = make_requests(4, seed=23, n_min=5000, n_max=20000, rate=50.0, leaf=300,
                         max_iter=4000, nugget=1e-2)
    runs = {}
    for prec in ("fp64", "fp32mix"):
        rs = [copy.copy(r) for r in base]
        Server(hbm_budget=64 << 30, max_active=4, prec=prec).run(rs, realtime=False)
        runs[prec] = rs
    for r64, r32 in zip(runs["fp64"], runs["fp32mix"]):
        if r64.kind == "matmul":
            continue
        orc = O.OracleH2(r64.h2, BD_JIT=1)
        nb = np.linalg.norm(r64.rhs)
        res64 = np.linalg.norm(r64.rhs - orc.matvec(r64.result, n_thread=16) - r64.shift * r64.result) / nb
        x = r32.result
        a64 = orc.matvec(x, n_thread=16) + r32.shift * x
        m32 = H2MatHIP(r32.h2, mode="jit", prec="fp32mix")
        a32 = m32.matvec(x) + r32.shift * x
        m32.close()
        res32 = np.linalg.norm(r32.rhs - a32) / nb
        res32_on64 = np.linalg.norm(r32.rhs - a64) / nb
        opdiff = np.linalg.norm(a32 - a64)
        print(f"request {r64.rid}: N={r64.N} nugget 1e-2: FP64 CG {r64.iters} it, true residual {res64:.2e}; "
              f"FP32-mixed CG {r32.iters} it, residual {res32:.2e} on its operator, {res32_on64:.2e} on the "
              f"FP64 one, operator difference {opdiff / np.linalg.norm(a64):.2e} of |A x|", flush=True)
        assert r64.relres <= r64.tol and res64 <= 2 * r64.tol
        # FP32-mixed with the server's residual replacements: the reported
        # residual is the FP64 operator's, and the oracle agrees with it
        assert r32.relres <= r32.tol and res32_on64 <= 2 * r32.tol
        assert abs(res32_on64 - r32.relres) <= 0.1 * r32.tol
        assert r32.iters <= 1.3 * r64.iters + 16
