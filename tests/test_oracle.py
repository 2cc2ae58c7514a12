"""Pinning the oracle (oracle/h2oracle.c) before trusting it.

The reference cannot be compiled in this image and holds no golden vectors
for the apply path, so the oracle is pinned by what the mathematics fixes:

* the H2 matrix assembled explicitly from nested bases (tests/problems.py
  ``dense_h2``, independent numpy code) -- catches dropped, transposed or
  mis-indexed U / B / D terms in every phase;
* closed forms of the four kernels (Coulomb, Gaussian, Matern-3/2 and the
  Rotne-Prager-Yamakawa tensor in its textbook form for equal and unequal
  radii, its continuity at r = a_i + a_j);
* the FP64 direct n-body product against the build tolerance
  (extra/test_H2_scalar.c:89-99);
* matmul == matvec per column (extra/test_H2_matmul.h:18-116), and AOT == JIT.
"""
import math

import numpy as np
import pytest

import oracle as O
from h2pack_amd import construct as C
from tests.problems import dense_h2, problem, rel


@pytest.mark.parametrize("name", ["tiny_coulomb", "tiny_gaussian", "tiny_rpy", "tiny_hss"])
@pytest.mark.parametrize("jit", [0, 1])
def test_oracle_matvec_equals_dense_assembly(name, jit):
    h2, _ = problem(name)
    A = dense_h2(h2)
    x = np.random.default_rng(1).random(h2.krnl_mat_size) - 0.5
    y = O.OracleH2(h2, BD_JIT=jit).matvec(x, n_thread=4)
    assert rel(y, A @ x) < 1e-13


def test_dense_assembly_catches_planted_mistakes():
    """Mistakes planted in the representation the oracle consumes (a B block
    used transposed, a dropped D block, a sign error in one U entry) are all
    visible to the dense-assembly check, so the check can see the kinds of
    error an apply path makes."""
    import copy
    h2, _ = problem("tiny_coulomb")
    A = dense_h2(h2)
    x = np.random.default_rng(2).random(h2.krnl_mat_size) - 0.5
    y_ok = O.OracleH2(h2, BD_JIT=0).matvec(x, n_thread=2)
    assert rel(y_ok, A @ x) < 1e-13
    # (1) a square B block stored transposed
    sq = np.nonzero(h2.B_nrow == h2.B_ncol)[0]
    i = int(sq[0])
    bad = copy.copy(h2)
    B = h2.B_data.copy()
    s, e, r = int(h2.B_ptr[i]), int(h2.B_ptr[i + 1]), int(h2.B_nrow[i])
    B[s:e] = B[s:e].reshape(r, r).T.reshape(-1)
    bad.B_data = B
    assert rel(O.OracleH2(bad, BD_JIT=0).matvec(x, n_thread=2), A @ x) > 1e-8
    # (2) a dropped off-diagonal D block
    bad = copy.copy(h2)
    D = h2.D_data.copy()
    D[int(h2.D_ptr[h2.n_leaf_node]):int(h2.D_ptr[h2.n_leaf_node + 1])] = 0.0
    bad.D_data = D
    assert rel(O.OracleH2(bad, BD_JIT=0).matvec(x, n_thread=2), A @ x) > 1e-6
    # (3) a sign error in one non-skeleton U entry of a leaf
    bad = copy.copy(h2)
    bad.U = [u.copy() for u in h2.U]
    leaf = next(n for n in h2.leaf_nodes if h2.U[n].shape[1] > 0)
    k = int(np.argmax(np.abs(bad.U[leaf]).sum(axis=1) * (np.abs(bad.U[leaf]).max(axis=1) < 0.999)))
    bad.U[leaf][k] *= -1.0
    assert rel(O.OracleH2(bad, BD_JIT=0).matvec(x, n_thread=2), A @ x) > 1e-8


def _rpy_textbook_equal(r_vec, a, eta):
    """RPY mobility for equal radii (Rotne & Prager 1969, Yamakawa 1970)."""
    r = np.linalg.norm(r_vec)
    if r == 0:
        return np.eye(3) / (6 * math.pi * eta * a)
    u = r_vec / r
    if r >= 2 * a:
        pre = 1.0 / (8 * math.pi * eta * r)
        return pre * ((1 + 2 * a * a / (3 * r * r)) * np.eye(3) + (1 - 2 * a * a / (r * r)) * np.outer(u, u))
    pre = 1.0 / (6 * math.pi * eta * a)
    return pre * ((1 - 9 * r / (32 * a)) * np.eye(3) + (3 * r / (32 * a)) * np.outer(u, u))


@pytest.mark.parametrize("r", [0.0, 0.7, 1.5, 2.0000001, 3.3, 10.0])
def test_oracle_rpy_equals_textbook(r):
    eta, a = 1.3, 1.0
    d = np.array([0.3, -0.4, 0.866]) / np.linalg.norm([0.3, -0.4, 0.866])
    c0 = np.array([[1.0], [2.0], [3.0], [a]])
    c1 = c0.copy()
    c1[:3, 0] -= r * d
    K = O.kernel_eval(3, [eta], c0, c1)
    assert np.allclose(K, _rpy_textbook_equal(r * d, a, eta), rtol=1e-12, atol=1e-14)


def test_oracle_rpy_unequal_radii_regimes():
    eta = 0.8
    C0 = 1.0 / (6 * math.pi * eta)
    # one sphere inside the other: C / max(a) I (Zuk et al. 2014)
    c0 = np.array([[0.0], [0.0], [0.0], [2.0]])
    c1 = np.array([[0.5], [0.0], [0.0], [0.5]])
    assert np.allclose(O.kernel_eval(3, [eta], c0, c1), C0 / 2.0 * np.eye(3))
    # continuity where the far and overlapping formulas meet, r = a_i + a_j
    for eps in (1e-9, -1e-9):
        c1 = np.array([[2.5 + eps], [0.0], [0.0], [0.5]])
        K = O.kernel_eval(3, [eta], c0, c1)
        c1b = np.array([[2.5 - eps], [0.0], [0.0], [0.5]])
        Kb = O.kernel_eval(3, [eta], c0, c1b)
        assert np.allclose(K, Kb, rtol=1e-6)
    # symmetry of the tensor kernel: K(x, y) = K(y, x)^T
    rng = np.random.default_rng(3)
    p = np.concatenate([rng.random((3, 7)) * 4, 0.5 + rng.random((1, 7))], axis=0)
    q = np.concatenate([rng.random((3, 5)) * 4, 0.5 + rng.random((1, 5))], axis=0)
    assert np.allclose(O.kernel_eval(3, [eta], p, q), O.kernel_eval(3, [eta], q, p).T, rtol=1e-13)


def test_oracle_scalar_kernels_closed_forms():
    c0 = np.array([[0.0], [0.0], [0.0]])
    c1 = np.array([[0.0, 1.0, 0.0], [0.0, 2.0, 0.0], [0.0, 2.0, 0.5]])   # r = 0, 3, 0.5
    r = np.array([0.0, 3.0, 0.5])
    assert np.allclose(O.kernel_eval(0, [], c0, c1), [[0.0, 1 / 3, 2.0]])
    assert np.allclose(O.kernel_eval(0, [4.0], c0, c1), [[4.0, 1 / 3, 2.0]])
    assert np.allclose(O.kernel_eval(1, [0.3], c0, c1), [np.exp(-0.3 * r * r)])
    t = math.sqrt(3) * 0.9 * r
    assert np.allclose(O.kernel_eval(2, [0.9], c0, c1), [(1 + t) * np.exp(-t)])


@pytest.mark.parametrize("name", ["coulomb", "gaussian", "matern32", "rpy"])
def test_oracle_matches_direct_nbody_to_build_tolerance(name):
    h2, coord = problem(name)
    kid, param = h2.krnl_id, list(h2.krnl_param)
    rng = np.random.default_rng(7)
    x = rng.random(h2.krnl_mat_size) - 0.5
    rows = np.sort(rng.choice(h2.n_point, 300, replace=False))
    y_ref = O.direct_nbody(kid, param, coord, x, rows, n_thread=4)
    # the direct product is itself checked against a numpy kernel matrix
    K = C.kernel_eval(kid, param, coord[:, rows], coord)
    assert rel(K @ x, y_ref) < 1e-13
    y = O.OracleH2(h2, BD_JIT=1).matvec(x, n_thread=4)
    kd = h2.krnl_dim
    yr = y.reshape(-1, kd)[rows].reshape(-1)
    assert rel(yr, y_ref) < 20 * h2.QR_stop_tol


@pytest.mark.parametrize("name", ["coulomb", "rpy"])
def test_oracle_aot_equals_jit_and_matmul_equals_matvec(name):
    h2, _ = problem(name)
    rng = np.random.default_rng(9)
    X = rng.random((h2.krnl_mat_size, 6)) - 0.5
    oa, oj = O.OracleH2(h2, BD_JIT=0), O.OracleH2(h2, BD_JIT=1)
    cols = np.stack([oa.matvec(X[:, i], n_thread=4) for i in range(6)], axis=1)
    assert rel(np.stack([oj.matvec(X[:, i], n_thread=4) for i in range(6)], 1), cols) < 1e-14
    assert rel(oa.matmul(X, "col", n_thread=4), cols) < 1e-14
    assert rel(oj.matmul(X, "row", n_thread=4), cols) < 1e-14


# ---------------------------------------------------------------------------
# PCG restatement (examples/PCG/pcg.c:11-152), pinned by what CG must do in
# exact arithmetic and by a direct solve.
def test_oracle_pcg_finite_termination_on_three_eigenvalues():
    """An SPD operator with 3 distinct eigenvalues: CG ends in 3 iterations."""
    rng = np.random.default_rng(1)
    d = np.repeat([1.0, 4.0, 9.0], 20)
    Q, _ = np.linalg.qr(rng.standard_normal((60, 60)))
    A = (Q * d) @ Q.T
    b = rng.standard_normal(60)
    x, flag, relres, it = O.pcg(lambda v: A @ v, b, 1e-10, 50)
    assert flag == 0 and it == 3
    assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) < 1e-10


def test_oracle_pcg_matches_direct_solve_and_edge_cases():
    rng = np.random.default_rng(2)
    M = rng.standard_normal((80, 80))
    A = M @ M.T / 80 + 0.5 * np.eye(80)
    b = rng.standard_normal(80)
    x, flag, relres, it = O.pcg(lambda v: A @ v, b, 1e-12, 500)
    assert flag == 0 and relres <= 1e-12
    assert np.linalg.norm(x - np.linalg.solve(A, b)) / np.linalg.norm(np.linalg.solve(A, b)) < 1e-10
    # max_iter = 0: no iteration, x0 returned, flag 1 (pcg.c:63, :125)
    x0, flag0, rr0, it0 = O.pcg(lambda v: A @ v, b, 1e-12, 0)
    assert it0 == 0 and flag0 == 1 and np.all(x0 == 0) and abs(rr0 - 1.0) < 1e-15
    # b = 0 with x0 = 0: r = 0 <= tol * 0 stops at once, flag 0
    xz, flagz, rrz, itz = O.pcg(lambda v: A @ v, np.zeros(80), 1e-6, 10)
    assert itz == 0 and flagz == 0 and math.isnan(rrz)


@pytest.mark.parametrize("name", ["hss_coulomb", "hss_gaussian"])
def test_hss_representation_accuracy(name):
    """HSS mode against the dense kernel matrix: weak admissibility keeps the
    error within a few hundred times the ID tolerance (reltol 1e-6)."""
    h2, coord = problem(name)
    x = np.random.default_rng(6).random(h2.krnl_mat_size) - 0.5
    y = O.OracleH2(h2, BD_JIT=1).matvec(x, n_thread=4)
    K = C.kernel_eval(h2.krnl_id, h2.krnl_param, coord, coord)
    err = rel(y, K @ x)
    print(f"[{name}] HSS rel err vs dense = {err:.2e}")
    assert err < 5e-4


# ---------------------------------------------------------------------------
# Periodic mode: the Ewald kernel (src/H2Pack_3D_kernels.h:961-1210) and the
# periodic matvec restatement (oracle.periodic_matvec)
def test_ewald_kernel_invariants():
    """The Ewald sum does not depend on its splitting parameter once both
    sums have converged; it is periodic in the cell and symmetric."""
    rng = np.random.default_rng(1)
    L = 12.0
    c = np.concatenate([rng.random((3, 40)) * L, np.full((1, 40), 0.5)])
    K1 = C.rpy_ewald_eval(c, c, L, 0.2, 4, 4)
    K2 = C.rpy_ewald_eval(c, c, L, 0.3, 6, 6)
    assert np.abs(K1 - K2).max() / np.abs(K1).max() < 1e-13
    assert np.abs(K1 - K1.T).max() < 1e-14
    c2 = c.copy(); c2[1] += L                      # every source moved one cell up
    K3 = C.rpy_ewald_eval(c, c2, L, 0.3, 6, 6)
    off = np.kron(~np.eye(40, dtype=bool), np.ones((3, 3), dtype=bool))
    assert np.abs(K3 - K2)[off].max() / np.abs(K2).max() < 1e-13
    # the reciprocal truncation depends on nk alone: with the real-space sum
    # already converged at nr = 2, nr = 6 changes nothing ...
    K62 = C.rpy_ewald_eval(c, c, L, 0.3, 6, 2)
    K22 = C.rpy_ewald_eval(c, c, L, 0.3, 2, 2)
    assert np.abs(K62 - K22).max() / np.abs(K22).max() < 1e-12
    # ... while a truncated reciprocal sum (a planted mistake) breaks the
    # splitting-parameter independence
    assert np.abs(K62 - K2).max() / np.abs(K2).max() > 1e-6


def test_ewald_self_mobility_matches_hasimoto():
    """One sphere of radius a per cubic cell of side L: the periodic
    self-mobility is (1/a) (1 - 2.8372974794806 a/L + (4 pi / 3) (a/L)^3)
    (Hasimoto, J. Fluid Mech. 5, 317 (1959), simple cubic array; the RPY
    Ewald sum reproduces these terms, Beenakker 1986), isotropic, and
    independent of where the sphere sits and of the splitting parameter."""
    H = 2.8372974794806
    for a, L in ((0.5, 10.0), (1.0, 10.0), (1.0, 6.0), (0.3, 20.0)):
        c = np.array([[1.3], [2.1], [0.7], [a]])
        for xi in (2.5 / L, 4.0 / L):
            K = C.rpy_ewald_eval(c, c, L, xi, 6, 8)
            want = (1.0 - H * a / L + 4.0 * np.pi / 3.0 * (a / L) ** 3) / a
            assert abs(K[0, 0] - want) < 1e-12 * want, (a, L, xi, K[0, 0], want)
            assert np.abs(K - K[0, 0] * np.eye(3)).max() < 1e-13 * want


def test_periodic_pairs_are_one_sided_images():
    from tests.problems import periodic_problem
    h2, _, _ = periodic_problem()
    assert h2.min_adm_level == 0 and h2.is_RPY_Ewald == 1
    for pairs, shifts in ((h2.r_adm_pairs, h2.per_adm_shifts), (h2.r_inadm_pairs, h2.per_inadm_shifts)):
        keys = {(int(a), int(b), tuple(np.round(s, 9))) for (a, b), s in zip(pairs.reshape(-1, 2), shifts)}
        assert len(keys) == pairs.size // 2
        for a, b, s in keys:                   # every interaction is present in both directions
            assert (b, a, tuple(-v + 0.0 for v in s)) in keys
    # B_p2i / D_p2i hold positive entries only (src/H2Pack_build.c:745-752)
    assert np.all(h2.B_p2i_val > 0) and np.all(h2.D_p2i_val > 0)


def test_periodic_oracle_against_dense_ewald():
    """The periodic representation applied by the oracle equals the dense
    Ewald kernel matrix to the build tolerance; dropping the root's periodic
    block (a plausible mistake) is caught."""
    from tests.problems import periodic_problem
    h2, coord, (L, xi, nr, nk) = periodic_problem()
    x = np.random.default_rng(3).standard_normal(h2.krnl_mat_size)
    K = C.rpy_ewald_eval(coord, coord, L, xi, nr, nk)
    yd = K @ x
    y = O.periodic_matvec(h2, x)
    assert rel(y, yd) < 1e-5
    saved = h2.per_blk
    try:
        h2.per_blk = np.zeros_like(saved)
        assert rel(O.periodic_matvec(h2, x), yd) > 1e-2
    finally:
        h2.per_blk = saved
