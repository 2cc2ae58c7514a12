"""Host construction: the representation has the reference's structure.

Structural invariants of H2P_partition_points (src/H2Pack_partition.c) and
H2P_build (src/H2Pack_build.c) that the apply path relies on.
"""
import numpy as np
import pytest

from h2pack_amd import construct as C
from tests.problems import problem


@pytest.mark.parametrize("name", ["tiny_coulomb", "tiny_rpy", "coulomb"])
def test_tree_partitions_points(name):
    h2, coord = problem(name)
    pc = h2.pt_cluster.reshape(-1, 2)
    ch = h2.children.reshape(h2.n_node, -1)
    # root covers everything, post order: root is last
    assert h2.root_idx == h2.n_node - 1
    assert tuple(pc[h2.root_idx]) == (0, h2.n_point - 1)
    for n in range(h2.n_node):
        kids = ch[n, :h2.n_child[n]]
        assert np.all(kids < n)                    # post order
        if kids.size:
            assert pc[kids[0], 0] == pc[n, 0] and pc[kids[-1], 1] == pc[n, 1]
            assert np.all(pc[kids[1:], 0] == pc[kids[:-1], 1] + 1)
            assert np.all(h2.node_level[kids] == h2.node_level[n] + 1)
    leaves = h2.leaf_nodes
    assert leaves.size == h2.n_leaf_node == int((h2.n_child == 0).sum())
    npts = pc[leaves, 1] - pc[leaves, 0] + 1
    assert npts.sum() == h2.n_point and np.all(npts <= h2.max_leaf_points)
    # sorted coordinates are a permutation of the input
    assert np.array_equal(np.sort(h2.coord_idx), np.arange(h2.n_point))
    assert np.allclose(h2.coord, coord[:, h2.coord_idx])
    # permutation indices are inverse of each other
    assert np.array_equal(h2.bwd_pmt_idx[h2.fwd_pmt_idx], np.arange(h2.krnl_mat_size))


@pytest.mark.parametrize("name", ["tiny_coulomb", "tiny_gaussian", "tiny_rpy", "tiny_hss"])
def test_block_partition_covers_matrix_once(name):
    """Diagonal leaf blocks, inadmissible pairs (both orientations) and the
    point sets of admissible pairs (both orientations) tile the matrix
    exactly once, as H2P_calc_reduced_adm_pairs guarantees."""
    h2, _ = problem(name)
    n = h2.n_point
    pc = h2.pt_cluster.reshape(-1, 2)
    cover = np.zeros((n, n), dtype=np.int32)
    for leaf in h2.leaf_nodes:
        s, e = pc[leaf]
        cover[s:e + 1, s:e + 1] += 1
    for pairs in (h2.r_inadm_pairs.reshape(-1, 2), h2.r_adm_pairs.reshape(-1, 2)):
        for a, b in pairs:
            (s0, e0), (s1, e1) = pc[a], pc[b]
            cover[s0:e0 + 1, s1:e1 + 1] += 1
            cover[s1:e1 + 1, s0:e0 + 1] += 1
    assert cover.min() == 1 and cover.max() == 1


@pytest.mark.parametrize("name", ["tiny_coulomb", "tiny_gaussian", "coulomb", "rpy", "tiny_hss"])
def test_U_has_identity_rows_on_skeleton(name):
    h2, _ = problem(name)
    kd = h2.krnl_dim
    pc = h2.pt_cluster.reshape(-1, 2)
    ch = h2.children.reshape(h2.n_node, -1)
    for n in range(h2.n_node):
        U = h2.U[n]
        if U.shape[1] == 0:
            continue
        if h2.n_child[n] == 0:
            cand = np.arange(pc[n, 0], pc[n, 1] + 1)
        else:
            cand = np.concatenate([h2.J[c] for c in ch[n, :h2.n_child[n]]])
        assert U.shape == (cand.size * kd, h2.J[n].size * kd)
        pos = np.searchsorted(cand, h2.J[n]) if h2.n_child[n] == 0 else \
            np.array([np.nonzero(cand == j)[0][0] for j in h2.J[n]])
        rows = (pos[:, None] * kd + np.arange(kd)).reshape(-1)
        assert np.allclose(U[rows], np.eye(rows.size), atol=1e-12)
        assert np.all(np.diff(h2.J[n]) > 0) or h2.n_child[n] > 0


def test_B_D_metadata_layout():
    h2, _ = problem("coulomb")
    kd = h2.krnl_dim
    assert h2.B_ptr[0] == 0 and h2.D_ptr[0] == 0
    assert np.all(np.diff(h2.B_ptr.astype(np.int64)) == h2.B_nrow.astype(np.int64) * h2.B_ncol)
    assert np.all(np.diff(h2.D_ptr.astype(np.int64)) == h2.D_nrow.astype(np.int64) * h2.D_ncol)
    # pair-to-index CSR holds each pair twice with +-(idx+1)
    v = h2.B_p2i_val
    assert sorted(v[v > 0]) == list(range(1, h2.n_B + 1))
    assert sorted(-v[v < 0]) == list(range(1, h2.n_B + 1))
    assert h2.D_p2i_val.size == h2.n_leaf_node + 2 * h2.n_r_inadm_pair
    pc = h2.pt_cluster.reshape(-1, 2)
    leaves = h2.leaf_nodes
    assert np.array_equal(h2.D_nrow[:h2.n_leaf_node], (pc[leaves, 1] - pc[leaves, 0] + 1) * kd)


def test_id_compress_relative_accuracy():
    """Row ID of a smooth kernel block reaches the requested tolerance
    (the criterion of extra/test_ID_compress.c:116-127)."""
    rng = np.random.default_rng(5)
    a = rng.random((3, 200))
    b = rng.random((3, 300)) + np.array([[3.0], [0.0], [0.0]])
    A = C.kernel_eval(C.KRNL_COULOMB, [], a, b)
    for tol in (1e-4, 1e-8):
        U, J = C.id_compress(A, C.QR_REL_NRM, tol)
        err = np.linalg.norm(A - U @ A[J]) / np.linalg.norm(A)
        assert err < 10 * tol
        assert J.size < 200


def test_host_kernel_matches_closed_forms():
    c0 = np.array([[0.0], [0.0], [0.0]])
    c1 = np.array([[2.0, 0.0], [0.0, 0.0], [0.0, 0.0]])
    assert np.allclose(C.kernel_eval(C.KRNL_COULOMB, [], c0, c1), [[0.5, 0.0]])
    assert np.allclose(C.kernel_eval(C.KRNL_COULOMB, [7.0], c0, c1), [[0.5, 7.0]])
    assert np.allclose(C.kernel_eval(C.KRNL_GAUSSIAN, [0.5], c0, c1), [[np.exp(-2.0), 1.0]])
    s3 = np.sqrt(3.0) * 2.0 * 0.7
    assert np.allclose(C.kernel_eval(C.KRNL_MATERN32, [0.7], c0, c1), [[(1 + s3) * np.exp(-s3), 1.0]])


@pytest.mark.parametrize("name", ["tiny_hss", "hss_coulomb"])
def test_hss_pairs_and_neighbour_lists(name):
    """HSS mode (H2P_HSS_calc_adm_inadm_pairs, H2P_calc_node_inadm_lists):
    weak admissibility makes every pair of distinct siblings admissible, so
    the minimum admissible level is 1 and only leaves keep dense blocks; a
    node's neighbours are the nodes of its level (and leaves above it) that
    touch it."""
    h2, _ = problem(name)
    assert h2.is_HSS == 1 and h2.min_adm_level == 1
    assert h2.n_r_inadm_pair == 0
    ch = h2.children.reshape(h2.n_node, -1)
    adm = {tuple(sorted(p)) for p in h2.r_adm_pairs.reshape(-1, 2).tolist()}
    for n in range(h2.n_node):
        kids = ch[n, :h2.n_child[n]].tolist()
        for i in range(len(kids)):
            for j in range(i + 1, len(kids)):
                assert tuple(sorted((kids[i], kids[j]))) in adm
    assert len(adm) == sum(k * (k - 1) // 2 for k in h2.n_child.tolist())
    box = h2.enbox.reshape(h2.n_node, -1)
    for n, lst in enumerate(h2.node_inadm_lists):
        for k in lst.tolist():
            assert k != n and n in h2.node_inadm_lists[k].tolist() or h2.n_child[k] == 0
            gap = np.abs((box[n, :3] + 0.5 * box[n, 3:]) - (box[k, :3] + 0.5 * box[k, 3:])) \
                - 0.5 * (box[n, 3:] + box[k, 3:])
            assert gap.max() < 1e-6 * box[n, 3:].max() + 1e-9     # touching, not separated
