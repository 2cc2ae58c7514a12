"""On-disk representation (h2pack_amd/h2file.py): store then load gives back
the same representation, field by field, and the same products (the oracle's
H2P_matvec restatement, bitwise); the file layout carries the reference's
field names (examples/meta_txt_to_json.py:14-145)."""
import json
import os

import numpy as np
import pytest

import oracle as O
from h2pack_amd import construct as C
from h2pack_amd import h2file
from tests.problems import problem

FIELDS = ("children", "n_child", "node_level", "pt_cluster", "mat_cluster", "level_n_node",
          "r_adm_pairs", "r_inadm_pairs", "node_n_r_adm", "U_nrow", "U_ncol", "coord",
          "fwd_pmt_idx", "bwd_pmt_idx", "B_nrow", "B_ncol", "B_ptr", "D_nrow", "D_ncol", "D_ptr",
          "leaf_nodes", "B_p2i_rowptr", "B_p2i_colidx", "B_p2i_val", "D_p2i_rowptr",
          "D_p2i_colidx", "D_p2i_val")


def _paths(tmp_path):
    return (str(tmp_path / "meta.json"), str(tmp_path / "aux.json"), str(tmp_path / "blocks.bin"))


@pytest.mark.parametrize("name", ["gaussian", "coulomb"])
def test_store_load_round_trip(name, tmp_path):
    h2, _ = problem(name)
    paths = _paths(tmp_path)
    h2file.store(h2, *paths)
    for jit in (1, 0):
        g = h2file.load(*paths, h2.krnl_id, h2.krnl_param, BD_JIT=jit)
        for f in FIELDS:
            assert np.array_equal(np.asarray(getattr(g, f)), np.asarray(getattr(h2, f))), f
        assert g.min_adm_level == h2.min_adm_level and g.max_level == h2.max_level
        for a, b in zip(g.U, h2.U):
            assert np.array_equal(a, b)
        for a, b in zip(g.J, h2.J):
            assert np.array_equal(a, b)
        if not jit:
            assert np.array_equal(g.B_data, h2.B_data) and np.array_equal(g.D_data, h2.D_data)
        x = np.random.default_rng(2).random(h2.krnl_mat_size) - 0.5
        # one thread: the oracle's threaded reduction order varies run to run
        ya = O.OracleH2(h2, BD_JIT=jit).matvec(x, n_thread=1)
        yb = O.OracleH2(g, BD_JIT=jit).matvec(x, n_thread=1)
        assert np.array_equal(ya, yb)


def test_jit_only_representation_is_stored_with_evaluated_blocks(tmp_path):
    """A JIT build holds no blocks; the file gets them evaluated (as
    H2P_get_Bij_block / H2P_get_Dij_block do), so an AOT load works."""
    h2, coord = problem("tiny_gaussian")
    g0 = C.build_h2(coord, h2.krnl_id, h2.krnl_param, 1e-8, 40, BD_JIT=1, host_blocks=False)
    assert g0.B_data is None
    paths = _paths(tmp_path)
    h2file.store(g0, *paths)
    g = h2file.load(*paths, g0.krnl_id, g0.krnl_param, BD_JIT=0)
    x = np.random.default_rng(3).random(g.krnl_mat_size) - 0.5
    y_jit = O.OracleH2(g0, BD_JIT=1).matvec(x, n_thread=4)
    y_aot = O.OracleH2(g, BD_JIT=0).matvec(x, n_thread=4)
    assert np.linalg.norm(y_aot - y_jit) / np.linalg.norm(y_jit) < 1e-14


def test_file_layout_and_refusals(tmp_path):
    h2, _ = problem("tiny_gaussian")
    paths = _paths(tmp_path)
    h2file.store(h2, *paths)
    meta = json.load(open(paths[0]))
    aux = json.load(open(paths[1]))
    assert list(meta) == ["nrow_matrix", "ncol_matrix", "is_symmetric", "num_node_row", "num_node_col",
                          "root_node_row", "root_node_col", "num_level_row", "num_level_col",
                          "num_inadmissible_blocks", "num_admissible_blocks", "has_partial_adm_blocks",
                          "nodes_row", "basis_matrices_row", "B_matrices", "D_matrices"]
    assert list(aux) == ["dim_point", "dim_kernel", "num_point", "is_HSS", "min_adm_level",
                         "max_leaf_points", "QR_stop_tol", "has_skeleton_points", "point_coordinate",
                         "permutation_array", "skeleton_points"]
    total = sum(u.size for u in h2.U) + int(h2.B_ptr[-1]) + int(h2.D_ptr[-1])
    assert os.path.getsize(paths[2]) == 8 * total
    # a truncated binary file is refused
    with open(paths[2], "r+b") as f:
        f.truncate(8 * (total - 1))
    with pytest.raises(ValueError):
        h2file.load(*paths, h2.krnl_id, h2.krnl_param)
    # RPY (extended coordinates) is not stored, as in the reference
    r, _ = problem("tiny_rpy")
    with pytest.raises(ValueError):
        h2file.store(r, *paths)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["jit", "aot"])
def test_upload_from_file_gives_the_same_device_product(mode, tmp_path):
    from h2pack_amd.h2mat import H2MatHIP
    h2, _ = problem("coulomb")
    paths = _paths(tmp_path)
    h2file.store(h2, *paths)
    g = h2file.load(*paths, h2.krnl_id, h2.krnl_param, BD_JIT=1 if mode == "jit" else 0)
    x = np.random.default_rng(4).random(h2.krnl_mat_size) - 0.5
    a = H2MatHIP(h2, mode=mode).matvec(x)
    b = H2MatHIP(g, mode=mode).matvec(x)
    assert np.array_equal(a, b)


def test_layout_matches_the_reference_converter(tmp_path):
    """Fixtures from tests/golden/make_file_io_vectors.py: the JSON files are
    the output of the reference's own converter (examples/meta_txt_to_json.py)
    for a 120-point Coulomb representation.  Loading them and storing again
    must reproduce both JSON files value for value and the binary bit for
    bit; the loaded representation must be the kernel matrix to its build
    tolerance."""
    gold = os.path.join(os.path.dirname(__file__), "golden", "file_io")
    blob = np.load(os.path.join(gold, "coulomb120_blocks.npy"))
    binf = str(tmp_path / "gold.bin")
    blob.tofile(binf)
    meta_f, aux_f = os.path.join(gold, "coulomb120_meta.json"), os.path.join(gold, "coulomb120_aux.json")
    g = h2file.load(meta_f, aux_f, binf, C.KRNL_COULOMB, [], BD_JIT=0)
    paths = _paths(tmp_path)
    h2file.store(g, *paths)
    assert json.load(open(paths[0])) == json.load(open(meta_f))
    assert json.load(open(paths[1])) == json.load(open(aux_f))
    assert np.array_equal(np.fromfile(paths[2], dtype=np.float64), blob)
    # the representation itself: within its build tolerance of the dense matrix
    aux = json.load(open(aux_f))
    coord0 = np.asarray(aux["point_coordinate"]).reshape(-1, 3).T
    x = np.random.default_rng(5).random(g.krnl_mat_size) - 0.5
    y = O.OracleH2(g, BD_JIT=0).matvec(x, n_thread=1)
    K = C.kernel_eval(C.KRNL_COULOMB, [], coord0, coord0)
    assert np.linalg.norm(y - K @ x) / np.linalg.norm(K @ x) < 1e-3
