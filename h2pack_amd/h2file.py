"""On-disk H2 representation in the reference's file layout.

``store`` / ``load`` mirror ``H2P_store_to_file`` / ``H2P_read_from_file``
(src/H2Pack_file_IO.c:17-219, :220-): a metadata JSON file and an auxiliary
JSON file, with the field names the reference's converter writes
(examples/meta_txt_to_json.py:14-145), plus one raw binary file of FP64
values: every U (node order, row-major), then every B block (admissible pair
order), then every D block (leaves in height_nodes order, then the
inadmissible pairs), each row-major (src/H2Pack_file_IO.c:104-160).

A representation read back uploads straight to the device
(``H2MatHIP(load(...))``) without rebuilding it on the host, in JIT mode from
the metadata alone or in AOT mode with the stored B and D blocks.
As in the reference, representations with extended coordinates (the RPY
radius) are not stored (src/H2Pack_file_IO.c:21-25).
"""
from __future__ import annotations

import json

import numpy as np

from . import construct as C


def _b_block(h2, i):
    if h2.B_data is not None:
        return h2.B_data[int(h2.B_ptr[i]):int(h2.B_ptr[i + 1])]
    c0, c1 = C.B_block_coords(h2, i)
    return C.kernel_eval(h2.krnl_id, h2.krnl_param, c0, c1).reshape(-1)


def _d_block(h2, i):
    if h2.D_data is not None:
        return h2.D_data[int(h2.D_ptr[i]):int(h2.D_ptr[i + 1])]
    c0, c1 = C.D_block_coords(h2, i)
    return C.kernel_eval(h2.krnl_id, h2.krnl_param, c0, c1).reshape(-1)


def store(h2, meta_json: str, aux_json: str, binary: str) -> None:
    """Write ``h2`` (a construct.H2Pack after build) to the three files."""
    if h2.xpt_dim > h2.pt_dim:
        raise ValueError("representations with extended coordinates (RPY) are not stored, "
                         "as in H2P_store_to_file")
    nn, mc = h2.n_node, h2.max_child
    ch = h2.children.reshape(nn, mc)
    pc = h2.pt_cluster.reshape(nn, 2)
    adm = h2.r_adm_pairs.reshape(-1, 2)
    inadm = h2.r_inadm_pairs.reshape(-1, 2)
    lvl = h2.node_level
    leaves = h2.leaf_nodes
    meta = {
        "nrow_matrix": int(h2.krnl_mat_size), "ncol_matrix": int(h2.krnl_mat_size),
        "is_symmetric": 1,
        "num_node_row": int(nn), "num_node_col": int(nn),
        "root_node_row": int(h2.root_idx), "root_node_col": int(h2.root_idx),
        "num_level_row": int(h2.max_level + 1), "num_level_col": int(h2.max_level + 1),
        "num_inadmissible_blocks": int(h2.n_r_inadm_pair + h2.n_leaf_node),
        "num_admissible_blocks": int(h2.n_r_adm_pair),
        "has_partial_adm_blocks": int(bool(np.any(lvl[adm[:, 0]] != lvl[adm[:, 1]]))) if adm.size else 0,
        "nodes_row": [{"index": i, "level": int(lvl[i]), "cluster_head": int(pc[i, 0]),
                       "cluster_tail": int(pc[i, 1]), "num_children": int(h2.n_child[i]),
                       "children": [int(c) for c in ch[i, :h2.n_child[i]]]} for i in range(nn)],
        "basis_matrices_row": [{"node": i, "num_row": int(h2.U_nrow[i]), "num_col": int(h2.U_ncol[i])}
                               for i in range(nn)],
        "B_matrices": [{"node_row": int(a), "node_col": int(b), "num_row": int(h2.B_nrow[i]),
                        "num_col": int(h2.B_ncol[i]), "is_part_adm": int(lvl[a] != lvl[b])}
                       for i, (a, b) in enumerate(adm)],
        "D_matrices": [{"node_row": int(n), "node_col": int(n), "num_row": int(h2.D_nrow[i]),
                        "num_col": int(h2.D_ncol[i])} for i, n in enumerate(leaves)]
                      + [{"node_row": int(a), "node_col": int(b),
                          "num_row": int(h2.D_nrow[h2.n_leaf_node + i]),
                          "num_col": int(h2.D_ncol[h2.n_leaf_node + i])} for i, (a, b) in enumerate(inadm)],
    }
    # coordinates back in the caller's order (src/H2Pack_file_IO.c:164-181)
    coord0 = np.empty((h2.n_point, h2.pt_dim))
    coord0[h2.coord_idx] = h2.coord[:h2.pt_dim].T
    aux = {
        "dim_point": int(h2.pt_dim), "dim_kernel": int(h2.krnl_dim), "num_point": int(h2.n_point),
        "is_HSS": 0, "min_adm_level": int(h2.min_adm_level),
        "max_leaf_points": int(h2.max_leaf_points), "QR_stop_tol": float(h2.QR_stop_tol),
        "has_skeleton_points": 1,
        "point_coordinate": coord0.reshape(-1).tolist(),
        "permutation_array": [int(v) for v in h2.coord_idx],
        "skeleton_points": [{"node": i, "num_skeleton_point": int(h2.J[i].size),
                             "skeleton_point_indices": [int(v) for v in h2.J[i]]} for i in range(nn)],
    }
    with open(binary, "wb") as f:
        for i in range(nn):
            f.write(np.ascontiguousarray(h2.U[i], dtype=np.float64).tobytes())
        for i in range(h2.n_r_adm_pair):
            f.write(np.ascontiguousarray(_b_block(h2, i), dtype=np.float64).tobytes())
        for i in range(h2.n_leaf_node + h2.n_r_inadm_pair):
            f.write(np.ascontiguousarray(_d_block(h2, i), dtype=np.float64).tobytes())
    with open(meta_json, "w") as f:
        json.dump(meta, f, indent=2)
    with open(aux_json, "w") as f:
        json.dump(aux, f, indent=2)


def load(meta_json: str, aux_json: str, binary: str, krnl_id: int, krnl_param,
         BD_JIT: int = 1) -> C.H2Pack:
    """Read a stored representation (src/H2Pack_file_IO.c:220-) for kernel
    ``krnl_id`` with ``krnl_param``; BD_JIT = 0 also reads the B and D blocks."""
    with open(meta_json) as f:
        meta = json.load(f)
    with open(aux_json) as f:
        aux = json.load(f)
    if meta["is_symmetric"] != 1:
        raise ValueError("only symmetric kernel matrices are supported (as in H2P_read_from_file)")
    if aux.get("is_HSS", 0):
        raise ValueError("HSS representations are not read by the H2 engine")
    pt_dim, kd, n_point = aux["dim_point"], aux["dim_kernel"], aux["num_point"]
    if C.KRNL_DIM[krnl_id] != kd:
        raise ValueError("kernel dimension does not match the file")
    h2 = C.H2Pack(pt_dim=pt_dim, krnl_dim=kd, QR_stop_type=C.QR_REL_NRM,
                  QR_stop_param=float(aux["QR_stop_tol"]))
    nn = meta["num_node_row"]
    mc = h2.max_child
    h2.n_point = n_point
    h2.krnl_mat_size = meta["nrow_matrix"]
    h2.n_node = nn
    h2.root_idx = meta["root_node_row"]
    h2.max_level = meta["num_level_row"] - 1
    h2.max_leaf_points = aux["max_leaf_points"]
    # ---- tree
    node_level = np.zeros(nn, np.int32)
    pt_cluster = np.zeros(2 * nn, np.int32)
    n_child = np.zeros(nn, np.int32)
    children = np.full(nn * mc, -1, np.int32)
    parent = np.full(nn, -1, np.int32)
    for nd in meta["nodes_row"]:
        i = nd["index"]
        node_level[i] = nd["level"]
        pt_cluster[2 * i], pt_cluster[2 * i + 1] = nd["cluster_head"], nd["cluster_tail"]
        n_child[i] = nd["num_children"]
        for j, c in enumerate(nd["children"]):
            children[i * mc + j] = c
            parent[c] = i
    h2.node_level, h2.pt_cluster, h2.n_child, h2.children, h2.parent = (
        node_level, pt_cluster, n_child, children, parent)
    pc = pt_cluster.reshape(nn, 2)
    h2.mat_cluster = np.stack([kd * pc[:, 0], kd * (pc[:, 1] + 1) - 1], axis=1).reshape(-1).astype(np.int32)
    n_leaf = int(np.sum(n_child == 0))
    h2.n_leaf_node = n_leaf
    nlev = h2.max_level + 1
    h2.level_n_node = np.zeros(nlev, np.int32)
    h2.level_nodes = np.full(nlev * n_leaf, -1, np.int32)
    for i in range(nn):           # nodes of a level in index order
        lv = node_level[i]
        h2.level_nodes[lv * n_leaf + h2.level_n_node[lv]] = i
        h2.level_n_node[lv] += 1
    # heights: leaves 0, parents one above their tallest child
    height = np.zeros(nn, np.int32)
    for lv in range(h2.max_level, -1, -1):
        for i in h2.level_nodes[lv * n_leaf: lv * n_leaf + h2.level_n_node[lv]]:
            if n_child[i]:
                height[i] = 1 + max(height[c] for c in children[i * mc: i * mc + n_child[i]])
    h2.node_height = height
    # ---- D order fixes the leaves' order (height_nodes row 0)
    D = meta["D_matrices"]
    leaves = np.asarray([d["node_row"] for d in D[:n_leaf]], np.int32)
    nh = int(height.max()) + 1
    h2.height_n_node = np.zeros(nh, np.int32)
    h2.height_nodes = np.full(nh * n_leaf, -1, np.int32)
    h2.height_nodes[:n_leaf] = leaves
    h2.height_n_node[0] = n_leaf
    for i in range(nn):
        if height[i] > 0:
            hh = height[i]
            h2.height_nodes[hh * n_leaf + h2.height_n_node[hh]] = i
            h2.height_n_node[hh] += 1
    # ---- pairs
    B = meta["B_matrices"]
    h2.n_r_adm_pair = len(B)
    h2.r_adm_pairs = np.asarray([[b["node_row"], b["node_col"]] for b in B], np.int32).reshape(-1)
    inad = D[n_leaf:]
    h2.n_r_inadm_pair = len(inad)
    h2.r_inadm_pairs = np.asarray([[d["node_row"], d["node_col"]] for d in inad], np.int32).reshape(-1)
    h2.min_adm_level = aux["min_adm_level"]
    # ---- points: sorted coordinates from the caller's order and the permutation
    coord0 = np.asarray(aux["point_coordinate"], np.float64).reshape(n_point, pt_dim)
    h2.coord_idx = np.asarray(aux["permutation_array"], np.int32)
    h2.coord = np.ascontiguousarray(coord0[h2.coord_idx].T)
    h2.xpt_dim = pt_dim
    # ---- U (binary) and skeletons
    raw = np.fromfile(binary, dtype=np.float64)
    off = 0
    U = [None] * nn
    U_nrow = np.zeros(nn, np.int32)
    U_ncol = np.zeros(nn, np.int32)
    for u in meta["basis_matrices_row"]:
        i, r, c = u["node"], u["num_row"], u["num_col"]
        U[i] = raw[off:off + r * c].reshape(r, c).copy()
        U_nrow[i], U_ncol[i] = r, c
        off += r * c
    h2.U, h2.U_nrow, h2.U_ncol = U, U_nrow, U_ncol
    J = [np.zeros(0, np.int64)] * nn
    for sk in aux["skeleton_points"]:
        J[sk["node"]] = np.asarray(sk["skeleton_point_indices"], np.int64)
    h2.J = J
    h2.J_coord = [np.ascontiguousarray(h2.coord[:, j]) for j in J]
    # ---- block metadata, then check it against the file's block sizes
    h2.krnl_id = int(krnl_id)
    h2.krnl_param = np.asarray(krnl_param if krnl_param is not None else [], dtype=np.float64)
    h2.krnl_bimv_flops = C.KRNL_BIMV_FLOPS[h2.krnl_id]
    h2.BD_JIT = int(BD_JIT)
    C._gen_B_metadata(h2)
    C._gen_D_metadata(h2)
    if (list(h2.B_nrow) != [b["num_row"] for b in B] or list(h2.B_ncol) != [b["num_col"] for b in B]
            or list(h2.D_nrow) != [d["num_row"] for d in D] or list(h2.D_ncol) != [d["num_col"] for d in D]):
        raise ValueError("block sizes in the file do not match the tree and skeletons")
    nb, nd = int(h2.B_ptr[-1]), int(h2.D_ptr[-1])
    if raw.size != off + nb + nd:
        raise ValueError(f"binary file holds {raw.size} values, expected {off + nb + nd}")
    if not BD_JIT:
        h2.B_data = raw[off:off + nb].copy()
        h2.D_data = raw[off + nb:off + nb + nd].copy()
    ci = h2.coord_idx.astype(np.int64)
    fwd = (ci[:, None] * kd + np.arange(kd)[None, :]).reshape(-1)
    bwd = np.empty_like(fwd)
    bwd[fwd] = np.arange(fwd.size)
    h2.fwd_pmt_idx = fwd.astype(np.int32)
    h2.bwd_pmt_idx = bwd.astype(np.int32)
    return h2
