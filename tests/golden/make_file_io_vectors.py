"""Writes the file-format fixtures of tests/test_h2file.py (run once, in the
build container, where the reference tree is present).

The JSON layout of a stored H2 representation is produced by the
reference's converter examples/meta_txt_to_json.py from the intermediate
text file that H2P_store_to_file writes (src/H2Pack_file_IO.c:36-199).  This
script builds a small representation with h2pack_amd.construct, writes that
intermediate text file in the reference's format (the fprintf sequence of
H2Pack_file_IO.c:36-199), runs the reference converter on it, and keeps the
converter's two JSON outputs plus the FP64 binary as fixtures.  Only the
outputs are kept, never the converter itself.
"""
import json
import os
import struct
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from h2pack_amd import construct as C   # noqa: E402

REF_EXAMPLES = "/root/reference/examples"
OUT = os.path.join(HERE, "file_io")
NAME = "coulomb120"


def write_intermediate_txt(h2, path):
    nn, mc = h2.n_node, h2.max_child
    ch = h2.children.reshape(nn, mc)
    pc = h2.pt_cluster.reshape(nn, 2)
    adm = h2.r_adm_pairs.reshape(-1, 2)
    inadm = h2.r_inadm_pairs.reshape(-1, 2)
    lvl = h2.node_level
    L = []
    L += [str(h2.pt_dim), str(h2.krnl_dim), str(h2.n_point), str(h2.krnl_mat_size), str(h2.krnl_mat_size),
          "1", str(nn), str(nn), str(h2.root_idx), str(h2.root_idx), str(h2.max_level + 1),
          str(h2.max_level + 1), "0", str(h2.min_adm_level)]
    part = int(any(lvl[a] != lvl[b] for a, b in adm))
    L += [str(h2.n_r_inadm_pair), str(h2.n_r_adm_pair), str(part)]
    for i in range(nn):
        kids = ch[i, :h2.n_child[i]].tolist()
        s = "%6d %2d %8d %8d %2d " % (i, lvl[i], pc[i, 0], pc[i, 1], h2.n_child[i])
        s += "".join("%8d " % k for k in kids) + "-1 " * (mc - len(kids))
        L.append(s)
    for i in range(nn):
        L.append("%6d %5d %5d" % (i, h2.U_nrow[i], h2.U_ncol[i]))
    for i, (a, b) in enumerate(adm):
        L.append("%6d %6d %5d %5d %d" % (a, b, h2.B_nrow[i], h2.B_ncol[i], int(lvl[a] != lvl[b])))
    for i, n in enumerate(h2.leaf_nodes):
        L.append("%6d %6d %5d %5d" % (n, n, h2.D_nrow[i], h2.D_ncol[i]))
    for i, (a, b) in enumerate(inadm):
        j = h2.n_leaf_node + i
        L.append("%6d %6d %5d %5d" % (a, b, h2.D_nrow[j], h2.D_ncol[j]))
    L += [str(h2.max_leaf_points), "%e" % h2.QR_stop_tol, "1"]
    coord0 = np.empty((h2.n_point, h2.pt_dim))
    coord0[h2.coord_idx] = h2.coord[:h2.pt_dim].T
    for i in range(h2.n_point):
        L.append("".join("%x " % struct.unpack("<Q", struct.pack("<d", v))[0] for v in coord0[i]))
    L += [str(v) for v in h2.coord_idx]
    for i in range(nn):
        L.append("%6d %6d " % (i, h2.J[i].size) + "".join("%d " % v for v in h2.J[i]))
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


def main():
    coord = C.uniform_cube_points(120, 1.0, seed=77)
    h2 = C.build_h2(coord, C.KRNL_COULOMB, [], 1e-4, 20, BD_JIT=0, pp_kind="surface", host_blocks=True)
    blob = [np.ascontiguousarray(u, dtype=np.float64).reshape(-1) for u in h2.U]
    blob += [h2.B_data, h2.D_data]
    os.makedirs(OUT, exist_ok=True)
    np.save(os.path.join(OUT, NAME + "_blocks.npy"), np.concatenate(blob))
    sys.path.insert(0, REF_EXAMPLES)
    from meta_txt_to_json import metadata_txt_to_json
    with tempfile.TemporaryDirectory() as d:
        txt = os.path.join(d, "tmp_metadata.txt")
        write_intermediate_txt(h2, txt)
        metadata_txt_to_json(txt, os.path.join(OUT, NAME + "_meta.json"), os.path.join(OUT, NAME + "_aux.json"))
    print("fixtures written to", OUT)


if __name__ == "__main__":
    main()
