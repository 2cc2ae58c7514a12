"""Host-side H2 construction that produces the H2Pack data layout.

The MI355X engine accelerates ``H2P_matvec`` / ``H2P_matmul`` on an
``H2Pack_p`` built on the host.  This module builds that representation with
the same algorithms the reference uses, written here in numpy/scipy:

* ``calc_enclosing_box``      -- src/H2Pack_gen_proxy_point.c:361-415
* ``partition_points``        -- src/H2Pack_partition.c:38-271 (octree
  bisection, post-order node numbering), :278-317 (tree arrays),
  :327-432 (reduced (in)admissible pairs), :544-714 (driver)
* ``generate_proxy_point_ID`` -- src/H2Pack_gen_proxy_point.c:56-356, 451-616
  (numerical proxy points, multi-layer algorithm 2)
* ``generate_proxy_point_surface`` -- src/H2Pack_gen_proxy_point.c:620-742
* ``build``                   -- src/H2Pack_build.c:32-270 (U/J by proxy ID),
  :652-782 (B metadata), :789-889 (B AOT), :896-1027 (D metadata),
  :1034-1140 (D AOT), :1143-1209 (driver + permutation indices)
* ``id_compress``             -- src/H2Pack_ID_compress.c:491-576 (row ID via
  column-pivoted QR of A^T with a relative-norm stop, skeleton sorted)

Arrays keep the reference's meaning and layout (row-major U with identity
rows on the skeleton, post-order nodes, ``level_nodes`` with stride
``n_leaf_node``, B/D blocks row-major with ``*_ptr`` offsets, the pair-to-block
CSR with negative values meaning "transposed").  Random streams differ from
the reference's drand48/rand, so the proxy points are not bit-identical to
what the reference would draw; the representation is an H2 matrix of the same
structure and accuracy.
"""
from __future__ import annotations

import math
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import scipy.linalg as sla
import scipy.sparse as sps

ALPHA_H2 = 0.999999          # src/H2Pack_config.h:103
ALPHA_HSS = -0.000001        # src/H2Pack_config.h:104
# the 27 lattice offsets of the periodic mode, in the order of
# per_lattices_3d (src/H2Pack_typedef.c:151-179)
PER_LATTICES_3D = np.array([[i, j, k] for i in (-1, 0, 1) for j in (-1, 0, 1) for k in (-1, 0, 1)],
                           dtype=np.float64)
QR_RANK, QR_REL_NRM, QR_ABS_NRM = 0, 1, 2   # src/H2Pack_config.h:97-101

KRNL_COULOMB, KRNL_GAUSSIAN, KRNL_MATERN32, KRNL_RPY = 0, 1, 2, 3
KERNEL_NAMES = {KRNL_COULOMB: "coulomb", KRNL_GAUSSIAN: "gaussian",
                KRNL_MATERN32: "matern32", KRNL_RPY: "rpy"}
# Effective flops per pair the reference charges a bi-matvec with
# (src/H2Pack_3D_kernels.h:50, 167, 384, 893).
KRNL_BIMV_FLOPS = {KRNL_COULOMB: 14, KRNL_GAUSSIAN: 14, KRNL_MATERN32: 17,
                   KRNL_RPY: 82}
KRNL_DIM = {KRNL_COULOMB: 1, KRNL_GAUSSIAN: 1, KRNL_MATERN32: 1, KRNL_RPY: 3}


def _nthreads() -> int:
    env = os.environ.get("H2P_BUILD_THREADS")
    if env:
        return max(1, int(env))
    return max(1, min(16, os.cpu_count() or 1))


# --------------------------------------------------------------------------
# Kernel evaluation on the host (construction only).  Formulas follow
# src/H2Pack_3D_kernels.h: Coulomb :52-96, Gaussian :169-209,
# Matern32 :388-432, RPY :893-959.
# --------------------------------------------------------------------------
def kernel_eval(kid: int, param, c0: np.ndarray, c1: np.ndarray,
                n_thread: int | None = None) -> np.ndarray:
    """Kernel block K(c0, c1); c0 is (xpt_dim, n0), c1 is (xpt_dim, n1).

    Returns a row-major (n0*kd, n1*kd) float64 array, laid out as the
    reference's ``krnl_eval`` writes it (kd x kd sub-blocks per point pair).
    Evaluated by ``h2p_host_kernel_eval`` (csrc/h2p_host.c).
    """
    from ._lib import host_lib
    kd = KRNL_DIM[kid]
    c0 = np.ascontiguousarray(c0, dtype=np.float64)
    c1 = np.ascontiguousarray(c1, dtype=np.float64)
    n0, n1 = c0.shape[1], c1.shape[1]
    out = np.empty((n0 * kd, n1 * kd))
    if out.size == 0:
        return out
    has_param = param is not None and len(param) > 0
    p = np.ascontiguousarray(param if has_param else [0.0], dtype=np.float64)
    pp = p.ctypes.data if has_param else None
    host_lib().h2p_host_kernel_eval(
        kid, pp, c0.ctypes.data, n0, n0, c1.ctypes.data, n1, n1,
        out.ctypes.data, n1 * kd, _nthreads() if n_thread is None else n_thread)
    return out


# --------------------------------------------------------------------------
# Interpolative decomposition
# --------------------------------------------------------------------------
def _qr_rank(R: np.ndarray, n_rows_t: int, stop_type: int, stop_param, kd: int) -> int:
    """Rank chosen by the partial-pivot QR stop rule.

    src/H2Pack_ID_compress.c:83-100 (stop_norm = max(sqrt(nrow)*eps, tol)
    scaled by the largest column norm for the relative rule) and :447-466
    (QR_REL_NRM multiplies the tolerance by sqrt(kdim)).  With column
    pivoting, |R_ii| is the norm of the pivot column at step i, which is the
    quantity the reference compares against stop_norm.
    """
    d = np.abs(np.diag(R))
    if d.size == 0:
        return 0
    max_iter = d.size
    if stop_type == QR_RANK:
        tol_rank, tol_norm, rel = int(stop_param), 1e-15, True
    elif stop_type == QR_REL_NRM:
        tol_rank, tol_norm, rel = max_iter, float(stop_param) * math.sqrt(kd), True
    else:
        tol_rank, tol_norm, rel = max_iter, float(stop_param), False
    stop_norm = max(math.sqrt(n_rows_t) * 1e-15, tol_norm)
    if rel:
        stop_norm *= d[0]
    below = np.nonzero(d < stop_norm)[0]
    r = int(below[0]) if below.size else max_iter
    return min(r, tol_rank, max_iter)


def id_compress(A: np.ndarray, stop_type: int, stop_param, kd: int = 1,
                need_U: bool = True):
    """Row interpolative decomposition A ~= U @ A[J, :].

    Follows H2P_ID_compress (src/H2Pack_ID_compress.c:491-576): column-pivoted
    QR of A^T, rank by the stop rule, E = R11^{-1} R12, U has identity rows at
    the skeleton rows, skeleton indices sorted ascending.  For kd > 1 the
    reference pivots kd-column blocks (:168-410); here a scalar pivoted QR
    selects rows and every point that owns a selected row contributes all kd
    of its rows, then U is the least-squares interpolation onto those rows.
    Returns (U or None, J) with J in point units.
    """
    m, n = A.shape
    if m == 0 or n == 0:
        return (np.zeros((m, 0)) if need_U else None), np.zeros(0, dtype=np.int64)
    R, P = sla.qr(A.T, mode="r", pivoting=True, check_finite=False)
    r = _qr_rank(R, n, stop_type, stop_param, kd)
    if r == 0:
        return (np.zeros((m, 0)) if need_U else None), np.zeros(0, dtype=np.int64)
    if kd == 1:
        skel = P[:r]
        order = np.argsort(skel, kind="stable")
        J = skel[order]
        if not need_U:
            return None, J.astype(np.int64)
        E = sla.solve_triangular(R[:r, :r], R[:r, r:], lower=False, check_finite=False)
        U = np.zeros((m, r))
        U[skel, np.arange(r)] = 1.0
        U[P[r:], :] = E.T
        U = U[:, order]
        return U, J.astype(np.int64)
    pts = np.unique(P[:r] // kd)
    rows = (pts[:, None] * kd + np.arange(kd)[None, :]).reshape(-1)
    if not need_U:
        return None, pts.astype(np.int64)
    AJ = A[rows, :]
    # U = A AJ^+ ; solve via least squares on AJ^T
    U = sla.lstsq(AJ.T, A.T, check_finite=False, lapack_driver="gelsy")[0].T
    U[rows, :] = np.eye(rows.size)
    return np.ascontiguousarray(U), pts.astype(np.int64)


# --------------------------------------------------------------------------
# H2 structure container
# --------------------------------------------------------------------------
class H2Pack:
    """Container mirroring the fields of ``struct H2Pack``
    (src/H2Pack_typedef.h:88-207) that the matvec/matmul path reads."""

    def __init__(self, pt_dim: int = 3, krnl_dim: int = 1,
                 QR_stop_type: int = QR_REL_NRM, QR_stop_param: float = 1e-6):
        # H2P_init, src/H2Pack_typedef.c:14-119
        self.pt_dim = pt_dim
        self.xpt_dim = pt_dim
        self.krnl_dim = krnl_dim
        self.QR_stop_type = QR_stop_type
        self.QR_stop_tol = float(QR_stop_param)
        self.max_child = 1 << pt_dim
        self.is_RPY = 0
        self.is_HSS = 0
        self.is_RPY_Ewald = 0
        self.BD_JIT = 1
        self.root_enbox = None
        self.B_data = None
        self.D_data = None
        self.krnl_id = None
        self.krnl_param = None
        self.timers = {}

    def run_RPY_Ewald(self):
        # H2P_run_RPY_Ewald, src/H2Pack_typedef.c:182-208: periodic RPY with
        # the radius as extended coordinate and the 27 near lattices
        self.is_RPY_Ewald = 1
        self.xpt_dim = self.pt_dim + 1
        self.per_lattices = PER_LATTICES_3D.copy()

    def run_RPY(self):
        # H2P_run_RPY, src/H2Pack_typedef.c:130-137: extended coordinate with radius
        self.is_RPY = 1
        self.xpt_dim = self.pt_dim + 1

    # ------------------------------------------------------------------
    @property
    def n_B(self):
        return self.n_r_adm_pair

    @property
    def n_D(self):
        return self.n_leaf_node + self.n_r_inadm_pair

    @property
    def leaf_nodes(self):
        return self.height_nodes[: self.height_n_node[0]]

    def level_list(self, lvl: int) -> np.ndarray:
        s = lvl * self.n_leaf_node
        return self.level_nodes[s: s + self.level_n_node[lvl]]


def calc_enclosing_box(pt_dim: int, coord: np.ndarray) -> np.ndarray:
    """src/H2Pack_gen_proxy_point.c:361-415 (no proxy-point file)."""
    center = coord[:pt_dim].mean(axis=1)
    radius = float(np.max(np.abs(coord[:pt_dim] - center[:, None])))
    enbox = np.empty(2 * pt_dim)
    enbox[:pt_dim] = center - radius
    enbox[pt_dim:] = 2.0 * radius
    return enbox


def partition_points(h2: H2Pack, coord: np.ndarray, max_leaf_points: int = 0,
                     max_leaf_size: float = 0.0) -> None:
    """Hierarchical octree partition, src/H2Pack_partition.c:544-714.

    ``coord`` is (xpt_dim, n_point), one coordinate per column as in the
    reference.  Produces the post-order tree arrays and the reduced
    (in)admissible pair lists.
    """
    pt_dim, xpt_dim = h2.pt_dim, h2.xpt_dim
    coord = np.ascontiguousarray(coord, dtype=np.float64)
    assert coord.shape[0] == xpt_dim
    n_point = coord.shape[1]
    if max_leaf_points <= 0:
        max_leaf_points = 200 if pt_dim == 2 else 400
    h2.n_point = n_point
    h2.max_leaf_points = max_leaf_points
    h2.max_leaf_size = max_leaf_size
    h2.coord0 = coord.copy()
    sc = coord.copy()
    idx = np.arange(n_point, dtype=np.int64)
    max_child = 1 << pt_dim
    if h2.root_enbox is None:
        h2.root_enbox = calc_enclosing_box(pt_dim, coord)

    # node records (filled in post order)
    rec_pt = []      # (s, e)
    rec_level = []
    rec_height = []
    rec_children = []
    rec_enbox = []
    pow2 = (1 << np.arange(pt_dim)).astype(np.int64)

    def bisect(level, s, e, enbox):
        # src/H2Pack_partition.c:38-271
        npts = e - s + 1
        if enbox is None:   # only for a NULL root box; not used by our drivers
            enbox = calc_enclosing_box(pt_dim, sc[:, s:e + 1])
        box_size = enbox[pt_dim]
        if npts <= max_leaf_points or box_size <= max_leaf_size:
            rec_pt.append((s, e)); rec_level.append(level); rec_height.append(0)
            rec_children.append([]); rec_enbox.append(enbox.copy())
            return len(rec_pt) - 1, 0
        rel = (2.0 * (sc[:pt_dim, s:e + 1] - enbox[:pt_dim, None])
               >= enbox[pt_dim:, None]).astype(np.int64)
        child = (rel * pow2[:, None]).sum(axis=0)
        order = np.argsort(child, kind="stable")
        sc[:, s:e + 1] = sc[:, s:e + 1][:, order]
        idx[s:e + 1] = idx[s:e + 1][order]
        counts = np.bincount(child, minlength=max_child)
        displs = np.concatenate(([0], np.cumsum(counts)))
        kids, max_h = [], 0
        for c in range(max_child):
            if counts[c] == 0:
                continue
            bits = (c >> np.arange(pt_dim)) & 1
            sub = np.empty(2 * pt_dim)
            sub[:pt_dim] = enbox[:pt_dim] + 0.5 * enbox[pt_dim:] * bits - 1e-12
            sub[pt_dim:] = 0.5 * enbox[pt_dim:] + 2e-12
            ci, ch = bisect(level + 1, s + displs[c], s + displs[c + 1] - 1, sub)
            kids.append(ci)
            max_h = max(max_h, ch)
        rec_pt.append((s, e)); rec_level.append(level); rec_height.append(max_h + 1)
        rec_children.append(kids); rec_enbox.append(enbox.copy())
        return len(rec_pt) - 1, max_h + 1

    bisect(0, 0, n_point - 1, np.asarray(h2.root_enbox, dtype=np.float64))
    n_node = len(rec_pt)
    h2.coord = sc
    h2.coord_idx = idx
    h2.n_node = n_node
    h2.root_idx = n_node - 1
    h2.node_level = np.asarray(rec_level, dtype=np.int32)
    h2.node_height = np.asarray(rec_height, dtype=np.int32)
    h2.max_level = int(h2.node_level.max())
    h2.pt_cluster = np.asarray(rec_pt, dtype=np.int32).reshape(-1)
    h2.n_child = np.asarray([len(c) for c in rec_children], dtype=np.int32)
    h2.children = np.full((n_node, max_child), -1, dtype=np.int32)
    h2.parent = np.full(n_node, -1, dtype=np.int32)
    for i, kids in enumerate(rec_children):
        h2.children[i, :len(kids)] = kids
        h2.parent[kids] = i
    h2.children = h2.children.reshape(-1)
    h2.enbox = np.asarray(rec_enbox).reshape(-1)
    h2.n_leaf_node = int((h2.n_child == 0).sum())
    nl = h2.max_level + 1
    h2.level_n_node = np.zeros(nl, dtype=np.int32)
    h2.level_nodes = np.zeros(nl * h2.n_leaf_node, dtype=np.int32)
    h2.height_n_node = np.zeros(nl, dtype=np.int32)
    h2.height_nodes = np.zeros(nl * h2.n_leaf_node, dtype=np.int32)
    for node in range(n_node):   # post order, as H2P_tree_to_array visits
        lv, ht = h2.node_level[node], h2.node_height[node]
        h2.level_nodes[lv * h2.n_leaf_node + h2.level_n_node[lv]] = node
        h2.height_nodes[ht * h2.n_leaf_node + h2.height_n_node[ht]] = node
        h2.level_n_node[lv] += 1
        h2.height_n_node[ht] += 1
    kd = h2.krnl_dim
    pc = h2.pt_cluster.reshape(-1, 2)
    mc = np.empty_like(pc)
    mc[:, 0] = kd * pc[:, 0]
    mc[:, 1] = kd * (pc[:, 1] + 1) - 1
    h2.mat_cluster = mc.reshape(-1).astype(np.int32)
    h2.krnl_mat_size = kd * n_point
    if getattr(h2, "is_RPY_Ewald", 0):
        _calc_reduced_adm_pairs_per(h2)
    else:
        _calc_reduced_adm_pairs(h2)


def _box_admissible(b0: np.ndarray, b1: np.ndarray, pt_dim: int, alpha: float) -> np.ndarray:
    """H2P_check_box_admissible (src/H2Pack_partition.c) on rows of boxes:
    admissible when along some dimension the centre distance is at least
    alpha * min(size) + (size0 + size1) / 2."""
    r0, r1 = b0[..., pt_dim:], b1[..., pt_dim:]
    c0 = b0[..., :pt_dim] + 0.5 * r0
    c1 = b1[..., :pt_dim] + 0.5 * r1
    return np.any(np.abs(c0 - c1) >= alpha * np.minimum(r0, r1) + 0.5 * (r0 + r1), axis=-1)


def _calc_reduced_adm_pairs(h2: H2Pack, alpha: float = ALPHA_H2) -> None:
    """Reduced (in)admissible pairs, src/H2Pack_partition.c:327-432.

    Same recursion rules, evaluated breadth-first with numpy.  The pair sets
    equal the reference's; their order is breadth-first instead of
    depth-first.  ``alpha`` is ALPHA_H2, or ALPHA_HSS for the HSS pairs.
    """
    pt_dim = h2.pt_dim
    ch = h2.children.reshape(h2.n_node, h2.max_child)
    nchild = h2.n_child
    box = h2.enbox.reshape(h2.n_node, 2 * pt_dim)
    level = h2.node_level
    adm0, adm1, inadm0, inadm1 = [], [], [], []
    selfs = np.array([h2.root_idx], dtype=np.int64)
    p0 = np.zeros(0, dtype=np.int64)
    p1 = np.zeros(0, dtype=np.int64)
    while selfs.size or p0.size:
        new0, new1 = [], []
        # self interaction of non-leaf nodes: children self pairs + child pairs i<j
        nl = selfs[nchild[selfs] > 0]
        new_selfs = []
        for n in nl:
            kids = ch[n, :nchild[n]]
            new_selfs.append(kids)
            ii, jj = np.triu_indices(kids.size, k=1)
            new0.append(kids[ii]); new1.append(kids[jj])
        # distinct pairs
        if p0.size:
            adm = _box_admissible(box[p0], box[p1], pt_dim, alpha)
            adm0.append(p0[adm]); adm1.append(p1[adm])
            q0, q1 = p0[~adm], p1[~adm]
            l0 = nchild[q0] == 0
            l1 = nchild[q1] == 0
            both = l0 & l1
            inadm0.append(q0[both]); inadm1.append(q1[both])
            # n0 leaf, n1 non-leaf
            m = l0 & ~l1
            a, b = q0[m], q1[m]
            kk = ch[b]
            aa = np.repeat(a[:, None], ch.shape[1], axis=1)
            v = kk >= 0
            new0.append(aa[v]); new1.append(kk[v])
            # n0 non-leaf, n1 leaf
            m = ~l0 & l1
            a, b = q0[m], q1[m]
            kk = ch[a]
            bb = np.repeat(b[:, None], ch.shape[1], axis=1)
            v = kk >= 0
            new0.append(kk[v]); new1.append(bb[v])
            # both non-leaf
            m = ~l0 & ~l1
            a, b = q0[m], q1[m]
            k0 = np.repeat(ch[a][:, :, None], ch.shape[1], axis=2)
            k1 = np.repeat(ch[b][:, None, :], ch.shape[1], axis=1)
            v = (k0 >= 0) & (k1 >= 0)
            new0.append(k0[v]); new1.append(k1[v])
        selfs = np.concatenate(new_selfs) if new_selfs else np.zeros(0, dtype=np.int64)
        p0 = np.concatenate(new0).astype(np.int64) if new0 else np.zeros(0, dtype=np.int64)
        p1 = np.concatenate(new1).astype(np.int64) if new1 else np.zeros(0, dtype=np.int64)
    cat = lambda L: np.concatenate(L).astype(np.int32) if L else np.zeros(0, dtype=np.int32)
    a0, a1, i0, i1 = cat(adm0), cat(adm1), cat(inadm0), cat(inadm1)
    h2.r_adm_pairs = np.stack([a0, a1], axis=1).reshape(-1) if a0.size else np.zeros(0, np.int32)
    h2.r_inadm_pairs = np.stack([i0, i1], axis=1).reshape(-1) if i0.size else np.zeros(0, np.int32)
    h2.n_r_adm_pair = int(a0.size)
    h2.n_r_inadm_pair = int(i0.size)
    if a0.size:
        h2.min_adm_level = int(np.min(np.maximum(level[a0], level[a1])))
    else:
        h2.min_adm_level = h2.max_level
    h2.r_adm_pairs = h2.r_adm_pairs.astype(np.int32)
    h2.r_inadm_pairs = h2.r_inadm_pairs.astype(np.int32)


def _calc_reduced_adm_pairs_per(h2: H2Pack) -> None:
    """Reduced pairs of the periodic mode, H2P_calc_reduced_adm_pairs_per and
    the lattice loop of H2P_partition_points_periodic
    (src/H2Pack_partition_periodic.c:25-145, 300-340): for every lattice
    offset l, the root is paired with its image shifted by l times the root
    box width, and the pair recursion runs with the second node shifted.
    Pairs are one-sided (node0 receives); an unshifted distinct pair is
    recorded in both orientations.  Depth-first, in the reference's order.
    min_adm_level is 0."""
    pt_dim = h2.pt_dim
    ch = h2.children.reshape(h2.n_node, h2.max_child)
    nchild = h2.n_child
    box = h2.enbox.reshape(h2.n_node, 2 * pt_dim)
    adm, inadm = [], []         # (n0, n1, lattice)

    def rec(n0, n1, shift, lid):
        has_shift = bool(np.any(np.abs(shift) > 1e-10))
        if n0 == n1 and not has_shift:
            if nchild[n0] == 0:
                return
            kids = ch[n0, :nchild[n0]]
            for c in kids:
                rec(c, c, shift, lid)
            for i in range(kids.size):
                for j in range(i + 1, kids.size):
                    rec(kids[i], kids[j], shift, lid)
            return
        b1 = box[n1].copy()
        b1[:pt_dim] += shift
        if _box_admissible(box[n0], b1, pt_dim, ALPHA_H2):
            adm.append((n0, n1, lid))
            if not has_shift:
                adm.append((n1, n0, lid))
            return
        l0, l1 = nchild[n0] == 0, nchild[n1] == 0
        if l0 and l1:
            inadm.append((n0, n1, lid))
            if not has_shift:
                inadm.append((n1, n0, lid))
            return
        if l0:
            for c in ch[n1, :nchild[n1]]:
                rec(n0, c, shift, lid)
            return
        if l1:
            for c in ch[n0, :nchild[n0]]:
                rec(c, n1, shift, lid)
            return
        for a in ch[n0, :nchild[n0]]:
            for b in ch[n1, :nchild[n1]]:
                rec(a, b, shift, lid)

    width = box[h2.root_idx, pt_dim:]
    shifts = h2.per_lattices * width[None, :]
    for lid in range(shifts.shape[0]):
        rec(h2.root_idx, h2.root_idx, shifts[lid], lid)
    A = np.asarray(adm, dtype=np.int64).reshape(-1, 3)
    I = np.asarray(inadm, dtype=np.int64).reshape(-1, 3)
    h2.r_adm_pairs = A[:, :2].reshape(-1).astype(np.int32)
    h2.r_inadm_pairs = I[:, :2].reshape(-1).astype(np.int32)
    h2.per_adm_shifts = np.ascontiguousarray(shifts[A[:, 2]]) if A.size else np.zeros((0, pt_dim))
    h2.per_inadm_shifts = np.ascontiguousarray(shifts[I[:, 2]]) if I.size else np.zeros((0, pt_dim))
    h2.n_r_adm_pair = int(A.shape[0])
    h2.n_r_inadm_pair = int(I.shape[0])
    h2.min_adm_level = 0


def rpy_ewald_eval(c0: np.ndarray, c1: np.ndarray, L: float, xi: float, nr: int, nk: int) -> np.ndarray:
    """Periodic RPY kernel block by Ewald summation (``h2p_host_rpy_ewald``:
    Beenakker's sum with unequal-radius overlap corrections, the role of
    RPY_Ewald_eval_std, src/H2Pack_3D_kernels.h:961-1210); c0, c1 are (4, n)
    with the radius in row 3; real-space images |n_i| <= nr, reciprocal
    vectors |h_i| <= nk.  No 1/(6 pi eta) factor."""
    from ._lib import host_lib
    c0 = np.ascontiguousarray(c0, dtype=np.float64)
    c1 = np.ascontiguousarray(c1, dtype=np.float64)
    n0, n1 = c0.shape[1], c1.shape[1]
    out = np.empty((3 * n0, 3 * n1))
    rc = host_lib().h2p_host_rpy_ewald(c0.ctypes.data, n0, n0, c1.ctypes.data, n1, n1, float(L), float(xi),
                                       int(nr), int(nk), out.ctypes.data, 3 * n1, _nthreads())
    if rc != 0:
        raise ValueError("bad Ewald parameters (L > 0, xi > 0, nr >= 0, nk >= 0)")
    return out


# --------------------------------------------------------------------------
# Proxy points
# --------------------------------------------------------------------------
def _coord_in_ring(rng, npt, pt_dim, L0, L1):
    """Uniform points in [-L1/2,L1/2]^d minus [-L0/2,L0/2]^d
    (src/H2Pack_utils.c:115-133, rejection sampling)."""
    out = np.empty((pt_dim, 0))
    while out.shape[1] < npt:
        m = max(2 * (npt - out.shape[1]), 64)
        c = rng.random((pt_dim, m)) * L1 - 0.5 * L1
        inside_inner = np.all(np.abs(c) <= 0.5 * L0, axis=0)
        out = np.concatenate([out, c[:, ~inside_inner]], axis=1)
    return out[:, :npt]


def _pad_ext(c, xpt_dim, ext_val):
    if c.shape[0] >= xpt_dim:
        return c
    e = np.full((xpt_dim - c.shape[0], c.shape[1]), ext_val)
    return np.concatenate([c, e], axis=0)


def _sketch(kid, param, Y, X, rows, vals, nnz):
    """(S^T K(Y, X))^T for a scalar kernel, S^T row c = (rows, vals)[c*nnz:(c+1)*nnz]
    (``h2p_host_sketch`` in csrc/h2p_host.c)."""
    from ._lib import host_lib
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    X = np.ascontiguousarray(X, dtype=np.float64)
    r = np.ascontiguousarray(rows, dtype=np.int32)
    v = np.ascontiguousarray(vals, dtype=np.float64)
    nx = X.shape[1]
    out = np.empty((nx, nx))
    has_param = param is not None and len(param) > 0
    p = np.ascontiguousarray(param if has_param else [0.0], dtype=np.float64)
    host_lib().h2p_host_sketch(kid, p.ctypes.data if has_param else None, Y.ctypes.data, Y.shape[1],
                               X.ctypes.data, nx, nx, r.ctypes.data, v.ctypes.data, int(nnz),
                               out.ctypes.data, _nthreads())
    return out


def _gen_pp_nlayer(kid, param, pt_dim, xpt_dim, kd, reltol, L1, L2, L3, rng,
                   X0_size=2000, Y0_lsize=4000, max_layer=8, ext_val=0.0):
    """src/H2Pack_gen_proxy_point.c:56-356 with alg = 2 (layer by layer)."""
    n_layer = min(int(round((L3 - L2) / L1)), max_layer)
    X0 = _pad_ext(_coord_in_ring(rng, X0_size, pt_dim, 0.0, L1), xpt_dim, ext_val)
    w = (L3 - L2) / n_layer
    layers = [_pad_ext(_coord_in_ring(rng, Y0_lsize, pt_dim, L2 + w * i, L2 + w * (i + 1)),
                       xpt_dim, ext_val) for i in range(n_layer)]
    Y0 = np.concatenate(layers, axis=1)
    # (1)-(2) kernel matrix and sparse random sketch with 32 +-1 per column.
    # The kernels are symmetric, so K(Y0, X0) is A^T already in row-major
    # order; the sketch is A1 = A S = (S^T A^T)^T.
    k, n = Y0.shape[1] * kd, X0.shape[1] * kd
    nnz_col = min(32, k)
    rows = np.concatenate([rng.choice(k, nnz_col, replace=False) for _ in range(n)])
    cols = np.repeat(np.arange(n), nnz_col)
    vals = rng.integers(0, 2, size=rows.size) * 2.0 - 1.0
    if kd == 1:
        # fused: the sketch straight from kernel values, K(Y0, X0) never formed
        A1 = _sketch(kid, param, Y0, X0, rows, vals, nnz_col)
    else:
        At = kernel_eval(kid, param, Y0, X0)
        St = sps.csr_matrix((vals, (cols, rows)), shape=(n, k))
        A1 = np.asarray(St @ At).T
        del At
    # (3) skeleton of X
    _, skel = id_compress(A1, QR_REL_NRM, reltol * 1e-2, kd, need_U=False)
    Xp = X0[:, skel]
    # (4) proxy points in Y, layer by layer
    Yp = np.zeros((xpt_dim, 0))
    for i in range(n_layer):
        T = np.concatenate([Yp, layers[i]], axis=1)
        A1 = kernel_eval(kid, param, T, Xp)
        _, skel = id_compress(A1, QR_REL_NRM, reltol * 1e-2, kd, need_U=False)
        Yp = T[:, skel]
    return Yp


def generate_proxy_point_ID(h2: H2Pack, kid: int, param, seed: int = 1112,
                            ext_val: float = 0.0):
    """Numerical proxy points per level, src/H2Pack_gen_proxy_point.c:451-616.

    Returns a list indexed by level; levels without admissible pairs get an
    empty set.
    """
    pt_dim, kd = h2.pt_dim, h2.krnl_dim
    reltol = h2.QR_stop_tol
    n_level = h2.max_level + 1
    pt_maxL = h2.root_enbox[pt_dim] * 0.5
    pt_minL = pt_maxL * 0.5 ** h2.max_level
    pt_maxL *= 0.25
    # a zero-size root box (all points coincide, e.g. N = 1) has no
    # admissible levels and needs no proxy points
    num_pp = int(round(math.log2(pt_maxL / pt_minL))) + 1 if pt_maxL >= pt_minL > 0.0 else 0
    rng = np.random.default_rng(seed)
    pp = [np.zeros((h2.xpt_dim, 0)) for _ in range(n_level)]
    for i in range(num_pp):
        L1 = 2.0 * pt_minL * 2.0 ** i
        L2 = (1.0 + 2.0 * ALPHA_H2) * L1
        L3 = (1.0 + 8.0 * ALPHA_H2) * L1
        lvl = (n_level - 1) - i
        if lvl < h2.min_adm_level:
            continue
        pp[lvl] = _gen_pp_nlayer(kid, param, pt_dim, h2.xpt_dim, kd, reltol,
                                 L1, L2, L3, rng, ext_val=ext_val)
    return pp


def generate_proxy_point_surface(h2: H2Pack, min_npt: int, ext_val: float = 0.0):
    """Proxy points on box surfaces, src/H2Pack_gen_proxy_point.c:620-742."""
    pt_dim = h2.pt_dim
    assert pt_dim == 3
    npt_axis = int(math.ceil(math.sqrt(min_npt / 6.0)))
    h = 2.0 / (npt_axis + 1)
    hs = -1.0 + h * np.arange(1, npt_axis + 1)
    hi, hj = np.meshgrid(hs, hs, indexing="ij")
    hi, hj = hi.reshape(-1), hj.reshape(-1)
    m1, p1 = -np.ones_like(hi), np.ones_like(hi)
    faces = [(hi, hj, m1), (hi, hj, p1), (hi, m1, hj), (hi, p1, hj), (m1, hi, hj), (p1, hi, hj)]
    # interleave faces per (i,j) as the reference's index += 6 loop does
    unit = np.stack([np.stack(f, axis=0) for f in faces], axis=2).reshape(3, -1)
    max_L = h2.root_enbox[pt_dim]
    pp = [np.zeros((h2.xpt_dim, 0)) for _ in range(h2.max_level + 1)]
    for level in range(h2.min_adm_level, h2.max_level + 1):
        box_width = max_L / 2.0 ** level * 0.5
        adm_width = (1.0 + 2.0 * ALPHA_H2) * box_width
        pp[level] = _pad_ext(adm_width * unit, h2.xpt_dim, ext_val)
    return pp


def surface_pp_count(reltol: float, pt_dim: int = 3) -> int:
    """Heuristic proxy count of examples/example_H2.c:95-101."""
    d = int(math.ceil(-math.log10(reltol)))
    d = min(max(d, 4), 10)
    return 2 * pt_dim * d * d if pt_dim == 3 else 2 * pt_dim * d


# --------------------------------------------------------------------------
# Build
# --------------------------------------------------------------------------
def _build_UJ(h2: H2Pack, pp) -> None:
    """U and skeleton sets by proxy-point ID, src/H2Pack_build.c:32-270.

    Nodes at levels >= min_adm_level, bottom-up; other nodes get empty U.
    """
    kid, param, kd = h2.krnl_id, h2.krnl_param, h2.krnl_dim
    n_node, pt_dim = h2.n_node, h2.pt_dim
    ch = h2.children.reshape(n_node, h2.max_child)
    box = h2.enbox.reshape(n_node, 2 * pt_dim)
    pc = h2.pt_cluster.reshape(-1, 2)
    U = [np.zeros((0, 0)) for _ in range(n_node)]
    J = [np.zeros(0, dtype=np.int64) for _ in range(n_node)]
    by_height = {}
    for node in range(n_node):
        if h2.node_level[node] >= h2.min_adm_level:
            by_height.setdefault(int(h2.node_height[node]), []).append(node)

    def one(node):
        if h2.n_child[node] == 0:
            cand = np.arange(pc[node, 0], pc[node, 1] + 1, dtype=np.int64)
        else:
            cand = np.concatenate([J[c] for c in ch[node, :h2.n_child[node]]])
        pts = h2.coord[:, cand].copy()
        center = box[node, :pt_dim] + 0.5 * box[node, pt_dim:]
        pts[:pt_dim] -= center[:, None]
        ppl = pp[h2.node_level[node]]
        if ppl.shape[1] == 0 or cand.size == 0:
            return node, np.zeros((cand.size * kd, 0)), np.zeros(0, dtype=np.int64)
        A = kernel_eval(kid, param, pts, ppl)
        Un, sub = id_compress(A, h2.QR_stop_type, h2.QR_stop_tol, kd)
        return node, Un, cand[sub]

    try:
        from threadpoolctl import threadpool_limits
    except ImportError:   # pragma: no cover
        threadpool_limits = None
    nt = _nthreads()
    ctx = threadpool_limits(limits=1) if threadpool_limits else None
    try:
        if ctx is not None:
            ctx.__enter__()
        with ThreadPoolExecutor(max_workers=nt) as ex:
            for ht in sorted(by_height):
                for node, Un, Jn in ex.map(one, by_height[ht]):
                    U[node] = np.ascontiguousarray(Un)
                    J[node] = Jn
    finally:
        if ctx is not None:
            ctx.__exit__(None, None, None)
    h2.U = U
    h2.J = J
    h2.J_coord = [np.ascontiguousarray(h2.coord[:, j]) for j in J]
    h2.U_nrow = np.asarray([u.shape[0] for u in U], dtype=np.int32)
    h2.U_ncol = np.asarray([u.shape[1] for u in U], dtype=np.int32)
    # U of nodes above min_adm_level stays 0 x 0, as the reference marks them
    for node in range(n_node):
        if h2.node_level[node] < h2.min_adm_level:
            h2.U_nrow[node] = 0
            h2.U_ncol[node] = 0


# --------------------------------------------------------------------------
# HSS mode (H2P_run_HSS): weak admissibility and hybrid bases
# --------------------------------------------------------------------------
def _calc_hss_pairs(h2: H2Pack) -> None:
    """H2P_HSS_calc_adm_inadm_pairs (src/H2Pack_partition.c:507-540): reduced
    pairs with ALPHA_HSS replace the pair lists, whose minimum admissible
    level must be 1.  The H2 pairs are kept as h2.H2_* for reference."""
    h2.H2_r_adm_pairs, h2.H2_r_inadm_pairs = h2.r_adm_pairs, h2.r_inadm_pairs
    h2.H2_min_adm_level = h2.min_adm_level
    _calc_reduced_adm_pairs(h2, ALPHA_HSS)
    if h2.min_adm_level != 1:
        raise ValueError("HSS minimal admissible level should be 1")


def _calc_node_inadm_lists(h2: H2Pack) -> list:
    """H2P_calc_node_inadm_lists (src/H2Pack_partition.c:434-505): for each
    node at a level >= 1, the nodes of its level and the leaves above it that
    are not H2-admissible but HSS-admissible (its near neighbours)."""
    pt_dim, nl = h2.pt_dim, h2.n_leaf_node
    box = h2.enbox.reshape(h2.n_node, 2 * pt_dim)
    leaves = h2.leaf_nodes
    lists = [np.zeros(0, np.int64) for _ in range(h2.n_node)]
    for lv in range(h2.max_level, h2.min_adm_level - 1, -1):
        nodes = h2.level_nodes[lv * nl: lv * nl + h2.level_n_node[lv]].astype(np.int64)
        targets = np.concatenate([nodes, leaves[h2.node_level[leaves] < lv].astype(np.int64)])
        bn, bt = box[nodes][:, None, :], box[targets][None, :, :]
        near = (~_box_admissible(bn, bt, pt_dim, ALPHA_H2) & _box_admissible(bn, bt, pt_dim, ALPHA_HSS)
                & (nodes[:, None] != targets[None, :]))
        for k, n in enumerate(nodes):
            lists[n] = targets[near[k]]
    return lists


def _build_UJ_hss(h2: H2Pack, pp, seed: int = 2718) -> None:
    """H2P_build_HSS_UJ_hybrid (src/H2Pack_build.c:272-600): bottom-up from
    the leaves, each node compresses K(its candidate points, [candidate points
    of its near neighbours, its level's proxy points]) (both centred on the
    node's box), sketched by a Gaussian random matrix when it has more than
    twice as many columns as rows, by ID.  Candidate points: a leaf's points,
    or the union of its children's skeletons.  A neighbour contributes its
    candidate points, fixed before any node of the level is compressed (the
    reference's parallel loop may read some neighbours already compressed;
    this order is deterministic)."""
    kid, param, kd = h2.krnl_id, h2.krnl_param, h2.krnl_dim
    nn, pt_dim, nl = h2.n_node, h2.pt_dim, h2.n_leaf_node
    ch = h2.children.reshape(nn, h2.max_child)
    box = h2.enbox.reshape(nn, 2 * pt_dim)
    pc = h2.pt_cluster.reshape(-1, 2)
    lists = _calc_node_inadm_lists(h2)
    h2.node_inadm_lists = lists
    h2.node_n_r_inadm = np.asarray([l.size for l in lists], dtype=np.int32)
    U = [np.zeros((0, 0)) for _ in range(nn)]
    J = [np.zeros(0, dtype=np.int64) for _ in range(nn)]
    cand = [None] * nn
    for n in h2.leaf_nodes:
        cand[n] = np.arange(pc[n, 0], pc[n, 1] + 1, dtype=np.int64)
    rng = np.random.default_rng(seed)
    for lv in range(h2.max_level, h2.min_adm_level - 1, -1):
        nodes = h2.level_nodes[lv * nl: lv * nl + h2.level_n_node[lv]]
        for n in nodes:
            if h2.n_child[n]:
                cand[n] = np.concatenate([J[c] for c in ch[n, :h2.n_child[n]]])
        for n in nodes:
            center = box[n, :pt_dim] + 0.5 * box[n, pt_dim:]
            pts = h2.coord[:, cand[n]].copy()
            pts[:pt_dim] -= center[:, None]
            nb = [cand[k] for k in lists[n]]
            nb = np.concatenate(nb) if nb else np.zeros(0, dtype=np.int64)
            far = h2.coord[:, nb].copy()
            far[:pt_dim] -= center[:, None]
            ppl = pp[lv] if lv < len(pp) else np.zeros((h2.xpt_dim, 0))
            far = np.concatenate([far, ppl], axis=1)
            if far.shape[1] == 0 or cand[n].size == 0:
                U[n] = np.zeros((cand[n].size * kd, 0))
                J[n] = np.zeros(0, dtype=np.int64)
                continue
            A = kernel_eval(kid, param, pts, far)
            if A.shape[1] > 2 * A.shape[0]:
                A = A @ rng.standard_normal((A.shape[1], A.shape[0]))
            Un, sub = id_compress(A, h2.QR_stop_type, h2.QR_stop_tol, kd)
            U[n] = np.ascontiguousarray(Un)
            J[n] = cand[n][sub]
    h2.U = U
    h2.J = J
    h2.J_coord = [np.ascontiguousarray(h2.coord[:, j]) for j in J]
    h2.U_nrow = np.asarray([u.shape[0] for u in U], dtype=np.int32)
    h2.U_ncol = np.asarray([u.shape[1] for u in U], dtype=np.int32)


def _coo_to_csr(n_node, pi, pj, pv):
    """H2P_int_COO_to_CSR (src/H2Pack_utils.c:265-290): columns sorted per row."""
    order = np.lexsort((pj, pi))
    pi, pj, pv = pi[order], pj[order], pv[order]
    rowptr = np.zeros(n_node + 1, dtype=np.int32)
    np.add.at(rowptr, pi + 1, 1)
    rowptr = np.cumsum(rowptr).astype(np.int32)
    return rowptr, pj.astype(np.int32), pv.astype(np.int32)


def _gen_B_metadata(h2: H2Pack) -> None:
    """src/H2Pack_build.c:652-782."""
    kd = h2.krnl_dim
    pairs = h2.r_adm_pairs.reshape(-1, 2)
    n0, n1 = pairs[:, 0], pairs[:, 1]
    l0, l1 = h2.node_level[n0], h2.node_level[n1]
    pc = h2.pt_cluster.reshape(-1, 2)
    jl = np.asarray([j.size for j in h2.J], dtype=np.int64)
    npt = (pc[:, 1] - pc[:, 0] + 1).astype(np.int64)
    a = np.where(l0 < l1, npt[n0], jl[n0])
    b = np.where(l0 > l1, npt[n1], jl[n1])
    h2.B_nrow = (a * kd).astype(np.int32)
    h2.B_ncol = (b * kd).astype(np.int32)
    sizes = h2.B_nrow.astype(np.int64) * h2.B_ncol.astype(np.int64)
    h2.B_ptr = np.concatenate(([0], np.cumsum(sizes))).astype(np.uint64)
    directed = bool(getattr(h2, "is_RPY_Ewald", 0))   # periodic: node0 receives only
    h2.node_n_r_adm = (np.bincount(n0, minlength=h2.n_node)
                       + (0 if directed else np.bincount(n1, minlength=h2.n_node))).astype(np.int32)
    m = n0.size
    if directed:
        pi, pj, pv = n0, n1, np.arange(1, m + 1)
    else:
        pi = np.concatenate([n0, n1]); pj = np.concatenate([n1, n0])
        pv = np.concatenate([np.arange(1, m + 1), -np.arange(1, m + 1)])
    h2.B_p2i_rowptr, h2.B_p2i_colidx, h2.B_p2i_val = _coo_to_csr(h2.n_node, pi, pj, pv)
    # statistics (src/H2Pack_build.c:755-758)
    h2.B_size = int(sizes.sum())
    h2.JIT_B_pairs = float((a * b).sum())


def _gen_D_metadata(h2: H2Pack) -> None:
    """src/H2Pack_build.c:896-1027."""
    kd = h2.krnl_dim
    pc = h2.pt_cluster.reshape(-1, 2)
    npt = (pc[:, 1] - pc[:, 0] + 1).astype(np.int64)
    leaves = h2.leaf_nodes.astype(np.int64)
    pairs = h2.r_inadm_pairs.reshape(-1, 2).astype(np.int64)
    r = np.concatenate([npt[leaves], npt[pairs[:, 0]]]) * kd
    c = np.concatenate([npt[leaves], npt[pairs[:, 1]]]) * kd
    h2.D_nrow = r.astype(np.int32)
    h2.D_ncol = c.astype(np.int32)
    sizes = r * c
    h2.D_ptr = np.concatenate(([0], np.cumsum(sizes))).astype(np.uint64)
    nl, m = leaves.size, pairs.shape[0]
    if getattr(h2, "is_RPY_Ewald", 0):   # periodic: one-sided pairs
        pi = np.concatenate([leaves, pairs[:, 0]])
        pj = np.concatenate([leaves, pairs[:, 1]])
        pv = np.concatenate([np.arange(1, nl + 1), nl + np.arange(1, m + 1)])
    else:
        pi = np.concatenate([leaves, pairs[:, 0], pairs[:, 1]])
        pj = np.concatenate([leaves, pairs[:, 1], pairs[:, 0]])
        pv = np.concatenate([np.arange(1, nl + 1), nl + np.arange(1, m + 1), -(nl + np.arange(1, m + 1))])
    h2.D_p2i_rowptr, h2.D_p2i_colidx, h2.D_p2i_val = _coo_to_csr(h2.n_node, pi, pj, pv)
    h2.D_size = int(sizes.sum())
    h2.JIT_D_pairs = float((npt[leaves] ** 2).sum() + (npt[pairs[:, 0]] * npt[pairs[:, 1]]).sum())


def B_block_coords(h2: H2Pack, i: int):
    """The two point sets whose kernel block is B_i (src/H2Pack_build.c:818-873)."""
    n0, n1 = h2.r_adm_pairs[2 * i], h2.r_adm_pairs[2 * i + 1]
    l0, l1 = h2.node_level[n0], h2.node_level[n1]
    pc = h2.pt_cluster
    c0 = h2.coord[:, pc[2 * n0]:pc[2 * n0 + 1] + 1] if l0 < l1 else h2.J_coord[n0]
    c1 = h2.coord[:, pc[2 * n1]:pc[2 * n1 + 1] + 1] if l0 > l1 else h2.J_coord[n1]
    return c0, c1


def D_block_coords(h2: H2Pack, i: int):
    """The two point sets whose kernel block is D_i (src/H2Pack_build.c:1067-1120)."""
    pc = h2.pt_cluster
    if i < h2.n_leaf_node:
        n0 = n1 = h2.leaf_nodes[i]
    else:
        k = i - h2.n_leaf_node
        n0, n1 = h2.r_inadm_pairs[2 * k], h2.r_inadm_pairs[2 * k + 1]
    return (h2.coord[:, pc[2 * n0]:pc[2 * n0 + 1] + 1],
            h2.coord[:, pc[2 * n1]:pc[2 * n1 + 1] + 1])


def _build_BD_AOT(h2: H2Pack) -> None:
    """B and D blocks on the host, src/H2Pack_build.c:789-889 and :1034-1140."""
    B = np.empty(int(h2.B_ptr[-1]))
    for i in range(h2.n_B):
        c0, c1 = B_block_coords(h2, i)
        B[int(h2.B_ptr[i]):int(h2.B_ptr[i + 1])] = kernel_eval(h2.krnl_id, h2.krnl_param, c0, c1).reshape(-1)
    D = np.empty(int(h2.D_ptr[-1]))
    for i in range(h2.n_D):
        c0, c1 = D_block_coords(h2, i)
        D[int(h2.D_ptr[i]):int(h2.D_ptr[i + 1])] = kernel_eval(h2.krnl_id, h2.krnl_param, c0, c1).reshape(-1)
    h2.B_data, h2.D_data = B, D


def build(h2: H2Pack, pp, BD_JIT: int, krnl_id: int, krnl_param,
          host_blocks: bool | None = None) -> None:
    """H2P_build, src/H2Pack_build.c:1143-1209.

    ``host_blocks`` controls whether AOT B/D blocks are materialised on the
    host (the reference always does for BD_JIT == 0).  With
    ``host_blocks=False`` an AOT representation carries only metadata and the
    device engine evaluates the blocks into HBM at upload.
    """
    import time
    h2.krnl_id = int(krnl_id)
    h2.krnl_param = np.asarray(krnl_param if krnl_param is not None else [], dtype=np.float64)
    h2.BD_JIT = int(BD_JIT)
    h2.krnl_bimv_flops = KRNL_BIMV_FLOPS[h2.krnl_id]
    t0 = time.perf_counter()
    if getattr(h2, "is_HSS", 0):
        _build_UJ_hss(h2, pp)
    else:
        _build_UJ(h2, pp)
    t1 = time.perf_counter()
    _gen_B_metadata(h2)
    _gen_D_metadata(h2)
    if host_blocks is None:
        host_blocks = (BD_JIT == 0)
    if host_blocks:
        _build_BD_AOT(h2)
    t2 = time.perf_counter()
    kd = h2.krnl_dim
    ci = h2.coord_idx.astype(np.int64)
    fwd = (ci[:, None] * kd + np.arange(kd)[None, :]).reshape(-1)
    bwd = np.empty_like(fwd)
    bwd[fwd] = np.arange(fwd.size)
    h2.fwd_pmt_idx = fwd.astype(np.int32)
    h2.bwd_pmt_idx = bwd.astype(np.int32)
    h2.timers["U_build"] = t1 - t0
    h2.timers["BD_build"] = t2 - t1


# --------------------------------------------------------------------------
# Convenience driver
# --------------------------------------------------------------------------
def uniform_cube_points(n: int, density: float, seed: int, xpt_dim: int = 3,
                        radius: float | None = None) -> np.ndarray:
    """i.i.d. U[0, L)^3 with L = (n / density)^(1/3); (xpt_dim, n)."""
    L = (n / density) ** (1.0 / 3.0)
    rng = np.random.default_rng(seed)
    c = rng.random((3, n)) * L
    if xpt_dim > 3:
        c = np.concatenate([c, np.full((xpt_dim - 3, n), 1.0 if radius is None else radius)], axis=0)
    return c


def build_h2(coord: np.ndarray, krnl_id: int, krnl_param, reltol: float,
             max_leaf_points: int, BD_JIT: int = 1, pp_kind: str = "auto",
             host_blocks: bool | None = None, max_leaf_size: float = 0.0,
             hss: bool = False) -> H2Pack:
    """Partition + proxy points + build in one call (the sequence of
    examples/example_H2.c:45-88; with ``hss`` the HSS sequence of
    examples/example_HSS.c:51-80, H2P_run_HSS before the partition).  An HSS
    representation carries its HSS pair lists and minimum admissible level in
    the ordinary fields, which is what H2P_matvec reads in HSS mode
    (src/H2Pack_matvec.c:101, 308, 592, 773, 904, 1009)."""
    import time
    kd = KRNL_DIM[krnl_id]
    h2 = H2Pack(pt_dim=3, krnl_dim=kd, QR_stop_type=QR_REL_NRM, QR_stop_param=reltol)
    if krnl_id == KRNL_RPY:
        h2.run_RPY()
    t0 = time.perf_counter()
    h2.root_enbox = calc_enclosing_box(3, coord)
    partition_points(h2, coord, max_leaf_points, max_leaf_size)
    t1 = time.perf_counter()
    if pp_kind == "auto":
        pp_kind = "surface" if krnl_id in (KRNL_RPY,) else "numerical"
    ext = float(np.max(coord[3])) if h2.xpt_dim > 3 else 0.0
    if pp_kind == "surface":
        pp = generate_proxy_point_surface(h2, surface_pp_count(reltol), ext_val=ext)
    else:
        pp = generate_proxy_point_ID(h2, krnl_id, krnl_param, ext_val=ext)
    t2 = time.perf_counter()
    h2.pp = pp
    h2.pp_kind = pp_kind
    if hss:
        h2.is_HSS = 1
        _calc_hss_pairs(h2)
    build(h2, pp, BD_JIT, krnl_id, krnl_param, host_blocks=host_blocks)
    h2.timers["partition"] = t1 - t0
    h2.timers["proxy_points"] = t2 - t1
    return h2


def build_periodic(coord: np.ndarray, unit_cell, eta: float, ewald, reltol: float,
                   max_leaf_points: int, max_leaf_size: float = 0.0) -> H2Pack:
    """Periodic RPY representation (H2P_run_RPY_Ewald, H2P_partition_points_
    periodic, surface proxy points, H2P_build_periodic; JIT only, as in the
    reference).  ``coord`` is (4, n) with the radius in row 3; ``unit_cell``
    = (corner x, y, z, side x, y, z), the root box; ``eta`` the RPY kernel
    parameter; ``ewald`` = (L, xi, nr, nk) of the Ewald kernel.  The root's
    periodic block (H2P_build_periodic_block, src/H2Pack_build_periodic.c:
    15-60) is the Ewald kernel on the root skeleton minus the plain kernel
    over the 27 near lattices; it lands in h2.per_blk.  As in the reference
    the Ewald kernel has no 1/(6 pi eta) factor, so the plain and periodic
    parts agree in scale for eta = 1/(6 pi)."""
    h2 = H2Pack(pt_dim=3, krnl_dim=3, QR_stop_type=QR_REL_NRM, QR_stop_param=reltol)
    h2.run_RPY_Ewald()
    h2.root_enbox = np.asarray(unit_cell, dtype=np.float64)
    partition_points(h2, coord, max_leaf_points, max_leaf_size)
    ext = float(np.max(coord[3]))
    pp = generate_proxy_point_surface(h2, surface_pp_count(reltol), ext_val=ext)
    h2.pp, h2.pp_kind = pp, "surface"
    build(h2, pp, 1, KRNL_RPY, [eta], host_blocks=False)
    # periodic block of the root
    root = h2.root_idx
    Jc = h2.J_coord[root]
    L, xi, nr, nk = ewald
    per = rpy_ewald_eval(Jc, Jc, L, xi, int(nr), int(nk))
    width = h2.enbox.reshape(h2.n_node, -1)[root, 3:6]
    for lat in h2.per_lattices:
        s = Jc.copy()
        s[:3] += (lat * width)[:, None]
        per -= kernel_eval(KRNL_RPY, [eta], Jc, s)
    h2.per_blk = per
    h2.ewald_param = (float(L), float(xi), int(nr), int(nk))
    return h2
