"""ORACLE -- test infrastructure only.

ctypes binding of ``oracle/_build/libh2oracle.so`` (the CPU restatement of the
reference's H2P_matvec / H2P_matmul, see h2oracle.h).  Imported only by
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, as the
checker; the product path never reaches it.
"""
from __future__ import annotations

import ctypes as ct
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "_build", "libh2oracle.so")

_IP = ct.POINTER(ct.c_int)
_DP = ct.POINTER(ct.c_double)


class H2OH2Pack(ct.Structure):
    _fields_ = [
        ("pt_dim", ct.c_int), ("xpt_dim", ct.c_int), ("krnl_dim", ct.c_int),
        ("n_point", ct.c_int), ("krnl_mat_size", ct.c_int),
        ("n_node", ct.c_int), ("n_leaf_node", ct.c_int), ("max_child", ct.c_int),
        ("max_level", ct.c_int), ("min_adm_level", ct.c_int),
        ("children", ct.c_void_p), ("n_child", ct.c_void_p), ("node_level", ct.c_void_p),
        ("pt_cluster", ct.c_void_p), ("mat_cluster", ct.c_void_p),
        ("level_n_node", ct.c_void_p), ("level_nodes", ct.c_void_p), ("leaf_nodes", ct.c_void_p),
        ("n_r_adm_pair", ct.c_int), ("n_r_inadm_pair", ct.c_int),
        ("r_adm_pairs", ct.c_void_p), ("r_inadm_pairs", ct.c_void_p), ("node_n_r_adm", ct.c_void_p),
        ("U_nrow", ct.c_void_p), ("U_ncol", ct.c_void_p), ("U", ct.c_void_p),
        ("J_npt", ct.c_void_p), ("J_coord", ct.c_void_p),
        ("coord", ct.c_void_p), ("fwd_pmt_idx", ct.c_void_p), ("bwd_pmt_idx", ct.c_void_p),
        ("B_nrow", ct.c_void_p), ("B_ncol", ct.c_void_p), ("B_ptr", ct.c_void_p), ("B_data", ct.c_void_p),
        ("D_nrow", ct.c_void_p), ("D_ncol", ct.c_void_p), ("D_ptr", ct.c_void_p), ("D_data", ct.c_void_p),
        ("B_p2i_rowptr", ct.c_void_p), ("B_p2i_colidx", ct.c_void_p), ("B_p2i_val", ct.c_void_p),
        ("D_p2i_rowptr", ct.c_void_p), ("D_p2i_colidx", ct.c_void_p), ("D_p2i_val", ct.c_void_p),
        ("krnl_id", ct.c_int), ("krnl_param", ct.c_double * 4), ("n_krnl_param", ct.c_int),
        ("BD_JIT", ct.c_int),
    ]


_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return SO


def lib() -> ct.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(SO):
            build()
        L = ct.CDLL(SO)
        L.h2o_matvec.argtypes = [ct.POINTER(H2OH2Pack), ct.c_void_p, ct.c_void_p, ct.c_int]
        L.h2o_matvec.restype = ct.c_int
        L.h2o_matmul.argtypes = [ct.POINTER(H2OH2Pack), ct.c_int, ct.c_int, ct.c_void_p,
                                 ct.c_int, ct.c_void_p, ct.c_int, ct.c_int]
        L.h2o_matmul.restype = ct.c_int
        L.h2o_kernel_eval.argtypes = [ct.c_int, ct.c_void_p, ct.c_int, ct.c_void_p, ct.c_int,
                                      ct.c_int, ct.c_void_p, ct.c_int, ct.c_int, ct.c_void_p, ct.c_int]
        L.h2o_kernel_eval.restype = None
        L.h2o_direct_nbody.argtypes = [ct.c_int, ct.c_void_p, ct.c_int, ct.c_int, ct.c_void_p,
                                       ct.c_int, ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_int,
                                       ct.c_int, ct.c_void_p, ct.c_int]
        L.h2o_direct_nbody.restype = None
        _lib = L
    return _lib


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class OracleH2:
    """Keeps the arrays alive and exposes the oracle's matvec/matmul on an
    H2 representation produced by ``h2pack_amd.construct``."""

    def __init__(self, h2, BD_JIT: int | None = None):
        self.h2 = h2
        jit = h2.BD_JIT if BD_JIT is None else int(BD_JIT)
        if not jit and (h2.B_data is None or h2.D_data is None):
            raise ValueError("AOT oracle needs host B/D blocks (build with host_blocks=True)")
        keep = []

        def arr(a, dt):
            a = np.ascontiguousarray(a, dtype=dt)
            keep.append(a)
            return a.ctypes.data

        s = H2OH2Pack()
        s.pt_dim, s.xpt_dim, s.krnl_dim = h2.pt_dim, h2.xpt_dim, h2.krnl_dim
        s.n_point, s.krnl_mat_size = h2.n_point, h2.krnl_mat_size
        s.n_node, s.n_leaf_node, s.max_child = h2.n_node, h2.n_leaf_node, h2.max_child
        s.max_level, s.min_adm_level = h2.max_level, h2.min_adm_level
        for f in ("children", "n_child", "node_level", "pt_cluster", "mat_cluster",
                  "level_n_node", "level_nodes", "r_adm_pairs", "r_inadm_pairs", "node_n_r_adm",
                  "U_nrow", "U_ncol", "fwd_pmt_idx", "bwd_pmt_idx", "B_nrow", "B_ncol",
                  "D_nrow", "D_ncol", "B_p2i_rowptr", "B_p2i_colidx", "B_p2i_val",
                  "D_p2i_rowptr", "D_p2i_colidx", "D_p2i_val"):
            setattr(s, f, arr(getattr(h2, f), np.int32))
        s.leaf_nodes = arr(h2.leaf_nodes, np.int32)
        s.n_r_adm_pair, s.n_r_inadm_pair = h2.n_r_adm_pair, h2.n_r_inadm_pair
        Us = [np.ascontiguousarray(u, dtype=np.float64) for u in h2.U]
        keep.extend(Us)
        uptr = (ct.c_void_p * h2.n_node)(*[u.ctypes.data if u.size else None for u in Us])
        keep.append(uptr)
        s.U = ct.cast(uptr, ct.c_void_p)
        jc = [np.ascontiguousarray(j, dtype=np.float64) for j in h2.J_coord]
        keep.extend(jc)
        jptr = (ct.c_void_p * h2.n_node)(*[j.ctypes.data if j.size else None for j in jc])
        keep.append(jptr)
        s.J_coord = ct.cast(jptr, ct.c_void_p)
        s.J_npt = arr([j.size for j in h2.J], np.int32)
        s.coord = arr(h2.coord, np.float64)
        s.B_ptr = arr(h2.B_ptr, np.uint64)
        s.D_ptr = arr(h2.D_ptr, np.uint64)
        s.B_data = arr(h2.B_data, np.float64) if not jit else None
        s.D_data = arr(h2.D_data, np.float64) if not jit else None
        s.krnl_id = h2.krnl_id
        p = list(h2.krnl_param)
        for i, v in enumerate(p[:4]):
            s.krnl_param[i] = float(v)
        s.n_krnl_param = len(p)
        s.BD_JIT = jit
        self.s = s
        self._keep = keep

    def matvec(self, x: np.ndarray, n_thread: int = 0) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float64)
        assert x.size == self.h2.krnl_mat_size
        y = np.empty_like(x)
        rc = lib().h2o_matvec(ct.byref(self.s), x.ctypes.data, y.ctypes.data, n_thread)
        if rc != 0:
            raise MemoryError("oracle matvec allocation failed")
        return y

    def matmul(self, X: np.ndarray, layout: str = "col", n_thread: int = 0) -> np.ndarray:
        """X is (N*kd, n_vec); layout selects how it is handed to the C code."""
        N = self.h2.krnl_mat_size
        n_vec = X.shape[1]
        if layout == "row":
            Xc = np.ascontiguousarray(X, dtype=np.float64)
            Y = np.empty_like(Xc)
            rc = lib().h2o_matmul(ct.byref(self.s), 101, n_vec, Xc.ctypes.data, n_vec,
                                  Y.ctypes.data, n_vec, n_thread)
            out = Y
        else:
            Xc = np.asfortranarray(X, dtype=np.float64)
            Y = np.empty((N, n_vec), order="F")
            rc = lib().h2o_matmul(ct.byref(self.s), 102, n_vec, Xc.ctypes.data, N,
                                  Y.ctypes.data, N, n_thread)
            out = Y
        if rc != 0:
            raise MemoryError("oracle matmul allocation failed")
        return np.ascontiguousarray(out)


def kernel_eval(kid: int, param, c0: np.ndarray, c1: np.ndarray) -> np.ndarray:
    kd = 3 if kid == 3 else 1
    c0 = np.ascontiguousarray(c0, dtype=np.float64)
    c1 = np.ascontiguousarray(c1, dtype=np.float64)
    n0, n1 = c0.shape[1], c1.shape[1]
    p = np.ascontiguousarray(list(param) + [0.0] * 4, dtype=np.float64)
    out = np.empty((n0 * kd, n1 * kd))
    lib().h2o_kernel_eval(kid, p.ctypes.data, len(param), c0.ctypes.data, n0, n0,
                          c1.ctypes.data, n1, n1, out.ctypes.data, n1 * kd)
    return out


def direct_nbody(kid: int, param, coord: np.ndarray, x: np.ndarray, rows: np.ndarray,
                 n_thread: int = 0) -> np.ndarray:
    """y[rows] of the dense kernel product K(coord, coord) x (FP64)."""
    kd = 3 if kid == 3 else 1
    coord = np.ascontiguousarray(coord, dtype=np.float64)
    dst = np.ascontiguousarray(coord[:, rows])
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty(rows.size * kd)
    p = np.ascontiguousarray(list(param) + [0.0] * 4, dtype=np.float64)
    lib().h2o_direct_nbody(kid, p.ctypes.data, len(param), kd, coord.ctypes.data,
                           coord.shape[1], coord.shape[1], x.ctypes.data,
                           dst.ctypes.data, rows.size, rows.size, y.ctypes.data, n_thread)
    return y


def pcg(matvec, b: np.ndarray, tol: float, max_iter: int, x0: np.ndarray | None = None):
    """Unpreconditioned CG of examples/PCG/pcg.c:11-152 (invMx = NULL, so
    z = r): r = b - A x; stop when ||r|| <= tol ||b|| or after max_iter
    iterations.  ``matvec`` applies the operator (for the shifted system of
    examples/SPDHSS-H2/pcg_tests.c:15-21 it adds shift * x itself).
    Returns (x, flag, relres, iter) with flag 0 on convergence."""
    n = b.size
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
    r = b - matvec(x)                                    # pcg.c:29-37
    b_2norm = math.sqrt(float(b @ b))                    # pcg.c:43-50
    r_2norm = math.sqrt(float(r @ r))
    rn_stop = b_2norm * tol
    it = 0
    rho = 1.0
    p = np.zeros(n)
    while it < max_iter and r_2norm > rn_stop:          # pcg.c:63
        z = r                                            # pcg.c:67-68
        rho0 = rho                                       # pcg.c:76-81
        rho = float(r @ z)
        beta = rho / rho0
        p = z.copy() if it == 0 else z + beta * p        # pcg.c:83-90
        s = matvec(p)                                    # pcg.c:95
        alpha = rho / float(p @ s)                       # pcg.c:101-105
        x = x + alpha * p                                # pcg.c:110-117
        r = r - alpha * s
        r_2norm = math.sqrt(float(r @ r))
        it += 1
    flag = 0 if r_2norm <= rn_stop else 1                # pcg.c:125-127
    relres = r_2norm / b_2norm if b_2norm > 0 else float("nan")   # C's 0 / 0
    return x, flag, relres, it


def periodic_matvec(h2, x: np.ndarray) -> np.ndarray:
    """H2P_matvec_periodic (src/H2Pack_matvec_periodic.c:394-525), JIT, in
    numpy over the oracle's kernel evaluation: forward permutation, upward
    sweep, the one-sided shifted pair products in their three level cases
    (:140-255), the root's periodic block y1[root] = per_blk y0[root]
    (:455-470), downward sweep, the one-sided shifted dense products
    (:300-370), backward permutation."""
    kid, par, kd = h2.krnl_id, list(h2.krnl_param), h2.krnl_dim
    nn, nl = h2.n_node, h2.n_leaf_node
    ch = h2.children.reshape(nn, h2.max_child)
    pc = h2.pt_cluster.reshape(-1, 2)
    mc = h2.mat_cluster.reshape(-1, 2)
    lvl = h2.node_level
    px = np.asarray(x, dtype=np.float64)[h2.fwd_pmt_idx]
    rows = lambda n: slice(mc[n, 0], mc[n, 1] + 1)                       # noqa: E731
    pts = lambda n: h2.coord[:, pc[n, 0]:pc[n, 1] + 1]                      # noqa: E731

    def shifted(c, s):
        c = np.array(c, dtype=np.float64)
        c[:len(s)] += np.asarray(s)[:, None]
        return c

    # upward sweep (H2P_matvec_fwd_transform)
    y0 = [np.zeros(0) for _ in range(nn)]
    for lv in range(h2.max_level, h2.min_adm_level - 1, -1):
        for n in h2.level_nodes[lv * nl: lv * nl + h2.level_n_node[lv]]:
            if h2.U_ncol[n] == 0:
                continue
            if h2.n_child[n] == 0:
                y0[n] = h2.U[n].T @ px[rows(n)]
            else:
                y0[n] = h2.U[n].T @ np.concatenate([y0[c] for c in ch[n, :h2.n_child[n]]])
    y1 = [np.zeros(h2.U_ncol[n]) for n in range(nn)]
    py = np.zeros_like(px)
    for i, (n0, n1) in enumerate(h2.r_adm_pairs.reshape(-1, 2)):
        s = h2.per_adm_shifts[i]
        if lvl[n0] == lvl[n1]:
            y1[n0] += kernel_eval(kid, par, h2.J_coord[n0], shifted(h2.J_coord[n1], s)) @ y0[n1]
        elif lvl[n0] > lvl[n1]:
            y1[n0] += kernel_eval(kid, par, h2.J_coord[n0], shifted(pts(n1), s)) @ px[rows(n1)]
        else:
            py[rows(n0)] += kernel_eval(kid, par, pts(n0), shifted(h2.J_coord[n1], s)) @ y0[n1]
    root = h2.root_idx
    y1[root] = h2.per_blk @ y0[root]
    # downward sweep (H2P_matvec_bwd_transform)
    for lv in range(h2.min_adm_level, h2.max_level + 1):
        for n in h2.level_nodes[lv * nl: lv * nl + h2.level_n_node[lv]]:
            if h2.U_ncol[n] == 0:
                continue
            v = h2.U[n] @ y1[n]
            if h2.n_child[n] == 0:
                py[rows(n)] += v
            else:
                off = 0
                for c in ch[n, :h2.n_child[n]]:
                    y1[c] = y1[c] + v[off:off + h2.U_ncol[c]]
                    off += h2.U_ncol[c]
    # dense blocks: the leaves, then the shifted inadmissible pairs
    for n in h2.leaf_nodes:
        py[rows(n)] += kernel_eval(kid, par, pts(n), pts(n)) @ px[rows(n)]
    for i, (n0, n1) in enumerate(h2.r_inadm_pairs.reshape(-1, 2)):
        s = h2.per_inadm_shifts[i]
        py[rows(n0)] += kernel_eval(kid, par, pts(n0), shifted(pts(n1), s)) @ px[rows(n1)]
    y = np.empty_like(py)
    y[h2.fwd_pmt_idx] = py
    return y
