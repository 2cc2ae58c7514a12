"""Host-side interface of the MI355X engine.

``H2MatHIP`` takes an H2 representation built on the host
(``h2pack_amd.construct``; the same fields as the reference's H2Pack struct),
uploads it once through ``h2p_hip_upload`` and offers the reference's apply
calls:

* ``matvec(x)``  -- H2P_matvec (src/H2Pack_matvec.c:1162), y := A_H2 x in the
  original point order, length krnl_mat_size;
* ``matmul(X, layout)`` -- H2P_matmul (src/H2Pack_matmul.c:455) on an
  (krnl_mat_size, n_vec) panel, row- or column-major as CBLAS_LAYOUT;
* device-pointer variants for torch tensors already in HBM.

Failures raise ``H2PHipError`` carrying the library's error code and message
(the C calls never abort).  There is no CPU path behind these calls.
"""
from __future__ import annotations

import ctypes as ct

import numpy as np

from . import abi
from ._lib import hip_lib


class H2PHipError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"h2p_hip error {code}: {msg}")
        self.code = code


def _check(rc: int):
    if rc != 0:
        raise H2PHipError(rc, hip_lib().h2p_hip_last_error().decode(errors="replace"))


def kernel_desc(krnl_id: int, param) -> abi.KernelDesc:
    k = abi.KernelDesc()
    k.id = int(krnl_id)
    p = [] if param is None else list(param)
    k.n_param = len(p)
    for i, v in enumerate(p[:4]):
        k.param[i] = float(v)
    return k


def describe(h2, with_blocks: bool = True):
    """Fill an ``H2PackDesc`` from a construct.H2Pack; returns (desc, keepalive)."""
    keep = []

    def arr(a, dt):
        a = np.ascontiguousarray(a, dtype=dt)
        keep.append(a)
        return a.ctypes.data

    d = abi.H2PackDesc()
    d.pt_dim, d.xpt_dim, d.krnl_dim = h2.pt_dim, h2.xpt_dim, h2.krnl_dim
    d.n_point, d.krnl_mat_size = h2.n_point, h2.krnl_mat_size
    d.n_node, d.n_leaf_node, d.max_child = h2.n_node, h2.n_leaf_node, h2.max_child
    d.max_level, d.min_adm_level = h2.max_level, h2.min_adm_level
    for f in ("children", "n_child", "node_level", "pt_cluster", "mat_cluster", "level_n_node",
              "level_nodes", "r_adm_pairs", "r_inadm_pairs", "node_n_r_adm", "U_nrow", "U_ncol",
              "fwd_pmt_idx", "bwd_pmt_idx", "B_nrow", "B_ncol", "D_nrow", "D_ncol"):
        setattr(d, f, arr(getattr(h2, f), np.int32))
    d.leaf_nodes = arr(h2.leaf_nodes, np.int32)
    d.n_r_adm_pair, d.n_r_inadm_pair = h2.n_r_adm_pair, h2.n_r_inadm_pair
    Us = [np.ascontiguousarray(u, dtype=np.float64) for u in h2.U]
    keep.extend(Us)
    up = (ct.c_void_p * h2.n_node)(*[u.ctypes.data if u.size else None for u in Us])
    keep.append(up)
    d.U = ct.cast(up, ct.c_void_p)
    jc = [np.ascontiguousarray(j, dtype=np.float64) for j in h2.J_coord]
    keep.extend(jc)
    jp = (ct.c_void_p * h2.n_node)(*[j.ctypes.data if j.size else None for j in jc])
    keep.append(jp)
    d.J_coord = ct.cast(jp, ct.c_void_p)
    d.J_npt = arr([j.size for j in h2.J], np.int32)
    d.coord = arr(h2.coord, np.float64)
    d.B_ptr = arr(h2.B_ptr, np.uint64)
    d.D_ptr = arr(h2.D_ptr, np.uint64)
    if with_blocks and h2.B_data is not None and h2.D_data is not None:
        d.B_data = arr(h2.B_data, np.float64)
        d.D_data = arr(h2.D_data, np.float64)
    else:
        d.B_data = None
        d.D_data = None
    if getattr(h2, "is_RPY_Ewald", 0):
        d.periodic = 1
        d.per_adm_shifts = arr(np.asarray(h2.per_adm_shifts).reshape(-1), np.float64)
        d.per_inadm_shifts = arr(np.asarray(h2.per_inadm_shifts).reshape(-1), np.float64)
        d.per_blk = arr(h2.per_blk, np.float64)
    else:
        d.periodic = 0
        d.per_adm_shifts = d.per_inadm_shifts = d.per_blk = None
    return d, keep


class H2MatHIP:
    """An H2 matrix resident in HBM (one ``h2p_hip_handle``)."""

    def __init__(self, h2, mode: str = "jit", prec: str = "fp64",
                 host_blocks: bool = True):
        self.lib = hip_lib()
        self.h2 = h2
        self.N = h2.krnl_mat_size
        self.mode = {"aot": abi.AOT, "jit": abi.JIT}[mode]
        self.prec = {"fp64": abi.FP64, "fp32mix": abi.FP32MIX}[prec]
        desc, keep = describe(h2, with_blocks=host_blocks and self.mode == abi.AOT)
        self._kdesc = kernel_desc(h2.krnl_id, h2.krnl_param)
        hnd = ct.c_void_p()
        _check(self.lib.h2p_hip_upload(ct.byref(desc), ct.byref(self._kdesc), self.mode,
                                       self.prec, ct.byref(hnd)))
        self.handle = hnd
        del keep

    def close(self):
        if getattr(self, "handle", None):
            self.lib.h2p_hip_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- host arrays -----------------------------------------------------
    def matvec(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float64)
        if x.size != self.N:
            raise ValueError(f"x has {x.size} entries, expected {self.N}")
        y = np.empty_like(x)
        _check(self.lib.h2p_hip_matvec(self.handle, x.ctypes.data, y.ctypes.data))
        return y

    def matmul(self, X: np.ndarray, layout: str = "col") -> np.ndarray:
        """Y = A X for X of shape (krnl_mat_size, n_vec)."""
        if X.shape[0] != self.N:
            raise ValueError("X must have krnl_mat_size rows")
        n_vec = X.shape[1]
        if layout == "row":
            Xc = np.ascontiguousarray(X, dtype=np.float64)
            Y = np.empty_like(Xc)
            _check(self.lib.h2p_hip_matmul(self.handle, abi.ROW_MAJOR, n_vec, Xc.ctypes.data,
                                           n_vec, Y.ctypes.data, n_vec))
            return Y
        Xf = np.asfortranarray(X, dtype=np.float64)
        Y = np.empty((self.N, n_vec), order="F")
        _check(self.lib.h2p_hip_matmul(self.handle, abi.COL_MAJOR, n_vec, Xf.ctypes.data,
                                       self.N, Y.ctypes.data, self.N))
        return np.ascontiguousarray(Y)

    # -- device arrays (torch tensors on the GPU) --------------------------
    def matvec_dev(self, x_ptr: int, y_ptr: int, stream: int = 0):
        _check(self.lib.h2p_hip_matvec_dev(self.handle, ct.c_void_p(x_ptr), ct.c_void_p(y_ptr),
                                           ct.c_void_p(stream)))

    def matmul_dev(self, layout: int, n_vec: int, x_ptr: int, ldx: int, y_ptr: int, ldy: int,
                   stream: int = 0):
        _check(self.lib.h2p_hip_matmul_dev(self.handle, layout, n_vec, ct.c_void_p(x_ptr), ldx,
                                           ct.c_void_p(y_ptr), ldy, ct.c_void_p(stream)))

    def matmul_dev_noalloc(self, layout: int, n_vec: int, x_ptr: int, ldx: int, y_ptr: int, ldy: int,
                           stream: int = 0):
        """matmul_dev with its panels in the library's static per-GPU scratch
        (h2p_hip_matmul_dev_noalloc): no device allocation in the call."""
        _check(self.lib.h2p_hip_matmul_dev_noalloc(self.handle, layout, n_vec, ct.c_void_p(x_ptr), ldx,
                                                   ct.c_void_p(y_ptr), ldy, ct.c_void_p(stream)))

    def autotune(self, n_vec: int) -> dict:
        """Time the kernel variants on this matrix and keep the fastest
        (h2p_hip_autotune): the matmul tiles per phase for ``n_vec`` columns,
        or with ``n_vec`` = 0 the matvec's grouped block launch, whether its
        D blocks run beside the sweeps (``d_concurrent``) and whether the
        levels above the leaf level run as a chain of their own beside the
        deep B blocks (``level_split``)."""
        t0, t1 = ct.c_double(0), ct.c_double(0)
        _check(self.lib.h2p_hip_autotune(self.handle, int(n_vec), ct.byref(t0), ct.byref(t1)))
        out = {"default_ms": t0.value, "tuned_ms": t1.value}
        if n_vec == 0:
            inf = self.info()
            out["variants"] = [inf["mv_variant"]]
            out["d_concurrent"] = inf["mv_d_concurrent"]
            out["level_split"] = inf["mv_level_split"]
        else:
            out["variants"] = [self.mm_variant(ph, n_vec) for ph in range(4)]
        return out

    def mm_variant(self, phase: int, n_vec: int, variant: int = -1) -> int:
        """Pin (variant >= 0) or read (variant -1) a matmul phase's tile variant."""
        rc = self.lib.h2p_hip_mm_variant(self.handle, int(phase), int(n_vec), int(variant))
        if rc < 0:
            _check(rc)
        return rc

    def set_profiling(self, on: bool):
        _check(self.lib.h2p_hip_set_profiling(self.handle, 1 if on else 0))

    def set_cg_replace(self, ratio: float):
        """Residual-replacement ratio of the next CG solve started on this
        handle (h2p_hip_cg_set_replace; 0 never pauses)."""
        _check(self.lib.h2p_hip_cg_set_replace(self.handle, float(ratio)))

    def info(self) -> dict:
        I = abi.Info()
        _check(self.lib.h2p_hip_get_info(self.handle, ct.byref(I)))
        return I.as_dict()


def direct_nbody(krnl_id: int, param, coord: np.ndarray, x: np.ndarray,
                 rows: np.ndarray) -> np.ndarray:
    """FP64 dense product rows on the device (examples/direct_nbody.h)."""
    lib = hip_lib()
    coord = np.ascontiguousarray(coord, dtype=np.float64)
    dst = np.ascontiguousarray(coord[:, rows])
    x = np.ascontiguousarray(x, dtype=np.float64)
    kd = 3 if krnl_id == abi.KRNL_RPY else 1
    y = np.empty(rows.size * kd)
    kdsc = kernel_desc(krnl_id, param)
    _check(lib.h2p_hip_direct_nbody(ct.byref(kdsc), coord.shape[0], coord.ctypes.data,
                                    coord.shape[1], coord.shape[1], x.ctypes.data,
                                    dst.ctypes.data, rows.size, rows.size, y.ctypes.data))
    return y


def footprint(h2, mode: str = "jit", prec: str = "fp64", n_vec: int = 0,
              with_cg: bool = False) -> int:
    """Device bytes an upload of ``h2`` would hold (plus matmul panels for
    ``n_vec`` columns and the CG state); allocates nothing on the device."""
    lib = hip_lib()
    m = {"aot": abi.AOT, "jit": abi.JIT}[mode]
    p = {"fp64": abi.FP64, "fp32mix": abi.FP32MIX}[prec]
    desc, keep = describe(h2, with_blocks=False)
    kd = kernel_desc(h2.krnl_id, h2.krnl_param)
    out = ct.c_size_t(0)
    _check(lib.h2p_hip_footprint(ct.byref(desc), ct.byref(kd), m, p, int(n_vec),
                                 1 if with_cg else 0, ct.byref(out)))
    del keep
    return int(out.value)


def _ptrs(vals):
    return (ct.c_void_p * max(len(vals), 1))(*[ct.c_void_p(int(v)) for v in vals])


class H2Batch:
    """Co-scheduled calls over several resident matrices (``h2p_hip_batch``).

    Each method takes lists of ``H2MatHIP`` objects and device addresses (for
    instance ``tensor.data_ptr()``) and enqueues on ``stream``; a matrix may
    appear once per call.  The batched solver follows the reference's
    unpreconditioned PCG (examples/PCG/pcg.c:11-152) on (A + shift I)."""

    def __init__(self):
        self.lib = hip_lib()
        b = ct.c_void_p()
        _check(self.lib.h2p_hip_batch_create(ct.byref(b)))
        self.handle = b

    def close(self):
        if getattr(self, "handle", None):
            self.lib.h2p_hip_batch_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _hs(mats):
        return _ptrs([m.handle.value for m in mats])

    def matvec(self, mats, x_ptrs, y_ptrs, shifts=None, stream: int = 0):
        n = len(mats)
        sh = None
        if shifts is not None:
            sh = (ct.c_double * max(n, 1))(*[float(v) for v in shifts])
        _check(self.lib.h2p_hip_matvec_batch(self.handle, n, self._hs(mats), _ptrs(x_ptrs),
                                             _ptrs(y_ptrs), sh, ct.c_void_p(stream)))

    def matmul(self, mats, layout: int, n_vecs, x_ptrs, ldxs, y_ptrs, ldys, stream: int = 0):
        n = len(mats)
        ia = lambda v: (ct.c_int * max(n, 1))(*[int(a) for a in v])   # noqa: E731
        _check(self.lib.h2p_hip_matmul_batch(self.handle, n, self._hs(mats), layout, ia(n_vecs),
                                             _ptrs(x_ptrs), ia(ldxs), _ptrs(y_ptrs), ia(ldys),
                                             ct.c_void_p(stream)))

    def cg_start(self, mats, b_ptrs, x_ptrs, shifts, tols, max_iters, stream: int = 0):
        n = len(mats)
        da = lambda v: (ct.c_double * max(n, 1))(*[float(a) for a in v])   # noqa: E731
        ia = (ct.c_int * max(n, 1))(*[int(a) for a in max_iters])
        _check(self.lib.h2p_hip_cg_start(self.handle, n, self._hs(mats), _ptrs(b_ptrs),
                                         _ptrs(x_ptrs), da(shifts), da(tols), ia,
                                         ct.c_void_p(stream)))

    def cg_step(self, mats, n_iter: int, stream: int = 0):
        _check(self.lib.h2p_hip_cg_step(self.handle, len(mats), self._hs(mats), int(n_iter),
                                        ct.c_void_p(stream)))

    def cg_replace(self, mats, accurate, stream: int = 0):
        """Residual replacement (h2p_hip_cg_replace): r = b - (A' + shift) x
        for each solve of ``mats``, the product by the matching handle of
        ``accurate`` (for instance the FP64 handle of the same matrix)."""
        _check(self.lib.h2p_hip_cg_replace(self.handle, len(mats), self._hs(mats), self._hs(accurate),
                                           ct.c_void_p(stream)))

    def cg_poll(self, mats, stream: int = 0):
        """Waits for ``stream``; returns (done, iters, relres) lists, done 0
        running, 1 finished, 2 paused for a residual replacement."""
        n = len(mats)
        done = (ct.c_int * max(n, 1))()
        it = (ct.c_int * max(n, 1))()
        rr = (ct.c_double * max(n, 1))()
        _check(self.lib.h2p_hip_cg_poll(self.handle, n, self._hs(mats), done, it, rr,
                                        ct.c_void_p(stream)))
        return [int(done[i]) for i in range(n)], [it[i] for i in range(n)], [rr[i] for i in range(n)]
