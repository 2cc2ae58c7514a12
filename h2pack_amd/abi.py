"""ctypes mirror of include/h2p_hip.h."""
from __future__ import annotations

import ctypes as ct

KRNL_COULOMB, KRNL_GAUSSIAN, KRNL_MATERN32, KRNL_RPY = 0, 1, 2, 3
AOT, JIT = 0, 1
FP64, FP32MIX = 0, 1
ROW_MAJOR, COL_MAJOR = 101, 102
EINVAL, ENOMEM, EDEVICE, ENOSYS = -1, -2, -3, -4

# every symbol include/h2p_hip.h declares
EXPORTS = ("h2p_hip_upload", "h2p_hip_matvec", "h2p_hip_matvec_dev", "h2p_hip_matmul",
           "h2p_hip_matmul_dev", "h2p_hip_set_profiling", "h2p_hip_get_info",
           "h2p_hip_free", "h2p_hip_last_error", "h2p_hip_direct_nbody",
           "h2p_hip_footprint", "h2p_hip_batch_create", "h2p_hip_batch_free",
           "h2p_hip_matvec_batch", "h2p_hip_matmul_batch", "h2p_hip_cg_start",
           "h2p_hip_cg_step", "h2p_hip_cg_poll", "h2p_hip_autotune", "h2p_hip_mm_variant",
           "h2p_hip_matmul_dev_noalloc", "h2p_hip_cg_set_replace", "h2p_hip_cg_replace")
MM_VARIANTS = 4
NOALLOC_BYTES = 256 << 20

_vp = ct.c_void_p


class KernelDesc(ct.Structure):
    _fields_ = [("id", ct.c_int), ("n_param", ct.c_int), ("param", ct.c_double * 4)]


class H2PackDesc(ct.Structure):
    _fields_ = [
        ("pt_dim", ct.c_int), ("xpt_dim", ct.c_int), ("krnl_dim", ct.c_int),
        ("n_point", ct.c_int), ("krnl_mat_size", ct.c_int),
        ("n_node", ct.c_int), ("n_leaf_node", ct.c_int), ("max_child", ct.c_int),
        ("max_level", ct.c_int), ("min_adm_level", ct.c_int),
        ("children", _vp), ("n_child", _vp), ("node_level", _vp), ("pt_cluster", _vp),
        ("mat_cluster", _vp), ("level_n_node", _vp), ("level_nodes", _vp), ("leaf_nodes", _vp),
        ("n_r_adm_pair", ct.c_int), ("n_r_inadm_pair", ct.c_int),
        ("r_adm_pairs", _vp), ("r_inadm_pairs", _vp), ("node_n_r_adm", _vp),
        ("U_nrow", _vp), ("U_ncol", _vp), ("U", _vp),
        ("J_npt", _vp), ("J_coord", _vp), ("coord", _vp),
        ("fwd_pmt_idx", _vp), ("bwd_pmt_idx", _vp),
        ("B_nrow", _vp), ("B_ncol", _vp), ("B_ptr", _vp), ("B_data", _vp),
        ("D_nrow", _vp), ("D_ncol", _vp), ("D_ptr", _vp), ("D_data", _vp),
        ("periodic", ct.c_int), ("per_adm_shifts", _vp), ("per_inadm_shifts", _vp), ("per_blk", _vp),
    ]


class Info(ct.Structure):
    _fields_ = [
        ("U_elems", ct.c_size_t), ("B_elems", ct.c_size_t), ("D_elems", ct.c_size_t),
        ("hbm_bytes_model", ct.c_size_t),
        ("B_pairs", ct.c_double), ("D_pairs", ct.c_double), ("jit_flops_model", ct.c_double),
        ("device_bytes", ct.c_size_t), ("n_block_tasks", ct.c_int), ("n_levels", ct.c_int),
        ("mode", ct.c_int), ("prec", ct.c_int),
        ("t_fwd", ct.c_double), ("t_blocks", ct.c_double), ("t_reduce", ct.c_double),
        ("t_bwd", ct.c_double), ("t_final", ct.c_double), ("t_total", ct.c_double),
        ("n_profiled", ct.c_int), ("mv_variant", ct.c_int),
        ("t_blocks_b", ct.c_double), ("t_blocks_d", ct.c_double),
        ("mv_d_concurrent", ct.c_int), ("mv_level_split", ct.c_int),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


def bind(lib: ct.CDLL) -> ct.CDLL:
    for name in EXPORTS:
        getattr(lib, name)   # raises AttributeError when a symbol is missing
    lib.h2p_hip_upload.argtypes = [ct.POINTER(H2PackDesc), ct.POINTER(KernelDesc), ct.c_int,
                                   ct.c_int, ct.POINTER(_vp)]
    lib.h2p_hip_upload.restype = ct.c_int
    lib.h2p_hip_matvec.argtypes = [_vp, _vp, _vp]
    lib.h2p_hip_matvec.restype = ct.c_int
    lib.h2p_hip_matvec_dev.argtypes = [_vp, _vp, _vp, _vp]
    lib.h2p_hip_matvec_dev.restype = ct.c_int
    lib.h2p_hip_matmul.argtypes = [_vp, ct.c_int, ct.c_int, _vp, ct.c_int, _vp, ct.c_int]
    lib.h2p_hip_matmul.restype = ct.c_int
    lib.h2p_hip_matmul_dev.argtypes = [_vp, ct.c_int, ct.c_int, _vp, ct.c_int, _vp, ct.c_int, _vp]
    lib.h2p_hip_matmul_dev.restype = ct.c_int
    lib.h2p_hip_matmul_dev_noalloc.argtypes = [_vp, ct.c_int, ct.c_int, _vp, ct.c_int, _vp, ct.c_int, _vp]
    lib.h2p_hip_matmul_dev_noalloc.restype = ct.c_int
    lib.h2p_hip_set_profiling.argtypes = [_vp, ct.c_int]
    lib.h2p_hip_set_profiling.restype = ct.c_int
    lib.h2p_hip_get_info.argtypes = [_vp, ct.POINTER(Info)]
    lib.h2p_hip_get_info.restype = ct.c_int
    lib.h2p_hip_free.argtypes = [_vp]
    lib.h2p_hip_free.restype = None
    lib.h2p_hip_last_error.argtypes = []
    lib.h2p_hip_last_error.restype = ct.c_char_p
    lib.h2p_hip_direct_nbody.argtypes = [ct.POINTER(KernelDesc), ct.c_int, _vp, ct.c_int, ct.c_int,
                                         _vp, _vp, ct.c_int, ct.c_int, _vp]
    lib.h2p_hip_direct_nbody.restype = ct.c_int
    lib.h2p_hip_footprint.argtypes = [ct.POINTER(H2PackDesc), ct.POINTER(KernelDesc), ct.c_int, ct.c_int,
                                      ct.c_int, ct.c_int, ct.POINTER(ct.c_size_t)]
    lib.h2p_hip_footprint.restype = ct.c_int
    lib.h2p_hip_batch_create.argtypes = [ct.POINTER(_vp)]
    lib.h2p_hip_batch_create.restype = ct.c_int
    lib.h2p_hip_batch_free.argtypes = [_vp]
    lib.h2p_hip_batch_free.restype = None
    # pointer arrays are passed as addresses of ctypes arrays
    lib.h2p_hip_matvec_batch.argtypes = [_vp, ct.c_int, _vp, _vp, _vp, _vp, _vp]
    lib.h2p_hip_matvec_batch.restype = ct.c_int
    lib.h2p_hip_matmul_batch.argtypes = [_vp, ct.c_int, _vp, ct.c_int, _vp, _vp, _vp, _vp, _vp, _vp]
    lib.h2p_hip_matmul_batch.restype = ct.c_int
    lib.h2p_hip_cg_start.argtypes = [_vp, ct.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
    lib.h2p_hip_cg_start.restype = ct.c_int
    lib.h2p_hip_cg_step.argtypes = [_vp, ct.c_int, _vp, ct.c_int, _vp]
    lib.h2p_hip_cg_step.restype = ct.c_int
    lib.h2p_hip_cg_poll.argtypes = [_vp, ct.c_int, _vp, _vp, _vp, _vp, _vp]
    lib.h2p_hip_cg_poll.restype = ct.c_int
    lib.h2p_hip_cg_set_replace.argtypes = [_vp, ct.c_double]
    lib.h2p_hip_cg_set_replace.restype = ct.c_int
    lib.h2p_hip_cg_replace.argtypes = [_vp, ct.c_int, _vp, _vp, _vp]
    lib.h2p_hip_cg_replace.restype = ct.c_int
    lib.h2p_hip_autotune.argtypes = [_vp, ct.c_int, ct.POINTER(ct.c_double), ct.POINTER(ct.c_double)]
    lib.h2p_hip_autotune.restype = ct.c_int
    lib.h2p_hip_mm_variant.argtypes = [_vp, ct.c_int, ct.c_int, ct.c_int]
    lib.h2p_hip_mm_variant.restype = ct.c_int
    return lib
