"""The H2Pack_p binding (h2pack_amd/glue/h2p_hip_h2pack.c) against the
reference's own struct H2Pack (src/H2Pack_typedef.h:88-207).

tests/h2pack_struct.c fills a struct H2Pack the way H2P_build leaves it (H2
pair lists in r_*_pairs and, in HSS mode, the HSS lists in HSS_r_*_pairs;
empty U / J_coord for nodes without a basis; every U and J_coord stored with
a padded leading dimension), and the binding's h2p_hip_h2pack must equal,
field by field and element by element, what h2mat.describe() uploads for the
same representation -- the description every GPU parity test runs through.
The library is compiled against the reference headers (build() does it where
they exist); without them these tests skip.  The GPU test uploads through
the binding and compares the products bitwise with the ordinary upload.
"""
import ctypes as ct
import os

import numpy as np
import pytest

from h2pack_amd import abi, _build
from h2pack_amd import construct as C
from tests.problems import problem, periodic_problem

SO = _build.H2PACK_CHECK_SO
_vp = ct.c_void_p


class Src(ct.Structure):
    _fields_ = ([(f, ct.c_int) for f in (
        "pt_dim", "xpt_dim", "krnl_dim", "n_point", "krnl_mat_size", "n_node", "n_leaf_node",
        "max_child", "max_level", "min_adm_level", "HSS_min_adm_level", "n_r_adm_pair",
        "n_r_inadm_pair", "HSS_n_r_adm_pair", "HSS_n_r_inadm_pair", "is_HSS", "is_RPY_Ewald",
        "BD_JIT", "ld_pad")]
        + [(f, _vp) for f in (
            "children", "n_child", "node_level", "pt_cluster", "mat_cluster", "level_n_node",
            "level_nodes", "height_nodes", "r_adm_pairs", "r_inadm_pairs", "HSS_r_adm_pairs",
            "HSS_r_inadm_pairs", "node_n_r_adm", "fwd_pmt_idx", "bwd_pmt_idx", "B_nrow", "B_ncol",
            "D_nrow", "D_ncol", "B_ptr", "D_ptr", "coord", "B_data", "D_data", "per_adm_shifts",
            "per_inadm_shifts", "per_blk", "U_nrow", "U_ncol", "J_n", "U", "J_coord", "krnl_param")])


def _lib():
    if not os.path.exists(SO):
        if _build.build_h2pack_check() is None:
            pytest.skip("reference headers (src/H2Pack_typedef.h) not available to build the binding check")
    L = ct.CDLL(SO)
    L.h2s_make.restype = _vp
    L.h2s_make.argtypes = [ct.POINTER(Src)]
    L.h2s_free.argtypes = [_vp]
    L.h2s_describe.argtypes = [_vp, ct.POINTER(abi.H2PackDesc), ct.POINTER(_vp)]
    L.h2s_release.argtypes = [_vp]
    L.h2s_kernel.argtypes = [_vp, ct.c_int, ct.POINTER(abi.KernelDesc)]
    L.h2s_upload.argtypes = [_vp, ct.c_int, ct.c_int, ct.POINTER(_vp)]
    L.h2s_offsetof_HSS_r_adm_pairs.restype = ct.c_size_t
    L.h2s_sizeof_H2Pack.restype = ct.c_size_t
    return L


def reference_struct(h2, BD_JIT, ld_pad=3):
    """An Src describing h2 as the reference's struct holds it; returns
    (Src, keepalive)."""
    keep = []

    def arr(a, dt):
        a = np.ascontiguousarray(a, dtype=dt)
        keep.append(a)
        return a.ctypes.data

    s = Src()
    for f in ("pt_dim", "xpt_dim", "krnl_dim", "n_point", "krnl_mat_size", "n_node", "n_leaf_node",
              "max_child", "max_level"):
        setattr(s, f, int(getattr(h2, f)))
    hss = int(getattr(h2, "is_HSS", 0))
    s.is_HSS, s.BD_JIT, s.ld_pad = hss, BD_JIT, ld_pad
    s.is_RPY_Ewald = int(getattr(h2, "is_RPY_Ewald", 0))
    if hss:   # H2P_run_HSS keeps the H2 lists and adds the HSS ones
        s.min_adm_level, s.HSS_min_adm_level = h2.H2_min_adm_level, h2.min_adm_level
        h2_adm, h2_inadm = h2.H2_r_adm_pairs, h2.H2_r_inadm_pairs
        s.r_adm_pairs, s.r_inadm_pairs = arr(h2_adm, np.int32), arr(h2_inadm, np.int32)
        s.n_r_adm_pair, s.n_r_inadm_pair = len(h2_adm) // 2, len(h2_inadm) // 2
        s.HSS_r_adm_pairs, s.HSS_r_inadm_pairs = arr(h2.r_adm_pairs, np.int32), arr(h2.r_inadm_pairs, np.int32)
        s.HSS_n_r_adm_pair, s.HSS_n_r_inadm_pair = h2.n_r_adm_pair, h2.n_r_inadm_pair
    else:
        s.min_adm_level, s.HSS_min_adm_level = h2.min_adm_level, -1
        s.r_adm_pairs, s.r_inadm_pairs = arr(h2.r_adm_pairs, np.int32), arr(h2.r_inadm_pairs, np.int32)
        s.n_r_adm_pair, s.n_r_inadm_pair = h2.n_r_adm_pair, h2.n_r_inadm_pair
        s.HSS_r_adm_pairs = s.HSS_r_inadm_pairs = None
    for f in ("children", "n_child", "node_level", "pt_cluster", "mat_cluster", "level_n_node",
              "level_nodes", "node_n_r_adm", "fwd_pmt_idx", "bwd_pmt_idx", "B_nrow", "B_ncol",
              "D_nrow", "D_ncol", "U_nrow", "U_ncol"):
        setattr(s, f, arr(getattr(h2, f), np.int32))
    hn = np.full(((h2.max_level + 1), h2.n_leaf_node), -1, dtype=np.int32)
    hn[0, :] = h2.leaf_nodes   # height_nodes row 0 = the leaves
    s.height_nodes = arr(hn, np.int32)
    s.B_ptr, s.D_ptr = arr(h2.B_ptr, np.uint64), arr(h2.D_ptr, np.uint64)
    s.coord = arr(h2.coord, np.float64)
    s.B_data = arr(h2.B_data, np.float64) if h2.B_data is not None else None
    s.D_data = arr(h2.D_data, np.float64) if h2.D_data is not None else None
    if s.is_RPY_Ewald:
        s.per_adm_shifts = arr(np.asarray(h2.per_adm_shifts).reshape(-1), np.float64)
        s.per_inadm_shifts = arr(np.asarray(h2.per_inadm_shifts).reshape(-1), np.float64)
        s.per_blk = arr(h2.per_blk, np.float64)
    Us = [np.ascontiguousarray(u, dtype=np.float64) for u in h2.U]
    Js = [np.ascontiguousarray(j, dtype=np.float64) for j in h2.J_coord]
    keep += Us + Js
    up = (_vp * h2.n_node)(*[u.ctypes.data if u.size else None for u in Us])
    jp = (_vp * h2.n_node)(*[j.ctypes.data if j.size else None for j in Js])
    keep += [up, jp]
    s.U, s.J_coord = ct.cast(up, _vp), ct.cast(jp, _vp)
    s.J_n = arr([j.size for j in h2.J], np.int32)
    param = np.asarray(h2.krnl_param, dtype=np.float64)
    s.krnl_param = arr(param, np.float64) if param.size else None
    return s, keep


def _ints(p, n):
    return np.ctypeslib.as_array(ct.cast(p, ct.POINTER(ct.c_int)), shape=(n,)).copy() if n else np.zeros(0, np.int32)


def _dbl(p, n):
    return np.ctypeslib.as_array(ct.cast(p, ct.POINTER(ct.c_double)), shape=(n,)).copy() if n else np.zeros(0)


def _u64(p, n):
    return np.ctypeslib.as_array(ct.cast(p, ct.POINTER(ct.c_uint64)), shape=(n,)).copy()


def contents(d):
    """Every field of an h2p_hip_h2pack, arrays read out by their lengths."""
    nn, nl, mc, ml = d.n_node, d.n_leaf_node, d.max_child, d.max_level
    nD = nl + d.n_r_inadm_pair
    out = {f: getattr(d, f) for f in ("pt_dim", "xpt_dim", "krnl_dim", "n_point", "krnl_mat_size",
                                      "n_node", "n_leaf_node", "max_child", "max_level",
                                      "min_adm_level", "n_r_adm_pair", "n_r_inadm_pair", "periodic")}
    for f, n in (("children", nn * mc), ("n_child", nn), ("node_level", nn), ("pt_cluster", 2 * nn),
                 ("mat_cluster", 2 * nn), ("level_n_node", ml + 1), ("level_nodes", (ml + 1) * nl),
                 ("leaf_nodes", nl), ("r_adm_pairs", 2 * d.n_r_adm_pair),
                 ("r_inadm_pairs", 2 * d.n_r_inadm_pair), ("node_n_r_adm", nn), ("U_nrow", nn),
                 ("U_ncol", nn), ("J_npt", nn), ("fwd_pmt_idx", d.krnl_mat_size),
                 ("bwd_pmt_idx", d.krnl_mat_size), ("B_nrow", d.n_r_adm_pair),
                 ("B_ncol", d.n_r_adm_pair), ("D_nrow", nD), ("D_ncol", nD)):
        out[f] = _ints(getattr(d, f), n)
    out["B_ptr"] = _u64(d.B_ptr, d.n_r_adm_pair + 1)
    out["D_ptr"] = _u64(d.D_ptr, nD + 1)
    out["coord"] = _dbl(d.coord, d.xpt_dim * d.n_point)
    for f, ptr in (("B_data", d.B_data), ("D_data", d.D_data)):
        tot = int((out["B_ptr"] if f == "B_data" else out["D_ptr"])[-1])
        out[f] = None if not ptr else _dbl(ptr, tot)
    Up = ct.cast(d.U, ct.POINTER(_vp))
    Jp = ct.cast(d.J_coord, ct.POINTER(_vp))
    for i in range(nn):
        r, c, j = out["U_nrow"][i], out["U_ncol"][i], out["J_npt"][i]
        out[f"U[{i}]"] = _dbl(Up[i], r * c) if (r * c and Up[i]) else None
        out[f"J_coord[{i}]"] = _dbl(Jp[i], d.xpt_dim * j) if (j and Jp[i]) else None
    if d.periodic:
        out["per_adm_shifts"] = _dbl(d.per_adm_shifts, 3 * d.n_r_adm_pair)
        out["per_inadm_shifts"] = _dbl(d.per_inadm_shifts, 3 * d.n_r_inadm_pair)
        out["per_blk"] = _dbl(d.per_blk, out["U_ncol"][nn - 1] ** 2)
    return out


def _same(a, b):
    if a is None or b is None:
        return a is None and b is None
    if isinstance(a, np.ndarray):
        return a.shape == b.shape and np.array_equal(a, b)
    return a == b


CASES = [("gaussian", 1), ("coulomb_diag", 0), ("rpy", 1), ("hss_gaussian", 1), ("hss_coulomb", 0)]


@pytest.mark.parametrize("name,BD_JIT", CASES)
def test_binding_fields_equal_the_uploaded_description(name, BD_JIT):
    L = _lib()
    from h2pack_amd.h2mat import describe, kernel_desc
    h2, _ = problem(name)
    src, keep = reference_struct(h2, BD_JIT)
    hp = L.h2s_make(ct.byref(src))
    try:
        d, view = abi.H2PackDesc(), _vp()
        assert L.h2s_describe(hp, ct.byref(d), ct.byref(view)) == 0
        mine = contents(d)
        ref_d, keep2 = describe(h2, with_blocks=(BD_JIT == 0))
        theirs = contents(ref_d)
        assert mine.keys() == theirs.keys()
        diff = [k for k in mine if not _same(mine[k], theirs[k])]
        assert not diff, f"fields differ: {diff[:10]}"
        if getattr(h2, "is_HSS", 0):   # the H2 lists were there too, and were not taken
            assert not np.array_equal(mine["r_adm_pairs"], np.asarray(h2.H2_r_adm_pairs, np.int32))
        k = abi.KernelDesc()
        assert L.h2s_kernel(hp, h2.krnl_id, ct.byref(k)) == 0
        kk = kernel_desc(h2.krnl_id, h2.krnl_param)
        assert (k.id, k.n_param, list(k.param)) == (kk.id, kk.n_param, list(kk.param))
        L.h2s_release(view)
    finally:
        L.h2s_free(hp)
        del keep


def test_binding_periodic_fields():
    L = _lib()
    from h2pack_amd.h2mat import describe
    h2, _, _ = periodic_problem()
    src, keep = reference_struct(h2, 1)
    hp = L.h2s_make(ct.byref(src))
    try:
        d, view = abi.H2PackDesc(), _vp()
        assert L.h2s_describe(hp, ct.byref(d), ct.byref(view)) == 0
        ref_d, keep2 = describe(h2, with_blocks=False)
        mine, theirs = contents(d), contents(ref_d)
        assert mine["periodic"] == 1
        diff = [k for k in mine if not _same(mine[k], theirs[k])]
        assert not diff, f"fields differ: {diff[:10]}"
        L.h2s_release(view)
    finally:
        L.h2s_free(hp)
        del keep


def test_binding_rejects_and_names_kernels():
    L = _lib()
    h2, _ = problem("coulomb")
    src, keep = reference_struct(h2, 1)
    hp = L.h2s_make(ct.byref(src))
    try:
        k = abi.KernelDesc()
        assert L.h2s_kernel(hp, abi.KRNL_COULOMB, ct.byref(k)) == 0 and k.n_param == 0   # NULL param
        assert L.h2s_kernel(hp, abi.KRNL_GAUSSIAN, ct.byref(k)) == abi.EINVAL         # needs l
        assert L.h2s_kernel(hp, 9, ct.byref(k)) == abi.EINVAL
        d, view = abi.H2PackDesc(), _vp()
        assert L.h2s_describe(None, ct.byref(d), ct.byref(view)) == abi.EINVAL
    finally:
        L.h2s_free(hp)
        del keep
    assert L.h2s_sizeof_H2Pack() > L.h2s_offsetof_HSS_r_adm_pairs() > 0


@pytest.mark.gpu
@pytest.mark.parametrize("name,BD_JIT", [("gaussian", 1), ("hss_coulomb", 0), ("rpy", 1)])
def test_upload_through_binding_is_the_same_matrix(name, BD_JIT):
    """h2p_hip_upload_h2pack on the filled struct gives bitwise the products
    of the ordinary upload of the same representation."""
    L = _lib()
    from h2pack_amd._lib import hip_lib
    from h2pack_amd.h2mat import H2MatHIP
    lib = hip_lib()
    h2, _ = problem(name)
    src, keep = reference_struct(h2, BD_JIT)
    hp = L.h2s_make(ct.byref(src))
    hnd = _vp()
    try:
        assert L.h2s_upload(hp, h2.krnl_id, abi.FP64, ct.byref(hnd)) == 0, lib.h2p_hip_last_error()
    finally:
        L.h2s_free(hp)   # the engine holds its own copies
    x = np.random.default_rng(3).random(h2.krnl_mat_size) - 0.5
    y = np.empty_like(x)
    assert lib.h2p_hip_matvec(hnd, x.ctypes.data, y.ctypes.data) == 0
    lib.h2p_hip_free(hnd)
    m = H2MatHIP(h2, mode="jit" if BD_JIT else "aot")
    y_ref = m.matvec(x)
    m.close()
    print(f"[binding {name}] max |y - y_ref| = {np.max(np.abs(y - y_ref)):.3e}", flush=True)
    assert np.array_equal(y, y_ref)
    del keep
