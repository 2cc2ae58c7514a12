"""The C-ABI library loads and exports every symbol include/h2p_hip.h declares.

No compute call is made here (no GPU in the CPU suite); only argument
validation that returns before touching the device.
"""
import ctypes as ct
import os
import re

import numpy as np
import pytest

from h2pack_amd import _build, abi

HEADER = os.path.join(_build.ROOT, "include", "h2p_hip.h")


def _declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(h2p_hip_\w+)\s*\(", txt)))


def _lib():
    if not os.path.exists(_build.HIP_SO):
        _build.build_hip()
    return abi.bind(ct.CDLL(_build.HIP_SO))


def test_header_and_binding_agree():
    assert _declared() == sorted(abi.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = _lib()
    for name in _declared():
        assert hasattr(lib, name), name


def test_struct_sizes_match_header_layout():
    # 10 ints, 8 pointers, 2 ints, 3 pointers, 2 + 1 + 2 + 1 + 1 + 2 + 4 + 4 pointers
    assert ct.sizeof(abi.KernelDesc) == 8 + 32
    assert ct.sizeof(abi.H2PackDesc) % 8 == 0


def test_invalid_arguments_return_error_codes():
    lib = _lib()
    out = ct.c_void_p()
    assert lib.h2p_hip_upload(None, None, 0, 0, ct.byref(out)) == abi.EINVAL
    assert b"NULL" in lib.h2p_hip_last_error()
    k = abi.KernelDesc()
    k.id = 9
    d = abi.H2PackDesc()
    assert lib.h2p_hip_upload(ct.byref(d), ct.byref(k), 0, 0, ct.byref(out)) == abi.EINVAL
    k.id = 0
    assert lib.h2p_hip_upload(ct.byref(d), ct.byref(k), 7, 0, ct.byref(out)) == abi.EINVAL
    assert lib.h2p_hip_matvec(None, None, None) == abi.EINVAL
    assert lib.h2p_hip_matmul(None, abi.COL_MAJOR, 1, None, 1, None, 1) == abi.EINVAL
    assert lib.h2p_hip_direct_nbody(None, 3, None, 0, 0, None, None, 0, 0, None) == abi.EINVAL


def test_upload_rejects_inconsistent_representation():
    """A representation whose U sizes disagree with its skeleton sets is
    refused on the host, before any device work."""
    from h2pack_amd import h2mat
    from tests.problems import problem
    h2, _ = problem("tiny_coulomb")
    d, keep = h2mat.describe(h2)
    bad_ncol = np.ascontiguousarray(h2.U_ncol, dtype=np.int32).copy()
    n = int(np.argmax(bad_ncol))
    bad_ncol[n] += 1
    d.U_ncol = bad_ncol.ctypes.data
    lib = _lib()
    out = ct.c_void_p()
    kd = h2mat.kernel_desc(h2.krnl_id, h2.krnl_param)
    rc = lib.h2p_hip_upload(ct.byref(d), ct.byref(kd), abi.JIT, abi.FP64, ct.byref(out))
    assert rc == abi.EINVAL and not out.value


def test_upload_validation_rejects_inconsistent_pair_data():
    """Upload validation (run without the device by h2p_hip_footprint):
    node_n_r_adm must be the per-node count of r_adm_pairs, as H2P_build
    computes it, and pair indices must name nodes; otherwise EINVAL."""
    from h2pack_amd.h2mat import H2PHipError, footprint
    from tests.problems import problem
    h2, _ = problem("gaussian")
    base = footprint(h2)
    cnt, pairs = h2.node_n_r_adm.copy(), h2.r_adm_pairs.copy()
    try:
        h2.node_n_r_adm[int(np.argmax(cnt))] += 1
        with pytest.raises(H2PHipError, match="node_n_r_adm"):
            footprint(h2)
        h2.node_n_r_adm[:] = cnt
        h2.r_adm_pairs[3] = h2.n_node + 5
        with pytest.raises(H2PHipError, match="out of range"):
            footprint(h2)
    finally:
        h2.node_n_r_adm[:] = cnt
        h2.r_adm_pairs[:] = pairs
    assert footprint(h2) == base
