"""ctypes loaders for the in-tree native libraries.

The engine library must exist; nothing falls back to a Python or CPU path
when it is missing.
"""
from __future__ import annotations

import ctypes as ct
import os

from . import _build

_host = None
_hip = None


def host_lib() -> ct.CDLL:
    """libh2p_host.so (construction helpers); compiled on first use if absent."""
    global _host
    if _host is None:
        if not os.path.exists(_build.HOST_SO):
            _build.build_host()
        lib = ct.CDLL(_build.HOST_SO)
        lib.h2p_host_kernel_eval.restype = None
        lib.h2p_host_kernel_eval.argtypes = [
            ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_int, ct.c_int,
            ct.c_void_p, ct.c_int, ct.c_int, ct.c_void_p, ct.c_long, ct.c_int]
        lib.h2p_host_sketch.restype = None
        lib.h2p_host_sketch.argtypes = [
            ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_int, ct.c_void_p, ct.c_int, ct.c_int,
            ct.c_void_p, ct.c_void_p, ct.c_int, ct.c_void_p, ct.c_int]
        lib.h2p_host_rpy_ewald.restype = ct.c_int
        lib.h2p_host_rpy_ewald.argtypes = [
            ct.c_void_p, ct.c_int, ct.c_int, ct.c_void_p, ct.c_int, ct.c_int, ct.c_double, ct.c_double,
            ct.c_int, ct.c_int, ct.c_void_p, ct.c_long, ct.c_int]
        _host = lib
    return _host


def hip_lib() -> ct.CDLL:
    """libh2p_hip.so, the MI355X engine.  Raises if it has not been built."""
    global _hip
    if _hip is None:
        if not os.path.exists(_build.HIP_SO):
            raise RuntimeError(
                f"{_build.HIP_SO} is missing: build it with "
                "`python -m h2pack_amd._build hip` (or __graft_entry__.build())")
        from . import abi
        _hip = abi.bind(ct.CDLL(_build.HIP_SO))
    return _hip
