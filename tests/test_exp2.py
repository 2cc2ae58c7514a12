"""The FP64 table-driven exponential of the JIT kernels (h2p_exp2.h), compiled
for the host from the same header and checked against extended-precision
exp2 (x86 long double, 64-bit significand) over the range the Gaussian and
Matern-3/2 evaluations use, at the boundaries of its reduction, in the
subnormal range and beyond, and through the clamped entry for any q >= 0."""
import ctypes as ct
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "h2pack_amd", "csrc", "h2p_exp2.h")

SRC = r"""
#include "h2p_exp2.h"
static const double tab[H2P_E2_N] = {H2P_E2_TABLE};
extern "C" void e2(const double *q, double *y, long n, int clamped) {   // 2^(-q)
    for (long i = 0; i < n; i++)
        y[i] = clamped ? h2p::exp2_neg128_clamped(128.0 * q[i], tab) : h2p::exp2_neg128(128.0 * q[i], tab);
}
extern "C" void ex(const double *x, double *y, long n, int clamped) {   // e^(-x), as the Matern kernel
    const double c = 128.0 * 1.4426950408889634;
    for (long i = 0; i < n; i++)
        y[i] = clamped ? h2p::exp2_neg128_clamped(x[i] * c, tab) : h2p::exp2_neg128(x[i] * c, tab);
}
extern "C" const double *e2tab(void) { return tab; }
"""


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    d = tmp_path_factory.mktemp("exp2")
    src = d / "e2.cpp"
    src.write_text(SRC)
    so = d / "libe2.so"
    subprocess.run(["g++", "-O2", "-mfma", "-ffp-contract=off", "-fPIC", "-shared", "-I",
                    os.path.dirname(HDR), "-o", str(so), str(src)], check=True)
    L = ct.CDLL(str(so))
    L.e2.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_long, ct.c_int]
    L.e2.restype = None
    L.ex.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_long, ct.c_int]
    L.ex.restype = None
    L.e2tab.restype = ct.POINTER(ct.c_double)
    return L


def run(lib, q, clamped=False):
    q = np.ascontiguousarray(q, dtype=np.float64)
    y = np.empty_like(q)
    lib.e2(q.ctypes.data, y.ctypes.data, q.size, 1 if clamped else 0)
    return y


def ref(q):
    return np.exp2(-np.asarray(q, dtype=np.longdouble))


def test_table_is_correctly_rounded(lib):
    t = np.ctypeslib.as_array(lib.e2tab(), shape=(128,))
    exact = np.exp2(np.arange(128, dtype=np.longdouble) / 128)
    assert np.array_equal(t, exact.astype(np.float64))


def test_relative_error_over_kernel_range(lib):
    rng = np.random.default_rng(5)
    q = np.concatenate([rng.random(200000) * 60.0,            # Gaussian / Matern range that matters
                        rng.random(200000) * 1000.0,          # far pairs
                        rng.random(20000) * 1e-3,             # near the diagonal
                        np.arange(0, 1022 * 128 + 1) / 128.0,  # every table entry and exponent
                        np.arange(0, 1022 * 128) / 128.0 + 1.0 / 256.0])   # rounding ties of k
    y = run(lib, q)
    e = np.abs((y.astype(np.longdouble) - ref(q)) / ref(q)).astype(np.float64)
    assert e.max() <= 3.5e-16, e.max()
    assert np.all(run(lib, np.zeros(3)) == 1.0)


def test_subnormal_and_underflow(lib):
    q = np.linspace(1021.0, 1100.0, 50001)
    y = run(lib, q)
    err = np.abs(y.astype(np.longdouble) - ref(q))
    # normal results to ~1 ulp; subnormal ones are rounded twice (the
    # significand, then ldexp to the subnormal grid)
    assert np.all(err <= ref(q) * 3.5e-16 + np.longdouble(2.0) ** -1074)
    assert np.all(run(lib, np.array([1076.0, 1099.0, 2e3, 1e6, 2.0 ** 30])) == 0.0)


def test_clamped_matches_and_extends(lib):
    rng = np.random.default_rng(6)
    q = rng.random(100000) * 1100.0
    assert np.array_equal(run(lib, q), run(lib, q, clamped=True))
    big = np.array([2.0 ** 31, 2.0 ** 40, 1e300, np.inf])
    assert np.all(run(lib, big, clamped=True) == 0.0)


def test_unclamped_range_limit_is_real(lib):
    """Above q = 2^31 (x = 2^38) the exponent field wraps, which is why
    callers that cannot bound the argument use the clamped entry (the planted
    case shows the check above would catch a missing clamp)."""
    y = run(lib, np.array([2.0 ** 31 + 1.0]))
    assert y[0] != 0.0


def test_scaled_argument_natural_exp(lib):
    """exp2_neg128(x * 128 log2(e)) is e^(-x), as the Matern-3/2 kernel uses
    it: the rounded constant and product add |x| * 2.2e-16 to the relative
    error."""
    rng = np.random.default_rng(7)
    x = np.concatenate([rng.random(200000) * 40.0, rng.random(20000) * 700.0, [0.0, 1e-300]])
    y = np.empty_like(x)
    lib.ex(x.ctypes.data, y.ctypes.data, x.size, 0)
    r = np.exp(-x.astype(np.longdouble))
    e = np.abs((y - r) / r).astype(np.float64)
    assert np.all(e <= 3.5e-16 + np.abs(x) * 2.5e-16), e.max()
    yc = np.empty_like(x)
    lib.ex(x.ctypes.data, yc.ctypes.data, x.size, 1)
    assert np.array_equal(y, yc)
