"""The reference's own point sets (examples/SPDHSS-H2/point_set: an 80,000-
point ball and an 80,000-point sphere surface), in the configuration of its
SPD HSS-H2 example (example_SPDHSSH2.c:20: Gaussian kernel, l = 0.1, reltol
1e-8, JIT).  The sphere is a 2D manifold in 3D, so its tree is far from the
uniform cubes of the other problems (empty boxes, uneven leaves).

CPU: the fixtures are the reference's files (shape and sums recorded by
tests/golden/make_point_set_fixtures.py) and have the geometry their names
say.  GPU: the FP64 matvec equals the oracle on the same representation to
1e-12 and the direct FP64 n-body product on 300 rows to the build tolerance;
the FP32-mixed matvec stays within 1e-6 of the FP64 one.
"""
import functools
import json
import os

import numpy as np
import pytest

import oracle as O
from h2pack_amd import construct as C
from h2pack_amd.h2mat import H2MatHIP, direct_nbody
from tests.problems import rel

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "point_sets")
SETS = ["3Dball_80000", "3Dsphere_80000"]
PARAM = [0.1]


def load(name):
    return np.load(os.path.join(HERE, name + ".npy"))


@functools.lru_cache(maxsize=None)
def built(name):
    coord = np.ascontiguousarray(load(name).T)
    h2 = C.build_h2(coord, C.KRNL_GAUSSIAN, PARAM, 1e-8, 300, BD_JIT=1, pp_kind="numerical", host_blocks=False)
    return h2, coord


@pytest.mark.parametrize("name", SETS)
def test_fixture_is_the_reference_file(name):
    with open(os.path.join(HERE, "meta.json")) as f:
        meta = json.load(f)[name]
    pts = load(name)
    assert list(pts.shape) == meta["shape"] and pts.dtype == np.float64
    assert pts.sum() == meta["sum"] and np.abs(pts).sum() == meta["abs_sum"]
    r = np.linalg.norm(pts, axis=1)               # both sets are centred on the origin
    if "sphere" in name:
        assert r.std() < 1e-12 * r.mean()         # every point on one sphere
    else:
        assert np.mean(r < 0.5 * r.max()) > 0.05  # a filled ball, not a shell


@pytest.mark.gpu
@pytest.mark.parametrize("name", SETS)
def test_matvec_on_reference_point_sets(name):
    h2, coord = built(name)
    x = np.random.default_rng(31).random(h2.krnl_mat_size) - 0.5
    y_ref = O.OracleH2(h2, BD_JIT=1).matvec(x, n_thread=8)
    m = H2MatHIP(h2, mode="jit")
    y = m.matvec(x)
    m.close()
    err_oracle = rel(y, y_ref)
    rows = np.sort(np.random.default_rng(32).choice(h2.n_point, 300, replace=False))
    y_dir = direct_nbody(C.KRNL_GAUSSIAN, PARAM, coord, x, rows)
    err_direct = rel(y[rows], y_dir)
    mf = H2MatHIP(h2, mode="jit", prec="fp32mix")
    yf = mf.matvec(x)
    mf.close()
    err_f = rel(yf, y)
    print(f"[{name}] levels {h2.max_level + 1} leaves {h2.n_leaf_node}: vs oracle {err_oracle:.2e}, "
          f"vs direct {err_direct:.2e}, FP32-mixed vs FP64 {err_f:.2e}", flush=True)
    assert err_oracle < 1e-12
    assert err_direct < 1e-6
    assert err_f < 1e-6
