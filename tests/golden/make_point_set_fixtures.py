"""Writes the reference's own point sets as fixtures.

The reference ships two 80,000-point sets for its SPD HSS-H2 examples
(examples/SPDHSS-H2/point_set/3Dball_80000.csv and 3Dsphere_80000.csv, one
point per line, x,y,z).  This script, run in the build container where the
reference tree is present, parses them into float64 arrays of shape
(80000, 3) and stores them under tests/golden/point_sets/ with their sums,
so the tests can use the same geometries without the reference tree.

    python tests/golden/make_point_set_fixtures.py [reference root]
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "point_sets")


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    os.makedirs(OUT, exist_ok=True)
    meta = {}
    for name in ("3Dball_80000", "3Dsphere_80000"):
        src = os.path.join(ref, "examples", "SPDHSS-H2", "point_set", name + ".csv")
        pts = np.loadtxt(src, delimiter=",", dtype=np.float64)
        assert pts.shape == (80000, 3), pts.shape
        np.save(os.path.join(OUT, name + ".npy"), pts)
        meta[name] = {"source": "examples/SPDHSS-H2/point_set/" + name + ".csv", "shape": list(pts.shape),
                      "sum": float(pts.sum()), "abs_sum": float(np.abs(pts).sum())}
    with open(os.path.join(OUT, "meta.json"), "w") as f:
        json.dump(meta, f, indent=1)


if __name__ == "__main__":
    main()
