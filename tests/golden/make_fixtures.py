"""Copies the data files that the reference's own integration tests read into
small committed fixtures.

Source: duxingyi-charles/Robust-Implicit-Surface-Networks, examples/tests/
(Apache-2.0).  The function lists are copied as JSON; the two tet meshes
(mesh.json, mesh_fail.json, each a JSON pair [points, tets]) are stored as
compressed .npz arrays.  The expected values these files are checked against
are the known answers of tests/test_implicit_networks.cpp, restated in
tests/golden/known_answers.json.

Run from the repository root:  python tests/golden/make_fixtures.py
"""
import json
import os
import sys

import numpy as np

SRC = "/root/reference/examples/tests"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_tests")


def main():
    if not os.path.isdir(SRC):
        sys.exit("reference examples not found at " + SRC)
    os.makedirs(DST, exist_ok=True)
    for name in sorted(os.listdir(SRC)):
        path = os.path.join(SRC, name)
        with open(path) as f:
            data = json.load(f)
        if name.startswith("mesh"):
            pts = np.asarray(data[0], dtype=np.float64)
            tets = np.asarray(data[1], dtype=np.int64)
            np.savez_compressed(os.path.join(DST, name.replace(".json", ".npz")), pts=pts, tets=tets)
        else:
            with open(os.path.join(DST, name), "w") as f:
                json.dump(data, f, indent=1)
                f.write("\n")
    print("wrote", sorted(os.listdir(DST)))


if __name__ == "__main__":
    main()
