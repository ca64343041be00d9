"""Pins the oracle (and the host graph stages behind it) to the reference's
known-answer integration tests, tests/test_implicit_networks.cpp:19-732:
patch / chain / cell / shell / corner counts and the patch and cell function
labels on generate_tet_mesh(101, [-1,1]^3) and on the unstructured mesh.json."""
import json
import os

import numpy as np
import pytest

import support as S

KA = json.load(open(os.path.join(S.GOLDEN, "known_answers.json")))


@pytest.fixture(scope="module")
def grid101():
    return S.grid(101)


def _check(case, out):
    assert out["ok"], out.get("error")
    for k in ("patches", "chains", "cells", "shells", "corners", "patch_label", "cell_label"):
        if k in case:
            assert out[k] == case[k], (case["name"], k, out[k], case[k])


@pytest.mark.parametrize("case", KA["ia"], ids=[c["name"] for c in KA["ia"]])
def test_ia_known_answer(case, grid101):
    pts, tets = grid101
    types, params = S.load_ref_funcs(case["funcs"])
    vals = S.evaluate(types, params, pts)
    for j in case.get("negate", []):
        vals[:, j] = -vals[:, j]
    _check(case, S.known_answer_run(0, pts, tets, vals))


@pytest.mark.parametrize("case", KA["mi"], ids=[c["name"] for c in KA["mi"]])
def test_mi_known_answer(case, grid101):
    if "mesh" in case:
        m = np.load(os.path.join(S.REF_TESTS, case["mesh"]))
        pts, tets = m["pts"], m["tets"].astype(np.uint64)
    else:
        pts, tets = grid101
    types, params = S.load_ref_funcs(case["funcs"])
    vals = S.evaluate(types, params, pts)
    _check(case, S.known_answer_run(1, pts, tets, vals))
