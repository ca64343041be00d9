"""The reference's own known-answer integration tests
(tests/test_implicit_networks.cpp:19-732) run through the drop-in
implicit_arrangement() entry point: per-tet path on the GPU, graph stages on
the host.  Also the -R robustness mode (src/implicit_arrangement.cpp:136-243)."""
import json
import os

import numpy as np
import pytest

import support as S

pytestmark = pytest.mark.gpu

KA = json.load(open(os.path.join(S.GOLDEN, "known_answers.json")))


@pytest.fixture(scope="module")
def grid101():
    return S.grid(101)


def _check(case, out):
    assert out["ok"], out
    for k in ("patches", "chains", "cells", "shells", "corners", "patch_label", "cell_label"):
        if k in case:
            assert out[k] == case[k], (case["name"], k, out[k], case[k])


@pytest.mark.parametrize("case", KA["ia"], ids=[c["name"] for c in KA["ia"]])
def test_dropin_ia_known_answer(case, grid101):
    pts, tets = grid101
    types, params = S.load_ref_funcs(case["funcs"])
    vals = S.evaluate(types, params, pts)
    for j in case.get("negate", []):
        vals[:, j] = -vals[:, j]
    out = S.dropin_run(0, pts, tets, vals)
    _check(case, out)
    # the stats the reference reports, against the oracle's counts
    ref = S.known_answer_run(0, pts, tets, vals)
    st = out["stats"]
    assert st["num_iso_verts"] == ref["num_verts"] and st["num_iso_faces"] == ref["num_faces"]
    assert st["num_intersecting_tet"] == ref["num_intersecting"]
    assert [st["num_1_func"], st["num_2_func"], st["num_more_func"]] == ref["buckets"]


def test_dropin_robust_mode(grid101):
    pts, tets = grid101
    types, params = S.load_ref_funcs("3-sphere-1.json")
    vals = S.evaluate(types, params, pts)
    out = S.dropin_run(0, pts, tets, vals, robust=True)
    assert out["ok"], out
    assert out["num_verts"] == 0 and out["patches"] == 0  # the -R run stops after the cut


@pytest.mark.parametrize("case", KA["mi"], ids=[c["name"] for c in KA["mi"]])
def test_dropin_mi_known_answer(case, grid101):
    if "mesh" in case:
        m = np.load(os.path.join(S.REF_TESTS, case["mesh"]))
        pts, tets = m["pts"], m["tets"].astype(np.uint64)
    else:
        pts, tets = grid101
    types, params = S.load_ref_funcs(case["funcs"])
    vals = S.evaluate(types, params, pts)
    _check(case, S.dropin_run(1, pts, tets, vals))
